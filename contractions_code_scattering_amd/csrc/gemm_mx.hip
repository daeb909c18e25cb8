// Mixed-precision block-complex GEMM on the BF16 matrix cores (gfx950).
//
// Operands are stored as complex FP32. Every real value x entering a product is split exactly
// into three BF16 pieces (x = x0 + x1 + x2, each piece the round-to-nearest BF16 of what the
// previous ones leave; an FP32 significand has 24 bits = 3 x 8, so the split loses nothing).
// A real product a*b is formed from the six piece products of weight >= 2^-16:
//   a0b0 + [a0b1 + a1b0 + a0b2 + a1b1 + a2b0]    (dropped terms are below 2^-24 |a||b|)
// on v_mfma_f32_16x16x32_bf16 with FP32 accumulation, and the complex product from four real
// products (re = ArBr - AiBi, im = ArBi + AiBr): 24 MFMAs per 16x16x32 complex tile. The leading
// a0b0 terms and the bracketed correction terms accumulate in separate registers, so the many
// small corrections never round the leading sum. Output is complex FP32; fused closing traces
// accumulate in FP64 with the same descriptors and per-tile partials as the FP64 kernel.
//
// Workgroup: 256 threads = 4 wave64 (2 x 2), 64 x 64 output tile of one spin block, K step 32.
// Each element is split once, when staged: a thread fetches 8 consecutive-K values of one row
// of A and of one column of B (the mapping follows the operand's memory order, so the global
// reads are coalesced or vectorized), splits them with packed two-at-a-time conversions while
// the matrix cores work on the previous step, and writes 16-byte BF16 vectors into six LDS
// planes [piece][re/im]. A workgroup processes `bpw` consecutive spin blocks of one product
// (bpw = 4 for Nv <= 64): prologue, elemental tile and trace reductions are amortized over four
// blocks, and the fetch pipeline runs across block boundaries.
#include "lc_device.h"

namespace lc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float fx4v __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ fx4v mfma_bf16(const bf16x8& a, const bf16x8& b, fx4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Result layout of v_mfma_f32_16x16x32_bf16: element q of lane l is row 4(l/16)+q, column l%16
// (measured with lc_mfma_bf16_selftest_kernel).
__device__ __forceinline__ int acc_row_bf16(int l, int q) { return ((l >> 4) << 2) + q; }

// Two FP32 -> packed round-to-nearest BF16 pair (one v_cvt_pk_bf16_f32).
__device__ __forceinline__ unsigned cvt2(float a, float b) {
  const bf16x2 v = __builtin_convertvector((f2v){a, b}, bf16x2);
  return __builtin_bit_cast(unsigned, v);
}
__device__ __forceinline__ float lo_f(unsigned u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float hi_f(unsigned u) { return __uint_as_float(u & 0xffff0000u); }

// Exact three-way split of a pair: p0,p1,p2 hold the pieces of (a, b), a in the low half.
__device__ __forceinline__ void split2(float a, float b, unsigned& p0, unsigned& p1, unsigned& p2) {
  p0 = cvt2(a, b);
  const float a1 = a - lo_f(p0), b1 = b - hi_f(p0);
  p1 = cvt2(a1, b1);
  const float a2 = a1 - lo_f(p1), b2 = b1 - hi_f(p1);
  p2 = cvt2(a2, b2);
}

constexpr int MX_TM = 64;
constexpr int MX_KT = 32;
constexpr int MX_LD = MX_KT + 8;  // padded LDS row length, in BF16

// 8 consecutive-K complex values of one line of an operand: global -> registers.
struct Fetch {
  float re[8], im[8];
};

template <bool IS_A>
__device__ __forceinline__ void fetch8(const BlockRef<float>& R, int line, int k, int nv, bool full,
                                       Fetch& f) {
  // A element (line, k) sits at line*nv + k unless LC_TV; B element (k, line) at k*nv + line
  // unless LC_TV. "contig": the 8 values are consecutive in memory.
  const bool tv = (R.flags & LC_TV) != 0;
  const bool contig = IS_A ? !tv : tv;
  const float2* p = reinterpret_cast<const float2*>(R.p);
  if (full) {
    const float2* c8 = p + (long long)line * nv + k;
    if (contig && (reinterpret_cast<uintptr_t>(c8) & 15) == 0) {
      const float4* q = reinterpret_cast<const float4*>(c8);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const float4 v = q[e];
        f.re[2 * e] = v.x; f.im[2 * e] = v.y; f.re[2 * e + 1] = v.z; f.im[2 * e + 1] = v.w;
      }
    } else if (contig) {
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float2 v = c8[e];
        f.re[e] = v.x; f.im[e] = v.y;
      }
    } else {
      const float2* q = p + (long long)k * nv + line;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float2 v = q[(long long)e * nv];
        f.re[e] = v.x; f.im[e] = v.y;
      }
    }
  } else {
#pragma unroll
    for (int e = 0; e < 8; e++) {
      float2 v = make_float2(0.f, 0.f);
      if (line < nv && k + e < nv)
        v = contig ? p[(long long)line * nv + k + e] : p[(long long)(k + e) * nv + line];
      f.re[e] = v.x; f.im[e] = v.y;
    }
  }
  if (R.flags & LC_CONJ) {
#pragma unroll
    for (int e = 0; e < 8; e++) f.im[e] = -f.im[e];
  }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
lc_gemm_mx_kernel(const LcGemmGroup* __restrict__ groups, int ng, int nv, long long total, int bpw) {
  typedef float T;
  // planes [piece][re/im][line][k]
  __shared__ __attribute__((aligned(16))) __bf16 As[3][2][MX_TM][MX_LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[3][2][MX_TM][MX_LD];
  __shared__ double red[8];

  const long long wg = xcd_order(total);
  if (wg < 0) return;
  const int gi = find_group(groups, ng, wg);
  const LcGemmGroup& g = groups[gi];
  const int tiles = g.tiles;
  const int tt = tiles * tiles;
  const long long local = wg - g.wg_begin;
  const long long per_item = 16LL * tt;            // output tiles per item
  const long long wg_per_item = per_item / bpw;    // workgroups per item
  const long long item = local / wg_per_item;
  const int r0 = (int)(local - item * wg_per_item) * bpw;  // first tile of this workgroup

  long long co[3];
  decode_item(item, g.n, co);
  const long long nv2 = (long long)nv * nv;
  const long long offA = voff(g.A.s, co), offB = voff(g.B.s, co);
  const bool akron = g.A.kind == LC_KRON, bkron = g.B.kind == LC_KRON;

  const int nterm = (akron || bkron) ? 1 : 4;
  const int nk = (nv + MX_KT - 1) / MX_KT;
  const int cpb = nterm * nk;  // chunks per tile
  const int nchunk = cpb * bpw;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  fx4v bre[2][2], bim[2][2], sre[2][2], sim[2][2];  // leading (a0b0) and correction sums
  auto zero_acc = [&]() {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) {
        bre[i][j] = fx4v{0, 0, 0, 0}; bim[i][j] = fx4v{0, 0, 0, 0};
        sre[i][j] = fx4v{0, 0, 0, 0}; sim[i][j] = fx4v{0, 0, 0, 0};
      }
  };
  zero_acc();

  // staging state of the chunk in flight
  Fetch fa, fb;
  unsigned pa[3][2][4], pb[3][2][4];  // split pieces [piece][re/im][pair]
  int a_line = 0, a_kg = 0, b_line = 0, b_kg = 0;

  auto tile_of = [&](int tl, int& a, int& c, int& tm, int& tn) {
    const int rr = r0 + tl;
    const int blk = rr / tt;
    const int rem = rr - blk * tt;
    tm = rem / tiles;
    tn = rem - tm * tiles;
    a = blk >> 2;
    c = blk & 3;
  };

  // Fetch chunk ch of both operands into registers; the view phase goes with A.
  auto fetch = [&](int ch) {
    const int tl = ch / cpb;
    const int rem = ch - tl * cpb;
    const int term = rem / nk;
    const int k0 = (rem - term * nk) * MX_KT;
    int a, c, tm, tn;
    tile_of(tl, a, c, tm, tn);
    int ab, bb, ba, bc;
    if (akron) { ab = a; bb = a; ba = a; bc = c; }
    else if (bkron) { ab = a; bb = c; ba = c; bc = c; }
    else { ab = a; bb = term; ba = term; bc = c; }
    BlockRef<T> A = view_block<T>(g.A, offA, ab, bb, nv2);
    const BlockRef<T> B = view_block<T>(g.B, offB, ba, bc, nv2);
    const T sr = A.sr * B.sr - A.si * B.si, si = A.sr * B.si + A.si * B.sr;
    // thread -> (line, k-group) following the memory order of each operand
    if (A.flags & LC_TV) { a_line = tid & 63; a_kg = tid >> 6; } else { a_line = tid >> 2; a_kg = tid & 3; }
    if (B.flags & LC_TV) { b_line = tid >> 2; b_kg = tid & 3; } else { b_line = tid & 63; b_kg = tid >> 6; }
    const bool fullA = (tm + 1) * MX_TM <= nv && k0 + MX_KT <= nv;
    const bool fullB = (tn + 1) * MX_TM <= nv && k0 + MX_KT <= nv;
    fetch8<true>(A, tm * MX_TM + a_line, k0 + a_kg * 8, nv, fullA, fa);
    fetch8<false>(B, tn * MX_TM + b_line, k0 + b_kg * 8, nv, fullB, fb);
    if (!(sr == 1.f && si == 0.f)) {
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float x = fa.re[e], y = fa.im[e];
        fa.re[e] = sr * x - si * y;
        fa.im[e] = sr * y + si * x;
      }
    }
  };

  auto split = [&]() {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      split2(fa.re[2 * q], fa.re[2 * q + 1], pa[0][0][q], pa[1][0][q], pa[2][0][q]);
      split2(fa.im[2 * q], fa.im[2 * q + 1], pa[0][1][q], pa[1][1][q], pa[2][1][q]);
      split2(fb.re[2 * q], fb.re[2 * q + 1], pb[0][0][q], pb[1][0][q], pb[2][0][q]);
      split2(fb.im[2 * q], fb.im[2 * q + 1], pb[0][1][q], pb[1][1][q], pb[2][1][q]);
    }
  };

  auto store = [&]() {
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int ri = 0; ri < 2; ri++) {
        *reinterpret_cast<u32x4*>(&As[p][ri][a_line][a_kg * 8]) =
            u32x4{pa[p][ri][0], pa[p][ri][1], pa[p][ri][2], pa[p][ri][3]};
        *reinterpret_cast<u32x4*>(&Bs[p][ri][b_line][b_kg * 8]) =
            u32x4{pb[p][ri][0], pb[p][ri][1], pb[p][ri][2], pb[p][ri][3]};
      }
  };

  // per-tile epilogue: store and fused trace accumulation (FP64, carried across tiles)
  double tsr[LC_MAX_FUSED_TRACES], tsi[LC_MAX_FUSED_TRACES];
#pragma unroll
  for (int t = 0; t < LC_MAX_FUSED_TRACES; t++) { tsr[t] = 0; tsi[t] = 0; }
  const T alr = T(g.alpha[0]), ali = T(g.alpha[1]);

  auto epilogue = [&](int tl) {
    int a, c, tm, tn;
    tile_of(tl, a, c, tm, tn);
    const int row0 = tm * MX_TM, col0 = tn * MX_TM;
    T zr[2][2][4], zi[2][2][4];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int n = 0; n < 2; n++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const T xr = bre[m][n][q] + sre[m][n][q], xi = bim[m][n][q] + sim[m][n][q];
          zr[m][n][q] = alr * xr - ali * xi;
          zi[m][n][q] = alr * xi + ali * xr;
        }
    if (g.C) {
      Cplx<T>* C = reinterpret_cast<Cplx<T>*>(g.C) + voff(g.cs, co) + (long long)(a * 4 + c) * nv2;
#pragma unroll
      for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int row = row0 + wm * 32 + m * 16 + acc_row_bf16(lane, q);
            const int col = col0 + wn * 32 + n * 16 + (lane & 15);
            if (row < nv && col < nv) {
              Cplx<T> z;
              z.x = zr[m][n][q];
              z.y = zi[m][n][q];
              C[(long long)row * nv + col] = z;
            }
          }
    }
    for (int t = 0; t < g.ntr; t++) {
      const BlockRef<T> Y = view_block<T>(g.Y[t], voff(g.Y[t].s, co), c, a, nv2);
      if (Y.zero) continue;
      double sr = 0, si = 0;
#pragma unroll
      for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int row = row0 + wm * 32 + m * 16 + acc_row_bf16(lane, q);
            const int col = col0 + wn * 32 + n * 16 + (lane & 15);
            if (row < nv && col < nv) {
              const Cplx<T> y = elem(Y, col, row, nv);
              const double xr = zr[m][n][q], xi = zi[m][n][q];
              sr += xr * (double)y.x - xi * (double)y.y;
              si += xr * (double)y.y + xi * (double)y.x;
            }
          }
      const double yr = Y.sr, yi = Y.si;  // the Y block phase is uniform over the tile
      tsr[t] += sr * yr - si * yi;
      tsi[t] += sr * yi + si * yr;
    }
    zero_acc();
  };

  fetch(0);
  split();
  store();
  __syncthreads();
  const int kq = (lane >> 4) * 8;
  for (int ch = 0; ch < nchunk; ch++) {
    const bool more = ch + 1 < nchunk;
    if (more) fetch(ch + 1);
    bf16x8 bf[2][3][2];  // [n][piece][re/im]
#pragma unroll
    for (int n = 0; n < 2; n++)
#pragma unroll
      for (int p = 0; p < 3; p++)
#pragma unroll
        for (int ri = 0; ri < 2; ri++)
          bf[n][p][ri] = *reinterpret_cast<const bf16x8*>(&Bs[p][ri][wn * 32 + n * 16 + (lane & 15)][kq]);
#pragma unroll
    for (int m = 0; m < 2; m++) {
      bf16x8 af[3][2], nai[3];
#pragma unroll
      for (int p = 0; p < 3; p++)
#pragma unroll
        for (int ri = 0; ri < 2; ri++)
          af[p][ri] = *reinterpret_cast<const bf16x8*>(&As[p][ri][wm * 32 + m * 16 + (lane & 15)][kq]);
#pragma unroll
      for (int p = 0; p < 3; p++)
        nai[p] = __builtin_bit_cast(bf16x8, __builtin_bit_cast(u32x4, af[p][1]) ^ 0x80008000u);
#pragma unroll
      for (int n = 0; n < 2; n++) {
        // correction terms (weights 2^-8 and 2^-16), then the leading term
        const int PI[5] = {2, 1, 0, 1, 0};
        const int PJ[5] = {0, 1, 2, 0, 1};
#pragma unroll
        for (int s = 0; s < 5; s++) {
          const int i = PI[s], j = PJ[s];
          sre[m][n] = mfma_bf16(af[i][0], bf[n][j][0], sre[m][n]);
          sre[m][n] = mfma_bf16(nai[i], bf[n][j][1], sre[m][n]);
          sim[m][n] = mfma_bf16(af[i][0], bf[n][j][1], sim[m][n]);
          sim[m][n] = mfma_bf16(af[i][1], bf[n][j][0], sim[m][n]);
        }
        bre[m][n] = mfma_bf16(af[0][0], bf[n][0][0], bre[m][n]);
        bre[m][n] = mfma_bf16(nai[0], bf[n][0][1], bre[m][n]);
        bim[m][n] = mfma_bf16(af[0][0], bf[n][0][1], bim[m][n]);
        bim[m][n] = mfma_bf16(af[0][1], bf[n][0][0], bim[m][n]);
      }
    }
    if (more) split();  // VALU work that can overlap the matrix-core work issued above
    if ((ch + 1) % cpb == 0) epilogue(ch / cpb);
    __syncthreads();
    if (more) {
      store();
      __syncthreads();
    }
  }

  for (int t = 0; t < g.ntr; t++) {
    double sr = tsr[t], si = tsi[t];
    block_sum2(sr, si, red);
    if (tid == 0) {
      const double ar = g.tr_alpha[t][0], ai = g.tr_alpha[t][1];
      double* out = g.part[t] + 2 * (item * wg_per_item + r0 / bpw);
      out[0] = ar * sr - ai * si;
      out[1] = ar * si + ai * sr;
    }
  }
}

// Layout self-test of v_mfma_f32_16x16x32_bf16 with A(16x32), B(32x16) given as FP32 values that
// are exact in BF16: lane l feeds A[l%16][8(l/16)..+7] and B[8(l/16)..+7][l%16]. d[0..255] = D
// row-major through acc_row_bf16, d[256..511] = raw accumulator q of lane l at 256 + 4l + q.
__global__ void lc_mfma_bf16_selftest_kernel(const float* a, const float* b, float* d) {
  const int l = threadIdx.x;
  bf16x8 av, bv;
  for (int e = 0; e < 8; e++) {
    av[e] = (__bf16)a[(l & 15) * 32 + (l >> 4) * 8 + e];
    bv[e] = (__bf16)b[((l >> 4) * 8 + e) * 16 + (l & 15)];
  }
  fx4v c = fx4v{0, 0, 0, 0};
  c = mfma_bf16(av, bv, c);
  for (int q = 0; q < 4; q++) {
    d[acc_row_bf16(l, q) * 16 + (l & 15)] = c[q];
    d[256 + l * 4 + q] = c[q];
  }
}

// ------------------------------------------------------------------------------------------
// Per-slice rescaling for the mixed path: y[i] = x[i] / ||x_i||_F, in FP32, with the scale
// ||x_i||_F of every slice i (nslice contiguous slices of len complex values) kept in FP64.
__global__ void __launch_bounds__(256) lc_slice_norm_kernel(const double2* __restrict__ x,
                                                            long long len, long long stride,
                                                            double* __restrict__ scale) {
  __shared__ double red[8];
  const double2* p = x + (long long)blockIdx.x * stride;
  double s = 0, dummy = 0;
  for (long long e = threadIdx.x; e < len; e += 256) {
    const double2 v = p[e];
    s += v.x * v.x + v.y * v.y;
  }
  block_sum2(s, dummy, red);
  if (threadIdx.x == 0) scale[blockIdx.x] = s > 0 ? sqrt(s) : 1.0;
}

__global__ void lc_slice_convert_kernel(const double2* __restrict__ x, float2* __restrict__ y,
                                        long long len, const double* __restrict__ scale,
                                        long long nslice) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len * nslice) return;
  const double inv = 1.0 / scale[i / len];
  const double2 v = x[i];
  y[i] = make_float2((float)(v.x * inv), (float)(v.y * inv));
}

}  // namespace lc

// Spin blocks per workgroup: 4 when one 64 x 64 tile covers a block.
extern "C" int lc_gemm_mx_bpw(int nv) { return nv <= 64 ? 4 : 1; }

extern "C" hipError_t lc_launch_gemm_mx(const LcGemmGroup* d_groups, int ng, long long total_wg,
                                        int nv, hipStream_t s) {
  if (ng <= 0 || total_wg <= 0) return hipSuccess;
  const long long per = (total_wg + 7) / 8;
  hipLaunchKernelGGL(lc::lc_gemm_mx_kernel, dim3((unsigned)(8 * per)), dim3(256), 0, s, d_groups, ng, nv,
                     total_wg, lc_gemm_mx_bpw(nv));
  return hipGetLastError();
}

extern "C" hipError_t lc_launch_mfma_bf16_selftest(const float* a, const float* b, float* d,
                                                   hipStream_t s) {
  hipLaunchKernelGGL(lc::lc_mfma_bf16_selftest_kernel, dim3(1), dim3(64), 0, s, a, b, d);
  return hipGetLastError();
}

// x: nslice slices of len complex doubles (contiguous); y: same layout in complex float.
extern "C" hipError_t lc_launch_rescale(const void* x, void* y, long long len, long long nslice,
                                        double* scale, hipStream_t s) {
  if (nslice <= 0 || len <= 0) return hipSuccess;
  hipLaunchKernelGGL(lc::lc_slice_norm_kernel, dim3((unsigned)nslice), dim3(256), 0, s,
                     (const double2*)x, len, len, scale);
  const long long n = len * nslice;
  hipLaunchKernelGGL(lc::lc_slice_convert_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     (const double2*)x, (float2*)y, len, (const double*)scale, nslice);
  return hipGetLastError();
}
