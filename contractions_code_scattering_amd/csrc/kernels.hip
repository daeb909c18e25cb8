// HIP kernels of the contraction engine (gfx950 / CDNA4).
//
// lc_gemm_kernel   grouped block-complex GEMM  C = alpha * A * B  on D-space views, on the
//                  matrix cores (v_mfma_f64_16x16x4_f64 for FP64, v_mfma_f32_16x16x4_f32 for FP32),
//                  complex product as four real MFMAs into separate real/imaginary accumulators.
// lc_trace_kernel  grouped closing trace  out = alpha * Tr[X * Y] = alpha * sum_ij X_ij Y_ji,
//                  so the final product of a chain is never formed; wavefront shuffle + LDS
//                  reduction, fixed summation order (deterministic).
//
// One workgroup = 256 threads = 4 wave64. A GEMM workgroup owns one TM x TM output tile of one
// spin block (a,c) of one batch item, loops over the spin index b of the product (4 terms for
// dense x dense, 1 when one operand is an elemental M (x) 1) and over K in chunks of KT,
// staging both operand tiles through LDS with the view's phase applied, and prefetching the
// next chunk into registers while the matrix cores work on the current one.
#include <hip/hip_runtime.h>

#include "lc_internal.h"

namespace lc {

typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));

template <typename T> struct Cplx { T x, y; };
template <typename T> struct Acc4;
template <> struct Acc4<double> { typedef dx4 type; };
template <> struct Acc4<float> { typedef fx4 type; };

__device__ __forceinline__ dx4 mfma16(double a, double b, dx4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ fx4 mfma16(float a, float b, fx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row of the 16x16 result tile held in accumulator element q of lane l (the column is l % 16),
// as measured on gfx950 with lc_mfma_selftest_kernel (raw accumulator dump):
//   v_mfma_f64_16x16x4_f64: row = l/16 + 4q       v_mfma_f32_16x16x4_f32: row = 4(l/16) + q
template <typename T> __device__ __forceinline__ int acc_row(int l, int q);
template <> __device__ __forceinline__ int acc_row<double>(int l, int q) { return (l >> 4) + 4 * q; }
template <> __device__ __forceinline__ int acc_row<float>(int l, int q) { return ((l >> 4) << 2) + q; }

// Raw block of a view for spin block (a,b): pointer, element flags, exact phase.
template <typename T> struct BlockRef {
  const Cplx<T>* p;
  int flags;
  T sr, si;
};

template <typename T>
__device__ __forceinline__ BlockRef<T> view_block(const LcView& v, long long off, int a, int b,
                                                  long long nv2) {
  BlockRef<T> r;
  const Cplx<T>* base = reinterpret_cast<const Cplx<T>*>(v.base) + off;
  r.flags = v.flags;
  if (v.kind == LC_KRON) {
    r.p = base;
    r.sr = T(1);
    r.si = T(0);
    return r;
  }
  const int ra = v.perm_r[a], cb = v.perm_c[b];
  const int blk = (v.flags & LC_TS) ? (cb * 4 + ra) : (ra * 4 + cb);
  r.p = base + (long long)blk * nv2;
  const double pr = v.ph_r[a][0] * v.ph_c[b][0] - v.ph_r[a][1] * v.ph_c[b][1];
  const double pi = v.ph_r[a][0] * v.ph_c[b][1] + v.ph_r[a][1] * v.ph_c[b][0];
  r.sr = T(pr);
  r.si = T(pi);
  return r;
}

// Element (i,j) of f(X) for a raw block.
template <typename T>
__device__ __forceinline__ Cplx<T> elem(const BlockRef<T>& r, int i, int j, int nv) {
  const long long idx = (r.flags & LC_TV) ? (long long)j * nv + i : (long long)i * nv + j;
  Cplx<T> z = r.p[idx];
  if (r.flags & LC_CONJ) z.y = -z.y;
  return z;
}

__device__ __forceinline__ void decode_item(long long item, const int* n, long long* c) {
  c[0] = item % n[0];
  const long long q = item / n[0];
  c[1] = q % n[1];
  c[2] = q / n[1];
}

__device__ __forceinline__ long long voff(const long long* s, const long long* c) {
  return s[0] * c[0] + s[1] * c[1] + s[2] * c[2];
}

template <typename G>
__device__ __forceinline__ int find_group(const G* groups, int ng, long long wg) {
  int lo = 0, hi = ng - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (groups[mid].wg_begin <= wg) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------------------------------------------
template <typename T, int TM>
__global__ void __launch_bounds__(256) lc_gemm_kernel(const LcGemmGroup* __restrict__ groups,
                                                      int ng, int nv) {
  constexpr int KT = 16;
  constexpr int E = TM * KT / 256;  // complex elements staged per thread per operand
  constexpr int WM = TM / 2;        // per-wave output tile (2 x 2 waves)
  constexpr int MI = WM / 16;       // 16x16 MFMA tiles per wave side
  constexpr int PAD = 2;
  typedef typename Acc4<T>::type acc_t;

  __shared__ T As_re[KT][TM + PAD], As_im[KT][TM + PAD];
  __shared__ T Bs_re[KT][TM + PAD], Bs_im[KT][TM + PAD];

  const long long wg = blockIdx.x;
  const int gi = find_group(groups, ng, wg);
  const LcGemmGroup& g = groups[gi];
  const int tiles = g.tiles;
  const long long local = wg - g.wg_begin;
  const long long per_item = 16LL * tiles * tiles;
  const long long item = local / per_item;
  const int r = (int)(local - item * per_item);
  const int blk = r / (tiles * tiles);
  const int tm = (r % (tiles * tiles)) / tiles;
  const int tn = r % tiles;
  const int a = blk >> 2, c = blk & 3;

  long long co[3];
  decode_item(item, g.n, co);
  const long long nv2 = (long long)nv * nv;
  const long long offA = voff(g.A.s, co), offB = voff(g.B.s, co);

  const int nterm = (g.A.kind == LC_KRON || g.B.kind == LC_KRON) ? 1 : 4;
  const int nk = (nv + KT - 1) / KT;
  const int nchunk = nterm * nk;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int row0 = tm * TM, col0 = tn * TM;

  acc_t cre[MI][MI], cim[MI][MI];
#pragma unroll
  for (int i = 0; i < MI; i++)
#pragma unroll
    for (int j = 0; j < MI; j++) {
      cre[i][j] = acc_t{0, 0, 0, 0};
      cim[i][j] = acc_t{0, 0, 0, 0};
    }

  Cplx<T> ra[E], rb[E];
  BlockRef<T> A, B;
  int cur_term = -1;

  // Stage chunk `ch` of both operands into registers (phase applied to A).
  auto load = [&](int ch) {
    const int term = ch / nk;
    const int k0 = (ch - term * nk) * KT;
    if (term != cur_term) {
      int ab, bb, ba, bc;  // A block (ab, bb), B block (ba, bc)
      if (g.A.kind == LC_KRON) { ab = a; bb = a; ba = a; bc = c; }
      else if (g.B.kind == LC_KRON) { ab = a; bb = c; ba = c; bc = c; }
      else { ab = a; bb = term; ba = term; bc = c; }
      A = view_block<T>(g.A, offA, ab, bb, nv2);
      B = view_block<T>(g.B, offB, ba, bc, nv2);
      // fold B's phase into A's scalar
      const T sr = A.sr * B.sr - A.si * B.si, si = A.sr * B.si + A.si * B.sr;
      A.sr = sr; A.si = si;
      cur_term = term;
    }
    // A tile: rows row0.., k k0..
    if (A.flags & LC_TV) {
      const int k = tid / (TM / E), i0 = (tid % (TM / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int i = row0 + i0 + e, kk = k0 + k;
        Cplx<T> z = {T(0), T(0)};
        if (i < nv && kk < nv) z = elem(A, i, kk, nv);
        ra[e].x = A.sr * z.x - A.si * z.y;
        ra[e].y = A.sr * z.y + A.si * z.x;
      }
    } else {
      const int i = tid / (KT / E), kq = (tid % (KT / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int ii = row0 + i, kk = k0 + kq + e;
        Cplx<T> z = {T(0), T(0)};
        if (ii < nv && kk < nv) z = elem(A, ii, kk, nv);
        ra[e].x = A.sr * z.x - A.si * z.y;
        ra[e].y = A.sr * z.y + A.si * z.x;
      }
    }
    // B tile: k k0.., cols col0..
    if (B.flags & LC_TV) {
      const int j = tid / (KT / E), kq = (tid % (KT / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int jj = col0 + j, kk = k0 + kq + e;
        Cplx<T> z = {T(0), T(0)};
        if (jj < nv && kk < nv) z = elem(B, kk, jj, nv);
        rb[e] = z;
      }
    } else {
      const int k = tid / (TM / E), j0 = (tid % (TM / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int jj = col0 + j0 + e, kk = k0 + k;
        Cplx<T> z = {T(0), T(0)};
        if (jj < nv && kk < nv) z = elem(B, kk, jj, nv);
        rb[e] = z;
      }
    }
  };

  auto store = [&]() {
    if (A.flags & LC_TV) {
      const int k = tid / (TM / E), i0 = (tid % (TM / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) { As_re[k][i0 + e] = ra[e].x; As_im[k][i0 + e] = ra[e].y; }
    } else {
      const int i = tid / (KT / E), kq = (tid % (KT / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) { As_re[kq + e][i] = ra[e].x; As_im[kq + e][i] = ra[e].y; }
    }
    if (B.flags & LC_TV) {
      const int j = tid / (KT / E), kq = (tid % (KT / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) { Bs_re[kq + e][j] = rb[e].x; Bs_im[kq + e][j] = rb[e].y; }
    } else {
      const int k = tid / (TM / E), j0 = (tid % (TM / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) { Bs_re[k][j0 + e] = rb[e].x; Bs_im[k][j0 + e] = rb[e].y; }
    }
  };

  load(0);
  for (int ch = 0; ch < nchunk; ch++) {
    __syncthreads();
    store();
    __syncthreads();
    if (ch + 1 < nchunk) load(ch + 1);
#pragma unroll
    for (int ks = 0; ks < KT / 4; ks++) {
      const int kr = ks * 4 + (lane >> 4);
      T are[MI], aim[MI], bre[MI], bim[MI];
#pragma unroll
      for (int m = 0; m < MI; m++) {
        const int rr = wm * WM + m * 16 + (lane & 15);
        are[m] = As_re[kr][rr];
        aim[m] = As_im[kr][rr];
        const int cc = wn * WM + m * 16 + (lane & 15);
        bre[m] = Bs_re[kr][cc];
        bim[m] = Bs_im[kr][cc];
      }
#pragma unroll
      for (int m = 0; m < MI; m++)
#pragma unroll
        for (int n = 0; n < MI; n++) {
          cre[m][n] = mfma16(are[m], bre[n], cre[m][n]);
          cre[m][n] = mfma16(-aim[m], bim[n], cre[m][n]);
          cim[m][n] = mfma16(are[m], bim[n], cim[m][n]);
          cim[m][n] = mfma16(aim[m], bre[n], cim[m][n]);
        }
    }
  }

  // Epilogue: scale by alpha, store block (a,c).
  const T alr = T(g.alpha[0]), ali = T(g.alpha[1]);
  Cplx<T>* C = reinterpret_cast<Cplx<T>*>(g.C) + voff(g.cs, co) + (long long)(a * 4 + c) * nv2;
#pragma unroll
  for (int m = 0; m < MI; m++)
#pragma unroll
    for (int n = 0; n < MI; n++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = row0 + wm * WM + m * 16 + acc_row<T>(lane, q);
        const int col = col0 + wn * WM + n * 16 + (lane & 15);
        if (row < nv && col < nv) {
          const T xr = cre[m][n][q], xi = cim[m][n][q];
          Cplx<T> z;
          z.x = alr * xr - ali * xi;
          z.y = alr * xi + ali * xr;
          C[(long long)row * nv + col] = z;
        }
      }
}

// ------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T shfl_xor_t(T v, int m) { return __shfl_xor(v, m, 64); }

template <typename T>
__global__ void __launch_bounds__(256) lc_trace_kernel(const LcTraceGroup* __restrict__ groups,
                                                       int ng, int nv) {
  __shared__ T red[2][4];
  const long long wg = blockIdx.x;
  const int gi = find_group(groups, ng, wg);
  const LcTraceGroup& g = groups[gi];
  const long long item = wg - g.wg_begin;
  long long co[3];
  decode_item(item, g.n, co);
  const long long nv2 = (long long)nv * nv;
  const long long offX = voff(g.X.s, co), offY = voff(g.Y.s, co);
  const bool xk = g.X.kind == LC_KRON, yk = g.Y.kind == LC_KRON;
  const int npair = (xk || yk) ? 4 : 16;

  T accr = 0, acci = 0;
  for (int pr = 0; pr < npair; pr++) {
    int a, b;
    if (npair == 4) { a = pr; b = pr; } else { a = pr >> 2; b = pr & 3; }
    const BlockRef<T> X = view_block<T>(g.X, offX, a, b, nv2);
    const BlockRef<T> Y = view_block<T>(g.Y, offY, b, a, nv2);
    const T sr = X.sr * Y.sr - X.si * Y.si, si = X.sr * Y.si + X.si * Y.sr;
    T pr_ = 0, pi_ = 0;
    for (long long e = threadIdx.x; e < nv2; e += 256) {
      const int i = (int)(e / nv), j = (int)(e - (long long)i * nv);
      const Cplx<T> x = elem(X, i, j, nv);
      const Cplx<T> y = elem(Y, j, i, nv);
      pr_ += x.x * y.x - x.y * y.y;
      pi_ += x.x * y.y + x.y * y.x;
    }
    accr += sr * pr_ - si * pi_;
    acci += sr * pi_ + si * pr_;
  }
  for (int m = 32; m >= 1; m >>= 1) {
    accr += shfl_xor_t(accr, m);
    acci += shfl_xor_t(acci, m);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { red[0][wave] = accr; red[1][wave] = acci; }
  __syncthreads();
  if (threadIdx.x == 0) {
    const T sr = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    const T si = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    const T alr = T(g.alpha[0]), ali = T(g.alpha[1]);
    Cplx<T>* out = reinterpret_cast<Cplx<T>*>(g.out) + voff(g.os, co);
    Cplx<T> z;
    z.x = alr * sr - ali * si;
    z.y = alr * si + ali * sr;
    if (g.accumulate) { z.x += out->x; z.y += out->y; }
    *out = z;
  }
}

// ------------------------------------------------------------------------------------------
// Self-test of the 16x16x4 MFMA operand/result layout: D = A(16x4) * B(4x16), A[i][k] = a[i*4+k],
// B[k][j] = b[k*16+j]. d[0..255]: D row-major through acc_row<T>(); d[256..511]: the raw
// accumulators, d[256 + l*4 + q] = element q of lane l. One wave.
template <typename T>
__global__ void lc_mfma_selftest_kernel(const T* a, const T* b, T* d) {
  typedef typename Acc4<T>::type acc_t;
  const int l = threadIdx.x;
  const T av = a[(l & 15) * 4 + (l >> 4)];
  const T bv = b[(l >> 4) * 16 + (l & 15)];
  acc_t c = acc_t{0, 0, 0, 0};
  c = mfma16(av, bv, c);
  for (int q = 0; q < 4; q++) {
    d[acc_row<T>(l, q) * 16 + (l & 15)] = c[q];
    d[256 + l * 4 + q] = c[q];
  }
}

}  // namespace lc

// ------------------------------------------------------------------------------------------
// Host launchers (called from the planner in capi.cpp).
extern "C" hipError_t lc_launch_gemm(int is_double, const LcGemmGroup* d_groups, int ng,
                                     long long total_wg, int nv, hipStream_t s) {
  if (ng <= 0 || total_wg <= 0) return hipSuccess;
  const dim3 grid((unsigned)total_wg), block(256);
  if (is_double) {
    if (nv <= 32) hipLaunchKernelGGL((lc::lc_gemm_kernel<double, 32>), grid, block, 0, s, d_groups, ng, nv);
    else hipLaunchKernelGGL((lc::lc_gemm_kernel<double, 64>), grid, block, 0, s, d_groups, ng, nv);
  } else {
    if (nv <= 32) hipLaunchKernelGGL((lc::lc_gemm_kernel<float, 32>), grid, block, 0, s, d_groups, ng, nv);
    else hipLaunchKernelGGL((lc::lc_gemm_kernel<float, 64>), grid, block, 0, s, d_groups, ng, nv);
  }
  return hipGetLastError();
}

extern "C" int lc_gemm_tile(int nv) { return nv <= 32 ? 32 : 64; }

extern "C" hipError_t lc_launch_trace(int is_double, const LcTraceGroup* d_groups, int ng,
                                      long long total_wg, int nv, hipStream_t s) {
  if (ng <= 0 || total_wg <= 0) return hipSuccess;
  const dim3 grid((unsigned)total_wg), block(256);
  if (is_double) hipLaunchKernelGGL((lc::lc_trace_kernel<double>), grid, block, 0, s, d_groups, ng, nv);
  else hipLaunchKernelGGL((lc::lc_trace_kernel<float>), grid, block, 0, s, d_groups, ng, nv);
  return hipGetLastError();
}

extern "C" hipError_t lc_launch_mfma_selftest(int is_double, const void* a, const void* b, void* d,
                                              hipStream_t s) {
  if (is_double)
    hipLaunchKernelGGL((lc::lc_mfma_selftest_kernel<double>), dim3(1), dim3(64), 0, s,
                       (const double*)a, (const double*)b, (double*)d);
  else
    hipLaunchKernelGGL((lc::lc_mfma_selftest_kernel<float>), dim3(1), dim3(64), 0, s,
                       (const float*)a, (const float*)b, (float*)d);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Engine epilogue (FP64): per (p, source k, sink tf) combine the raw traces of the rho / pi-pi
// set into correlators, then average over sources on the device.
//   raw  [np][nsrc][nt][10] = {rho_0, rho_1, rho_2, square, dT, dTHor, st1, st2, fi1, fi2}
//   corr [np][nsrc][nt][8]  = {rho_0, rho_1, rho_2, square, dT, dTHor, st1*st2, fi1*fi2}
//   avg  [np][8][nt]        C(dt) = 1/nsrc sum_k corr[p][k][(t0_k + dt) % nt]
namespace lc {
__global__ void lc_engine_combine_kernel(const double2* __restrict__ raw, double2* __restrict__ corr,
                                         long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2* r = raw + i * 10;
  double2* c = corr + i * 8;
#pragma unroll
  for (int d = 0; d < 6; d++) c[d] = r[d];
  c[6].x = r[6].x * r[7].x - r[6].y * r[7].y;
  c[6].y = r[6].x * r[7].y + r[6].y * r[7].x;
  c[7].x = r[8].x * r[9].x - r[8].y * r[9].y;
  c[7].y = r[8].x * r[9].y + r[8].y * r[9].x;
}

__global__ void lc_engine_average_kernel(const double2* __restrict__ corr, double2* __restrict__ avg,
                                         int np, int nsrc, int nt, int ncorr, int src0, int src_step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np * ncorr * nt) return;
  const int dt = i % nt, d = (i / nt) % ncorr, p = i / (nt * ncorr);
  double sr = 0, si = 0;
  for (int k = 0; k < nsrc; k++) {
    const int t0 = src0 + k * src_step;
    const int tf = (t0 + dt) % nt;
    const double2 v = corr[(((long long)p * nsrc + k) * nt + tf) * ncorr + d];
    sr += v.x;
    si += v.y;
  }
  avg[i] = make_double2(sr / nsrc, si / nsrc);
}
}  // namespace lc

extern "C" hipError_t lc_launch_engine_epilogue(const void* raw, void* corr, void* avg, int np,
                                                int nsrc, int nt, int src0, int src_step,
                                                hipStream_t s) {
  const long long n = (long long)np * nsrc * nt;
  hipLaunchKernelGGL(lc::lc_engine_combine_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     (const double2*)raw, (double2*)corr, n);
  const int m = np * 8 * nt;
  hipLaunchKernelGGL(lc::lc_engine_average_kernel, dim3((m + 255) / 256), dim3(256), 0, s,
                     (const double2*)corr, (double2*)avg, np, nsrc, nt, 8, src0, src_step);
  return hipGetLastError();
}
