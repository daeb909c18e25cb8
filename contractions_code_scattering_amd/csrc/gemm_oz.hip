// Exact-accumulation block-complex GEMM on the int8 matrix cores (gfx950), and the conversions
// into and out of digit storage.
//
// Number format (lc_internal.h). Every real value x of a stored line (a row of an A operand, a
// column of a B operand) is four balanced base-128 digits D0..D3 in [-64, 64] under the line's
// power-of-two exponent E: x = 2^(E-27) * v, v = D0*128^3 + D1*128^2 + D2*128 + D3, with E the
// smallest integer such that every |x| of the line is below 2^E. The representation keeps 27 bits
// below the largest magnitude of the line (truncation error <= 2^(E-28)).
//
// Product. For C = A B every digit pair (a, b) of total order o = a + b <= 4 is one
// v_mfma_i32_32x32x32_i8 per real product (13 pairs); the pairs of one order share an int32
// accumulator. Integer accumulation is exact (|acc_o| <= 2 (pairs of order o) K 2^12 < 2^25 for
// K <= 512), so the only rounding of a product is the digit truncation of its operands and the
// dropped orders 5 and 6. Order 4 is needed: a typical element sits a few binades below its line's
// bound, so its leading digit is small while the lower digits span their whole range, and the
// order-4 cross terms are of the size of the truncation itself (a numerical model of the format at
// D = 256 gives 7e-8 of max |C| with orders <= 3, 1.5e-8 with orders <= 4). The epilogue forms,
// per element,
//   S = 2^28 acc_0 + 2^21 acc_1 + 2^14 acc_2 + 2^7 acc_3 + acc_4    (exact in FP64, < 2^51)
//   C_ij = alpha * 2^(E_A,i + E_B,j - 40) * S
// and either stores it in FP64, traces it in FP64 against FP64 closing factors, or writes it
// back as digits with a fresh exponent per D-row (the largest |C_ij| of the row over all four
// blocks, which the workgroup strip covers). The complex product takes four real products
// (re += Ar Br - Ai Bi, im += Ar Bi + Ai Br): 52 MFMAs per 32 x 32 output tile and K step 32.
//
// Workgroup: NW wave64s own a strip of 32 rows x 32*NW columns of one item; wave w the 32 x 32
// tile at columns 32 w. Per K step of 32 the operand chunks (32 lines x 32 K x 8 digit planes,
// one contiguous kilobyte per plane in memory) go global -> LDS by 16-byte direct loads, in the
// order of the MFMA fragments (lane l: line l % 32, K bytes 16 (l / 32) .. +15), so every
// fragment read is a linear 16-byte-per-lane LDS read. Two stage buffers, one barrier per step.
// The A chunk of a dense product is shared by all waves (their rows are the same); each wave has
// its own B chunk. View phases (powers of i) and conjugation are applied to the fragments as
// digit negations (exact: the digits are bounded by 64) and a choice of planes.
//
// The MFMA is issued with the B fragment first, so lane l holds row l % 32 of C and the 16
// columns 8 g + 4 (l / 32) + r (g, r < 4): runs of four consecutive columns, stored as one dword
// per digit plane.
#include <climits>

#include "lc_device.h"

namespace lc {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));
constexpr int NORD = 5;  // digit-pair orders 0..4 accumulated

__device__ __forceinline__ i32x16 mfma_i8(const i32x4& x, const i32x4& y, const i32x16& c) {
  return __builtin_amdgcn_mfma_i32_32x32x32_i8(x, y, c, 0, 0, 0);
}

// Negation of four packed int8 digits in [-64, 64]: in the sign-flipped (biased) domain the
// digits are 64..192, whose negation 256 - b is ~b + 1 with no carry between bytes.
__device__ __forceinline__ unsigned neg4(unsigned x) { return ((x ^ 0x7F7F7F7Fu) + 0x01010101u) ^ 0x80808080u; }
__device__ __forceinline__ i32x4 neg16(const i32x4& v) {
  return i32x4{(int)neg4((unsigned)v[0]), (int)neg4((unsigned)v[1]), (int)neg4((unsigned)v[2]),
               (int)neg4((unsigned)v[3])};
}

// Exponent E of a nonzero value: |x| < 2^E (the frexp exponent); INT_MIN for zero, so that zeros
// never raise a line's exponent.
__device__ __forceinline__ int exp_of(double x) { return x == 0.0 ? INT_MIN : __builtin_amdgcn_frexp_exp(x); }

// Digits of x under exponent E (|x| < 2^E): v = rint(x 2^(27-E)), |v| <= 2^27; with the bias
// 64 (1 + 128 + 128^2) the three low digits are plain 7-bit fields minus 64, the top digit the
// arithmetic remainder (in [-64, 64]). d[0] is the most significant.
__device__ __forceinline__ void digits_of(double x, int E, int (&d)[4]) {
  const int v = (int)__builtin_rint(__builtin_ldexp(x, 27 - E));
  const int w = v + 1056832;
  d[3] = (w & 127) - 64;
  d[2] = ((w >> 7) & 127) - 64;
  d[1] = ((w >> 14) & 127) - 64;
  d[0] = w >> 21;
}
__device__ __forceinline__ unsigned pack4(int a, int b, int c, int d) {
  return (unsigned)(a & 255) | ((unsigned)(b & 255) << 8) | ((unsigned)(c & 255) << 16) | ((unsigned)d << 24);
}

// Stored block of view block (x, y); its stored block row is the result / 4.
__device__ __forceinline__ int stored_block(const LcView& v, int x, int y) {
  if (v.kind == LC_KRON) return 0;
  const int ra = v.perm_r[x], cb = v.perm_c[y];
  return (v.flags & LC_TS) ? (cb * 4 + ra) : (ra * 4 + cb);
}
__device__ __forceinline__ void block_phase(const LcView& v, int x, int y, double& pr, double& pi) {
  if (v.kind == LC_KRON) {
    pr = 1.0;
    pi = 0.0;
    return;
  }
  pr = v.ph_r[x][0] * v.ph_c[y][0] - v.ph_r[x][1] * v.ph_c[y][1];
  pi = v.ph_r[x][0] * v.ph_c[y][1] + v.ph_r[x][1] * v.ph_c[y][0];
}

template <int NW>
__global__ void __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(2, 2)))
lc_gemm_oz_kernel(const LcOzGroup* __restrict__ groups, int ng, int nv, int P, long long total) {
  constexpr int NSLOT = NW + 1;
  __shared__ __attribute__((aligned(16))) unsigned char stage[2][NSLOT][8][1024];

  const long long wg = xcd_order(total);
  if (wg < 0) return;
  const int gi = find_group(groups, ng, wg);
  const LcOzGroup& g = groups[gi];
  const int nseg = g.nseg;
  const int nstrip = (4 * P) >> 5;
  const long long wpi = (long long)nstrip * nseg;  // workgroups per item
  const long long local = wg - g.wg_begin;
  const long long item = local / wpi;
  const int rr = (int)(local - item * wpi);
  const int strip = rr / nseg, seg = rr - strip * nseg;
  const int a = (strip * 32) / P, i0 = (strip * 32) % P;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wcol0 = seg * NW * 32;          // first D-column of the workgroup
  const int c = (wcol0 + wave * 32) / P;    // this wave's block column
  const int j0 = (wcol0 + wave * 32) % P;   // and first column inside it

  long long co[3];
  decode_item(item, g.n, co);
  const long long P2 = (long long)P * P;
  const long long nv2 = (long long)nv * nv;
  const int kind = g.kind;
  const int ksteps = P >> 5;
  const int nstep = (kind == 0 ? 4 : 1) * ksteps;
  // operand chunks per step: A chunks (shared by the waves of one block column, or by all) then
  // B chunks (one per wave, or one per 32-column group of an elemental)
  const int nA = kind == 2 ? (NW * 32) / P : 1;
  const int nB = kind == 2 ? ksteps : NW;
  const int sA = kind == 2 ? (wave * 32) / P : 0;
  const int sB = kind == 2 ? (j0 >> 5) : wave;
  const int cfirst = wcol0 / P;
  const long long offA = voff(g.A.s, co), offB = voff(g.B.s, co);

  // view blocks of slot s at spin index b of the product
  auto a_block = [&](int s, int b, int& x, int& y) {
    x = a;
    y = kind == 2 ? cfirst + s : (kind == 1 ? a : b);
  };
  auto b_block = [&](int s, int b, int& x, int& y) {
    if (kind == 2) { x = 0; y = 0; return; }
    x = kind == 1 ? a : b;
    y = (wcol0 + s * 32) / P;
  };
  auto b_line0 = [&](int s) { return kind == 2 ? s * 32 : (wcol0 + s * 32) % P; };

  // direct global -> LDS loads of step st into stage buffer bf: one wave-wide 16-byte load per
  // (chunk, plane), distributed over the waves
  const int lane_off = 32 * (lane & 31) + 16 * (lane >> 5);
  auto issue = [&](int st, int bf) {
    const int b = st / ksteps, kk = st - b * ksteps;
    const int nload = 8 * (nA + nB);
    for (int f = wave; f < nload; f += NW) {
      const int s = f >> 3, q = f & 7;
      const char* src;
      if (s < nA) {
        int x, y;
        a_block(s, b, x, y);
        src = static_cast<const char*>(g.A.base) + 8 * (offA + stored_block(g.A, x, y) * P2) + q * P2 +
              ((long long)kk * P + i0) * 32;
      } else {
        const int sb = s - nA;
        int x, y;
        b_block(sb, b, x, y);
        src = static_cast<const char*>(g.B.base) + 8 * (offB + stored_block(g.B, x, y) * P2) + q * P2 +
              ((long long)kk * P + b_line0(sb)) * 32;
      }
      __builtin_amdgcn_global_load_lds((const void*)(src + lane_off),
                                       (__attribute__((address_space(3))) void*)&stage[bf][s][q][0], 16, 0, 0);
    }
  };

  i32x16 re[NORD], im[NORD];
#pragma unroll
  for (int o = 0; o < NORD; o++) {
    re[o] = i32x16{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    im[o] = re[o];
  }

  const bool conjA = (g.A.flags & LC_CONJ) != 0, conjB = (g.B.flags & LC_CONJ) != 0;
  issue(0, 0);
  for (int st = 0; st < nstep; st++) {
    const int bf = st & 1;
    __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0): this wave's loads of step st have landed
    __syncthreads();                    // and every other wave's; buffer bf^1 is free again
    if (st + 1 < nstep) issue(st + 1, bf ^ 1);
    // phase of this wave's product term: B' = phi * conj?(B), A' = conj?(A)
    const int b = st / ksteps;
    int ax, ay, bx, by;
    a_block(sA, b, ax, ay);
    b_block(sB, b, bx, by);
    double p1r, p1i, p2r, p2i;
    block_phase(g.A, ax, ay, p1r, p1i);
    block_phase(g.B, bx, by, p2r, p2i);
    const double phr = p1r * p2r - p1i * p2i, phi = p1r * p2i + p1i * p2r;
    // B'r = nr ? -X : X, B'i = ni ? -Y : Y with (X, Y) = (Br, Bi), or (Bi, Br) when swapped
    bool sw, nr, ni;
    if (phi == 0.0) {
      sw = false;
      nr = phr < 0.0;
      ni = (phr < 0.0) != conjB;
    } else if (phi > 0.0) {  // i (Br + i c Bi) = -c Bi + i Br
      sw = true;
      nr = !conjB;
      ni = false;
    } else {                 // -i (Br + i c Bi) = c Bi - i Br
      sw = true;
      nr = conjB;
      ni = true;
    }
    const unsigned char* bsrc = &stage[bf][nA + sB][0][16 * lane];
    const unsigned char* asrc = &stage[bf][sA][0][16 * lane];
    i32x4 br[4], bi[4];
#pragma unroll
    for (int d = 0; d < 4; d++) {
      const i32x4 x = *reinterpret_cast<const i32x4*>(bsrc + d * 1024);
      const i32x4 y = *reinterpret_cast<const i32x4*>(bsrc + (4 + d) * 1024);
      br[d] = sw ? y : x;
      bi[d] = sw ? x : y;
    }
    if (nr) {
#pragma unroll
      for (int d = 0; d < 4; d++) br[d] = neg16(br[d]);
    }
    if (ni) {
#pragma unroll
      for (int d = 0; d < 4; d++) bi[d] = neg16(bi[d]);
    }
#pragma unroll
    for (int da = 0; da < 4; da++) {
      const i32x4 ar = *reinterpret_cast<const i32x4*>(asrc + da * 1024);
      const i32x4 ai0 = *reinterpret_cast<const i32x4*>(asrc + (4 + da) * 1024);
      const i32x4 ain = neg16(ai0);
      const i32x4 ai = conjA ? ain : ai0;   // imaginary part of A'
      const i32x4 nai = conjA ? ai0 : ain;  // and its negation
#pragma unroll
      for (int db = 0; db < 4 && db + da < NORD; db++) {
        const int o = da + db;
        re[o] = mfma_i8(br[db], ar, re[o]);
        im[o] = mfma_i8(bi[db], ar, im[o]);
        re[o] = mfma_i8(bi[db], nai, re[o]);
        im[o] = mfma_i8(br[db], ai, im[o]);
      }
    }
  }

  // ---- epilogue
  // the stage buffer of step nstep (not the one of the last step, which every wave left at the
  // last barrier) holds the cross-wave reductions
  int (*rowe)[32] = reinterpret_cast<int (*)[32]>(&stage[nstep & 1][0][0][0]);
  double (*tred)[2][NW] = reinterpret_cast<double (*)[2][NW]>(&stage[nstep & 1][1][0][0]);
  // exponents: row i0 + lane % 32 of A; columns b_line0 + 8 g + 4 (lane / 32) + r of B
  int ea;
  {
    int x, y;
    a_block(sA, 0, x, y);
    const int R = stored_block(g.A, x, y) >> 2;
    ea = g.A.ex[voff(g.A.es, co) + (long long)R * g.A.ebs + i0 + (lane & 31)];
  }
  int eb[16];
  {
    int x, y;
    b_block(sB, 0, x, y);
    const int R = stored_block(g.B, x, y) >> 2;
    const int* pe = g.B.ex + voff(g.B.es, co) + (long long)R * g.B.ebs + b_line0(sB) + 4 * (lane >> 5);
#pragma unroll
    for (int gg = 0; gg < 4; gg++) {
      const i32x4 v = gload<i32x4>(pe + 8 * gg);
#pragma unroll
      for (int r = 0; r < 4; r++) eb[4 * gg + r] = v[r];
    }
  }
  const double alr = g.alpha[0], ali = g.alpha[1];
  double vr[16], vi[16];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    // 2^21 (128 acc_0 + acc_1) + 2^7 (128 acc_2 + acc_3) + acc_4: the int32 partial combinations
    // stay below 2^31 for K <= 512
    const int hr = re[0][q] * 128 + re[1][q], mr = re[2][q] * 128 + re[3][q];
    const int hi = im[0][q] * 128 + im[1][q], mi = im[2][q] * 128 + im[3][q];
    const int e = ea + eb[q] - 40;
    const double xr = __builtin_ldexp((double)hr * 2097152.0 + (double)mr * 128.0 + (double)re[4][q], e);
    const double xi = __builtin_ldexp((double)hi * 2097152.0 + (double)mi * 128.0 + (double)im[4][q], e);
    vr[q] = alr * xr - ali * xi;
    vi[q] = alr * xi + ali * xr;
  }
  const int row = i0 + (lane & 31);  // row of C inside block (a, c)
  auto col_of = [&](int q) { return j0 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3); };

  if (g.C) {
    Cplx<double>* C = reinterpret_cast<Cplx<double>*>(g.C) + voff(g.cs, co) + (long long)(a * 4 + c) * nv2;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int col = col_of(q);
      if (row < nv && col < nv) gstore_c(C + (long long)row * nv + col, Cplx<double>{vr[q], vi[q]});
    }
  }

  if (g.Cd) {
    // exponent of the D-row: largest |value| over every wave of the strip (the whole row)
    int em = INT_MIN;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e1 = exp_of(vr[q]), e2 = exp_of(vi[q]);
      em = e1 > em ? e1 : em;
      em = e2 > em ? e2 : em;
    }
    const int eo = __shfl_xor(em, 32, 64);
    em = eo > em ? eo : em;
    if (lane < 32) rowe[wave][lane] = em;
    __syncthreads();
    int E = INT_MIN;
#pragma unroll
    for (int w = 0; w < NW; w++) {
      const int x = rowe[w][lane & 31];
      E = x > E ? x : E;
    }
    if (E == INT_MIN) E = 0;
    if (wave == 0 && lane < 32) g.Ce[voff(g.ces, co) + (long long)a * P + i0 + lane] = E;
    unsigned char* base = static_cast<unsigned char*>(g.Cd) + 8 * (voff(g.cds, co) + (long long)(a * 4 + c) * P2) +
                          ((long long)(j0 >> 5) * P + row) * 32 + 4 * (lane >> 5);
#pragma unroll
    for (int gg = 0; gg < 4; gg++) {
      int dr[4][4], di[4][4];  // [column r][digit]
#pragma unroll
      for (int r = 0; r < 4; r++) {
        digits_of(vr[4 * gg + r], E, dr[r]);
        digits_of(vi[4 * gg + r], E, di[r]);
      }
#pragma unroll
      for (int d = 0; d < 4; d++) {
        gstore(base + d * P2 + 8 * gg, pack4(dr[0][d], dr[1][d], dr[2][d], dr[3][d]));
        gstore(base + (4 + d) * P2 + 8 * gg, pack4(di[0][d], di[1][d], di[2][d], di[3][d]));
      }
    }
  }

  if (g.ntr > 0) {
    double tr_re[LC_MAX_FUSED_TRACES], tr_im[LC_MAX_FUSED_TRACES];
#pragma unroll
    for (int t = 0; t < LC_MAX_FUSED_TRACES; t++) {
      tr_re[t] = 0;
      tr_im[t] = 0;
      if (t >= g.ntr) continue;
      // block (c, a) of f(Y), element (col, row)
      const BlockRef<double> Y = view_block<double>(g.Y[t], voff(g.Y[t].s, co), c, a, nv2);
      if (Y.zero) continue;
      const bool tv = (Y.flags & LC_TV) != 0;
      const double cj = (Y.flags & LC_CONJ) ? -1.0 : 1.0;
      Cplx<double> yv[16];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int col = col_of(q);
        const long long idx = tv ? (long long)row * nv + col : (long long)col * nv + row;
        yv[q] = (row < nv && col < nv) ? gload_c(Y.p + idx) : Cplx<double>{0.0, 0.0};
      }
      double sr = 0, si = 0;
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const double yr = yv[q].x, yi = cj * yv[q].y;
        sr += vr[q] * yr - vi[q] * yi;
        si += vr[q] * yi + vi[q] * yr;
      }
      const double pr = sr * Y.sr - si * Y.si, pi = sr * Y.si + si * Y.sr;
      const double ar = g.tr_alpha[t][0], ai = g.tr_alpha[t][1];
      tr_re[t] = ar * pr - ai * pi;
      tr_im[t] = ar * pi + ai * pr;
    }
#pragma unroll
    for (int t = 0; t < LC_MAX_FUSED_TRACES; t++) {
      tr_re[t] = wave_sum(tr_re[t]);
      tr_im[t] = wave_sum(tr_im[t]);
    }
    if (lane == 0) {
#pragma unroll
      for (int t = 0; t < LC_MAX_FUSED_TRACES; t++) {
        tred[t][0][wave] = tr_re[t];
        tred[t][1][wave] = tr_im[t];
      }
    }
    __syncthreads();
    if (tid < g.ntr) {
      double sr = 0, si = 0;
#pragma unroll
      for (int w = 0; w < NW; w++) {
        sr += tred[tid][0][w];
        si += tred[tid][1][w];
      }
      typedef double d2v __attribute__((ext_vector_type(2)));
      gstore(g.part[tid] + 2 * (item * wpi + rr), d2v{sr, si});
    }
  }
}

// ------------------------------------------------------------------------------------------
// Conversion FP64 -> digits.
//
// Exponents. mode 0 (inputs): per matrix, one exponent per row index (over every block and
// column) into ex_r and one per column index (over every block and row) into ex_t, P ints each.
// mode 1 (products written in FP64): one exponent per D-row (block row R, row i over the four
// blocks of R) into ex_r[R * P + i]. nblk = 16 (D-space) or 1 (elemental); matrices of nblk
// blocks of nv x nv complex doubles, `stride` complex elements apart. One workgroup per matrix.
__global__ void __launch_bounds__(256) lc_oz_exps_kernel(const double2* __restrict__ x, long long stride, int nv,
                                                         int P, int nblk, int mode, int* __restrict__ ex_r,
                                                         int* __restrict__ ex_t) {
  __shared__ int er[512], et[128];
  const long long m = blockIdx.x;
  const int nr = mode ? 4 * P : P;
  for (int i = threadIdx.x; i < nr; i += 256) er[i] = INT_MIN;
  for (int i = threadIdx.x; i < P; i += 256) et[i] = INT_MIN;
  __syncthreads();
  const double2* p = x + m * stride;
  const long long nv2 = (long long)nv * nv;
  const long long n = nblk * nv2;
  for (long long e = threadIdx.x; e < n; e += 256) {
    const int blk = (int)(e / nv2);
    const int r = (int)((e / nv) % nv), cc = (int)(e % nv);
    const double2 v = p[e];
    const int e1 = exp_of(v.x), e2 = exp_of(v.y);
    const int ev = e1 > e2 ? e1 : e2;
    if (ev == INT_MIN) continue;
    atomicMax(&er[mode ? (blk >> 2) * P + r : r], ev);
    if (!mode) atomicMax(&et[cc], ev);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nr; i += 256) ex_r[m * nr + i] = er[i] == INT_MIN ? 0 : er[i];
  if (!mode)
    for (int i = threadIdx.x; i < P; i += 256) ex_t[m * P + i] = et[i] == INT_MIN ? 0 : et[i];
}

// Digits. One workgroup per 32 x 32 tile of a P x P block (zero outside nv x nv): the row-line
// planes (line = row, K = column) always, the column-line planes (line = column, K = row) in mode
// 0. Both are one contiguous kilobyte per plane and tile; the bytes are gathered in LDS first.
__global__ void __launch_bounds__(256) lc_oz_digits_kernel(const double2* __restrict__ x, long long stride, int nv,
                                                           int P, int nblk, int mode, const int* __restrict__ ex_r,
                                                           const int* __restrict__ ex_t, unsigned char* __restrict__ yr,
                                                           unsigned char* __restrict__ yt) {
  __shared__ __attribute__((aligned(16))) unsigned char tr_[8][32][32], tt_[8][32][32];
  const int tpb = P >> 5;
  const long long t = blockIdx.x;
  const long long bi = t / (tpb * tpb);  // global block index: matrix * nblk + block
  const long long m = bi / nblk;
  const int blk = (int)(bi - m * nblk);
  const int tr = (int)((t / tpb) % tpb), tc = (int)(t % tpb);
  const double2* p = x + m * stride + (long long)blk * nv * nv;
  const long long P2 = (long long)P * P;
  const int nr = mode ? 4 * P : P;
#pragma unroll
  for (int pass = 0; pass < 4; pass++) {
    const int r = (threadIdx.x >> 5) + 8 * pass, cc = threadIdx.x & 31;
    const int row = tr * 32 + r, col = tc * 32 + cc;
    double2 v = make_double2(0.0, 0.0);
    if (row < nv && col < nv) v = p[(long long)row * nv + col];
    int dr[4], di[4];
    const int er = ex_r[m * nr + (mode ? (blk >> 2) * P : 0) + row];
    digits_of(v.x, er, dr);
    digits_of(v.y, er, di);
#pragma unroll
    for (int d = 0; d < 4; d++) {
      tr_[d][r][cc] = (unsigned char)dr[d];
      tr_[4 + d][r][cc] = (unsigned char)di[d];
    }
    if (!mode) {
      const int et = ex_t[m * P + col];
      digits_of(v.x, et, dr);
      digits_of(v.y, et, di);
#pragma unroll
      for (int d = 0; d < 4; d++) {
        tt_[d][cc][r] = (unsigned char)dr[d];
        tt_[4 + d][cc][r] = (unsigned char)di[d];
      }
    }
  }
  __syncthreads();
  // 8 planes x 1 KB each: thread k writes 16 bytes of plane k / 32 ... in two passes
  unsigned char* ob = yr + bi * 8 * P2;
#pragma unroll
  for (int pass = 0; pass < 2; pass++) {
    const int k = threadIdx.x + 256 * pass;  // 512 16-byte pieces = 8 planes x 64
    const int q = k >> 6, piece = k & 63;
    const int ln = piece >> 1, half = piece & 1;
    const uint4 v = *reinterpret_cast<const uint4*>(&tr_[q][ln][16 * half]);
    *reinterpret_cast<uint4*>(ob + q * P2 + ((long long)tc * P + tr * 32 + ln) * 32 + 16 * half) = v;
    if (!mode) {
      const uint4 w = *reinterpret_cast<const uint4*>(&tt_[q][ln][16 * half]);
      *reinterpret_cast<uint4*>(yt + bi * 8 * P2 + q * P2 + ((long long)tr * P + tc * 32 + ln) * 32 + 16 * half) = w;
    }
  }
}

// Digits back to FP64 (tests and diagnostics): D-space or elemental digit matrices with
// exponents in the mode-1 layout (mode 1) or per line index (mode 0), row-line planes.
__global__ void __launch_bounds__(256) lc_oz_decode_kernel(const unsigned char* __restrict__ y, const int* __restrict__ ex,
                                                           int nv, int P, int nblk, int mode, long long nmat,
                                                           double2* __restrict__ out) {
  const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long nv2 = (long long)nv * nv;
  if (gid >= nmat * nblk * nv2) return;
  const long long bi = gid / nv2;
  const long long m = bi / nblk;
  const int blk = (int)(bi - m * nblk);
  const int row = (int)((gid / nv) % nv), col = (int)(gid % nv);
  const long long P2 = (long long)P * P;
  const unsigned char* b = y + bi * 8 * P2 + ((long long)(col >> 5) * P + row) * 32 + (col & 31);
  const int nr = mode ? 4 * P : P;
  const int E = ex[m * nr + (mode ? (blk >> 2) * P : 0) + row];
  double v[2];
  for (int part = 0; part < 2; part++) {
    long long s = 0;
    for (int d = 0; d < 4; d++) s = s * 128 + (signed char)b[(part * 4 + d) * P2];
    v[part] = __builtin_ldexp((double)s, E - 27);
  }
  out[gid] = make_double2(v[0], v[1]);
}

}  // namespace lc

// NW = 4 for P = 32 (strips of 128 columns), 8 otherwise (256 columns).
extern "C" hipError_t lc_launch_gemm_oz(const LcOzGroup* d_groups, int ng, long long total_wg, int nv, int P,
                                        hipStream_t s) {
  if (ng <= 0 || total_wg <= 0) return hipSuccess;
  const long long per = (total_wg + 7) / 8;
  const dim3 grid((unsigned)(8 * per));
  if (P == 32)
    hipLaunchKernelGGL((lc::lc_gemm_oz_kernel<4>), grid, dim3(256), 0, s, d_groups, ng, nv, P, total_wg);
  else
    hipLaunchKernelGGL((lc::lc_gemm_oz_kernel<8>), grid, dim3(512), 0, s, d_groups, ng, nv, P, total_wg);
  return hipGetLastError();
}

// x: nmat matrices of nblk blocks (nv x nv complex doubles), `stride` complex elements apart.
// mode 0: y_rows / y_trans (nmat * nblk * 8 P^2 bytes each), ex_r / ex_t (nmat * P ints each).
// mode 1: y_rows and ex_r (nmat * 4 P ints) only.
extern "C" hipError_t lc_launch_oz_convert(const void* x, long long stride, long long nmat, int nv, int P, int nblk,
                                           int mode, int* ex_r, int* ex_t, void* y_rows, void* y_trans,
                                           hipStream_t s) {
  if (nmat <= 0) return hipSuccess;
  hipLaunchKernelGGL(lc::lc_oz_exps_kernel, dim3((unsigned)nmat), dim3(256), 0, s, (const double2*)x, stride, nv, P,
                     nblk, mode, ex_r, ex_t);
  const long long tiles = nmat * nblk * (long long)(P / 32) * (P / 32);
  hipLaunchKernelGGL(lc::lc_oz_digits_kernel, dim3((unsigned)tiles), dim3(256), 0, s, (const double2*)x, stride, nv,
                     P, nblk, mode, (const int*)ex_r, (const int*)ex_t, (unsigned char*)y_rows,
                     (unsigned char*)y_trans);
  return hipGetLastError();
}

extern "C" hipError_t lc_launch_oz_decode(const void* y, const int* ex, int nv, int P, int nblk, int mode,
                                          long long nmat, void* out, hipStream_t s) {
  const long long n = nmat * nblk * (long long)nv * nv;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(lc::lc_oz_decode_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     (const unsigned char*)y, ex, nv, P, nblk, mode, nmat, (double2*)out);
  return hipGetLastError();
}
