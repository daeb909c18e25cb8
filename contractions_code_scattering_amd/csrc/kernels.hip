// HIP kernels of the contraction engine (gfx950 / CDNA4), FP64 and FP32.
//
// lc_gemm_kernel   grouped block-complex GEMM  C = alpha * A * B  on D-space views, on the
//                  matrix cores (v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32), complex
//                  product as four real MFMAs into separate real and imaginary accumulators;
//                  optionally fused closing traces Tr[C * Y_t] so C need not be stored.
// lc_trace_kernel  grouped stand-alone closing trace  out = alpha * Tr[X * Y] = sum_ij X_ij Y_ji.
// lc_reduce_kernel fixed-order sum of the per-tile trace pieces (deterministic results).
//
// A GEMM workgroup (256 threads = 4 wave64 in a 2x2 arrangement) owns one TM x TM tile of one
// spin block (a,c) of one batch item. It loops over the spin index b of the product (4 terms
// for dense x dense, 1 when an operand is an elemental M (x) 1) and over K in chunks of KT,
// staging both operand tiles through double-buffered LDS with the view's phase applied; the
// next chunk is fetched into registers while the matrix cores work on the current one, so one
// barrier per chunk suffices. Workgroups are renumbered so that every tile of a product runs
// on the same XCD (shared L2).
#include "lc_device.h"

namespace lc {

// K chunks of 8 keep the LDS stages at 34 KB (FP64, TM = 64) and the registers within three waves
// per SIMD, so that three workgroups share a CU: one workgroup's first loads and epilogue then
// overlap the others' matrix-core work (chunks of 16 allowed two: 13% slower).
template <typename T, int TM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
lc_gemm_kernel(const LcGemmGroup* __restrict__ groups, int ng, int nv, long long total) {
  constexpr int KT = 8;
  constexpr int E = TM * KT / 256;  // complex elements staged per thread per operand
  constexpr int WM = TM / 2;        // per-wave output tile (2 x 2 waves)
  constexpr int MI = WM / 16;       // 16x16 MFMA tiles per wave side
  constexpr int PAD = 2;
  typedef typename Acc4<T>::type acc_t;

  __shared__ T As_re[2][KT][TM + PAD], As_im[2][KT][TM + PAD];
  __shared__ T Bs_re[2][KT][TM + PAD], Bs_im[2][KT][TM + PAD];
  __shared__ double tred[LC_MAX_FUSED_TRACES][2][4];

  const long long wg = xcd_order(total);
  if (wg < 0) return;
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

  // Fetch chunk `ch` of both operands into registers as stored; the view's conjugation and the
  // term's phase are applied when the chunk goes to LDS, so that the loads are waited for only
  // after the matrix cores have worked through the current chunk.
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
      const T sr = A.sr * B.sr - A.si * B.si, si = A.sr * B.si + A.si * B.sr;
      A.sr = sr;
      A.si = si;
      cur_term = term;
    }
    if (A.flags & LC_TV) {
      const int k = tid / (TM / E), i0 = (tid % (TM / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int i = row0 + i0 + e, kk = k0 + k;
        ra[e] = (i < nv && kk < nv) ? gload_c(A.p + (long long)kk * nv + i) : Cplx<T>{T(0), T(0)};
      }
    } else {
      const int i = tid / (KT / E), kq = (tid % (KT / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int ii = row0 + i, kk = k0 + kq + e;
        ra[e] = (ii < nv && kk < nv) ? gload_c(A.p + (long long)ii * nv + kk) : Cplx<T>{T(0), T(0)};
      }
    }
    if (B.flags & LC_TV) {
      const int j = tid / (KT / E), kq = (tid % (KT / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int jj = col0 + j, kk = k0 + kq + e;
        rb[e] = (jj < nv && kk < nv) ? gload_c(B.p + (long long)jj * nv + kk) : Cplx<T>{T(0), T(0)};
      }
    } else {
      const int k = tid / (TM / E), j0 = (tid % (TM / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int jj = col0 + j0 + e, kk = k0 + k;
        rb[e] = (jj < nv && kk < nv) ? gload_c(B.p + (long long)kk * nv + jj) : Cplx<T>{T(0), T(0)};
      }
    }
  };

  auto store = [&](int buf) {
    // A: conjugation, then the term's phase; B: conjugation
    const T ca = (A.flags & LC_CONJ) ? T(-1) : T(1), cb = (B.flags & LC_CONJ) ? T(-1) : T(1);
#pragma unroll
    for (int e = 0; e < E; e++) {
      const T zx = ra[e].x, zy = ca * ra[e].y;
      ra[e].x = A.sr * zx - A.si * zy;
      ra[e].y = A.sr * zy + A.si * zx;
      rb[e].y = cb * rb[e].y;
    }
    if (A.flags & LC_TV) {
      const int k = tid / (TM / E), i0 = (tid % (TM / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) { As_re[buf][k][i0 + e] = ra[e].x; As_im[buf][k][i0 + e] = ra[e].y; }
    } else {
      const int i = tid / (KT / E), kq = (tid % (KT / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) { As_re[buf][kq + e][i] = ra[e].x; As_im[buf][kq + e][i] = ra[e].y; }
    }
    if (B.flags & LC_TV) {
      const int j = tid / (KT / E), kq = (tid % (KT / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) { Bs_re[buf][kq + e][j] = rb[e].x; Bs_im[buf][kq + e][j] = rb[e].y; }
    } else {
      const int k = tid / (TM / E), j0 = (tid % (TM / E)) * E;
#pragma unroll
      for (int e = 0; e < E; e++) { Bs_re[buf][k][j0 + e] = rb[e].x; Bs_im[buf][k][j0 + e] = rb[e].y; }
    }
  };

  load(0);
  store(0);
  __syncthreads();
  for (int ch = 0; ch < nchunk; ch++) {
    const int buf = ch & 1;
    if (ch + 1 < nchunk) load(ch + 1);
#pragma unroll
    for (int ks = 0; ks < KT / 4; ks++) {
      const int kr = ks * 4 + (lane >> 4);
      T are[MI], aim[MI], bre[MI], bim[MI];
#pragma unroll
      for (int m = 0; m < MI; m++) {
        const int rr = wm * WM + m * 16 + (lane & 15);
        are[m] = As_re[buf][kr][rr];
        aim[m] = As_im[buf][kr][rr];
        const int cc = wn * WM + m * 16 + (lane & 15);
        bre[m] = Bs_re[buf][kr][cc];
        bim[m] = Bs_im[buf][kr][cc];
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
    // buffer buf^1 was last read in iteration ch-1, which every wave finished before the
    // barrier that closed it: it can be refilled now.
    if (ch + 1 < nchunk) store(buf ^ 1);
    __syncthreads();
  }

  // Epilogue: optionally block (a,c) of alpha C, and the fused trace pieces, which are linear in C:
  // alpha Tr[C Y] is formed from the accumulators as they are and scaled once (no second copy of
  // the tile in registers, which would keep the occupancy at two waves per SIMD).
  const T alr = T(g.alpha[0]), ali = T(g.alpha[1]);
  Cplx<T>* C = g.C ? reinterpret_cast<Cplx<T>*>(g.C) + voff(g.cs, co) + (long long)(a * 4 + c) * nv2
                   : nullptr;
  if (C) {
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
            gstore_c(C + (long long)row * nv + col, z);
          }
        }
  }
  // Fused trace pieces sum_ij C_ij Y_ji: per trace, all 4 * MI * MI loads of this lane are issued
  // before the first is used (one memory round trip per trace), then one combined reduction.
  double tr_re[LC_MAX_FUSED_TRACES], tr_im[LC_MAX_FUSED_TRACES];
#pragma unroll
  for (int t = 0; t < LC_MAX_FUSED_TRACES; t++) {
    tr_re[t] = 0;
    tr_im[t] = 0;
    if (t >= g.ntr) continue;
    const BlockRef<T> Y = view_block<T>(g.Y[t], voff(g.Y[t].s, co), c, a, nv2);
    if (Y.zero) continue;
    const bool tv = (Y.flags & LC_TV) != 0;
    Cplx<T> yv[MI][MI][4];
#pragma unroll
    for (int m = 0; m < MI; m++)
#pragma unroll
      for (int n = 0; n < MI; n++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int row = row0 + wm * WM + m * 16 + acc_row<T>(lane, q);
          const int col = col0 + wn * WM + n * 16 + (lane & 15);
          const long long idx = tv ? (long long)row * nv + col : (long long)col * nv + row;
          yv[m][n][q] = (row < nv && col < nv) ? gload_c(Y.p + idx) : Cplx<T>{T(0), T(0)};
        }
    const double cj = (Y.flags & LC_CONJ) ? -1.0 : 1.0;
    double sr = 0, si = 0;
#pragma unroll
    for (int m = 0; m < MI; m++)
#pragma unroll
      for (int n = 0; n < MI; n++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const double yr = yv[m][n][q].x, yi = cj * (double)yv[m][n][q].y;
          const double xr = cre[m][n][q], xi = cim[m][n][q];
          sr += xr * yr - xi * yi;
          si += xr * yi + xi * yr;
        }
    // alpha, the Y block phase, then the trace coefficient
    const double ur = alr * sr - ali * si, ui = alr * si + ali * sr;
    const double yr = Y.sr, yi = Y.si;
    const double pr = ur * yr - ui * yi, pi = ur * yi + ui * yr;
    const double ar = g.tr_alpha[t][0], ai = g.tr_alpha[t][1];
    tr_re[t] = ar * pr - ai * pi;
    tr_im[t] = ar * pi + ai * pr;
  }
  if (g.ntr > 0) {
    // one reduction for every trace: waves -> LDS -> threads 0..ntr-1
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
      const int t = tid;
      typedef double d2v __attribute__((ext_vector_type(2)));
      double* out = g.part[t] + 2 * (item * per_item + r);
      gstore(out, d2v{tred[t][0][0] + tred[t][0][1] + tred[t][0][2] + tred[t][0][3],
                      tred[t][1][0] + tred[t][1][1] + tred[t][1][2] + tred[t][1][3]});
    }
  }
}

// ------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) lc_trace_kernel(const LcTraceGroup* __restrict__ groups,
                                                       int ng, int nv) {
  __shared__ double red[8];
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

  double accr = 0, acci = 0;
  for (int pr = 0; pr < npair; pr++) {
    int a, b;
    if (npair == 4) { a = pr; b = pr; } else { a = pr >> 2; b = pr & 3; }
    const BlockRef<T> X = view_block<T>(g.X, offX, a, b, nv2);
    const BlockRef<T> Y = view_block<T>(g.Y, offY, b, a, nv2);
    const double sr = (double)X.sr * Y.sr - (double)X.si * Y.si;
    const double si = (double)X.sr * Y.si + (double)X.si * Y.sr;
    T pr_ = 0, pi_ = 0;
    for (int e = threadIdx.x; e < (int)nv2; e += 256) {
      const int i = e / nv, j = e - i * nv;
      const Cplx<T> x = elem(X, i, j, nv);
      const Cplx<T> y = elem(Y, j, i, nv);
      pr_ += x.x * y.x - x.y * y.y;
      pi_ += x.x * y.y + x.y * y.x;
    }
    accr += sr * pr_ - si * pi_;
    acci += sr * pi_ + si * pr_;
  }
  block_sum2(accr, acci, red);
  if (threadIdx.x == 0) {
    const double alr = g.alpha[0], ali = g.alpha[1];
    Cplx<T>* out = reinterpret_cast<Cplx<T>*>(g.out) + voff(g.os, co);
    double zx = alr * accr - ali * acci, zy = alr * acci + ali * accr;
    if (g.accumulate) { zx += out->x; zy += out->y; }
    out->x = T(zx);
    out->y = T(zy);
  }
}

// One thread per item: out[item] (+)= sum_k part[item][k], in a fixed order.
__global__ void lc_reduce_kernel(const LcReduceJob* __restrict__ jobs, int njobs,
                                 const long long* __restrict__ begins) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int j = 0;
  while (j + 1 < njobs && begins[j + 1] <= gid) j++;
  const LcReduceJob& J = jobs[j];
  const long long item = gid - begins[j];
  if (item >= (long long)J.n[0] * J.n[1] * J.n[2]) return;
  long long co[3];
  decode_item(item, J.n, co);
  const double* p = J.part + 2 * item * J.nk;
  double sr = 0, si = 0;
  for (int k = 0; k < J.nk; k++) {
    sr += p[2 * k];
    si += p[2 * k + 1];
  }
  if (J.conj) si = -si;
  const long long o = voff(J.os, co);
  if (J.out_double) {
    double* out = reinterpret_cast<double*>(J.out) + 2 * o;
    if (J.accumulate) { sr += out[0]; si += out[1]; }
    out[0] = sr;
    out[1] = si;
  } else {
    float* out = reinterpret_cast<float*>(J.out) + 2 * o;
    if (J.accumulate) { sr += out[0]; si += out[1]; }
    out[0] = (float)sr;
    out[1] = (float)si;
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

// ------------------------------------------------------------------------------------------
// Engine epilogue (FP64): per (p, source k, sink tf) combine the raw traces of the rho / pi-pi
// set into correlators, then average over sources on the device.
//   raw  [np][nsrc][nt][10] = {rho_0, rho_1, rho_2, square, dT, dTHor, st1, st2, fi1, fi2}
//   corr [np][nsrc][nt][8]  = {rho_0, rho_1, rho_2, square, dT, dTHor, st1*st2, fi1*fi2}
//   avg  [np][8][nt]        C(dt) = 1/nsrc sum_k corr[p][k][(t0_k + dt) % nt]
// With per-slice rescaled inputs (mixed path) every trace is multiplied back by the FP64 scale
// factors of the matrices it contains: s(tf,t0) of tau, s(t,t) of D(t) = tau(t,t), s(p,t) of
// Phi. rho, star and fish traces hold tau^2 Phi(tf) Phi(t0); square and doubleTriangle
// tau^2 D(tf) D(t0) Phi(tf)^2 Phi(t0)^2; doubleTriangleHor tau^4 Phi(tf)^2 Phi(t0)^2.
// Scale factors (mixed form) of each raw value: per-slice scales of tau (s_tau [t][t0]) and of the
// four elemental roles P1..P4 (s_p[k][item * ps + t]); which factors a diagram carries follows
// its chain: rho and star1 P4(tf) tau P1(t0) tau^dag, star2 P3 tau P2 tau^dag, fish1 P3 . P1,
// fish2 P4 . P2, square and dT all four with D(tf) and D(t0), dTHor all four with four tau.
struct LcRoleScales {
  const double* s[4];
  long long ps;
};

__global__ void lc_engine_combine_kernel(const double2* __restrict__ raw, double2* __restrict__ corr,
                                         long long n, int nsrc, int nt, int src0, int src_step,
                                         const double* __restrict__ s_tau, LcRoleScales sp) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2* r = raw + i * 10;
  double2* c = corr + i * 8;
  double f[10];
#pragma unroll
  for (int d = 0; d < 10; d++) f[d] = 1.0;
  if (s_tau) {
    const int tf = (int)(i % nt);
    const int k = (int)((i / nt) % nsrc);
    const long long p = i / ((long long)nt * nsrc);
    const int t0 = src0 + k * src_step;
    const double st = s_tau[(long long)tf * nt + t0];
    const double dtf = s_tau[(long long)tf * nt + tf], dt0 = s_tau[(long long)t0 * nt + t0];
    const double s1 = sp.s[0][p * sp.ps + t0], s2 = sp.s[1][p * sp.ps + t0];
    const double s3 = sp.s[2][p * sp.ps + tf], s4 = sp.s[3][p * sp.ps + tf];
    const double t2 = st * st, all4 = s1 * s2 * s3 * s4;
    f[0] = f[1] = f[2] = t2 * s4 * s1;
    f[3] = f[4] = t2 * dtf * dt0 * all4;
    f[5] = t2 * t2 * all4;
    f[6] = t2 * s4 * s1;
    f[7] = t2 * s3 * s2;
    f[8] = t2 * s3 * s1;
    f[9] = t2 * s4 * s2;
  }
  double2 v[10];
#pragma unroll
  for (int d = 0; d < 10; d++) v[d] = make_double2(r[d].x * f[d], r[d].y * f[d]);
#pragma unroll
  for (int d = 0; d < 6; d++) c[d] = v[d];
  c[6].x = v[6].x * v[7].x - v[6].y * v[7].y;
  c[6].y = v[6].x * v[7].y + v[6].y * v[7].x;
  c[7].x = v[8].x * v[9].x - v[8].y * v[9].y;
  c[7].y = v[8].x * v[9].y + v[8].y * v[9].x;
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

// ------------------------------------------------------------------------------------------
// Perambulator file slab -> device layout. The file holds complex double in the order
// [s1][v1][t1][s2][v2][t2] (lib/contract.cpp:111-128); one (s1, v1) slab is [t1][s2][v2][t2].
// Element (t1, s2, v2, t2) goes to tau[t1][t2][s1][s2][v1][v2] (lib/contract.cpp:140,152), in FP64
// or rounded to FP32. Threads run over the destination order (v2 fastest) so the writes coalesce.
template <typename T>
__global__ void lc_prop_scatter_kernel(const double2* __restrict__ slab, T* __restrict__ tau, int nt, int nv,
                                       int s1, int v1) {
  const long long n = (long long)nt * nt * 4 * nv;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v2 = (int)(i % nv);
  const int s2 = (int)((i / nv) % 4);
  const int t2 = (int)((i / (4LL * nv)) % nt);
  const int t1 = (int)(i / (4LL * nv * nt));
  const double2 z = slab[(((long long)t1 * 4 + s2) * nv + v2) * nt + t2];
  const long long dst = (((((long long)t1 * nt + t2) * 4 + s1) * 4 + s2) * nv + v1) * nv + v2;
  tau[2 * dst] = (T)z.x;
  tau[2 * dst + 1] = (T)z.y;
}

}  // namespace lc

// ------------------------------------------------------------------------------------------
// Mixed form of the sigma / sigma-pi / rho-pi engine: multiply every raw trace of
// raw [np][nsrc][nt][10] back by the FP64 per-slice scales of the matrices it contains:
// s(tf,t0) of tau, s(t,t) of D(t), s(p,t,j) of elemental type j (types as in engine_sigma.cpp).
//   sigma(l, r):  tau^2 X_l(tf) X_r(t0)
//   triangle(ip), rho-pi(iop): tau^2 D(t0) P1(t0) P2(t0) times X_ip(tf) or P12_ip(tf)
__global__ void lc_sigma_scale_kernel(double2* __restrict__ raw, long long n, int nsrc, int nt, int src0,
                                      int src_step, int idir, const double* __restrict__ s_tau,
                                      const double* __restrict__ s_el) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int tf = (int)(i % nt);
  const int k = (int)((i / nt) % nsrc);
  const long long p = i / ((long long)nt * nsrc);
  const int t0 = src0 + k * src_step;
  const double st = s_tau[(long long)tf * nt + t0], d0 = s_tau[(long long)t0 * nt + t0];
  const double* ef = s_el + (p * nt + tf) * 11;
  const double* e0 = s_el + (p * nt + t0) * 11;
  const double base = st * st * d0 * e0[9] * e0[10];
  const int ip_of_op[4] = {0, 0, 1, 2 + idir};
  double f[10];
  for (int l = 0; l < 2; l++)
    for (int r = 0; r < 2; r++) f[l * 2 + r] = st * st * ef[l] * e0[r];
  f[4] = base * ef[2];
  f[5] = base * ef[3];
  for (int op = 0; op < 4; op++) f[6 + op] = base * ef[4 + ip_of_op[op]];
  double2* r = raw + i * 10;
  for (int d = 0; d < 10; d++) {
    const double2 v = r[d];
    r[d] = make_double2(v.x * f[d], v.y * f[d]);
  }
}

// ------------------------------------------------------------------------------------------
// Host launchers (called from the planner).
extern "C" int lc_gemm_tile(int nv) { return nv <= 32 ? 32 : 64; }

extern "C" hipError_t lc_launch_gemm(int is_double, const LcGemmGroup* d_groups, int ng,
                                     long long total_wg, int nv, hipStream_t s) {
  if (ng <= 0 || total_wg <= 0) return hipSuccess;
  const long long per = (total_wg + 7) / 8;
  const dim3 grid((unsigned)(8 * per)), block(256);
  if (is_double) {
    if (nv <= 32) hipLaunchKernelGGL((lc::lc_gemm_kernel<double, 32>), grid, block, 0, s, d_groups, ng, nv, total_wg);
    else hipLaunchKernelGGL((lc::lc_gemm_kernel<double, 64>), grid, block, 0, s, d_groups, ng, nv, total_wg);
  } else {
    if (nv <= 32) hipLaunchKernelGGL((lc::lc_gemm_kernel<float, 32>), grid, block, 0, s, d_groups, ng, nv, total_wg);
    else hipLaunchKernelGGL((lc::lc_gemm_kernel<float, 64>), grid, block, 0, s, d_groups, ng, nv, total_wg);
  }
  return hipGetLastError();
}

extern "C" hipError_t lc_launch_trace(int is_double, const LcTraceGroup* d_groups, int ng,
                                      long long total_wg, int nv, hipStream_t s) {
  if (ng <= 0 || total_wg <= 0) return hipSuccess;
  const dim3 grid((unsigned)total_wg), block(256);
  if (is_double) hipLaunchKernelGGL((lc::lc_trace_kernel<double>), grid, block, 0, s, d_groups, ng, nv);
  else hipLaunchKernelGGL((lc::lc_trace_kernel<float>), grid, block, 0, s, d_groups, ng, nv);
  return hipGetLastError();
}

extern "C" hipError_t lc_launch_reduce(const LcReduceJob* d_jobs, const long long* d_begins, int njobs,
                                       long long total, hipStream_t s) {
  if (njobs <= 0 || total <= 0) return hipSuccess;
  hipLaunchKernelGGL(lc::lc_reduce_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     d_jobs, njobs, d_begins);
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

// s_p: scales of the roles P1..P4 (item p, time t at s_p[k][p * ps + t]); unused without s_tau.
extern "C" hipError_t lc_launch_engine_epilogue(const void* raw, void* corr, void* avg, int np,
                                                int nsrc, int nt, int src0, int src_step,
                                                const double* s_tau, const double* const* s_p,
                                                long long ps, hipStream_t s) {
  const long long n = (long long)np * nsrc * nt;
  lc::LcRoleScales sp;
  for (int k = 0; k < 4; k++) sp.s[k] = s_tau ? s_p[k] : nullptr;
  sp.ps = ps;
  hipLaunchKernelGGL(lc::lc_engine_combine_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     (const double2*)raw, (double2*)corr, n, nsrc, nt, src0, src_step, s_tau, sp);
  const int m = np * 8 * nt;
  hipLaunchKernelGGL(lc::lc_engine_average_kernel, dim3((m + 255) / 256), dim3(256), 0, s,
                     (const double2*)corr, (double2*)avg, np, nsrc, nt, 8, src0, src_step);
  return hipGetLastError();
}

extern "C" hipError_t lc_launch_sigma_scale(void* raw, int np, int nsrc, int nt, int src0, int src_step,
                                            int idir, const double* s_tau, const double* s_el, hipStream_t s) {
  const long long n = (long long)np * nsrc * nt;
  hipLaunchKernelGGL(lc_sigma_scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (double2*)raw, n,
                     nsrc, nt, src0, src_step, idir, s_tau, s_el);
  return hipGetLastError();
}

extern "C" hipError_t lc_launch_prop_scatter(const void* slab, void* tau, int is_double, int nt, int nv, int s1,
                                             int v1, hipStream_t s) {
  const long long n = (long long)nt * nt * 4 * nv;
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (is_double)
    hipLaunchKernelGGL(lc::lc_prop_scatter_kernel<double>, dim3(grid), dim3(256), 0, s, (const double2*)slab,
                       (double*)tau, nt, nv, s1, v1);
  else
    hipLaunchKernelGGL(lc::lc_prop_scatter_kernel<float>, dim3(grid), dim3(256), 0, s, (const double2*)slab,
                       (float*)tau, nt, nv, s1, v1);
  return hipGetLastError();
}
