// Elementals from Laplacian eigenvectors on the FP64 matrix cores:
//
//   Phi(p, t)[r][c] = sum_k conj(V_t[r][k]) * ph_p[k] * V_t[c][k]
//
// the plain momentum matrix that the reference's drivers build on the host, one scalar product per
// (r, c) of eigenvector r with the phase-multiplied eigenvector c (tests/compute_rho.cpp:116-129,
// ph_p[k] = exp(-2 pi i p.x / L) at the site of entry k). V is [nt][nv][nsite] complex double (k runs
// over site and colour in any order the phases follow), ph is [np][nsite], Phi is [np][nt][nv][nv].
//
// Per (p, t) this is a complex GEMM with a long K (3 L^3): A[r][k] = conj(V_t[r][k]), B[k][c] =
// ph_p[k] V_t[c][k]. A workgroup (4 wave64, 2 x 2) owns a 64 x 64 tile of Phi and one K slice,
// staging 16-wide K chunks through double-buffered LDS (the phase is applied while staging), and
// writes its partial tile; a second kernel sums the K slices in a fixed order.
//
// The kernels take the left and right vector sets separately, X^dag ph Y, so that the derivative
// types of the drivers (tests/compute_rho.cpp:131-202) are sums of such bilinears over the
// covariantly differentiated eigenvectors; the reduction adds alpha * (X^dag ph Y) into, or in place
// of, its destination.
//
// The covariant derivative (lc_covderiv_kernel) is a nearest-neighbour stencil on the spatial
// lattice of one time slice: (nabla_j v)(x) = 1/2 [U_j(x) v(x + j) - U_j(x - j)^dag v(x - j)],
// periodic, one thread per (time, direction, site) with its three colours, looping over the vectors.
#include "lc_device.h"

namespace lc {

constexpr int PH_TM = 64, PH_KT = 16;

__global__ void __launch_bounds__(256) lc_phi_partial_kernel(const double2* __restrict__ X,
                                                             const double2* __restrict__ Y,
                                                             const double2* __restrict__ ph, int nt, int nv,
                                                             long long nsite, int tiles, int nks,
                                                             long long kslice, double2* __restrict__ part) {
  __shared__ double As_re[2][PH_KT][PH_TM], As_im[2][PH_KT][PH_TM];
  __shared__ double Bs_re[2][PH_KT][PH_TM], Bs_im[2][PH_KT][PH_TM];
  const long long wg = blockIdx.x;
  const int ks = (int)(wg % nks);
  const long long rest = wg / nks;
  const int tile = (int)(rest % (tiles * tiles));
  const long long item = rest / (tiles * tiles);  // p * nt + t
  const int t = (int)(item % nt);
  const long long p = item / nt;
  const int row0 = (tile / tiles) * PH_TM, col0 = (tile % tiles) * PH_TM;
  const long long k_begin = (long long)ks * kslice;
  const long long k_end = k_begin + kslice < nsite ? k_begin + kslice : nsite;
  const int nchunk = k_end > k_begin ? (int)((k_end - k_begin + PH_KT - 1) / PH_KT) : 0;

  const double2* Xt = X + (long long)t * nv * nsite;
  const double2* Yt = Y + (long long)t * nv * nsite;
  const double2* php = ph + p * nsite;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  // staging: line (row of A, column of B) tid / 4, four consecutive k from 4 * (tid % 4)
  const int s_line = tid >> 2, s_k = (tid & 3) * 4;

  double2 ra[4], rb[4];
  auto load = [&](int ch) {
    const long long k0 = k_begin + (long long)ch * PH_KT + s_k;
    const int r = row0 + s_line, c = col0 + s_line;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const long long k = k0 + e;
      const bool kin = k < k_end;
      ra[e] = (kin && r < nv) ? Xt[(long long)r * nsite + k] : double2{0.0, 0.0};
      if (kin && c < nv) {
        const double2 v = Yt[(long long)c * nsite + k], f = php[k];
        rb[e] = double2{f.x * v.x - f.y * v.y, f.x * v.y + f.y * v.x};
      } else {
        rb[e] = double2{0.0, 0.0};
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      As_re[buf][s_k + e][s_line] = ra[e].x;
      As_im[buf][s_k + e][s_line] = -ra[e].y;  // conj(X)
      Bs_re[buf][s_k + e][s_line] = rb[e].x;
      Bs_im[buf][s_k + e][s_line] = rb[e].y;
    }
  };

  dx4 cre[2][2], cim[2][2];
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 2; n++) { cre[m][n] = dx4{0, 0, 0, 0}; cim[m][n] = dx4{0, 0, 0, 0}; }

  if (nchunk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int ch = 0; ch < nchunk; ch++) {
    const int buf = ch & 1;
    if (ch + 1 < nchunk) load(ch + 1);
#pragma unroll
    for (int kq = 0; kq < PH_KT / 4; kq++) {
      const int kr = kq * 4 + (lane >> 4);
      double are[2], aim[2], bre[2], bim[2];
#pragma unroll
      for (int m = 0; m < 2; m++) {
        const int rr = wm * 32 + m * 16 + (lane & 15);
        are[m] = As_re[buf][kr][rr];
        aim[m] = As_im[buf][kr][rr];
        const int cc = wn * 32 + m * 16 + (lane & 15);
        bre[m] = Bs_re[buf][kr][cc];
        bim[m] = Bs_im[buf][kr][cc];
      }
#pragma unroll
      for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++) {
          cre[m][n] = mfma16(are[m], bre[n], cre[m][n]);
          cre[m][n] = mfma16(-aim[m], bim[n], cre[m][n]);
          cim[m][n] = mfma16(are[m], bim[n], cim[m][n]);
          cim[m][n] = mfma16(aim[m], bre[n], cim[m][n]);
        }
    }
    // buffer buf^1 was last read before the barrier that closed the previous iteration
    if (ch + 1 < nchunk) store(buf ^ 1);
    __syncthreads();
  }
  // the partial tile, 64 x 64 row-major
  double2* out = part + (rest * nks + ks) * (long long)(PH_TM * PH_TM);
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 2; n++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = wm * 32 + m * 16 + acc_row<double>(lane, q);
        const int col = wn * 32 + n * 16 + (lane & 15);
        out[row * PH_TM + col] = double2{cre[m][n][q], cim[m][n][q]};
      }
}

// Phi[dst(item)][r][c] (+)= alpha * sum over K slices of the partial tiles, in slice order; item =
// p * nt + t goes to matrix p * pstride + t of phi.
__global__ void __launch_bounds__(256) lc_phi_reduce_kernel(const double2* __restrict__ part, int nv, int tiles,
                                                            int nks, long long items, int nt, long long pstride,
                                                            double alpha, int accumulate,
                                                            double2* __restrict__ phi) {
  const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long nv2 = (long long)nv * nv;
  if (gid >= items * nv2) return;
  const long long item = gid / nv2;
  const int e = (int)(gid - item * nv2);
  const int r = e / nv, c = e - r * nv;
  const int tile = (r / PH_TM) * tiles + c / PH_TM;
  const double2* p = part + ((item * tiles * tiles + tile) * nks) * (long long)(PH_TM * PH_TM) +
                     (r % PH_TM) * PH_TM + (c % PH_TM);
  double sr = 0, si = 0;
  for (int k = 0; k < nks; k++) {
    const double2 v = p[(long long)k * PH_TM * PH_TM];
    sr += v.x;
    si += v.y;
  }
  const long long ip = item / nt;
  double2* d = phi + (ip * pstride + (item - ip * nt)) * nv2 + e;
  if (accumulate) {
    const double2 o = *d;
    *d = double2{o.x + alpha * sr, o.y + alpha * si};
  } else {
    *d = double2{alpha * sr, alpha * si};
  }
}

// dV[j][t][c][3 * site + a] = (nabla_j V[t][c])(site)_a = cn [U_j(x) v(x + j) - U_j(x - j)^dag v(x - j)]_a
// (cn = 1/2: the central difference); U [nt][3][nsite][3][3] row-major SU(3) links U_j(x) (colour
// index a, b), V [nt][nv][3 * nsite], sites x + nx (y + ny z).
// One thread per (t, j, site) keeps U_j(x) and U_j(x - j) in registers and walks the nv vectors,
// so the links are read once; the three directions of a site chunk are neighbouring workgroups,
// so the vector reads of one direction find the others' in L2.
__global__ void __launch_bounds__(256) lc_covderiv_kernel(const double2* __restrict__ U,
                                                          const double2* __restrict__ V, int nt, int nv, int nx,
                                                          int ny, int nz, double cn, double2* __restrict__ dV) {
  const long long nsite = (long long)nx * ny * nz;
  const long long nchunk = (nsite + 255) / 256;
  const long long b = blockIdx.x;
  const int j = (int)(b % 3);
  const long long chunk = (b / 3) % nchunk;
  const int t = (int)(b / (3 * nchunk));
  const long long x = chunk * 256 + threadIdx.x;
  if (x >= nsite) return;
  const int ix = (int)(x % nx), iy = (int)((x / nx) % ny), iz = (int)(x / ((long long)nx * ny));
  int fx = ix, fy = iy, fz = iz, bx = ix, by = iy, bz = iz;
  if (j == 0) { fx = ix + 1 == nx ? 0 : ix + 1; bx = ix == 0 ? nx - 1 : ix - 1; }
  if (j == 1) { fy = iy + 1 == ny ? 0 : iy + 1; by = iy == 0 ? ny - 1 : iy - 1; }
  if (j == 2) { fz = iz + 1 == nz ? 0 : iz + 1; bz = iz == 0 ? nz - 1 : iz - 1; }
  const long long xf = fx + (long long)nx * (fy + (long long)ny * fz);
  const long long xb = bx + (long long)nx * (by + (long long)ny * bz);
  const double2* Uf = U + (((long long)t * 3 + j) * nsite + x) * 9;   // U_j(x)
  const double2* Ub = U + (((long long)t * 3 + j) * nsite + xb) * 9;  // U_j(x - j)
  double2 f[9], g[9];
#pragma unroll
  for (int e = 0; e < 9; e++) {
    f[e] = Uf[e];
    g[e] = Ub[e];
  }
  const long long vstride = 3 * nsite;
  const double2* v = V + (long long)t * nv * vstride;
  double2* out = dV + ((long long)j * nt + t) * nv * vstride + 3 * x;
  for (int c = 0; c < nv; c++) {
    const double2* vc = v + (long long)c * vstride;
    double2 vf[3], vb[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      vf[q] = vc[3 * xf + q];
      vb[q] = vc[3 * xb + q];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
      double re = 0, im = 0;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const double2 u = f[3 * a + q], w = g[3 * q + a];  // U_j(x)_aq, U_j(x - j)_qa
        re += u.x * vf[q].x - u.y * vf[q].y;
        im += u.x * vf[q].y + u.y * vf[q].x;
        re -= w.x * vb[q].x + w.y * vb[q].y;  // conj(w) * vb
        im -= w.x * vb[q].y - w.y * vb[q].x;
      }
      out[(long long)c * vstride + a] = double2{cn * re, cn * im};
    }
  }
}

}  // namespace lc

// Partial buffer size for lc_launch_phi, in bytes.
extern "C" long long lc_phi_partial_bytes(int nt, int nv, int np, int nks) {
  const long long tiles = (nv + lc::PH_TM - 1) / lc::PH_TM;
  return (long long)np * nt * tiles * tiles * nks * lc::PH_TM * lc::PH_TM * 16;
}

// phi[p * pstride + t] (+)= alpha * X_t^dag diag(ph_p) Y_t for every (p < np, t < nt).
extern "C" hipError_t lc_launch_phi_bilinear(const void* X, const void* Y, const void* ph, int nt, int nv,
                                             long long nsite, int np, int nks, void* part, double alpha,
                                             int accumulate, void* phi, long long pstride, hipStream_t s) {
  const int tiles = (nv + lc::PH_TM - 1) / lc::PH_TM;
  const long long items = (long long)np * nt;
  long long kslice = (nsite + nks - 1) / nks;
  kslice = (kslice + lc::PH_KT - 1) / lc::PH_KT * lc::PH_KT;
  const long long nwg = items * tiles * tiles * nks;
  hipLaunchKernelGGL(lc::lc_phi_partial_kernel, dim3((unsigned)nwg), dim3(256), 0, s, (const double2*)X,
                     (const double2*)Y, (const double2*)ph, nt, nv, nsite, tiles, nks, kslice, (double2*)part);
  const long long n = items * nv * nv;
  hipLaunchKernelGGL(lc::lc_phi_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     (const double2*)part, nv, tiles, nks, items, nt, pstride, alpha, accumulate, (double2*)phi);
  return hipGetLastError();
}

extern "C" hipError_t lc_launch_phi(const void* V, const void* ph, int nt, int nv, long long nsite, int np, int nks,
                                    void* part, void* phi, hipStream_t s) {
  return lc_launch_phi_bilinear(V, V, ph, nt, nv, nsite, np, nks, part, 1.0, 0, phi, nt, s);
}

extern "C" hipError_t lc_launch_covderiv(const void* U, const void* V, int nt, int nv, int nx, int ny, int nz,
                                         double c, void* dV, hipStream_t s) {
  const long long nchunk = ((long long)nx * ny * nz + 255) / 256;
  hipLaunchKernelGGL(lc::lc_covderiv_kernel, dim3((unsigned)(3LL * nt * nchunk)), dim3(256), 0, s,
                     (const double2*)U, (const double2*)V, nt, nv, nx, ny, nz, c, (double2*)dV);
  return hipGetLastError();
}
