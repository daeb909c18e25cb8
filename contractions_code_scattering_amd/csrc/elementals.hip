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
#include "lc_device.h"

namespace lc {

constexpr int PH_TM = 64, PH_KT = 16;

__global__ void __launch_bounds__(256) lc_phi_partial_kernel(const double2* __restrict__ V,
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

  const double2* Vt = V + (long long)t * nv * nsite;
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
      ra[e] = (kin && r < nv) ? Vt[(long long)r * nsite + k] : double2{0.0, 0.0};
      if (kin && c < nv) {
        const double2 v = Vt[(long long)c * nsite + k], f = php[k];
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
      As_im[buf][s_k + e][s_line] = -ra[e].y;  // conj(V)
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

// Phi[item][r][c] = sum over K slices of the partial tiles, in slice order.
__global__ void __launch_bounds__(256) lc_phi_reduce_kernel(const double2* __restrict__ part, int nv, int tiles,
                                                            int nks, long long items, double2* __restrict__ phi) {
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
  phi[gid] = double2{sr, si};
}

}  // namespace lc

// Partial buffer size for lc_launch_phi, in bytes.
extern "C" long long lc_phi_partial_bytes(int nt, int nv, int np, int nks) {
  const long long tiles = (nv + lc::PH_TM - 1) / lc::PH_TM;
  return (long long)np * nt * tiles * tiles * nks * lc::PH_TM * lc::PH_TM * 16;
}

extern "C" hipError_t lc_launch_phi(const void* V, const void* ph, int nt, int nv, long long nsite, int np, int nks,
                                    void* part, void* phi, hipStream_t s) {
  const int tiles = (nv + lc::PH_TM - 1) / lc::PH_TM;
  const long long items = (long long)np * nt;
  long long kslice = (nsite + nks - 1) / nks;
  kslice = (kslice + lc::PH_KT - 1) / lc::PH_KT * lc::PH_KT;
  const long long nwg = items * tiles * tiles * nks;
  hipLaunchKernelGGL(lc::lc_phi_partial_kernel, dim3((unsigned)nwg), dim3(256), 0, s, (const double2*)V,
                     (const double2*)ph, nt, nv, nsite, tiles, nks, kslice, (double2*)part);
  const long long n = items * nv * nv;
  hipLaunchKernelGGL(lc::lc_phi_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     (const double2*)part, nv, tiles, nks, items, (double2*)phi);
  return hipGetLastError();
}
