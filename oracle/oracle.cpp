// Oracle: CPU restatement of the reference's contraction cores, in FP64.
//
// TEST INFRASTRUCTURE ONLY. Loaded by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker; the product library never links or calls it.
//
// Each orc_* function restates one reference host launcher together with its kernel body,
// keeping the reference's per-table-row structure (no sharing of sub-products between rows),
// the operand order of every product, the conj placement of the closing trace, the host-side
// spin-row sums and the diagram weights:
//
//   orc_rho_row              lib/contract_kernels_rho.cu:69-124      + lib/calculate_rho_core.h
//   orc_rhopi_row            lib/contract_kernels_rhopi.cu:64-123    + lib/calculate_rhopi_core.h
//   orc_pipi_row             lib/contract_kernels_pipi.cu:170-323 (I=1), :327-502 (I=0)
//                            + lib/calculate_pipi_{square,doubleTriangle,doubleTriangleHor,
//                              starfish}_core.h
//   orc_loop_unity           lib/contract_kernels_loop_unity.cu:45-92 + lib/calculate_loop_unity_core.h
//   orc_sigma_row            lib/contract_kernels_sigma.cu:68-132    + lib/calculate_sigma_one_fish_core.h
//   orc_sigmapi_fish_hor     lib/contract_kernels_sigmapi.cu:97-137  + lib/calculate_sigmapi_hor_fish_core.h
//   orc_sigmapi_triangle_row lib/contract_kernels_sigmapi.cu:139-188 + lib/calculate_sigmapi_triangle_core.h
//
// Layouts (complex = interleaved re,im doubles), as in lib/contract.cpp:171-219:
//   row   = tau(tf, .)      [Nt][4][4][Nv][Nv]   (P[t]_{ab} = row + (t*16 + a*4 + b)*Nv*Nv)
//   diag  = tau(t, t)       [Nt][4][4][Nv][Nv]
//   mom5  = Mom_matrices    [Nt][5][Nv][Nv]
//   x2    = Mom_matrices_x2 [Nt][2][Nv][Nv]
//   mom   = Mom_matrix      [Nt][Nv][Nv]
// The spin tables are passed in by the caller (oracle/tables.py regenerates them).
#include <complex>
#include <cstring>
#include <vector>
#include <omp.h>

typedef std::complex<double> cd;

namespace {

// C = A * B for n x n row-major complex matrices.
void mm(const cd* A, const cd* B, cd* C, int n) {
  for (int i = 0; i < n * n; i++) C[i] = 0.0;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < n; k++) {
      const cd a = A[i * n + k];
      const cd* b = B + k * n;
      cd* c = C + i * n;
      for (int j = 0; j < n; j++) c[j] += a * b[j];
    }
}

// C = A * conj(B) elementwise-conjugated right operand (no transpose).
void mm_conjB(const cd* A, const cd* B, cd* C, int n) {
  for (int i = 0; i < n * n; i++) C[i] = 0.0;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < n; k++) {
      const cd a = A[i * n + k];
      const cd* b = B + k * n;
      cd* c = C + i * n;
      for (int j = 0; j < n; j++) c[j] += a * std::conj(b[j]);
    }
}

// sum_ij X_ij * conj(Y_ij)
cd dot_conjY(const cd* X, const cd* Y, int n) {
  cd s = 0.0;
  for (int i = 0; i < n * n; i++) s += X[i] * std::conj(Y[i]);
  return s;
}

inline const cd* blk(const double* base, long idx, int nv) {
  return reinterpret_cast<const cd*>(base) + idx * (long)nv * nv;
}

struct Scratch {
  std::vector<cd> a, b;
  explicit Scratch(int nv) : a((size_t)nv * nv), b((size_t)nv * nv) {}
};

}  // namespace

extern "C" {

// lib/calculate_rho_core.h: tmp1 = P[ti]_{s3s4} M_lp(ti); tmp2 = M_rp(tf) P[ti]_{s1s2};
// out[ti][bx] = coef * sum conj(tmp2) o tmp1. Host: corr[ti][ico][igi] = sum_is out.
void orc_rho_row_range(const double* row, const double* mom5, int tf, int Nt, int Nv,
                       const short* idx, const short* lprp, const float* coef, int r0, int r1,
                       double* corr);

void orc_rho_row(const double* row, const double* mom5, int tf, int Nt, int Nv,
                 const short* idx, const short* lprp, const float* coef, double* corr) {
  orc_rho_row_range(row, mom5, tf, Nt, Nv, idx, lprp, coef, 0, 768, corr);
}

// Same, restricted to table rows [r0, r1) (other rows contribute zero): lets the CPU baseline
// time exactly the op pairs a workload asks for.
void orc_rho_row_range(const double* row, const double* mom5, int tf, int Nt, int Nv,
                       const short* idx, const short* lprp, const float* coef, int r0, int r1,
                       double* corr) {
  const int nrow = 768;
  std::vector<cd> part((size_t)Nt * nrow, cd(0.0));
#pragma omp parallel
  {
    Scratch s(Nv);
#pragma omp for collapse(2) schedule(dynamic, 4)
    for (int ti = 0; ti < Nt; ti++)
      for (int bx = r0; bx < r1; bx++) {
        const int s1 = idx[bx * 4 + 0], s2 = idx[bx * 4 + 1], s3 = idx[bx * 4 + 2], s4 = idx[bx * 4 + 3];
        const int lp = lprp[bx * 2 + 0], rp = lprp[bx * 2 + 1];
        mm(blk(row, ti * 16 + s3 * 4 + s4, Nv), blk(mom5, ti * 5 + lp, Nv), s.a.data(), Nv);
        mm(blk(mom5, tf * 5 + rp, Nv), blk(row, ti * 16 + s1 * 4 + s2, Nv), s.b.data(), Nv);
        cd v = dot_conjY(s.a.data(), s.b.data(), Nv);  // sum conj(tmp2)*tmp1
        part[(size_t)ti * nrow + bx] = cd(coef[bx * 2], coef[bx * 2 + 1]) * v;
      }
  }
  cd* out = reinterpret_cast<cd*>(corr);
  for (int ti = 0; ti < Nt; ti++)
    for (int ico = 0; ico < 16; ico++)
      for (int igi = 0; igi < 3; igi++) {
        cd acc = 0.0;
        for (int is = 0; is < 16; is++) acc += part[(size_t)ti * nrow + (ico * 3 + igi) * 16 + is];
        out[(ti * 16 + ico) * 3 + igi] = acc;
      }
}

// lib/calculate_rhopi_core.h (table row idir*256+bx):
// X = M12_ip(tf) P[ti]_{s5s6} M2(ti) D[ti]_{s1s2} M1(ti); out = coef * sum X o conj(P[ti]_{s3s2}).
// Host: corr[ti][ico] = sum_{is<64} out[ti][ico*64+is].
void orc_rhopi_row(const double* row, const double* diag, const double* p1, const double* p2,
                   const double* p12x5, int tf, int Nt, int Nv, int idir, const short* idx,
                   const short* ip, const float* coef, double* corr) {
  const int nrow = 256;
  std::vector<cd> part((size_t)Nt * nrow);
#pragma omp parallel
  {
    Scratch s(Nv);
#pragma omp for collapse(2) schedule(dynamic, 4)
    for (int ti = 0; ti < Nt; ti++)
      for (int bx = 0; bx < nrow; bx++) {
        const int r = idir * nrow + bx;
        const int s1 = idx[r * 6 + 0], s2 = idx[r * 6 + 1], s3 = idx[r * 6 + 2];
        const int s5 = idx[r * 6 + 4], s6 = idx[r * 6 + 5];
        mm(blk(p12x5, tf * 5 + ip[r], Nv), blk(row, ti * 16 + s5 * 4 + s6, Nv), s.a.data(), Nv);
        mm(s.a.data(), blk(p2, ti, Nv), s.b.data(), Nv);
        mm(s.b.data(), blk(diag, ti * 16 + s1 * 4 + s2, Nv), s.a.data(), Nv);
        mm(s.a.data(), blk(p1, ti, Nv), s.b.data(), Nv);
        cd v = dot_conjY(s.b.data(), blk(row, ti * 16 + s3 * 4 + s2, Nv), Nv);
        part[(size_t)ti * nrow + bx] = cd(coef[r * 2], coef[r * 2 + 1]) * v;
      }
  }
  cd* out = reinterpret_cast<cd*>(corr);
  for (int ti = 0; ti < Nt; ti++)
    for (int ico = 0; ico < 4; ico++) {
      cd acc = 0.0;
      for (int is = 0; is < 64; is++) acc += part[(size_t)ti * nrow + ico * 64 + is];
      out[ti * 4 + ico] = acc;
    }
}

// ---------------------------------------------------------------------------------------------
// pi-pi. One sink row: square / doubleTriangle / doubleTriangleHor per (ti, table row),
// star and fish as products of two spin-summed traces, diagram weights
// lib/contract_kernels_pipi.cu:288-303 (I=1) and :463-478 (I=0).
// iso: 1 -> I=1, 0 -> I=0. corr layout [Nt][5] complex.
void orc_pipi_row(int iso, const double* row, const double* diag, const double* m1,
                  const double* m2, const double* m3, const double* m4, int tf, int Nt, int Nv,
                  const short* sq_idx, const float* sq_coef, const short* dt_idx,
                  const float* dt_coef, const short* hor_idx, const float* hor_coef,
                  double* corr) {
  const int nrow = 256;
  const int ndiag = iso == 0 ? 3 : 2;
  std::vector<cd> part((size_t)3 * Nt * nrow, cd(0.0));
  std::vector<cd> star((size_t)2 * Nt * 16), fish((size_t)2 * Nt * 16);
#pragma omp parallel
  {
    Scratch s(Nv);
#pragma omp for collapse(3) schedule(dynamic, 2)
    for (int d = 0; d < ndiag; d++)
      for (int ti = 0; ti < Nt; ti++)
        for (int bx = 0; bx < nrow; bx++) {
          cd v;
          if (d == 0) {  // lib/calculate_pipi_square_core.h
            const short* q = sq_idx + bx * 6;
            const int s1 = q[0], s2 = q[1], s3 = q[2], s4 = q[3], s5 = q[4], s6 = q[5];
            mm(blk(m3, tf, Nv), blk(diag, tf * 16 + s3 * 4 + s4, Nv), s.a.data(), Nv);
            mm(s.a.data(), blk(m4, tf, Nv), s.b.data(), Nv);
            mm(s.b.data(), blk(row, ti * 16 + s5 * 4 + s6, Nv), s.a.data(), Nv);
            mm(s.a.data(), blk(m2, ti, Nv), s.b.data(), Nv);
            mm(s.b.data(), blk(diag, ti * 16 + s1 * 4 + s2, Nv), s.a.data(), Nv);
            mm(s.a.data(), blk(m1, ti, Nv), s.b.data(), Nv);
            v = cd(sq_coef[bx * 2], sq_coef[bx * 2 + 1]) *
                dot_conjY(s.b.data(), blk(row, ti * 16 + s3 * 4 + s2, Nv), Nv);
          } else if (d == 1) {  // lib/calculate_pipi_doubleTriangle_core.h
            const short* q = dt_idx + bx * 6;
            const int s1 = q[0], s2 = q[1], s3 = q[2], s4 = q[3], s5 = q[4], s6 = q[5];
            mm(blk(m3, tf, Nv), blk(diag, tf * 16 + s5 * 4 + s6, Nv), s.a.data(), Nv);
            mm(s.a.data(), blk(m4, tf, Nv), s.b.data(), Nv);
            mm(s.b.data(), blk(row, ti * 16 + s1 * 4 + s2, Nv), s.a.data(), Nv);
            mm(s.a.data(), blk(m1, ti, Nv), s.b.data(), Nv);
            mm(s.b.data(), blk(diag, ti * 16 + s3 * 4 + s4, Nv), s.a.data(), Nv);
            mm(s.a.data(), blk(m2, ti, Nv), s.b.data(), Nv);
            v = cd(dt_coef[bx * 2], dt_coef[bx * 2 + 1]) *
                dot_conjY(s.b.data(), blk(row, ti * 16 + s5 * 4 + s4, Nv), Nv);
          } else {  // lib/calculate_pipi_doubleTriangleHor_core.h (element-wise conj at :126)
            const short* q = hor_idx + bx * 4;
            const int s1 = q[0], s2 = q[1], s3 = q[2], s4 = q[3];
            mm(blk(m3, tf, Nv), blk(row, ti * 16 + s1 * 4 + s2, Nv), s.a.data(), Nv);
            mm(s.a.data(), blk(m1, ti, Nv), s.b.data(), Nv);
            mm_conjB(s.b.data(), blk(row, ti * 16 + s3 * 4 + s2, Nv), s.a.data(), Nv);
            mm(s.a.data(), blk(m4, tf, Nv), s.b.data(), Nv);
            mm(s.b.data(), blk(row, ti * 16 + s3 * 4 + s4, Nv), s.a.data(), Nv);
            mm(s.a.data(), blk(m2, ti, Nv), s.b.data(), Nv);
            v = cd(hor_coef[bx * 2], hor_coef[bx * 2 + 1]) *
                dot_conjY(s.b.data(), blk(row, ti * 16 + s1 * 4 + s4, Nv), Nv);
          }
          part[((size_t)d * Nt + ti) * nrow + bx] = v;
        }
    // lib/calculate_pipi_starfish_core.h: first trace (Ma at tf, P1 at ti) -> out[bx],
    // second trace (Mb at tf, P2 at ti) -> out[grid + bx]. Star passes (P3,P4) as (Mb,Ma),
    // fish swaps them (lib/contract_kernels_pipi.cu:222-223): star Ma=P4, Mb=P3; fish Ma=P3, Mb=P4.
#pragma omp for collapse(3) schedule(dynamic, 2)
    for (int w = 0; w < 2; w++)
      for (int ti = 0; ti < Nt; ti++)
        for (int is = 0; is < 16; is++) {
          const int s1 = is / 4, s2 = is % 4;
          const double* ma = w == 0 ? m4 : m3;
          const double* mb = w == 0 ? m3 : m4;
          const cd* P = blk(row, ti * 16 + s1 * 4 + s2, Nv);
          mm(blk(ma, tf, Nv), P, s.a.data(), Nv);
          mm(s.a.data(), blk(m1, ti, Nv), s.b.data(), Nv);
          cd t1 = dot_conjY(s.b.data(), P, Nv);
          mm(blk(mb, tf, Nv), P, s.a.data(), Nv);
          mm(s.a.data(), blk(m2, ti, Nv), s.b.data(), Nv);
          cd t2 = dot_conjY(s.b.data(), P, Nv);
          std::vector<cd>& o = w == 0 ? star : fish;
          o[(size_t)0 * Nt * 16 + ti * 16 + is] = t1;
          o[(size_t)1 * Nt * 16 + ti * 16 + is] = t2;
        }
  }
  cd* out = reinterpret_cast<cd*>(corr);
  for (int ti = 0; ti < Nt; ti++) {
    cd red[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; d++)
      for (int is = 0; is < nrow; is++) red[d] += part[((size_t)d * Nt + ti) * nrow + is];
    cd sr[2] = {0.0, 0.0}, fr[2] = {0.0, 0.0};
    for (int is = 0; is < 16; is++)
      for (int tr = 0; tr < 2; tr++) {
        sr[tr] += star[(size_t)tr * Nt * 16 + ti * 16 + is];
        fr[tr] += fish[(size_t)tr * Nt * 16 + ti * 16 + is];
      }
    const cd st = sr[0] * sr[1], ft = fr[0] * fr[1];
    if (iso == 1) {
      out[ti * 5 + 0] = -2.0 * red[0];
      out[ti * 5 + 1] = 2.0 * red[1];
      out[ti * 5 + 2] = -st;
      out[ti * 5 + 3] = ft;
      out[ti * 5 + 4] = 0.0;
    } else {
      out[ti * 5 + 0] = (-1. / 3.) * red[0];
      out[ti * 5 + 1] = -(1. / 3.) * red[1];
      out[ti * 5 + 2] = st;
      out[ti * 5 + 3] = ft;
      out[ti * 5 + 4] = -(5. / 3.) * red[2];
    }
  }
}

// lib/calculate_loop_unity_core.h: the full product D[t]_{s1s1} M(t), then its diagonal sum;
// host: loop[t] = sum_{s1} out[t*4+s1].
void orc_loop_unity(const double* diag, const double* mom, int Nt, int Nv, double* loop) {
  cd* out = reinterpret_cast<cd*>(loop);
#pragma omp parallel
  {
    Scratch s(Nv);
#pragma omp for
    for (int t = 0; t < Nt; t++) {
      cd acc = 0.0;
      for (int s1 = 0; s1 < 4; s1++) {
        mm(blk(diag, t * 16 + s1 * 4 + s1, Nv), blk(mom, t, Nv), s.a.data(), Nv);
        cd tr = 0.0;
        for (int i = 0; i < Nv; i++) tr += s.a[(size_t)i * Nv + i];
        acc += tr;
      }
      out[t] = acc;
    }
  }
}

// lib/calculate_sigma_one_fish_core.h: tmp1 = P[ti]_{s3s4} X_rp(ti); tmp2 = X_lp(tf) P[ti]_{s2s1};
// out = coef * sum conj(tmp2) o tmp1. Host: corr[ti][ico] = sum_{is<16} out[ti][ico*16+is].
void orc_sigma_row(const double* row, const double* x2, int tf, int Nt, int Nv, const short* idx,
                   const short* lprp, const float* coef, double* corr) {
  const int nrow = 64;
  std::vector<cd> part((size_t)Nt * nrow);
#pragma omp parallel
  {
    Scratch s(Nv);
#pragma omp for collapse(2) schedule(dynamic, 4)
    for (int ti = 0; ti < Nt; ti++)
      for (int bx = 0; bx < nrow; bx++) {
        const int s1 = idx[bx * 4 + 0], s2 = idx[bx * 4 + 1], s3 = idx[bx * 4 + 2], s4 = idx[bx * 4 + 3];
        const int lp = lprp[bx * 2 + 0], rp = lprp[bx * 2 + 1];
        mm(blk(row, ti * 16 + s3 * 4 + s4, Nv), blk(x2, ti * 2 + rp, Nv), s.a.data(), Nv);
        mm(blk(x2, tf * 2 + lp, Nv), blk(row, ti * 16 + s2 * 4 + s1, Nv), s.b.data(), Nv);
        part[(size_t)ti * nrow + bx] =
            cd(coef[bx * 2], coef[bx * 2 + 1]) * dot_conjY(s.a.data(), s.b.data(), Nv);
      }
  }
  cd* out = reinterpret_cast<cd*>(corr);
  for (int ti = 0; ti < Nt; ti++)
    for (int ico = 0; ico < 4; ico++) {
      cd acc = 0.0;
      for (int is = 0; is < 16; is++) acc += part[(size_t)ti * nrow + ico * 16 + is];
      out[ti * 4 + ico] = acc;
    }
}

// lib/calculate_sigmapi_hor_fish_core.h: X = M2(t) D[t]_{s1s2} M1(t); out = coef * sum X o conj(D[t]_{s1s2}).
// Host: corr[t] = sum_is out[t*16+is].
void orc_sigmapi_fish_hor(const double* diag, const double* p1, const double* p2, int Nt, int Nv,
                          const short* idx, const float* coef, double* corr) {
  cd* out = reinterpret_cast<cd*>(corr);
  std::vector<cd> part((size_t)Nt * 16);
#pragma omp parallel
  {
    Scratch s(Nv);
#pragma omp for collapse(2)
    for (int t = 0; t < Nt; t++)
      for (int bx = 0; bx < 16; bx++) {
        const int s1 = idx[bx * 2], s2 = idx[bx * 2 + 1];
        const cd* Dm = blk(diag, t * 16 + s1 * 4 + s2, Nv);
        mm(blk(p2, t, Nv), Dm, s.a.data(), Nv);
        mm(s.a.data(), blk(p1, t, Nv), s.b.data(), Nv);
        part[(size_t)t * 16 + bx] = cd(coef[bx * 2], coef[bx * 2 + 1]) * dot_conjY(s.b.data(), Dm, Nv);
      }
  }
  for (int t = 0; t < Nt; t++) {
    cd acc = 0.0;
    for (int is = 0; is < 16; is++) acc += part[(size_t)t * 16 + is];
    out[t] = acc;
  }
}

// lib/calculate_sigmapi_triangle_core.h:
// X = X2_ip(tf) P[ti]_{s2s3} M2(ti) D[ti]_{s4s5} M1(ti); out = coef * sum X o conj(P[ti]_{s1s5}).
// Host: corr[ti][iop] = sum_{is<64} out[ti][iop*64+is].
void orc_sigmapi_triangle_row(const double* row, const double* diag, const double* p1,
                              const double* p2, const double* x2, int tf, int Nt, int Nv,
                              const short* idx, const short* ip, const float* coef, double* corr) {
  const int nrow = 128;
  std::vector<cd> part((size_t)Nt * nrow);
#pragma omp parallel
  {
    Scratch s(Nv);
#pragma omp for collapse(2) schedule(dynamic, 4)
    for (int ti = 0; ti < Nt; ti++)
      for (int bx = 0; bx < nrow; bx++) {
        const short* q = idx + bx * 5;
        const int s1 = q[0], s2 = q[1], s3 = q[2], s4 = q[3], s5 = q[4];
        mm(blk(x2, tf * 2 + ip[bx], Nv), blk(row, ti * 16 + s2 * 4 + s3, Nv), s.a.data(), Nv);
        mm(s.a.data(), blk(p2, ti, Nv), s.b.data(), Nv);
        mm(s.b.data(), blk(diag, ti * 16 + s4 * 4 + s5, Nv), s.a.data(), Nv);
        mm(s.a.data(), blk(p1, ti, Nv), s.b.data(), Nv);
        part[(size_t)ti * nrow + bx] = cd(coef[bx * 2], coef[bx * 2 + 1]) *
                                       dot_conjY(s.b.data(), blk(row, ti * 16 + s1 * 4 + s5, Nv), Nv);
      }
  }
  cd* out = reinterpret_cast<cd*>(corr);
  for (int ti = 0; ti < Nt; ti++)
    for (int iop = 0; iop < 2; iop++) {
      cd acc = 0.0;
      for (int is = 0; is < 64; is++) acc += part[(size_t)ti * nrow + iop * 64 + is];
      out[ti * 2 + iop] = acc;
    }
}

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
