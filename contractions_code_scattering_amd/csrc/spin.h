// Spin algebra of the planner: 4x4 gamma matrices in the chiral basis used by the reference's
// tables (gamma_5 = diag(1,1,-1,-1)), the rho interpolator matrices of devTools/comb_rho.cpp:19-49,
// and the view operations that fold a spin matrix with one non-zero per row and column into
// block renames + exact phases (see lc_internal.h).
#pragma once
#include <complex>
#include <stdexcept>

#include "lc_internal.h"

namespace lc {

typedef std::complex<double> cd;

struct Spin {
  cd m[4][4];
  Spin() {
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) m[i][j] = 0.0;
  }
  static Spin id() {
    Spin s;
    for (int i = 0; i < 4; i++) s.m[i][i] = 1.0;
    return s;
  }
  Spin operator*(const Spin& o) const {
    Spin r;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++)
        for (int k = 0; k < 4; k++) r.m[i][j] += m[i][k] * o.m[k][j];
    return r;
  }
  Spin adjoint() const {
    Spin r;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) r.m[i][j] = std::conj(m[j][i]);
    return r;
  }
  Spin scaled(cd f) const {
    Spin r;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) r.m[i][j] = m[i][j] * f;
    return r;
  }
};

// gamma_mu, mu = 1..5 (5 = gamma_5).
inline Spin gamma(int mu) {
  const cd I(0.0, 1.0);
  Spin g;
  switch (mu) {
    case 1: g.m[0][3] = -I; g.m[1][2] = -I; g.m[2][1] = I; g.m[3][0] = I; break;
    case 2: g.m[0][3] = 1; g.m[1][2] = -1; g.m[2][1] = -1; g.m[3][0] = 1; break;
    case 3: g.m[0][2] = -I; g.m[1][3] = I; g.m[2][0] = I; g.m[3][1] = -I; break;
    case 4: g.m[0][2] = -1; g.m[1][3] = -1; g.m[2][0] = -1; g.m[3][1] = -1; break;
    case 5: g.m[0][0] = 1; g.m[1][1] = 1; g.m[2][2] = -1; g.m[3][3] = -1; break;
    default: throw std::invalid_argument("gamma index");
  }
  return g;
}

// Rho interpolator j = 1..4 in direction i = 0..2 (devTools/comb_rho.cpp:19-49):
// j=1,3: gamma_i (dag: -gamma_i); j=2: gamma_4 gamma_i; j=4: 1/2 gamma_4 gamma_4 (dag: -1/2 ...).
inline Spin rhoj(int j, int i, bool dag) {
  const Spin gi = gamma(i + 1);
  switch (j) {
    case 1:
    case 3: return dag ? gi.scaled(-1.0) : gi;
    case 2: return gamma(4) * gi;
    case 4: return (gamma(4) * gamma(4)).scaled(dag ? -0.5 : 0.5);
    default: throw std::invalid_argument("rho interpolator index");
  }
}

// Elemental types at source (lp, from k) and sink (rp, from l) of the rho-rho op pair
// (devTools/comb_rho.cpp:52-74).
inline void lp_rp(int l, int k, int i, int* lp, int* rp) {
  const int sel[4] = {0, 0, 1, 2 + i};
  *lp = sel[k];
  *rp = sel[l];
}

inline int rhopi_ip(int idir, int iop) {
  const int sel[4] = {0, 0, 1, 2 + idir};
  return sel[iop];
}

// Column of the single non-zero in row r (throws if the matrix is not a phased permutation).
inline int nz_col(const Spin& s, int r) {
  int c = -1;
  for (int j = 0; j < 4; j++)
    if (s.m[r][j] != 0.0) {
      if (c >= 0) throw std::invalid_argument("spin matrix is not a phased permutation");
      c = j;
    }
  if (c < 0) throw std::invalid_argument("singular spin matrix");
  return c;
}
inline int nz_row(const Spin& s, int c) {
  int r = -1;
  for (int i = 0; i < 4; i++)
    if (s.m[i][c] != 0.0) {
      if (r >= 0) throw std::invalid_argument("spin matrix is not a phased permutation");
      r = i;
    }
  if (r < 0) throw std::invalid_argument("singular spin matrix");
  return r;
}

inline LcView make_view(const void* base, long long s0, long long s1, long long s2, int kind,
                        int flags) {
  LcView v;
  v.base = base;
  v.s[0] = s0; v.s[1] = s1; v.s[2] = s2;
  v.kind = kind;
  v.flags = flags;
  v.ex = nullptr;
  v.es[0] = v.es[1] = v.es[2] = 0;
  v.ebs = 0;
  for (int a = 0; a < 4; a++) {
    v.perm_r[a] = a; v.perm_c[a] = a;
    v.ph_r[a][0] = 1; v.ph_r[a][1] = 0;
    v.ph_c[a][0] = 1; v.ph_c[a][1] = 0;
  }
  return v;
}

// Gamma * V: new block (a,b) = G[a][pi(a)] * V(pi(a), b).
inline LcView left(LcView v, const Spin& G) {
  if (v.kind != LC_DENSE) throw std::invalid_argument("spin matrix on an elemental view");
  LcView r = v;
  for (int a = 0; a < 4; a++) {
    const int p = nz_col(G, a);
    const cd ph = G.m[a][p] * cd(v.ph_r[p][0], v.ph_r[p][1]);
    r.perm_r[a] = v.perm_r[p];
    r.ph_r[a][0] = ph.real(); r.ph_r[a][1] = ph.imag();
  }
  return r;
}

// V * Gamma: new block (a,b) = V(a, sigma(b)) * G[sigma(b)][b].
inline LcView right(LcView v, const Spin& G) {
  if (v.kind != LC_DENSE) throw std::invalid_argument("spin matrix on an elemental view");
  LcView r = v;
  for (int b = 0; b < 4; b++) {
    const int q = nz_row(G, b);
    const cd ph = cd(v.ph_c[q][0], v.ph_c[q][1]) * G.m[q][b];
    r.perm_c[b] = v.perm_c[q];
    r.ph_c[b][0] = ph.real(); r.ph_c[b][1] = ph.imag();
  }
  return r;
}

// (X (x) 1) Gamma for an elemental view X: the elemental view X (x) Gamma, block (a, b) = G[a][b] X.
// Only the exact kernel's closing factors read the spin of an elemental view.
inline LcView kron_right(LcView v, const Spin& G) {
  if (v.kind != LC_KRON) throw std::invalid_argument("kron_right of a dense view");
  LcView r = v;
  for (int b = 0; b < 4; b++) {
    const int q = nz_row(G, b);
    const cd ph = cd(v.ph_c[q][0], v.ph_c[q][1]) * G.m[q][b];
    r.perm_c[b] = v.perm_c[q];
    r.ph_c[b][0] = ph.real(); r.ph_c[b][1] = ph.imag();
  }
  return r;
}

// An elemental view read as X (x) 1: identity perms and unit phases.
inline bool kron_plain(const LcView& v) {
  if (v.kind != LC_KRON) return true;
  for (int a = 0; a < 4; a++)
    if (v.perm_r[a] != a || v.perm_c[a] != a || v.ph_r[a][0] != 1 || v.ph_r[a][1] != 0 || v.ph_c[a][0] != 1 ||
        v.ph_c[a][1] != 0)
      return false;
  return true;
}

}  // namespace lc
