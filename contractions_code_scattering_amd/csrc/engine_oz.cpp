// Exact form of the batched engines: the rho + pi-pi chains of engine.cpp and the sigma / sigma-pi
// / rho-pi chains of engine_sigma.cpp, with every product on the int8 matrix cores through digit
// storage and exact integer accumulation (gemm_oz.hip), FP64 outputs, closing factors and traces.
#include <climits>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "engine_oz.h"
#include "runtime.h"
#include "spin.h"

extern "C" hipError_t lc_launch_engine_epilogue(const void* raw, void* corr, void* avg, int np,
                                                int nsrc, int nt, int src0, int src_step,
                                                const double* s_tau, const double* const* s_p,
                                                long long ps, hipStream_t s);

using namespace lc;

namespace {

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return LC_OK;
  } catch (const Error& e) {
    last_error() = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return LC_ERR_ARG;
  }
}

// Device allocation freed at scope exit (test hook only: the engines use the context's buffers).
struct Scoped {
  void* p = nullptr;
  explicit Scoped(size_t n) {
    if (hipMalloc(&p, n ? n : 256) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      fail(LC_ERR_NOMEM, "hipMalloc of %zu bytes failed", n);
    }
  }
  ~Scoped() {
    if (p) (void)hipFree(p);
  }
  char* c() const { return static_cast<char*>(p); }
};

// ------------------------------------------------------------------------------------------
// Buffers and views of the exact form. Digit matrices are counted in complex digit elements
// (8 bytes each, P^2 per block); exponents in ints.
struct OzShape {
  int nt, nv, P, src0, step, nsrc;
  long long m, b;      // FP64 elemental / D-space matrix, complex elements (nv^2, 16 nv^2)
  long long md, bd;    // digit elemental / D-space matrix, digit elements (P^2, 16 P^2)
};

// A digit matrix array: planes and exponents (exponents per D-row: 4 P ints per D-space matrix,
// or per line index: P ints per matrix); at P = 128, h receives a product's half-row exponents
// (8 P ints per matrix) until the level's fix-up makes e.
struct DArr {
  char* d;
  int* e;
  int* h = nullptr;
};

inline char* bump(char*& cur, size_t bytes) {
  char* p = cur;
  cur += (bytes + 255) & ~size_t(255);
  return p;
}

inline LcView with_ex(LcView v, const int* ex, long long e0, long long e1, long long e2, int ebs) {
  v.ex = ex;
  v.es[0] = e0; v.es[1] = e1; v.es[2] = e2;
  v.ebs = ebs;
  return v;
}

// One elemental factor P1..P4 of the chains in digit storage: row-line planes r (exponents er)
// and column-line planes t (et) of each matrix, whether the chain uses S or its adjoint, and the
// stride between momentum items in matrices.
struct OzRole {
  const char* r;
  const char* t;
  const int* er;
  const int* et;
  bool dag;
  long long stride;
};

// The stored products of one level whose half-row exponents are made whole after it (P = 128: a
// workgroup strip covers half a D-row, lc_oz_fix_kernel).
struct PendingFix {
  DArr y;
  long long nmat;
  bool cf;
};

// Digit output of a product into `a` (items laid out with strides st, in matrices): the D-row
// exponents directly (4 P <= 256), or the half-row ones, with a fix-up queued for the level's end.
inline DigitOut digit_out(const DArr& a, const long long* st, long long bd, int P, bool cf) {
  const bool direct = 4 * P <= 256;
  DigitOut d;
  d.d = a.d;
  d.e = direct ? a.e : a.h;
  d.cf = cf;
  for (int k = 0; k < 3; k++) {
    d.ds[k] = st[k] * bd;
    d.es[k] = st[k] * (direct ? 4LL : 8LL) * P;
  }
  return d;
}

// rho + pi-pi chain of one momentum chunk in the exact form. Per (tf, t0, p), with P1 = P1(t0),
// P2 = P2(t0), P3 = P3(tf), P4 = P4(tf) and tau = tau(tf, t0):
//   1  B = P4 tau, L1 = P3 tau (digits); sets only: B4 = P4^dag tau. Per sink: Qa^T = P4^T D(tf)^T;
//      per source: T1^T = P1^T D(t0)^T, T1'^T = P2^T D(t0)^T (digits, transposed so that the
//      consumers read them along their stored lines)
//   2  L = L1 P1 (digits)   -> fish1 = Tr[L tau^dag]; back-to-back also rho_0,1 = -Tr[L K' tau^dag L']
//      Y = L1 P2            -> star2 = Tr[Y tau^dag]
//      Rr = B P2 (digits)   -> fish2 = Tr[Rr tau^dag]; back-to-back also
//                              rho_2 = -conj Tr[Rr K'^dag tau^dag L'^dag]
//      X = B P1             -> star1 = Tr[X tau^dag]
//      sets: Lr = B4 P1     -> rho_i = -Tr[Lr K'_i tau^dag L'_i]
//      Q = P3 Qa (digits, per sink); R^T = T1^T P2^T, R'^T = T1'^T P1^T (FP64, per source)
//   3  Zh^T = conj(tau) Rr^T (FP64), U^T = (g5 tau^T g5) Q^T (digits)
//   4  M = tau^dag U        -> square = Tr[M R], dT = Tr[M R']
//      LW = L W             -> dTHor = Tr[LW Zh]
// These are the traces of the reference's pi-pi cores (lib/calculate_pipi_*_core.h,
// lib/contract_kernels_pipi.cu) reassociated as in the D-space rows (capi.cpp): with P (x) 1
// commuting with spin matrices, Tr[S1 B] = Tr[X tau^dag] with S1 = P1 tau^dag, and so on; the
// closing factor of the short diagrams is the FP64 perambulator itself.
void run_chunk_oz(const OzShape& E, bool sets, const char* tau_cf, const char* tau_r, const char* tau_t,
                  const int* tex_r, const int* tex_t, const OzRole* R, int p0, int pc, char* scratch,
                  void* raw) {
  const int nt = E.nt, ns = E.nsrc, P = E.P;
  const long long bd = E.bd, md = E.md;
  const long long npair = (long long)nt * ns;
  const long long P4i = 4LL * P;  // exponents per D-space matrix
  const bool direct = 4 * P <= 256;  // digit outputs written by the GEMMs themselves
  char* cur = scratch;
  Program pg(E.nv, Program::OZ);
  auto darr = [&](long long nmat) {
    DArr a;
    a.d = bump(cur, (size_t)nmat * bd * 8);
    a.e = reinterpret_cast<int*>(bump(cur, (size_t)nmat * P4i * 4));
    if (!direct) a.h = reinterpret_cast<int*>(bump(cur, (size_t)nmat * P4i * 8));
    return a;
  };
  const long long pairs = (long long)pc * npair;
  // B, L1 (and B4) are stored only when the GEMMs cannot chain them (P = 128): at P <= 64 each is
  // formed on the chip as the A operand of its level-2 products
  const DArr none{nullptr, nullptr};
  DArr Bd = direct ? none : darr(pairs), L1d = direct ? none : darr(pairs);
  DArr Ld = darr(pairs), Rrd = darr(pairs), UTd = darr(pairs);
  DArr B4d = sets && !direct ? darr(pairs) : none;
  DArr Qat = darr((long long)pc * nt), Qd = darr((long long)pc * nt);
  DArr T1t = darr((long long)pc * ns), T1pt = darr((long long)pc * ns);
  DArr ZhTd = darr(pairs), RTd = darr((long long)pc * ns), RpTd = darr((long long)pc * ns);
  std::vector<PendingFix> pending;

  // batch strides of the item (axis 0 sink tf, axis 1 source k, axis 2 momentum p) for the pair,
  // per-sink and per-source buffers, in matrices
  const long long PR[3] = {1, nt, npair}, SK[3] = {1, 0, nt}, SR[3] = {0, 1, ns};
  auto dview = [&](const DArr& a, const long long* st, int flags) {
    LcView v = make_view(a.d, st[0] * bd, st[1] * bd, st[2] * bd, LC_DENSE, flags);
    return with_ex(v, a.e, st[0] * P4i, st[1] * P4i, st[2] * P4i, P);
  };
  pg.set_last_axis_fastest(true);  // momenta of one (tf, t0) pair next to each other
  // digit output of a product into `a` (items laid out with strides st); with twin_B, the twin
  // product A twin_B traced by twin_tr, formed by the same kernel from A on the chip
  auto dgemm = [&](const LcView& A, const LcView& Bv, const DArr& a, const long long* st, int n0, int n1, int n2,
                   const std::vector<FusedTrace>& tr = {}, bool cf = false, const LcView* twin_B = nullptr,
                   const std::vector<FusedTrace>* twin_tr = nullptr) {
    const DigitOut d = digit_out(a, st, bd, P, cf);
    pg.gemm_oz(A, Bv, nullptr, 0, 0, 0, &d, n0, n1, n2, 1.0, 0.0, tr, twin_B, twin_tr);
    if (!direct) pending.push_back(PendingFix{a, (long long)n0 * n1 * n2, cf});
  };
  // close a level: a barrier, then (P = 128) the fix-ups of its stored products
  auto close_level = [&] {
    pg.barrier();
    for (const PendingFix& q : pending) pg.fix_digits(q.y.d, q.y.h, q.y.e, q.nmat, q.cf);
    pending.clear();
  };

  // perambulator views: tau(tf, t0_k) from the row-line (A role) or column-line (B role) planes
  auto tauv = [&](bool tplanes, int flags) {
    LcView v = make_view((tplanes ? tau_t : tau_r) + (size_t)E.src0 * bd * 8, nt * bd, E.step * bd, 0, LC_DENSE, flags);
    return with_ex(v, (tplanes ? tex_t : tex_r) + (long long)E.src0 * P, (long long)nt * P, (long long)E.step * P, 0, 0);
  };
  const LcView tau_B = tauv(true, LC_TV);
  const LcView tau_dag_A = tauv(true, LC_CONJ | LC_TS);
  const LcView tau_cj_A = tauv(false, LC_CONJ);
  // doubleTriangleHor's W as a B operand: the reference's element-wise conj(tau_ba) from the
  // column-line planes, or tau^dag (lc_set_dthor_dagger)
  const LcView W_B = ctx().dthor_dagger ? tauv(false, LC_DAG) : tauv(true, LC_CONJ | LC_TS | LC_TV);
  // D(t)^T as a B operand (its columns are the rows of D): per sink and per source
  LcView Dtf_TB = make_view(tau_r, (nt + 1) * bd, 0, 0, LC_DENSE, LC_TV | LC_TS);
  Dtf_TB = with_ex(Dtf_TB, tex_r, (long long)(nt + 1) * P, 0, 0, 0);
  LcView Dt0_TB = make_view(tau_r + (size_t)E.src0 * (nt + 1) * bd * 8, 0, (long long)E.step * (nt + 1) * bd, 0,
                            LC_DENSE, LC_TV | LC_TS);
  Dt0_TB = with_ex(Dt0_TB, tex_r + (long long)E.src0 * (nt + 1) * P, 0, (long long)E.step * (nt + 1) * P, 0, 0);
  // tau^dag, the closing factor of the short diagrams, in the closing-factor form under the row
  // exponents: block (c, a) element (j, i) of tau^dag is conj tau_ac(i, j), read along the rows of
  // tau like the product tile
  LcView tau_dag_Y = make_view(tau_cf + (size_t)E.src0 * bd * 8, nt * bd, E.step * bd, 0, LC_DENSE, LC_DAG);
  tau_dag_Y = with_ex(tau_dag_Y, tex_r + (long long)E.src0 * P, (long long)nt * P, (long long)E.step * P, 0, 0);
  // role k (P1..P4) as an operand: X or X^T with X = S or S^dag (adj toggles the adjoint); A
  // reads the operand by rows, B by columns
  auto role = [&](int k, bool sink, bool brole, bool transposed, bool adj = false) {
    const OzRole& Rk = R[k];
    const bool dag = Rk.dag != adj;
    const int op = dag ? (transposed ? 3 : 2) : (transposed ? 1 : 0);
    const bool tl = (op == 1 || op == 2) != brole;  // read the column-line planes
    const int fl = ((op >= 2) ? LC_CONJ : 0) | (brole ? LC_TV : 0);
    const long long m0 = (long long)p0 * Rk.stride + (sink ? 0 : E.src0);
    const char* base = (tl ? Rk.t : Rk.r) + (size_t)m0 * md * 8;
    const int* ex = (tl ? Rk.et : Rk.er) + m0 * P;
    LcView v = sink ? make_view(base, md, 0, Rk.stride * md, LC_KRON, fl)
                    : make_view(base, 0, E.step * md, Rk.stride * md, LC_KRON, fl);
    return sink ? with_ex(v, ex, P, 0, Rk.stride * P, 0) : with_ex(v, ex, 0, (long long)E.step * P, Rk.stride * P, 0);
  };
  char* rw = static_cast<char*>(raw) + (size_t)p0 * ns * nt * 10 * 16;
  const long long O0 = 10, O1 = nt * 10LL, O2 = (long long)ns * nt * 10;
  auto slot = [&](int k, const LcView& Y, double ar = 1.0) {
    FusedTrace f = lc::fused(Y, rw + (size_t)k * 16, O0, O1, O2, ar);
    f.out_double = 1;
    return f;
  };
  const Spin g5 = gamma(5);
  auto rho_traces = [&](bool conj_form, std::vector<int> pols) {
    std::vector<FusedTrace> v;
    for (int i : pols) {
      const Spin Lp = g5 * rhoj(1, i, false);
      const Spin Kp = rhoj(1, i, true) * g5;
      if (!conj_form) {
        v.push_back(slot(i, right(left(tau_dag_Y, Kp), Lp), -1.0));
      } else {
        v.push_back(slot(i, right(left(tau_dag_Y, Kp.adjoint()), Lp.adjoint()), -1.0));
        v.back().conj = true;
      }
    }
    return v;
  };

  // level 1 (B, L1, B4: chained into level 2 at P <= 64)
  if (!direct) {
    dgemm(role(3, true, false, false), tau_B, Bd, PR, nt, ns, pc);
    dgemm(role(2, true, false, false), tau_B, L1d, PR, nt, ns, pc);
    if (sets) dgemm(role(3, true, false, false, true), tau_B, B4d, PR, nt, ns, pc);
  }
  dgemm(role(3, true, false, true), Dtf_TB, Qat, SK, nt, 1, pc);
  dgemm(role(0, false, false, true), Dt0_TB, T1t, SR, 1, ns, pc);
  dgemm(role(1, false, false, true), Dt0_TB, T1pt, SR, 1, ns, pc);
  close_level();
  // level 2
  {
    // L = L1 P1 with its twin Y = L1 P2 (star2), Rr = B P2 with its twin X = B P1 (star1): each
    // pair shares its A operand
    std::vector<FusedTrace> tl = sets ? std::vector<FusedTrace>{} : rho_traces(false, {0, 1});
    tl.push_back(slot(8, tau_dag_Y));
    const LcView P1B = role(0, false, true, false), P2B = role(1, false, true, false);
    const std::vector<FusedTrace> ty = {slot(7, tau_dag_Y)}, tx = {slot(6, tau_dag_Y)};
    std::vector<FusedTrace> tr = {slot(9, tau_dag_Y)};
    if (!sets) {
      const std::vector<FusedTrace> r2 = rho_traces(true, {2});
      tr.push_back(r2[0]);
    }
    if (direct) {
      // chained: L1 = P3 tau, B = P4 tau (and B4 = P4^dag tau) formed per strip on the chip
      const DigitOut dL = digit_out(Ld, PR, bd, P, false), dR = digit_out(Rrd, PR, bd, P, false);
      pg.gemm_oz_chain(role(2, true, false, false), tau_B, P1B, &dL, nt, ns, pc, tl, &P2B, &ty);
      pg.gemm_oz_chain(role(3, true, false, false), tau_B, P2B, &dR, nt, ns, pc, tr, &P1B, &tx);
      if (sets)
        pg.gemm_oz_chain(role(3, true, false, false, true), tau_B, role(0, false, true, false), nullptr, nt, ns, pc,
                         rho_traces(false, {0, 1, 2}));
    } else {
      dgemm(dview(L1d, PR, 0), P1B, Ld, PR, nt, ns, pc, tl, false, &P2B, &ty);
      dgemm(dview(Bd, PR, 0), P2B, Rrd, PR, nt, ns, pc, tr, false, &P1B, &tx);
      if (sets)
        pg.gemm_oz(dview(B4d, PR, 0), role(0, false, true, false), nullptr, 0, 0, 0, nullptr, nt, ns, pc, 1.0, 0.0,
                   rho_traces(false, {0, 1, 2}));
    }
    dgemm(role(2, true, false, false), dview(Qat, SK, LC_TV | LC_TS), Qd, SK, nt, 1, pc);
    dgemm(dview(T1t, SR, 0), role(1, false, true, true), RTd, SR, 1, ns, pc, {}, true);
    dgemm(dview(T1pt, SR, 0), role(0, false, true, true), RpTd, SR, 1, ns, pc, {}, true);
  }
  close_level();
  // level 3; U walks the sources before the sinks, so that Q(tf, p) of the momenta in flight is
  // shared by neighbouring workgroups
  {
    dgemm(tau_cj_A, dview(Rrd, PR, LC_TV | LC_TS), ZhTd, PR, nt, ns, pc, {}, true);
    pg.set_second_axis(1);
    dgemm(right(left(tauv(true, LC_TS), g5), g5), dview(Qd, SK, LC_TV | LC_TS), UTd, PR, nt, ns, pc);
    pg.set_second_axis(0);
  }
  close_level();
  // level 4
  pg.gemm_oz(tau_dag_A, dview(UTd, PR, LC_TV | LC_TS), nullptr, 0, 0, 0, nullptr, nt, ns, pc, 1.0, 0.0,
             {slot(3, dview(RTd, SR, LC_TV | LC_TS)), slot(4, dview(RpTd, SR, LC_TV | LC_TS))});
  pg.gemm_oz(dview(Ld, PR, 0), W_B, nullptr, 0, 0, 0, nullptr, nt, ns, pc, 1.0, 0.0,
             {slot(5, dview(ZhTd, PR, LC_TV | LC_TS))});
  pg.run(ctx().stream);
}

// One digit matrix with its exponents (and, at P = 128, its half-row exponents), with the
// alignment of each piece.
size_t oz_digit_matrix_bytes(const OzShape& E) {
  return (size_t)E.bd * 8 + (4 * E.P <= 256 ? 4 : 12) * (size_t)E.P * 4 + 768;
}

// Scratch of one momentum (pc = 1): digit pair buffers L, Rr, Zh^T, U^T (and, for P = 128, B, L1
// and for sets B4), per sink Qa^T, Q, per source T1^T, T1'^T, R^T, R'^T.
size_t oz_bytes_per_momentum(const OzShape& E, bool sets) {
  const long long npair = (long long)E.nt * E.nsrc;
  const bool direct = 4 * E.P <= 256;
  const long long npb = direct ? 4 : (sets ? 7 : 6);
  return (size_t)(npair * npb + 2LL * E.nt + 4LL * E.nsrc) * oz_digit_matrix_bytes(E) + 16 * 256;
}

// ------------------------------------------------------------------------------------------
// sigma / sigma-pi / rho-pi set. Elementals el [np][nt][11][nv][nv] by type (engine_sigma.cpp):
// 0, 1 X_l (sigma), 2, 3 X_ip (triangle), 4..8 P12 (rho-pi), 9, 10 P1, P2; raw slots as there.
constexpr int SG_NEL = 11;
enum { SG_SIG = 0, SG_TRI = 2, SG_P12 = 4, SG_P1 = 9, SG_P2 = 10 };

// Per (tf, t0, p), with tau = tau(tf, t0), X_l = X_l(tf) at the sink, X_r = X_r(t0), P1, P2 at the
// source and D(t) = tau(t, t), the traces of lib/contract_kernels_sigma.cu (one_fish),
// lib/contract_kernels_sigmapi.cu (triangle) and lib/contract_kernels_rhopi.cu in the row forms of
// engine_sigma.cpp:
//   1  B_l = X_l tau (closing-factor form, l = 0, 1); per source T1^T = P1^T D(t0)^T (digits)
//   2  A_r = tau X_r (r = 0, 1: twins sharing tau) -> sigma(l, r) = Tr[A_r (g5 B_l^dag g5)];
//      per source R = P2 T1 (digits)
//   3  Y^T = conj(tau) R^T g5, the transpose of Y = (g5 R) tau^dag (digits; a run over the momenta
//      of a (sink, source) pair: A is the perambulator)
//   4  V = tau Y (a run likewise) -> triangle(ip) = Tr[V (X_ip(tf) (x) g5)],
//      rhopi(iop) = Tr[V (P12_ip(tf) (x) G_iop)], G_iop = g5 rho_{iop+1}(idir)
// Level 4 is the reference's Tr[(g5 X tau)(g5 R tau^dag)] and Tr[(G P12 tau)(g5 R tau^dag)] with
// X (x) 1 commuting with the spin matrices: one product per item, six traces against the elementals
// (x) spin in the closing-factor form (transposed planes), instead of one product per elemental.
struct OzSigmaIn {
  const char *tau_r, *tau_t, *el_r, *el_t, *el_cft;
  const int *tex_r, *tex_t, *eex_r, *eex_t;
};

void run_chunk_oz_sigma(const OzShape& E, int idir, const OzSigmaIn& in, int p0, int pc, char* scratch,
                        void* raw) {
  const int nt = E.nt, ns = E.nsrc, P = E.P;
  const long long bd = E.bd, md = E.md;
  const long long npair = (long long)nt * ns;
  const long long P4i = 4LL * P;
  const bool direct = 4 * P <= 256;
  char* cur = scratch;
  Program pg(E.nv, Program::OZ);
  auto darr = [&](long long nmat) {
    DArr a;
    a.d = bump(cur, (size_t)nmat * bd * 8);
    a.e = reinterpret_cast<int*>(bump(cur, (size_t)nmat * P4i * 4));
    if (!direct) a.h = reinterpret_cast<int*>(bump(cur, (size_t)nmat * P4i * 8));
    return a;
  };
  const long long pairs = (long long)pc * npair;
  const DArr Bl[2] = {darr(pairs), darr(pairs)};
  const DArr YTd = darr(pairs);
  const DArr T1t = darr((long long)pc * ns), Rd = darr((long long)pc * ns);
  std::vector<PendingFix> pending;

  const long long PR[3] = {1, nt, npair}, SR[3] = {0, 1, ns};
  auto dview = [&](const DArr& a, const long long* st, int flags) {
    LcView v = make_view(a.d, st[0] * bd, st[1] * bd, st[2] * bd, LC_DENSE, flags);
    return with_ex(v, a.e, st[0] * P4i, st[1] * P4i, st[2] * P4i, P);
  };
  pg.set_last_axis_fastest(true);  // momenta of one (tf, t0) pair next to each other
  auto dgemm = [&](const LcView& A, const LcView& Bv, const DArr& a, const long long* st, int n0, int n1, int n2,
                   bool cf) {
    const DigitOut d = digit_out(a, st, bd, P, cf);
    pg.gemm_oz(A, Bv, nullptr, 0, 0, 0, &d, n0, n1, n2);
    if (!direct) pending.push_back(PendingFix{a, (long long)n0 * n1 * n2, cf});
  };
  // trace-only pair products, with a twin on the same A
  auto tgemm = [&](const LcView& A, const LcView& Bv, const std::vector<FusedTrace>& tr, const LcView* twin_B = nullptr,
                   const std::vector<FusedTrace>* twin_tr = nullptr) {
    pg.gemm_oz(A, Bv, nullptr, 0, 0, 0, nullptr, nt, ns, pc, 1.0, 0.0, tr, twin_B, twin_tr);
  };
  auto close_level = [&] {
    pg.barrier();
    for (const PendingFix& q : pending) pg.fix_digits(q.y.d, q.y.h, q.y.e, q.nmat, q.cf);
    pending.clear();
  };

  // perambulator tau(tf, t0_k): row-line planes (A role) or column-line planes (B role)
  auto tauv = [&](bool tplanes, int flags) {
    LcView v = make_view((tplanes ? in.tau_t : in.tau_r) + (size_t)E.src0 * bd * 8, nt * bd, E.step * bd, 0,
                         LC_DENSE, flags);
    return with_ex(v, (tplanes ? in.tex_t : in.tex_r) + (long long)E.src0 * P, (long long)nt * P,
                   (long long)E.step * P, 0, 0);
  };
  const LcView tau_B = tauv(true, LC_TV);
  LcView Dt0_TB = make_view(in.tau_r + (size_t)E.src0 * (nt + 1) * bd * 8, 0, (long long)E.step * (nt + 1) * bd, 0,
                            LC_DENSE, LC_TV | LC_TS);
  Dt0_TB = with_ex(Dt0_TB, in.tex_r + (long long)E.src0 * (nt + 1) * P, 0, (long long)E.step * (nt + 1) * P, 0, 0);
  // elemental type j as an operand (A reads rows, B columns), X or X^T, at the sink (per tf) or
  // the source (per k); items of momentum p0 + axis-2 index
  const long long eP = (long long)nt * SG_NEL;  // matrices per momentum
  auto elv = [&](int j, bool sink, bool brole, bool transposed) {
    const bool tl = transposed != brole;  // the column-line planes
    const long long m0 = (long long)p0 * eP + (sink ? 0 : (long long)E.src0 * SG_NEL) + j;
    const char* base = (tl ? in.el_t : in.el_r) + (size_t)m0 * md * 8;
    const int* ex = (tl ? in.eex_t : in.eex_r) + m0 * P;
    const int fl = brole ? LC_TV : 0;
    if (sink) return with_ex(make_view(base, SG_NEL * md, 0, eP * md, LC_KRON, fl), ex, SG_NEL * P, 0, eP * P, 0);
    const long long ks = (long long)E.step * SG_NEL;
    return with_ex(make_view(base, 0, ks * md, eP * md, LC_KRON, fl), ex, 0, ks * P, eP * P, 0);
  };
  char* rw = static_cast<char*>(raw) + (size_t)p0 * ns * nt * 10 * 16;
  const long long O0 = 10, O1 = nt * 10LL, O2 = (long long)ns * nt * 10;
  auto slot = [&](int k, const LcView& Y) {
    FusedTrace f = lc::fused(Y, rw + (size_t)k * 16, O0, O1, O2);
    f.out_double = 1;
    return f;
  };
  const Spin g5 = gamma(5);

  // level 1
  for (int l = 0; l < 2; l++) dgemm(elv(SG_SIG + l, true, false, false), tau_B, Bl[l], PR, nt, ns, pc, true);
  dgemm(elv(SG_P1, false, false, true), Dt0_TB, T1t, SR, 1, ns, pc, false);
  close_level();
  // level 2
  {
    std::vector<FusedTrace> ts[2];
    for (int r = 0; r < 2; r++)
      for (int l = 0; l < 2; l++) ts[r].push_back(slot(l * 2 + r, right(left(dview(Bl[l], PR, LC_DAG), g5), g5)));
    const LcView X1 = elv(SG_SIG + 1, false, true, false);
    tgemm(tauv(false, 0), elv(SG_SIG + 0, false, true, false), ts[0], &X1, &ts[1]);
    dgemm(elv(SG_P2, false, false, false), dview(T1t, SR, LC_TV | LC_TS), Rd, SR, 1, ns, pc, false);
  }
  close_level();
  // level 3
  dgemm(tauv(false, LC_CONJ), right(dview(Rd, SR, LC_TV | LC_TS), g5), YTd, PR, nt, ns, pc, false);
  close_level();
  // level 4: the elementals X(tf) as closing factors X (x) S, read through the closing-factor form
  // of X^T (its rows are the columns of X, under the column exponents)
  auto elcf = [&](int j) {
    const long long m0 = (long long)p0 * eP + j;
    const LcView v = make_view(in.el_cft + (size_t)m0 * md * 8, SG_NEL * md, 0, eP * md, LC_KRON, LC_TV);
    return with_ex(v, in.eex_t + m0 * P, SG_NEL * P, 0, eP * P, 0);
  };
  std::vector<FusedTrace> t6;
  for (int ip = 0; ip < 2; ip++) t6.push_back(slot(4 + ip, kron_right(elcf(SG_TRI + ip), g5)));
  for (int iop = 0; iop < 4; iop++)
    t6.push_back(slot(6 + iop, kron_right(elcf(SG_P12 + rhopi_ip(idir, iop)), g5 * rhoj(iop + 1, idir, false))));
  tgemm(tauv(false, 0), dview(YTd, PR, LC_TV | LC_TS), t6);
  pg.run(ctx().stream);
}

// Scratch of one momentum set: pair buffers B_0, B_1, Y^T, per source T1^T, R.
size_t oz_sigma_bytes_per_momentum_(const OzShape& E) {
  const long long npair = (long long)E.nt * E.nsrc;
  return (size_t)(3 * npair + 2LL * E.nsrc) * oz_digit_matrix_bytes(E) + 8 * 256;
}

OzShape oz_shape(int nt, int nv, int src0, int src_step, int nsrc) {
  const int P = oz_pad(nv);
  return OzShape{nt, nv, P, src0, src_step, nsrc, (long long)nv * nv, 16LL * nv * nv, (long long)P * P,
                 16LL * P * P};
}

// Inputs of the sigma chain in digits, ahead of the scratch area: tau (row-line and column-line
// planes, exponents per line index) and the elementals (likewise, and the closing-factor form of
// their transposes).
struct OzSigmaHead {
  size_t tau_d, tau_e, el_d, el_e;
  size_t total() const { return 2 * (tau_d + tau_e + el_d + el_e) + el_d; }
};
OzSigmaHead oz_sigma_head(const OzShape& E, int np) {
  const long long nel = (long long)np * E.nt * SG_NEL;
  OzSigmaHead h;
  h.tau_d = ((size_t)E.nt * E.nt * E.bd * 8 + 255) & ~size_t(255);
  h.tau_e = ((size_t)E.nt * E.nt * E.P * 4 + 255) & ~size_t(255);
  h.el_d = ((size_t)nel * E.md * 8 + 255) & ~size_t(255);
  h.el_e = ((size_t)nel * E.P * 4 + 255) & ~size_t(255);
  return h;
}

}  // namespace

namespace lc {

void oz_engine_rho_pipi(bool sets, const void* d_tau, const void* d_phi, int nt, int nv, int np, int src0,
                        int src_step, int nsrc, long long max_scratch_bytes, void* d_raw, void* d_corr, void* d_avg) {
  if (nv > 128) fail(LC_ERR_ARG, "the exact mixed form supports nv <= 128 (got %d)", nv);
  const int P = oz_pad(nv);
  OzShape E{nt, nv, P, src0, src_step, nsrc, (long long)nv * nv, 16LL * nv * nv, (long long)P * P, 16LL * P * P};
  const int nel = sets ? 4 : 1;
  const long long nphi = (long long)np * nel * nt;
  // inputs in digits, row-line and column-line planes, exponents per line index
  const size_t tau_d = ((size_t)nt * nt * E.bd * 8 + 255) & ~size_t(255);
  const size_t tau_e = ((size_t)nt * nt * P * 4 + 255) & ~size_t(255);
  const size_t phi_d = ((size_t)nphi * E.md * 8 + 255) & ~size_t(255);
  const size_t phi_e = ((size_t)nphi * P * 4 + 255) & ~size_t(255);
  const size_t head = 3 * tau_d + 2 * (tau_e + phi_d + phi_e);  // tau: row, column and closing-factor form
  const size_t per_p = oz_bytes_per_momentum(E, sets);
  if (max_scratch_bytes <= 0) max_scratch_bytes = default_scratch_budget();
  const long long avail = max_scratch_bytes - (long long)head;
  int pc = avail > 0 ? (int)(avail / (long long)per_p) : 0;
  if (pc < 1) fail(LC_ERR_NOMEM, "scratch budget %lld B below one momentum (%zu B)", max_scratch_bytes, per_p + head);
  if (pc > np) pc = np;
  ctx().scratch.ensure(head + per_p * pc + 4096);
  char* base = static_cast<char*>(ctx().scratch.p);
  char* tau_r = base;
  char* tau_t = tau_r + tau_d;
  int* tex_r = reinterpret_cast<int*>(tau_t + tau_d);
  int* tex_t = reinterpret_cast<int*>(tau_t + tau_d + tau_e);
  char* phi_r = tau_t + tau_d + 2 * tau_e;
  char* phi_t = phi_r + phi_d;
  int* pex_r = reinterpret_cast<int*>(phi_t + phi_d);
  int* pex_t = reinterpret_cast<int*>(phi_t + phi_d + phi_e);
  char* tau_cf = phi_t + phi_d + 2 * phi_e;
  char* scratch = base + head;
  auto call = [&] {
    const long long b = E.b, m = E.m;
    launch([=](hipStream_t st) {
             return lc_launch_oz_convert(d_tau, b, (long long)nt * nt, nv, P, 16, tex_r, tex_t, tau_r, tau_t, tau_cf, nullptr,
                                         st);
           },
           "tau to digits");
    launch([=](hipStream_t st) {
             return lc_launch_oz_convert(d_phi, m, nphi, nv, P, 1, pex_r, pex_t, phi_r, phi_t, nullptr, nullptr, st);
           },
           "elementals to digits");
    OzRole R[4];
    for (int k = 0; k < 4; k++) {
      const long long m0 = sets ? (long long)k * nt : 0;
      R[k] = OzRole{phi_r + (size_t)m0 * E.md * 8, phi_t + (size_t)m0 * E.md * 8, pex_r + m0 * P, pex_t + m0 * P,
                    !sets && (k == 1 || k == 2), (long long)nel * nt};
    }
    for (int p0 = 0; p0 < np; p0 += pc) {
      const int n = (np - p0) < pc ? (np - p0) : pc;
      run_chunk_oz(E, sets, tau_cf, tau_r, tau_t, tex_r, tex_t, R, p0, n, scratch, d_raw);
    }
    const double* sp[4] = {nullptr, nullptr, nullptr, nullptr};
    const long long ps = (long long)nel * nt;
    launch([=](hipStream_t st) {
             return lc_launch_engine_epilogue(d_raw, d_corr, d_avg, np, nsrc, nt, src0, src_step, nullptr, sp, ps, st);
           },
           "engine epilogue");
    hip_check(hipStreamSynchronize(ctx().stream), "hipStreamSynchronize");
  };
  char gkey[320];
  std::snprintf(gkey, sizeof gkey, "oz %s nt=%d nv=%d np=%d src0=%d step=%d nsrc=%d chunk=%d tau=%p phi=%p raw=%p "
                "corr=%p avg=%p", sets ? "pipi_sets" : "rho_pipi", nt, nv, np, src0, src_step, nsrc, pc, d_tau, d_phi,
                d_raw, d_corr, d_avg);
  run_engine_call(std::string(gkey), call);
}

size_t oz_sigma_head_bytes(int nt, int nv, int np) {
  if (nv > 128) fail(LC_ERR_ARG, "the exact mixed form supports nv <= 128 (got %d)", nv);
  return oz_sigma_head(oz_shape(nt, nv, 0, 1, 1), np).total();
}

size_t oz_sigma_bytes_per_momentum(int nt, int nv, int nsrc) {
  return oz_sigma_bytes_per_momentum_(oz_shape(nt, nv, 0, 1, nsrc));
}

void oz_sigma_set(const void* d_tau, const void* d_el, int nt, int nv, int np, int idir, int src0, int src_step,
                  int nsrc, char* head, char* scratch, int pc, void* d_raw) {
  const OzShape E = oz_shape(nt, nv, src0, src_step, nsrc);
  const OzSigmaHead h = oz_sigma_head(E, np);
  const int P = E.P;
  char* tau_r = head;
  char* tau_t = tau_r + h.tau_d;
  int* tex_r = reinterpret_cast<int*>(tau_t + h.tau_d);
  int* tex_t = reinterpret_cast<int*>(tau_t + h.tau_d + h.tau_e);
  char* el_r = tau_t + h.tau_d + 2 * h.tau_e;
  char* el_t = el_r + h.el_d;
  int* eex_r = reinterpret_cast<int*>(el_t + h.el_d);
  int* eex_t = reinterpret_cast<int*>(el_t + h.el_d + h.el_e);
  char* el_cft = el_t + h.el_d + 2 * h.el_e;
  const long long b = E.b, m = E.m, nel = (long long)np * nt * SG_NEL;
  launch([=](hipStream_t st) {
           return lc_launch_oz_convert(d_tau, b, (long long)nt * nt, nv, P, 16, tex_r, tex_t, tau_r, tau_t, nullptr, nullptr,
                                       st);
         },
         "tau to digits");
  launch([=](hipStream_t st) {
           return lc_launch_oz_convert(d_el, m, nel, nv, P, 1, eex_r, eex_t, el_r, el_t, nullptr, el_cft, st);
         },
         "elementals to digits");
  const OzSigmaIn in{tau_r, tau_t, el_r, el_t, el_cft, tex_r, tex_t, eex_r, eex_t};
  for (int p0 = 0; p0 < np; p0 += pc) {
    const int n = (np - p0) < pc ? (np - p0) : pc;
    run_chunk_oz_sigma(E, idir, in, p0, n, scratch, d_raw);
  }
}

}  // namespace lc

namespace {

// One exact-form product of host operands (the test entry points below): C in FP64 into h_c and,
// with h_out, its digit output decoded (cf false) or its closing-factor output as stored (cf true,
// with the D-row exponents into h_e).
int selftest_oz(int nv, int ka, int kb, int fa, int fb, const int* gam4, const void* h_a, const void* h_b, void* h_c,
                void* h_out, bool cf, int* h_e) {
  return guarded([&] {
    require_init();
    if (nv < 1 || nv > 128) fail(LC_ERR_ARG, "nv=%d out of range", nv);
    if (ka && kb) fail(LC_ERR_ARG, "elemental x elemental product");
    if (!h_a || !h_b || !h_c || !gam4 || (cf && (!h_out || !h_e))) fail(LC_ERR_ARG, "null pointer");
    if ((fa | fb) & ~(LC_CONJ | LC_TV | LC_TS)) fail(LC_ERR_ARG, "bad view flags");
    const int P = oz_pad(nv);
    const long long nv2 = (long long)nv * nv, P2 = (long long)P * P;
    const int nba = ka ? 1 : 16, nbb = kb ? 1 : 16;
    hipStream_t s = ctx().stream;
    Scoped xa(nba * nv2 * 16), xb(nbb * nv2 * 16);
    Scoped ya(2 * nba * 8 * P2), yb(2 * nbb * 8 * P2), ea(2 * P * 4), eb(2 * P * 4);
    Scoped c(16 * nv2 * 16), cd(16 * 8 * P2), ce(12 * P * 4), dec(16 * nv2 * 16);
    hip_check(hipMemcpyAsync(xa.p, h_a, nba * nv2 * 16, hipMemcpyHostToDevice, s), "copy a");
    hip_check(hipMemcpyAsync(xb.p, h_b, nbb * nv2 * 16, hipMemcpyHostToDevice, s), "copy b");
    int* exa = static_cast<int*>(ea.p);
    int* exb = static_cast<int*>(eb.p);
    hip_check(lc_launch_oz_convert(xa.p, 0, 1, nv, P, nba, exa, exa + P, ya.p, ya.c() + nba * 8 * P2, nullptr, nullptr, s),
              "convert a");
    hip_check(lc_launch_oz_convert(xb.p, 0, 1, nv, P, nbb, exb, exb + P, yb.p, yb.c() + nbb * 8 * P2, nullptr, nullptr, s),
              "convert b");
    // A is read by rows: its row-line planes, or the column-line ones for an eigenvector transpose;
    // B is read by columns: the column-line planes, or the row-line ones for a transpose
    const bool ta = (fa & LC_TV) != 0, tb = (fb & LC_TV) != 0;
    LcView A = make_view(ya.c() + (ta ? nba * 8 * P2 : 0), 0, 0, 0, ka ? LC_KRON : LC_DENSE, fa & ~LC_TV);
    A.ex = exa + (ta ? P : 0);
    LcView B = make_view(yb.c() + (tb ? 0 : nbb * 8 * P2), 0, 0, 0, kb ? LC_KRON : LC_DENSE, fb | LC_TV);
    B.ex = exb + (tb ? 0 : P);
    if (!ka) {
      if (gam4[0]) A = left(A, gamma(gam4[0]));
      if (gam4[1]) A = right(A, gamma(gam4[1]));
    }
    if (!kb) {
      if (gam4[2]) B = left(B, gamma(gam4[2]));
      if (gam4[3]) B = right(B, gamma(gam4[3]));
    }
    // the digit or closing-factor output (P = 128: with half-row exponents, then the fix-up)
    const bool out = h_out != nullptr;
    DArr da;
    da.d = cd.c();
    da.e = static_cast<int*>(ce.p);
    da.h = static_cast<int*>(ce.p) + 4 * P;
    const long long one[3] = {0, 0, 0};
    const DigitOut d = digit_out(da, one, 16 * P2, P, cf);
    Program pg(nv, Program::OZ);
    pg.gemm_oz(A, B, c.p, 0, 0, 0, out ? &d : nullptr, 1, 1, 1);
    if (out && 4 * P > 256) {
      pg.barrier();
      pg.fix_digits(da.d, da.h, da.e, 1, cf);
    }
    pg.run(s);
    hip_check(hipMemcpyAsync(h_c, c.p, 16 * nv2 * 16, hipMemcpyDeviceToHost, s), "copy c");
    if (out && cf) {
      hip_check(hipMemcpyAsync(h_out, cd.p, 16 * 8 * P2, hipMemcpyDeviceToHost, s), "copy cf");
      hip_check(hipMemcpyAsync(h_e, da.e, 4 * P * 4, hipMemcpyDeviceToHost, s), "copy exponents");
    } else if (out) {
      hip_check(lc_launch_oz_decode(cd.p, da.e, nv, P, 16, 1, 1, dec.p, s), "decode");
      hip_check(hipMemcpyAsync(h_out, dec.p, 16 * nv2 * 16, hipMemcpyDeviceToHost, s), "copy cd");
    }
    hip_check(hipStreamSynchronize(s), "sync");
  });
}

}  // namespace

extern "C" {

int lc_selftest_oz_gemm(int nv, int ka, int kb, int fa, int fb, const int* gam4, const void* h_a, const void* h_b,
                        void* h_c, void* h_cd) {
  return selftest_oz(nv, ka, kb, fa, fb, gam4, h_a, h_b, h_c, h_cd, false, nullptr);
}

int lc_selftest_oz_gemm_cf(int nv, int ka, int kb, int fa, int fb, const int* gam4, const void* h_a, const void* h_b,
                           void* h_c, void* h_cf, int* h_e) {
  return selftest_oz(nv, ka, kb, fa, fb, gam4, h_a, h_b, h_c, h_cf, true, h_e);
}

}  // extern "C"
