// Batched engine: every (sink tf, source t0, momentum p) chain of the rho + pi-pi correlator set
// over a perambulator resident in HBM, as a few grouped launches per momentum chunk.
//
// Inputs (FP64 complex):
//   tau  [nt][nt][4][4][nv][nv]   tau(t, t0), the reference's host layout (lib/contract.cpp:140)
//   phi  [np][nt][nv][nv]         Phi(p, t); Phi(-p, t) = Phi(p, t)^dag is read through a view
// Sources t0_k = src0 + k * src_step, k < nsrc.
//
// Per (p, tf, t0), with P1 = P4 = Phi(p), P2 = P3 = Phi(p)^dag (back-to-back pions, the driver
// convention of tests/compute_pipi_I0.cpp:143-146), rho from the op pair (l,k) = (0,0) of
// lib/contract_kernels_rho.cu with Gamma = gamma_i:
//   rho_i  = -Tr[ B^dag (g5 g_i) A (-g_i g5) ]     A = tau Phi(t0),  B = Phi(tf) tau
//   square = Tr[ M R ],  dT = Tr[ M R' ]            M = tau^dag U, U = Q (g5 tau g5),
//                                                   Q = Phi^dag D(tf) Phi, R = Phi^dag D(t0) Phi,
//                                                   R' = Phi D(t0) Phi^dag
//     (the same traces as Tr[U (R tau^dag)] and Tr[U (R' tau^dag)], reassociated so that the two
//      share one dense product instead of needing R tau^dag and R' tau^dag separately)
//   dTHor  = Tr[ (L W) (Rr tau^dag) ]               L = Phi^dag tau Phi, Rr = Phi tau Phi^dag
//   star   = Tr[B S1] * Tr[L1 S2], fish = Tr[L1 S1] * Tr[B S2]
//                                                   S1 = Phi tau^dag, S2 = Phi^dag tau^dag,
//                                                   L1 = Phi^dag tau
// Per-time prefixes (Q, R, R') are formed once per (t, p) and shared by every pair.
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

struct EngineShape {
  int nt, nv, np, src0, step, nsrc;
  long long m, blk;
};

// One elemental factor P1..P4 of the pi-pi chains (the m1..m4 of run_ContractPiPi,
// lib/contract_kernels_pipi.cu): the stored matrices S (FP64 array, or the row-major `r` and
// block-transposed `t` split planes of the mixed form), whether the chain uses S or its adjoint,
// and the stride between momentum items in complex elements. Back-to-back pions (the
// rho + pi-pi engine) use one array Phi for all four: P1 = P4 = Phi, P2 = P3 = Phi^dag.
struct Role {
  const char* r;
  const char* t;
  bool dag;
  long long stride;
};

inline LcView dense(const void* p, long long s0, long long s1, long long s2, int flags = 0) {
  return make_view(p, s0, s1, s2, LC_DENSE, flags);
}
inline LcView kron(const void* p, long long s0, long long s1, long long s2, int flags = 0) {
  return make_view(p, s0, s1, s2, LC_KRON, flags);
}

// Byte offsets of complex arrays: esz = 16 (complex double) or 8 (complex float).
inline const char* off(const void* p, long long elems, size_t esz) {
  return static_cast<const char*>(p) + (size_t)elems * esz;
}
inline char* off(char* p, long long elems, size_t esz) { return p + (size_t)elems * esz; }

// Number of D-space scratch matrices per (tf, t0, p) item and per (t, p) item.
const int PAIR_BUFS = 6;

// Levels (each one grouped launch; closing traces fused into the GEMM that forms their left
// factor, so A, S1, S2, M and LW are never stored):
//   1  B = Phi(tf) tau, L1 = Phi(tf)^dag tau;  Qa = D(tf) Phi, T1 = D(t0) Phi, T1' = D(t0) Phi^dag
//   2  A = tau Phi(t0)          -> rho_i = -Tr[A (K'_i B^dag L'_i)]
//      S1 = Phi(t0) tau^dag     -> star1 = Tr[S1 B], fish1 = Tr[S1 L1]
//      S2 = Phi(t0)^dag tau^dag -> star2 = Tr[S2 L1], fish2 = Tr[S2 B]
//      L = L1 Phi(t0), Rr = B Phi(t0)^dag;  Q = Phi^dag Qa, R = Phi^dag T1, R' = Phi T1'
//   3  Zh = Rr tau^dag, U = Q (g5 tau g5)
//   4  M = tau^dag U            -> square = Tr[M R], dT = Tr[M R']
//      LW = L W                 -> dTHor = Tr[LW Zh]
// Dense D x D products per (tf, t0, p): Zh, U, M, LW (4 x 64 block GEMMs).
// tau / phi / scratch hold complex double (mode F64) or complex float (mode MX); raw is always
// complex double.
void run_chunk(const EngineShape& E, Program::Mode mode, const void* tau, const Role* P, int p0,
               int pc, char* scratch, void* raw) {
  const size_t es = mode == Program::F64 ? 16 : 8;
  auto off = [es](auto p, long long elems) { return ::off(p, elems, es); };
  const int nt = E.nt, ns = E.nsrc;
  const long long m = E.m, b = E.blk;
  const long long npair = (long long)nt * ns;
  // pair buffers [pc][nsrc][nt] -> strides {b, nt*b, nsrc*nt*b}
  char* pb[PAIR_BUFS];
  for (int i = 0; i < PAIR_BUFS; i++) pb[i] = off(scratch, (long long)i * pc * npair * b);
  char* tb0 = off(scratch, (long long)PAIR_BUFS * pc * npair * b);
  char *B = pb[0], *L1 = pb[1], *L = pb[2], *Rr = pb[3], *U = pb[4], *Zh = pb[5];
  // per-sink buffers [pc][nt]: strides {b, 0, nt*b}; per-source buffers [pc][nsrc]: {0, b, ns*b}
  char* Qa = tb0;
  char* Q = off(tb0, (long long)pc * nt * b);
  char* T1 = off(tb0, (long long)2 * pc * nt * b);
  char* T1p = off(T1, (long long)pc * ns * b);
  char* R = off(T1, (long long)2 * pc * ns * b);
  char* Rp = off(T1, (long long)3 * pc * ns * b);

  const long long P0 = b, P1 = nt * b, P2 = (long long)ns * nt * b;  // pair strides
  const LcView tau_v = dense(off(tau, (long long)E.src0 * b), nt * b, E.step * b, 0);
  const LcView tau_dag = dense(off(tau, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_DAG);
  // doubleTriangleHor's W: the reference's element-wise conj(tau_ba), or tau^dag (lc_set_dthor_dagger)
  const LcView W = dense(off(tau, (long long)E.src0 * b), nt * b, E.step * b, 0,
                         ctx().dthor_dagger ? LC_DAG : (LC_CONJ | LC_TS));
  // role k as a view over the sinks (P3, P4) or the sources (P1, P2) of the chunk's items
  auto role = [&](int k, bool sink) {
    const Role& R = P[k];
    const char* base = off(R.r, (long long)p0 * R.stride + (sink ? 0 : (long long)E.src0 * m));
    const int fl = R.dag ? (LC_CONJ | LC_TV) : 0;
    return sink ? kron(base, m, 0, R.stride, fl) : kron(base, 0, E.step * m, R.stride, fl);
  };
  const LcView phi_tf = role(3, true);       // P4(tf)
  const LcView phi_tf_dag = role(2, true);   // P3(tf)
  const LcView phi_t0 = role(0, false);      // P1(t0)
  const LcView phi_t0_dag = role(1, false);  // P2(t0)
  const LcView D_tf = dense(tau, (nt + 1) * b, 0, 0);
  const LcView D_t0 = dense(off(tau, (long long)E.src0 * (nt + 1) * b), 0, E.step * (nt + 1) * b, 0);
  auto pair = [&](char* p, int flags = 0) { return dense(p, P0, P1, P2, flags); };
  // raw [np][nsrc][nt][10]
  char* rw = ::off(static_cast<char*>(raw), (long long)p0 * ns * nt * 10, 16);
  const long long O0 = 10, O1 = nt * 10LL, O2 = (long long)ns * nt * 10;
  auto slot = [&](int k, const LcView& Y, double ar = 1.0) {
    FusedTrace f = fused(Y, ::off(rw, k, 16), O0, O1, O2, ar);
    f.out_double = 1;
    return f;
  };

  Program pg(E.nv, mode);
  pg.set_last_axis_fastest(true);  // momenta of one (tf, t0) pair next to each other
  // level 1
  pg.gemm(phi_tf, tau_v, B, P0, P1, P2, nt, ns, pc);
  pg.gemm(phi_tf_dag, tau_v, L1, P0, P1, P2, nt, ns, pc);
  pg.gemm(D_tf, phi_tf, Qa, b, 0, nt * b, nt, 1, pc);
  pg.gemm(D_t0, phi_t0, T1, 0, b, ns * b, 1, ns, pc);
  pg.gemm(D_t0, phi_t0_dag, T1p, 0, b, ns * b, 1, ns, pc);
  pg.barrier();
  // level 2
  const Spin g5 = gamma(5);
  std::vector<FusedTrace> rho;
  for (int i = 0; i < 3; i++) {
    const Spin Lp = g5 * rhoj(1, i, false);
    const Spin Kp = rhoj(1, i, true) * g5;
    rho.push_back(slot(i, right(left(pair(B, LC_DAG), Kp), Lp), -1.0));
  }
  pg.gemm(tau_v, phi_t0, nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, rho);
  pg.gemm(phi_t0, tau_dag, nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, {slot(6, pair(B)), slot(8, pair(L1))});
  pg.gemm(phi_t0_dag, tau_dag, nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, {slot(7, pair(L1)), slot(9, pair(B))});
  pg.gemm(pair(L1), phi_t0, L, P0, P1, P2, nt, ns, pc);
  pg.gemm(pair(B), phi_t0_dag, Rr, P0, P1, P2, nt, ns, pc);
  pg.gemm(phi_tf_dag, dense(Qa, b, 0, nt * b), Q, b, 0, nt * b, nt, 1, pc);
  pg.gemm(phi_t0_dag, dense(T1, 0, b, ns * b), R, 0, b, ns * b, 1, ns, pc);
  pg.gemm(phi_t0, dense(T1p, 0, b, ns * b), Rp, 0, b, ns * b, 1, ns, pc);
  pg.barrier();
  // level 3
  pg.gemm(pair(Rr), tau_dag, Zh, P0, P1, P2, nt, ns, pc);
  pg.set_second_axis(1);  // sources before sinks: Q(tf, p) shared by neighbouring workgroups
  pg.gemm(dense(Q, b, 0, nt * b), right(left(tau_v, g5), g5), U, P0, P1, P2, nt, ns, pc);
  pg.set_second_axis(0);
  pg.barrier();
  // level 4
  pg.gemm(tau_dag, pair(U), nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0,
          {slot(3, dense(R, 0, b, ns * b)), slot(4, dense(Rp, 0, b, ns * b))});
  pg.gemm(pair(L), W, nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, {slot(5, pair(Zh))});
  pg.run(ctx().stream);
}

// Per-momentum-chunk scratch of the mixed form, in bytes per (pc = 1): FP32 pair buffers B, L1,
// Zh (8 B per complex), split pair buffers B, L1, L, Rr, U (12 B), split per-sink Qa, Q and
// per-source T1, T1' buffers (12 B), FP32 per-source R, R' (8 B).
size_t mx_bytes_per_momentum(const EngineShape& E) {
  const long long npair = (long long)E.nt * E.nsrc;
  return (size_t)E.blk * (size_t)(npair * (3 * 8 + 5 * 12) + (2LL * E.nt + 2LL * E.nsrc) * 12 +
                                  2LL * E.nsrc * 8);
}

// The same chain as run_chunk in the mixed form (Program::MX). tau and Phi are held split, twice:
// row-major planes (_r) and with every block transposed (_t); a GEMM operand reads whichever of
// the two has its K dimension contiguous (A: rows, B: columns), i.e. a view of the transposed
// copy carries LC_TV toggled. Intermediates are written by their producer directly in the form
// their consumer reads: split row-major when used as A, split transposed when used as B, FP32
// when used as a closing-trace factor.
void run_chunk_mx(const EngineShape& E, const char* tau_r, const char* tau_t, const Role* P, int p0, int pc,
                  char* scratch, void* raw) {
  const int nt = E.nt, ns = E.nsrc;
  const long long m = E.m, b = E.blk;
  const long long npair = (long long)nt * ns;
  auto f32 = [](auto p, long long e) { return ::off(p, e, 8); };
  auto spl = [](auto p, long long e) { return ::off(p, e, 12); };
  char* cur = scratch;
  auto take = [&](long long elems, size_t es) {
    char* p = cur;
    cur += ((size_t)elems * es + 255) & ~size_t(255);
    return p;
  };
  char *B32 = take(pc * npair * b, 8), *L1_32 = take(pc * npair * b, 8), *ZhT = take(pc * npair * b, 8);
  char *Bs = take(pc * npair * b, 12), *L1s = take(pc * npair * b, 12), *Ls = take(pc * npair * b, 12),
       *Rrs = take(pc * npair * b, 12), *Us = take(pc * npair * b, 12);
  char *Qa = take((long long)pc * nt * b, 12), *Q = take((long long)pc * nt * b, 12);
  char *T1 = take((long long)pc * ns * b, 12), *T1p = take((long long)pc * ns * b, 12),
       *RT = take((long long)pc * ns * b, 8), *RpT = take((long long)pc * ns * b, 8);

  const long long P0 = b, P1 = nt * b, P2 = (long long)ns * nt * b;  // pair strides
  // tau(tf, t0_k): A role from the row-major planes, B role from the transposed planes
  const LcView tau_A = dense(spl(tau_r, (long long)E.src0 * b), nt * b, E.step * b, 0);
  const LcView tau_B = dense(spl(tau_t, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_TV);
  const LcView tau_dag_A = dense(spl(tau_t, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_CONJ | LC_TS);
  const LcView tau_dag_B = dense(spl(tau_r, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_DAG);
  // doubleTriangleHor's W as a B operand: the reference's element-wise conj(tau_ba) from the
  // transposed planes, or tau^dag from the row-major ones (lc_set_dthor_dagger)
  const LcView W_B = ctx().dthor_dagger
                         ? dense(spl(tau_r, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_DAG)
                         : dense(spl(tau_t, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_CONJ | LC_TS | LC_TV);
  // Role k as an operand over the sinks (P3, P4) or the sources (P1, P2). The operand is X or X^T
  // with X = S or S^dag: S, S^T, S^dag = conj(S^T) or conj(S). The A role reads the operand by
  // rows, the B role reads its transpose by rows (flag TV); S by rows is the row-major planes,
  // S^T by rows the transposed ones.
  auto role = [&](int k, bool sink, bool brole, bool transposed) {
    const Role& R = P[k];
    const int op = R.dag ? (transposed ? 3 : 2) : (transposed ? 1 : 0);
    const bool t_rows = (op == 1 || op == 2) != brole;  // read the transposed planes
    const int fl = ((op >= 2) ? LC_CONJ : 0) | (brole ? LC_TV : 0);
    const char* base = spl(t_rows ? R.t : R.r, (long long)p0 * R.stride + (sink ? 0 : (long long)E.src0 * m));
    return sink ? kron(base, m, 0, R.stride, fl) : kron(base, 0, E.step * m, R.stride, fl);
  };
  const LcView phi_tf_A = role(3, true, false, false);       // P4(tf)
  const LcView phi_tf_B = role(3, true, true, false);
  const LcView phi_tf_dag_A = role(2, true, false, false);   // P3(tf)
  const LcView phi_t0_A = role(0, false, false, false);      // P1(t0)
  const LcView phi_t0_B = role(0, false, true, false);
  const LcView phi_t0_dag_A = role(1, false, false, false);  // P2(t0)
  const LcView phi_t0_dag_B = role(1, false, true, false);
  const LcView D_tf = dense(tau_r, (nt + 1) * b, 0, 0);
  const LcView D_t0 = dense(spl(tau_r, (long long)E.src0 * (nt + 1) * b), 0, E.step * (nt + 1) * b, 0);
  auto pair = [&](char* p, int flags = 0) { return dense(p, P0, P1, P2, flags); };
  char* rw = ::off(static_cast<char*>(raw), (long long)p0 * ns * nt * 10, 16);
  const long long O0 = 10, O1 = nt * 10LL, O2 = (long long)ns * nt * 10;
  auto slot = [&](int k, const LcView& Y, double ar = 1.0) {
    FusedTrace f = fused(Y, ::off(rw, k, 16), O0, O1, O2, ar);
    f.out_double = 1;
    return f;
  };
  (void)f32;

  Program pg(E.nv, Program::MX);
  pg.set_last_axis_fastest(true);  // momenta of one (tf, t0) pair next to each other
  // level 1
  pg.gemm(phi_tf_A, tau_B, B32, P0, P1, P2, nt, ns, pc, 1.0, 0.0, {}, Bs);
  pg.gemm(phi_tf_dag_A, tau_B, L1_32, P0, P1, P2, nt, ns, pc, 1.0, 0.0, {}, L1s);
  pg.gemm(D_tf, phi_tf_B, nullptr, b, 0, nt * b, nt, 1, pc, 1.0, 0.0, {}, Qa, true);
  pg.gemm(D_t0, phi_t0_B, nullptr, 0, b, ns * b, 1, ns, pc, 1.0, 0.0, {}, T1, true);
  pg.gemm(D_t0, phi_t0_dag_B, nullptr, 0, b, ns * b, 1, ns, pc, 1.0, 0.0, {}, T1p, true);
  pg.barrier();
  // level 2
  const Spin g5 = gamma(5);
  std::vector<FusedTrace> rho;
  for (int i = 0; i < 3; i++) {
    const Spin Lp = g5 * rhoj(1, i, false);
    const Spin Kp = rhoj(1, i, true) * g5;
    rho.push_back(slot(i, right(left(pair(B32, LC_DAG), Kp), Lp), -1.0));
  }
  pg.gemm(tau_A, phi_t0_B, nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, rho);
  pg.gemm(phi_t0_A, tau_dag_B, nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, {slot(6, pair(B32)), slot(8, pair(L1_32))});
  pg.gemm(phi_t0_dag_A, tau_dag_B, nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, {slot(7, pair(L1_32)), slot(9, pair(B32))});
  pg.gemm(pair(L1s), phi_t0_B, nullptr, P0, P1, P2, nt, ns, pc, 1.0, 0.0, {}, Ls);
  pg.gemm(pair(Bs), phi_t0_dag_B, nullptr, P0, P1, P2, nt, ns, pc, 1.0, 0.0, {}, Rrs);
  pg.gemm(phi_tf_dag_A, dense(Qa, b, 0, nt * b, LC_TV), nullptr, b, 0, nt * b, nt, 1, pc, 1.0, 0.0, {}, Q);
  // R = Phi^dag T1 and R' = Phi T1' are only closing-trace factors: formed transposed,
  // RT = T1^T conj(Phi) and RpT = T1'^T Phi^T, so that their fused traces read 16-byte vectors
  // (T1, T1' are stored block-transposed: a spin-transposed view of them is T1^T read by rows)
  pg.gemm(dense(T1, 0, b, ns * b, LC_TS), role(1, false, true, true), RT, 0, b, ns * b, 1, ns, pc);
  pg.gemm(dense(T1p, 0, b, ns * b, LC_TS), role(0, false, true, true), RpT, 0, b, ns * b, 1, ns, pc);
  pg.barrier();
  // level 3
  // Zh = Rr tau^dag, a closing-trace factor, formed transposed: ZhT = conj(tau) Rr^T
  pg.gemm(dense(spl(tau_r, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_CONJ), pair(Rrs, LC_TV | LC_TS), ZhT, P0,
          P1, P2, nt, ns, pc);
  pg.set_second_axis(1);  // sources before sinks: Q(tf, p) shared by neighbouring workgroups
  pg.gemm(dense(Q, b, 0, nt * b), right(left(tau_B, g5), g5), nullptr, P0, P1, P2, nt, ns, pc, 1.0, 0.0,
          {}, Us, true);
  pg.set_second_axis(0);
  pg.barrier();
  // level 4
  pg.gemm(tau_dag_A, pair(Us, LC_TV), nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0,
          {slot(3, dense(RT, 0, b, ns * b, LC_TV | LC_TS)), slot(4, dense(RpT, 0, b, ns * b, LC_TV | LC_TS))});
  pg.gemm(pair(Ls), W_B, nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, {slot(5, pair(ZhT, LC_TV | LC_TS))});
  pg.run(ctx().stream);
}

// Scratch per momentum of the sandwich form: split pair buffers L, Rr, U (12 B per complex),
// FP32 pair buffer Zh (8 B), split per-sink Q (12 B), FP32 per-source R, R' (8 B).
size_t mx_sw_bytes_per_momentum(const EngineShape& E) {
  const long long npair = (long long)E.nt * E.nsrc;
  return (size_t)E.blk * (size_t)(npair * (3 * 12 + 8) + E.nt * 12LL + 2LL * E.nsrc * 8);
}

// The mixed form with fused sandwiches (Nv = 64): every elemental-perambulator-elemental chain is
// one pass, Phi^(+-)(tf) tau Phi^(+-)(t0), with the middle product kept on the chip, and the
// closing traces of the short diagrams taken against tau^dag itself (FP32 copy tau32):
//   1  L  = Phi(tf)^dag tau Phi(t0)      -> rho_i = -Tr[L (K'_i tau^dag L'_i)] (i = 0, 1),
//                                           fish1 = Tr[L tau^dag]
//      Rr = Phi(tf) tau Phi(t0)^dag      -> fish2 = Tr[Rr tau^dag],
//                                           rho_2 = -conj Tr[Rr (K'_2^dag tau^dag L'_2^dag)]
//      X  = Phi(tf) tau Phi(t0)          -> star1 = Tr[X tau^dag]
//      Y  = Phi(tf)^dag tau Phi(t0)^dag  -> star2 = Tr[Y tau^dag]
//      Q = Phi^dag D(tf) Phi, R = Phi^dag D(t0) Phi, R' = Phi D(t0) Phi^dag
//   2  Zh = Rr tau^dag, U = Q (g5 tau g5)
//   3  M = tau^dag U -> square = Tr[M R], dT = Tr[M R'];  LW = L W -> dTHor = Tr[LW Zh]
// These are the traces of run_chunk reassociated: with Phi (x) 1 commuting with spin matrices,
// Tr[A (K' B^dag L')] = Tr[tau Phi(t0) K' tau^dag Phi(tf)^dag L'] = Tr[L K' tau^dag L'],
// Tr[S1 B] = Tr[X tau^dag], Tr[S1 L1] = Tr[L tau^dag], Tr[S2 L1] = Tr[Y tau^dag], Tr[S2 B] =
// Tr[Rr tau^dag], and conj Tr[L K' tau^dag L'] = Tr[L'^dag tau K'^dag L^dag] = Tr[Rr K'^dag tau^dag
// L'^dag] (one rho polarisation moves to Rr so that no product carries more than three traces).
// B and L1 are never formed.
void run_chunk_mx_sw(const EngineShape& E, const char* tau_r, const char* tau_t, const char* tau32,
                     const char* phi_r, const char* phi_t, int p0, int pc, char* scratch, void* raw) {
  const int nt = E.nt, ns = E.nsrc;
  const long long m = E.m, b = E.blk;
  const long long npair = (long long)nt * ns;
  auto f32 = [](auto p, long long e) { return ::off(p, e, 8); };
  auto spl = [](auto p, long long e) { return ::off(p, e, 12); };
  char* cur = scratch;
  auto take = [&](long long elems, size_t es) {
    char* p = cur;
    cur += ((size_t)elems * es + 255) & ~size_t(255);
    return p;
  };
  // Zh, R and R' are only closing-trace factors: they are stored transposed, so that the fused
  // traces read four consecutive values per 16-byte load (ZhT = conj(tau) Rr^T, RT = Phi^T D^T conj(Phi),
  // RpT = conj(Phi) D^T Phi^T)
  char* ZhT = take(pc * npair * b, 8);
  char *Ls = take(pc * npair * b, 12), *Rrs = take(pc * npair * b, 12), *Us = take(pc * npair * b, 12);
  char* Q = take((long long)pc * nt * b, 12);
  char *RT = take((long long)pc * ns * b, 8), *RpT = take((long long)pc * ns * b, 8);

  const long long P0 = b, P1 = nt * b, P2 = (long long)ns * nt * b;  // pair strides
  const LcView tau_B = dense(spl(tau_t, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_TV);
  const LcView tau_dag_A = dense(spl(tau_t, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_CONJ | LC_TS);
  const LcView tau_dag_B = dense(spl(tau_r, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_DAG);
  // doubleTriangleHor's W as a B operand: the reference's element-wise conj(tau_ba) from the
  // transposed planes, or tau^dag from the row-major ones (lc_set_dthor_dagger)
  const LcView W_B = ctx().dthor_dagger
                         ? tau_dag_B
                         : dense(spl(tau_t, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_CONJ | LC_TS | LC_TV);
  const LcView tau_dag_Y = dense(f32(tau32, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_DAG);
  const char* pr = spl(phi_r, (long long)p0 * nt * m);
  const char* pt = spl(phi_t, (long long)p0 * nt * m);
  const LcView phi_tf_A = kron(pr, m, 0, nt * m);
  const LcView phi_tf_B = kron(pt, m, 0, nt * m, LC_TV);
  const LcView phi_tf_dag_A = kron(pt, m, 0, nt * m, LC_CONJ);
  const LcView phi_t0_B = kron(spl(pt, (long long)E.src0 * m), 0, E.step * m, nt * m, LC_TV);
  const LcView phi_t0_dag_B = kron(spl(pr, (long long)E.src0 * m), 0, E.step * m, nt * m, LC_CONJ | LC_TV);
  const LcView D_tf_B = dense(tau_t, (nt + 1) * b, 0, 0, LC_TV);
  // transposes as operands: Phi^T (A) and conj(Phi) (A), D^T (B), conj(Phi) and Phi^T (B), conj(tau) (A)
  const LcView phi_t0_T_A = kron(spl(pt, (long long)E.src0 * m), 0, E.step * m, nt * m);
  const LcView phi_t0_cj_A = kron(spl(pr, (long long)E.src0 * m), 0, E.step * m, nt * m, LC_CONJ);
  const LcView D_t0_T_B = dense(spl(tau_r, (long long)E.src0 * (nt + 1) * b), 0, E.step * (nt + 1) * b, 0, LC_TS | LC_TV);
  const LcView phi_t0_cj_B = kron(spl(pt, (long long)E.src0 * m), 0, E.step * m, nt * m, LC_CONJ | LC_TV);
  const LcView phi_t0_T_B = kron(spl(pr, (long long)E.src0 * m), 0, E.step * m, nt * m, LC_TV);
  const LcView tau_cj_A = dense(spl(tau_r, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_CONJ);
  auto pair = [&](char* p, int flags = 0) { return dense(p, P0, P1, P2, flags); };
  char* rw = ::off(static_cast<char*>(raw), (long long)p0 * ns * nt * 10, 16);
  const long long O0 = 10, O1 = nt * 10LL, O2 = (long long)ns * nt * 10;
  auto slot = [&](int k, const LcView& Y, double ar = 1.0) {
    FusedTrace f = fused(Y, ::off(rw, k, 16), O0, O1, O2, ar);
    f.out_double = 1;
    return f;
  };

  Program pg(E.nv, Program::MX);
  pg.set_last_axis_fastest(true);  // momenta of one (tf, t0) pair next to each other
  // level 1: the sandwiches
  const Spin g5 = gamma(5);
  std::vector<FusedTrace> tl, tr;
  for (int i = 0; i < 3; i++) {
    const Spin Lp = g5 * rhoj(1, i, false);
    const Spin Kp = rhoj(1, i, true) * g5;
    if (i < 2) {
      tl.push_back(slot(i, right(left(tau_dag_Y, Kp), Lp), -1.0));
    } else {
      tr.push_back(slot(i, right(left(tau_dag_Y, Kp.adjoint()), Lp.adjoint()), -1.0));
      tr.back().conj = true;
    }
  }
  tl.push_back(slot(8, tau_dag_Y));
  tr.push_back(slot(9, tau_dag_Y));
  // L and Y share Phi(tf)^dag tau, Rr and X share Phi(tf) tau: one pass each
  const FusedTrace star1 = slot(6, tau_dag_Y), star2 = slot(7, tau_dag_Y);
  pg.sandwich(phi_tf_dag_A, tau_B, phi_t0_B, nullptr, P0, P1, P2, nt, ns, pc, 1.0, 0.0, tl, Ls, false,
              &phi_t0_dag_B, &star2);
  pg.sandwich(phi_tf_A, tau_B, phi_t0_dag_B, nullptr, P0, P1, P2, nt, ns, pc, 1.0, 0.0, tr, Rrs, false,
              &phi_t0_B, &star1);
  pg.sandwich(phi_tf_dag_A, D_tf_B, phi_tf_B, nullptr, b, 0, nt * b, nt, 1, pc, 1.0, 0.0, {}, Q);
  pg.sandwich(phi_t0_T_A, D_t0_T_B, phi_t0_cj_B, RT, 0, b, ns * b, 1, ns, pc);
  pg.sandwich(phi_t0_cj_A, D_t0_T_B, phi_t0_T_B, RpT, 0, b, ns * b, 1, ns, pc);
  pg.barrier();
  // level 2; U walks the sources before the sinks, so that Q(tf, p) of the momenta in flight is
  // shared by neighbouring workgroups
  pg.gemm(tau_cj_A, pair(Rrs, LC_TV | LC_TS), ZhT, P0, P1, P2, nt, ns, pc);
  // U = Q (g5 tau g5) is formed transposed, UT = (g5 tau^T g5) Q^T, so that its split store is
  // row-major (tau^T read by rows is the transposed planes with the spin blocks exchanged)
  pg.set_second_axis(1);
  pg.gemm(right(left(dense(spl(tau_t, (long long)E.src0 * b), nt * b, E.step * b, 0, LC_TS), g5), g5),
          dense(Q, b, 0, nt * b, LC_TV | LC_TS), nullptr, P0, P1, P2, nt, ns, pc, 1.0, 0.0, {}, Us, false);
  pg.set_second_axis(0);
  pg.barrier();
  // level 3
  pg.gemm(tau_dag_A, pair(Us, LC_TV | LC_TS), nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0,
          {slot(3, dense(RT, 0, b, ns * b, LC_TV | LC_TS)), slot(4, dense(RpT, 0, b, ns * b, LC_TV | LC_TS))});
  pg.gemm(pair(Ls), W_B, nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, {slot(5, pair(ZhT, LC_TV | LC_TS))});
  pg.run(ctx().stream);
}

}  // namespace

extern "C" {

// Engine entry points; see include/lapH_contract.h.
// sets = false: d_phi [np][nt][nv][nv], back-to-back pions P1 = P4 = Phi, P2 = P3 = Phi^dag.
// sets = true:  d_phi [np][4][nt][nv][nv], the four elementals P1..P4 of each of np momentum sets.
static int engine_impl(bool mixed, bool sets, const void* d_tau, const void* d_phi, int nt, int nv, int np,
                       int src0, int src_step, int nsrc, long long max_scratch_bytes, void* d_raw,
                       void* d_corr, void* d_avg) {
  try {
    require_init();
    if (nt < 1 || nv < 1 || nv > 1024 || np < 1) fail(LC_ERR_ARG, "bad sizes nt=%d nv=%d np=%d", nt, nv, np);
    if (nsrc < 1 || src_step < 1 || src0 < 0 || src0 + (long long)(nsrc - 1) * src_step >= nt)
      fail(LC_ERR_ARG, "bad source list src0=%d step=%d nsrc=%d (nt=%d)", src0, src_step, nsrc, nt);
    if (!d_tau || !d_phi || !d_raw || !d_corr || !d_avg) fail(LC_ERR_ARG, "null pointer");
    if (mixed && ctx().mixed_form == 0) {
      oz_engine_rho_pipi(sets, d_tau, d_phi, nt, nv, np, src0, src_step, nsrc, max_scratch_bytes, d_raw, d_corr,
                         d_avg);
      return LC_OK;
    }
    EngineShape E{nt, nv, np, src0, src_step, nsrc, (long long)nv * nv, 16LL * nv * nv};
    const long long npair = (long long)nt * nsrc;
    // mixed form: tau and Phi split twice (row-major and transposed planes, 12 B per complex
    // each) + FP64 per-slice scale factors, ahead of the scratch area
    // the fused-sandwich form (Nv = 64) also keeps tau in FP32 for its closing traces
    // (the sandwich chain pairs Phi with Phi^dag through shared passes: back-to-back only)
    const bool sw = mixed && nv == 64 && !sets;
    const int nel = sets ? 4 : 1;  // elemental arrays per momentum item
    const long long nphi = (long long)np * nel * nt;  // elemental matrices
    const size_t tau_split = mixed ? (((size_t)nt * nt * E.blk * 12 + 255) & ~size_t(255)) : 0;
    const size_t tau_f32 = sw ? (((size_t)nt * nt * E.blk * 8 + 255) & ~size_t(255)) : 0;
    const size_t phi_split = mixed ? (((size_t)nphi * E.m * 12 + 255) & ~size_t(255)) : 0;
    const size_t scale_bytes = mixed ? (((size_t)(nt * nt + nphi) * 8 + 255) & ~size_t(255)) : 0;
    const size_t head = 2 * tau_split + tau_f32 + 2 * phi_split + scale_bytes;
    const size_t per_p = sw      ? mx_sw_bytes_per_momentum(E) + 8 * 256
                         : mixed ? mx_bytes_per_momentum(E) + 20 * 256
                                 : (size_t)(PAIR_BUFS * npair + 2LL * nt + 4LL * nsrc) * E.blk * 16;
    if (max_scratch_bytes <= 0) max_scratch_bytes = default_scratch_budget();
    long long avail = max_scratch_bytes - (long long)head;
    int pc = avail > 0 ? (int)(avail / (long long)per_p) : 0;
    if (pc < 1) fail(LC_ERR_NOMEM, "scratch budget %lld B below one momentum (%zu B)", max_scratch_bytes, per_p + head);
    if (pc > np) pc = np;
    ctx().scratch.ensure(head + per_p * pc + 4096);
    char* base = static_cast<char*>(ctx().scratch.p);
    hipStream_t s = ctx().stream;
    char* scratch = base + head;
    // roles P1..P4 over arrays r (and t) whose item p sits at p * stride elements
    auto roles = [&](const char* r, const char* t, size_t esz, Role* P) {
      for (int k = 0; k < 4; k++) {
        if (sets)
          P[k] = Role{r + (size_t)k * nt * E.m * esz, t ? t + (size_t)k * nt * E.m * esz : nullptr, false,
                      4LL * nt * E.m};
        else
          P[k] = Role{r, t, k == 1 || k == 2, (long long)nt * E.m};
      }
    };
    auto call = [&] {
      double* scales = nullptr;
      if (mixed) {
        char* tau_r = base;
        char* tau_t = tau_r + tau_split;
        char* tau32 = sw ? tau_t + tau_split : nullptr;
        char* phi_r = tau_t + tau_split + tau_f32;
        char* phi_t = phi_r + phi_split;
        scales = reinterpret_cast<double*>(phi_t + phi_split);
        const long long blk = E.blk, m = E.m;
        launch([=](hipStream_t st) { return lc_launch_split_rescale(d_tau, tau_r, tau_t, tau32, blk, nv, (long long)nt * nt, scales, st); },
               "split tau");
        launch([=](hipStream_t st) {
                 return lc_launch_split_rescale(d_phi, phi_r, phi_t, nullptr, m, nv, nphi, scales + (size_t)nt * nt,
                                                st);
               },
               "split phi");
        Role P[4];
        roles(phi_r, phi_t, 12, P);
        for (int p0 = 0; p0 < np; p0 += pc) {
          const int n = (np - p0) < pc ? (np - p0) : pc;
          if (sw) run_chunk_mx_sw(E, tau_r, tau_t, tau32, phi_r, phi_t, p0, n, scratch, d_raw);
          else run_chunk_mx(E, tau_r, tau_t, P, p0, n, scratch, d_raw);
        }
      } else {
        Role P[4];
        roles(static_cast<const char*>(d_phi), nullptr, 16, P);
        for (int p0 = 0; p0 < np; p0 += pc) {
          const int n = (np - p0) < pc ? (np - p0) : pc;
          run_chunk(E, Program::F64, d_tau, P, p0, n, scratch, d_raw);
        }
      }
      // per-role scale factors of the elementals: item p, time t at s_p[k][p * ps + t]
      const double* s_phi = scales ? scales + (size_t)nt * nt : nullptr;
      const double* sp[4];
      for (int k = 0; k < 4; k++) sp[k] = s_phi ? s_phi + (sets ? (size_t)k * nt : 0) : nullptr;
      const long long ps = (long long)nel * nt;
      launch([=](hipStream_t st) {
               return lc_launch_engine_epilogue(d_raw, d_corr, d_avg, np, nsrc, nt, src0, src_step, scales, sp, ps,
                                                st);
             },
             "engine epilogue");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    };
    char key[160], gkey[320];
    std::snprintf(key, sizeof key, "%s nt=%d nv=%d np=%d src0=%d step=%d nsrc=%d chunk=%d",
                  sets ? "pipi_sets" : "rho_pipi", nt, nv, np, src0, src_step, nsrc, pc);
    std::snprintf(gkey, sizeof gkey, "%s mx=%d tau=%p phi=%p raw=%p corr=%p avg=%p", key, mixed ? 1 : 0, d_tau, d_phi,
                  d_raw, d_corr, d_avg);
    const std::string graph_key(gkey);
    run_with_tiles(key, mixed, [&] { run_engine_call(graph_key, call); });
    return LC_OK;
  } catch (const Error& e) {
    last_error() = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return LC_ERR_ARG;
  }
}

int lc_engine_rho_pipi(const void* d_tau, const void* d_phi, int nt, int nv, int np, int src0,
                       int src_step, int nsrc, long long max_scratch_bytes, void* d_raw,
                       void* d_corr, void* d_avg) {
  return engine_impl(false, false, d_tau, d_phi, nt, nv, np, src0, src_step, nsrc, max_scratch_bytes, d_raw,
                     d_corr, d_avg);
}

int lc_engine_rho_pipi_mixed(const void* d_tau, const void* d_phi, int nt, int nv, int np, int src0,
                             int src_step, int nsrc, long long max_scratch_bytes, void* d_raw,
                             void* d_corr, void* d_avg) {
  return engine_impl(true, false, d_tau, d_phi, nt, nv, np, src0, src_step, nsrc, max_scratch_bytes, d_raw,
                     d_corr, d_avg);
}

int lc_engine_pipi_sets(const void* d_tau, const void* d_elem, int nt, int nv, int nset, int src0, int src_step,
                        int nsrc, long long max_scratch_bytes, void* d_raw, void* d_corr, void* d_avg) {
  return engine_impl(false, true, d_tau, d_elem, nt, nv, nset, src0, src_step, nsrc, max_scratch_bytes, d_raw,
                     d_corr, d_avg);
}

int lc_engine_pipi_sets_mixed(const void* d_tau, const void* d_elem, int nt, int nv, int nset, int src0,
                              int src_step, int nsrc, long long max_scratch_bytes, void* d_raw, void* d_corr,
                              void* d_avg) {
  return engine_impl(true, true, d_tau, d_elem, nt, nv, nset, src0, src_step, nsrc, max_scratch_bytes, d_raw,
                     d_corr, d_avg);
}

int lc_set_mixed_form(int form) {
  if (form != LC_MIXED_EXACT && form != LC_MIXED_BF16X3) {
    last_error() = "mixed form must be LC_MIXED_EXACT or LC_MIXED_BF16X3";
    return LC_ERR_ARG;
  }
  ctx().mixed_form = form;
  return LC_OK;
}

int lc_profile(int enable) {
  ctx().profile = enable != 0;
  ctx().times.clear();
  return LC_OK;
}

int lc_profile_read(double* out /* [6] */) {
  if (!out) return LC_ERR_ARG;
  const KernelTimes& t = ctx().times;
  out[0] = t.gemm_ms;
  out[1] = (double)t.gemm_launches;
  out[2] = t.gemm_flops;
  out[3] = t.trace_ms;
  out[4] = (double)t.trace_launches;
  out[5] = t.trace_flops;
  return LC_OK;
}

}  // extern "C"
