// Batched engine for the sigma / sigma-pi / rho-pi set: every (sink tf, source t0, momentum set p)
// of the sigma one_fish, sigma-pi triangle and rho-pi correlators over a perambulator resident in
// HBM, plus the per-time loops and sigma-pi hor_fish, as a few grouped launches per momentum chunk.
//
// Inputs (FP64 complex):
//   tau  [nt][nt][4][4][nv][nv]   tau(t, t0), the reference's host layout (lib/contract.cpp:140)
//   el   [np][nt][11][nv][nv]     elementals of momentum set p at time t, by type:
//                                  0, 1   X_l   sigma operators (Mom_matrices_x2 of run_ContractSigma)
//                                  2, 3   X_ip  sigma-pi triangle operators (P1P2 Mom_matrices_x2)
//                                  4..8   P12   rho-pi Mom_matrices of P1P2 (5 types)
//                                  9, 10  P1, P2 pion elementals
// Per (p, tf, t0), with tau = tau(tf, t0) and D(t) = tau(t, t), the row forms of
// lib/contract_kernels_sigma.cu, lib/contract_kernels_sigmapi.cu and lib/contract_kernels_rhopi.cu:
//   sigma(l, r)   = Tr[ (X_l(tf) tau)^dag g5 (tau X_r(t0)) g5 ]           raw slots 0..3 (l*2 + r)
//   triangle(ip)  = Tr[ (g5 X_ip(tf) tau) (g5 R tau^dag) ]                slots 4, 5
//   rhopi(iop)    = Tr[ (G_iop P12_ip(tf) tau) (g5 R tau^dag) ]           slots 6..9
//                   R = P2(t0) D(t0) P1(t0), G_iop = g5 rho_{iop+1}(idir), ip = {0, 0, 1, 2+idir}
// and per (p, t): loop_l = Tr[D(t) X_l(t)] (l = 0, 1), hor_fish = Tr[P2 D P1 D^dag].
// The raw traces are unweighted, as the row launchers return them.
//
// Levels (closing traces fused into the GEMM of their left factor):
//   1  B_l = X_l(tf) tau (stored);  T1 = D(t0) P1 (per source)
//   2  A_r = tau X_r(t0) -> sigma(l, r) = Tr[A_r (g5 B_l^dag g5)];  R = P2 T1 (per source)
//   3  Y = (g5 R) tau^dag (stored)
//   4  Z = X_ip(tf) tau -> triangle(ip) = Tr[Z (Y g5)];  Z' = P12_ip(tf) tau -> rhopi = Tr[Z' (Y G)]
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "engine_oz.h"
#include "runtime.h"
#include "spin.h"

extern "C" hipError_t lc_launch_sigma_scale(void* raw, int np, int nsrc, int nt, int src0, int src_step,
                                            int idir, const double* s_tau, const double* s_el, hipStream_t s);

using namespace lc;

namespace {

constexpr int NT_EL = 11;
enum { EL_SIG = 0, EL_TRI = 2, EL_P12 = 4, EL_P1 = 9, EL_P2 = 10 };
constexpr int N_RAW = 10, N_LOOP = 3;

struct Shape {
  int nt, nv, np, src0, step, nsrc, idir;
  long long m, blk;
};

// Storage of one operand: FP64 data (rows == trans), or the mixed form's split planes, row-major
// and with every block transposed.
struct Store {
  const char* rows;
  const char* trans;
};

class Builder {
 public:
  Builder(bool mixed, int nv) : mx_(mixed), es_(mixed ? 12 : 16), pg_(nv, mixed ? Program::MX : Program::F64) {}
  Program& pg() { return pg_; }
  // A operand: read along rows (the mixed kernel needs K contiguous: a view with LC_TV reads the
  // transposed planes instead)
  LcView A(const Store& st, long long off, long long s0, long long s1, long long s2, int kind, int flags = 0) const {
    if (mx_ && (flags & LC_TV)) return make_view(st.trans + off * es_, s0, s1, s2, kind, flags ^ LC_TV);
    return make_view(st.rows + off * es_, s0, s1, s2, kind, flags);
  }
  LcView B(const Store& st, long long off, long long s0, long long s1, long long s2, int kind, int flags = 0) const {
    if (mx_ && !(flags & LC_TV)) return make_view(st.trans + off * es_, s0, s1, s2, kind, flags ^ LC_TV);
    return make_view(st.rows + off * es_, s0, s1, s2, kind, flags);
  }
  bool mixed() const { return mx_; }

 private:
  bool mx_;
  size_t es_;
  Program pg_;
};

// Scratch per momentum: pair buffers B_0, B_1, Y (trace operands: FP64, or FP32 in the mixed
// form) and per-source T1, R (FP64, or split).
size_t bytes_per_momentum(const Shape& E, bool mixed) {
  const long long npair = (long long)E.nt * E.nsrc;
  const size_t tr = mixed ? 8 : 16, op = mixed ? 12 : 16;
  return (size_t)E.blk * ((size_t)(3 * npair) * tr + (size_t)(2LL * E.nsrc) * op) + 8 * 256;
}

void run_chunk(const Shape& E, bool mixed, const Store& tau, const Store& el, int p0, int pc, char* scratch,
               void* raw) {
  const int nt = E.nt, ns = E.nsrc;
  const long long m = E.m, b = E.blk;
  const long long npair = (long long)nt * ns;
  const size_t tr = mixed ? 8 : 16, op = mixed ? 12 : 16;
  char* cur = scratch;
  auto take = [&](long long elems, size_t es) {
    char* p = cur;
    cur += ((size_t)elems * es + 255) & ~size_t(255);
    return p;
  };
  char* Bsig = take(2 * pc * npair * b, tr);
  char* Y = take(pc * npair * b, tr);
  char* T1 = take((long long)pc * ns * b, op);
  char* R = take((long long)pc * ns * b, op);

  Builder bl(mixed, E.nv);
  Program& pg = bl.pg();
  pg.set_last_axis_fastest(true);  // momentum sets of one (tf, t0) pair next to each other
  const long long P0 = b, P1 = nt * b, P2 = (long long)ns * nt * b;  // pair strides
  const long long tau0 = (long long)E.src0 * b;
  // elemental type j of momentum p0 + co2: at sink (per tf), at source (per k)
  const long long eP = (long long)nt * NT_EL * m;
  auto el_tf = [&](int j) { return (long long)p0 * eP + (long long)j * m; };
  auto el_t0 = [&](int j) { return (long long)p0 * eP + (long long)E.src0 * NT_EL * m + (long long)j * m; };
  const long long eTF = NT_EL * m, eK = (long long)E.step * NT_EL * m;
  auto pair = [&](char* p, int flags = 0) { return make_view(p, P0, P1, P2, LC_DENSE, flags); };
  char* rw = static_cast<char*>(raw) + (size_t)p0 * ns * nt * N_RAW * 16;
  auto slot = [&](int k, const LcView& Yv) {
    FusedTrace f = fused(Yv, rw + (size_t)k * 16, N_RAW, (long long)nt * N_RAW, (long long)ns * nt * N_RAW);
    f.out_double = 1;
    return f;
  };
  const Spin g5 = gamma(5);

  // level 1
  for (int l = 0; l < 2; l++)
    pg.gemm(bl.A(el, el_tf(EL_SIG + l), eTF, 0, eP, LC_KRON), bl.B(tau, tau0, nt * b, E.step * b, 0, LC_DENSE),
            Bsig + (size_t)l * pc * npair * b * tr, P0, P1, P2, nt, ns, pc);
  {
    const LcView D0 = bl.A(tau, (long long)E.src0 * (nt + 1) * b, 0, E.step * (nt + 1) * b, 0, LC_DENSE);
    const LcView P1v = bl.B(el, el_t0(EL_P1), 0, eK, eP, LC_KRON);
    if (mixed) pg.gemm(D0, P1v, nullptr, 0, b, ns * b, 1, ns, pc, 1.0, 0.0, {}, T1, true);
    else pg.gemm(D0, P1v, T1, 0, b, ns * b, 1, ns, pc);
  }
  pg.barrier();
  // level 2
  for (int r = 0; r < 2; r++) {
    std::vector<FusedTrace> tr2;
    for (int l = 0; l < 2; l++) {
      char* Bl = Bsig + (size_t)l * pc * npair * b * tr;
      tr2.push_back(slot(l * 2 + r, right(left(pair(Bl, LC_DAG), g5), g5)));
    }
    pg.gemm(bl.A(tau, tau0, nt * b, E.step * b, 0, LC_DENSE), bl.B(el, el_t0(EL_SIG + r), 0, eK, eP, LC_KRON),
            nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, tr2);
  }
  {
    const LcView P2v = bl.A(el, el_t0(EL_P2), 0, eK, eP, LC_KRON);
    // T1 as the B operand: FP64 as stored; split planes stored transposed (K contiguous)
    const LcView T1v = make_view(T1, 0, b, ns * b, LC_DENSE, mixed ? LC_TV : 0);
    if (mixed) pg.gemm(P2v, T1v, nullptr, 0, b, ns * b, 1, ns, pc, 1.0, 0.0, {}, R, false);
    else pg.gemm(P2v, T1v, R, 0, b, ns * b, 1, ns, pc);
  }
  pg.barrier();
  // level 3
  pg.gemm(left(make_view(R, 0, b, ns * b, LC_DENSE, 0), g5), bl.B(tau, tau0, nt * b, E.step * b, 0, LC_DENSE, LC_DAG),
          Y, P0, P1, P2, nt, ns, pc);
  pg.barrier();
  // level 4
  for (int ip = 0; ip < 2; ip++)
    pg.gemm(bl.A(el, el_tf(EL_TRI + ip), eTF, 0, eP, LC_KRON), bl.B(tau, tau0, nt * b, E.step * b, 0, LC_DENSE),
            nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, {slot(4 + ip, right(pair(Y), g5))});
  const int ips[3] = {0, 1, 2 + E.idir};
  for (int j = 0; j < 3; j++) {
    std::vector<FusedTrace> t4;
    for (int iop = 0; iop < 4; iop++)
      if (rhopi_ip(E.idir, iop) == ips[j]) t4.push_back(slot(6 + iop, right(pair(Y), g5 * rhoj(iop + 1, E.idir, false))));
    pg.gemm(bl.A(el, el_tf(EL_P12 + ips[j]), eTF, 0, eP, LC_KRON), bl.B(tau, tau0, nt * b, E.step * b, 0, LC_DENSE),
            nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, t4);
  }
  pg.run(ctx().stream);
}

// Scratch per momentum set of the sandwich form: the pair buffer Y (FP32) and the per-source R
// (split).
size_t sw_bytes_per_momentum(const Shape& E) {
  const long long npair = (long long)E.nt * E.nsrc;
  return (size_t)E.blk * ((size_t)npair * 8 + (size_t)E.nsrc * 12) + 4 * 256;
}

// The mixed form with fused sandwiches (Nv = 64). With X_l (x) 1 commuting with g5,
//   sigma(l, r) = Tr[tau^dag X_l^dag g5 tau X_r g5] = Tr[S_lr tau^dag],  S_lr = X_l(tf)^dag (g5 tau g5) X_r(t0)
// and one pass per l forms X_l^dag (g5 tau g5) once and both right factors X_0, X_1 (traces only,
// against tau^dag in FP32); R = P2 D(t0) P1 is one sandwich per source. B_l is never formed.
//   1  sandwiches S_l0, S_l1 (l = 0, 1) -> sigma;  R (per source, split)
//   2  Y = (g5 R) tau^dag
//   3  Z = X_ip(tf) tau -> triangle(ip) = Tr[Z (Y g5)];  Z' = P12_ip(tf) tau -> rhopi = Tr[Z' (Y G)]
void run_chunk_sw(const Shape& E, const Store& tau, const char* tau32, const Store& el, int p0, int pc,
                  char* scratch, void* raw) {
  const int nt = E.nt, ns = E.nsrc;
  const long long m = E.m, b = E.blk;
  const long long npair = (long long)nt * ns;
  char* cur = scratch;
  auto take = [&](long long elems, size_t es) {
    char* p = cur;
    cur += ((size_t)elems * es + 255) & ~size_t(255);
    return p;
  };
  // Y = (g5 R) tau^dag is only a closing-trace factor: it is formed transposed, YT = conj(tau) R^T g5,
  // so that the fused traces read four consecutive values per 16-byte load
  char* YT = take(pc * npair * b, 8);
  char* R = take((long long)pc * ns * b, 12);

  Builder bl(true, E.nv);
  Program& pg = bl.pg();
  pg.set_last_axis_fastest(true);  // momentum sets of one (tf, t0) pair next to each other
  const long long P0 = b, P1 = nt * b, P2 = (long long)ns * nt * b;  // pair strides
  const long long tau0 = (long long)E.src0 * b;
  const long long eP = (long long)nt * NT_EL * m;
  auto el_tf = [&](int j) { return (long long)p0 * eP + (long long)j * m; };
  auto el_t0 = [&](int j) { return (long long)p0 * eP + (long long)E.src0 * NT_EL * m + (long long)j * m; };
  const long long eTF = NT_EL * m, eK = (long long)E.step * NT_EL * m;
  auto pair = [&](char* p, int flags = 0) { return make_view(p, P0, P1, P2, LC_DENSE, flags); };
  char* rw = static_cast<char*>(raw) + (size_t)p0 * ns * nt * N_RAW * 16;
  auto slot = [&](int k, const LcView& Yv) {
    FusedTrace f = fused(Yv, rw + (size_t)k * 16, N_RAW, (long long)nt * N_RAW, (long long)ns * nt * N_RAW);
    f.out_double = 1;
    return f;
  };
  const Spin g5 = gamma(5);
  const LcView tau_dag_Y = make_view(tau32 + (size_t)tau0 * 8, nt * b, E.step * b, 0, LC_DENSE, LC_DAG);
  const LcView g5tg5 = right(left(bl.B(tau, tau0, nt * b, E.step * b, 0, LC_DENSE), g5), g5);

  // level 1
  const LcView X1_t0 = bl.B(el, el_t0(EL_SIG + 1), 0, eK, eP, LC_KRON);
  for (int l = 0; l < 2; l++) {
    const FusedTrace s1 = slot(l * 2 + 1, tau_dag_Y);
    pg.sandwich(bl.A(el, el_tf(EL_SIG + l), eTF, 0, eP, LC_KRON, LC_CONJ | LC_TV), g5tg5,
                bl.B(el, el_t0(EL_SIG + 0), 0, eK, eP, LC_KRON), nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0,
                {slot(l * 2 + 0, tau_dag_Y)}, nullptr, false, &X1_t0, &s1);
  }
  pg.sandwich(bl.A(el, el_t0(EL_P2), 0, eK, eP, LC_KRON),
              bl.B(tau, (long long)E.src0 * (nt + 1) * b, 0, E.step * (nt + 1) * b, 0, LC_DENSE),
              bl.B(el, el_t0(EL_P1), 0, eK, eP, LC_KRON), nullptr, 0, b, ns * b, 1, ns, pc, 1.0, 0.0, {}, R);
  pg.barrier();
  // level 2
  pg.gemm(bl.A(tau, tau0, nt * b, E.step * b, 0, LC_DENSE, LC_CONJ),
          right(make_view(R, 0, b, ns * b, LC_DENSE, LC_TV | LC_TS), g5), YT, P0, P1, P2, nt, ns, pc);
  pg.barrier();
  // level 3
  for (int ip = 0; ip < 2; ip++)
    pg.gemm(bl.A(el, el_tf(EL_TRI + ip), eTF, 0, eP, LC_KRON), bl.B(tau, tau0, nt * b, E.step * b, 0, LC_DENSE),
            nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, {slot(4 + ip, right(pair(YT, LC_TV | LC_TS), g5))});
  const int ips[3] = {0, 1, 2 + E.idir};
  for (int j = 0; j < 3; j++) {
    std::vector<FusedTrace> t4;
    for (int iop = 0; iop < 4; iop++)
      if (rhopi_ip(E.idir, iop) == ips[j])
        t4.push_back(slot(6 + iop, right(pair(YT, LC_TV | LC_TS), g5 * rhoj(iop + 1, E.idir, false))));
    pg.gemm(bl.A(el, el_tf(EL_P12 + ips[j]), eTF, 0, eP, LC_KRON), bl.B(tau, tau0, nt * b, E.step * b, 0, LC_DENSE),
            nullptr, 0, 0, 0, nt, ns, pc, 1.0, 0.0, t4);
  }
  pg.run(ctx().stream);
}

// Per-time loops and hor_fish, FP64 on the FP64 inputs (a few products per time slice).
void run_loops(const Shape& E, const void* tau, const void* el, int p0, int pc, char* scratch, void* loops) {
  const int nt = E.nt;
  const long long m = E.m, b = E.blk;
  const long long eP = (long long)nt * NT_EL * m;
  char* Tf = scratch;
  char* Tf2 = scratch + (((size_t)pc * nt * b * 16 + 255) & ~size_t(255));
  const char* e = static_cast<const char*>(el) + (size_t)p0 * eP * 16;
  auto at = [&](long long elems) { return e + (size_t)elems * 16; };
  char* lo = static_cast<char*>(loops) + (size_t)p0 * nt * N_LOOP * 16;
  Program pg(E.nv, Program::F64);
  const LcView D = make_view(tau, (nt + 1) * b, 0, 0, LC_DENSE, 0);
  for (int l = 0; l < 2; l++)
    pg.trace(D, make_view(at((long long)(EL_SIG + l) * m), NT_EL * m, 0, eP, LC_KRON, 0), lo + (size_t)l * 16, N_LOOP, 0,
             (long long)nt * N_LOOP, nt, 1, pc);
  pg.gemm(make_view(at((long long)EL_P2 * m), NT_EL * m, 0, eP, LC_KRON, 0), D, Tf, b, 0, (long long)nt * b, nt, 1, pc);
  pg.barrier();
  pg.gemm(make_view(Tf, b, 0, (long long)nt * b, LC_DENSE, 0), make_view(at((long long)EL_P1 * m), NT_EL * m, 0, eP, LC_KRON, 0),
          Tf2, b, 0, (long long)nt * b, nt, 1, pc);
  pg.barrier();
  pg.trace(make_view(Tf2, b, 0, (long long)nt * b, LC_DENSE, 0), make_view(tau, (nt + 1) * b, 0, 0, LC_DENSE, LC_DAG),
           lo + 2 * 16, N_LOOP, 0, (long long)nt * N_LOOP, nt, 1, pc);
  pg.run(ctx().stream);
}

}  // namespace

extern "C" {

static int sigma_impl(bool mixed, const void* d_tau, const void* d_el, int nt, int nv, int np, int idir,
                      int src0, int src_step, int nsrc, long long max_scratch_bytes, void* d_raw,
                      void* d_loops) {
  try {
    require_init();
    if (nt < 1 || nv < 1 || nv > 1024 || np < 1) fail(LC_ERR_ARG, "bad sizes nt=%d nv=%d np=%d", nt, nv, np);
    if (idir < 0 || idir > 2) fail(LC_ERR_ARG, "idir=%d out of range", idir);
    if (nsrc < 1 || src_step < 1 || src0 < 0 || src0 + (long long)(nsrc - 1) * src_step >= nt)
      fail(LC_ERR_ARG, "bad source list src0=%d step=%d nsrc=%d (nt=%d)", src0, src_step, nsrc, nt);
    if (!d_tau || !d_el || !d_raw || !d_loops) fail(LC_ERR_ARG, "null pointer");
    Shape E{nt, nv, np, src0, src_step, nsrc, idir, (long long)nv * nv, 16LL * nv * nv};
    hipStream_t s = ctx().stream;
    // the exact form (mixed, LC_MIXED_EXACT): tau and the elementals in digits ahead of the scratch
    const bool exact = mixed && ctx().mixed_form == 0;
    if (exact) mixed = false;  // (the split-storage form below is the BF16x3 one)
    // mixed form: tau and the elementals split twice + FP64 per-slice scales, ahead of the scratch
    // the fused-sandwich form (Nv = 64) also keeps tau in FP32 for its closing traces
    const bool sw = mixed && nv == 64;
    const size_t tau_split = mixed ? (((size_t)nt * nt * E.blk * 12 + 255) & ~size_t(255)) : 0;
    const size_t tau_f32 = sw ? (((size_t)nt * nt * E.blk * 8 + 255) & ~size_t(255)) : 0;
    const size_t el_split = mixed ? (((size_t)np * nt * NT_EL * E.m * 12 + 255) & ~size_t(255)) : 0;
    const size_t scale_bytes = mixed ? (((size_t)nt * nt + (size_t)np * nt * NT_EL) * 8 + 255) & ~size_t(255) : 0;
    const size_t head = exact ? oz_sigma_head_bytes(nt, nv, np) : 2 * tau_split + tau_f32 + 2 * el_split + scale_bytes;
    const size_t loop_per_p = 2 * ((((size_t)nt * E.blk * 16) + 255) & ~size_t(255));
    const size_t per_p = exact ? oz_sigma_bytes_per_momentum(nt, nv, nsrc)
                         : sw  ? sw_bytes_per_momentum(E)
                               : bytes_per_momentum(E, mixed);
    const size_t need1 = per_p > loop_per_p ? per_p : loop_per_p;
    if (max_scratch_bytes <= 0) max_scratch_bytes = default_scratch_budget();
    const long long avail = max_scratch_bytes - (long long)head;
    int pc = avail > 0 ? (int)(avail / (long long)need1) : 0;
    if (pc < 1) fail(LC_ERR_NOMEM, "scratch budget %lld B below one momentum set (%zu B)", max_scratch_bytes, need1 + head);
    if (pc > np) pc = np;
    ctx().scratch.ensure(head + need1 * pc + 4096);
    char* base = static_cast<char*>(ctx().scratch.p);
    char* scratch = base + head;
    auto call = [&] {
      double* scales = nullptr;
      char* tau32 = nullptr;
      Store tau{static_cast<const char*>(d_tau), static_cast<const char*>(d_tau)};
      Store el{static_cast<const char*>(d_el), static_cast<const char*>(d_el)};
      if (mixed) {
        char* tr = base;
        char* tt = tr + tau_split;
        tau32 = sw ? tt + tau_split : nullptr;
        char* er = tt + tau_split + tau_f32;
        char* et = er + el_split;
        scales = reinterpret_cast<double*>(et + el_split);
        const long long blk = E.blk, m = E.m;
        char* t32 = tau32;
        double* sc = scales;
        launch([=](hipStream_t st) { return lc_launch_split_rescale(d_tau, tr, tt, t32, blk, nv, (long long)nt * nt, sc, st); },
               "split tau");
        launch([=](hipStream_t st) {
                 return lc_launch_split_rescale(d_el, er, et, nullptr, m, nv, (long long)np * nt * NT_EL,
                                                sc + (size_t)nt * nt, st);
               },
               "split elementals");
        tau = Store{tr, tt};
        el = Store{er, et};
      }
      if (exact) {
        oz_sigma_set(d_tau, d_el, nt, nv, np, idir, src0, src_step, nsrc, base, scratch, pc, d_raw);
      } else {
        for (int p0 = 0; p0 < np; p0 += pc) {
          const int n = (np - p0) < pc ? (np - p0) : pc;
          if (sw) run_chunk_sw(E, tau, tau32, el, p0, n, scratch, d_raw);
          else run_chunk(E, mixed, tau, el, p0, n, scratch, d_raw);
        }
      }
      if (mixed) {
        double* sc = scales;
        launch([=](hipStream_t st) {
                 return lc_launch_sigma_scale(d_raw, np, nsrc, nt, src0, src_step, idir, sc, sc + (size_t)nt * nt, st);
               },
               "rescale traces");
      }
      for (int p0 = 0; p0 < np; p0 += pc) {
        const int n = (np - p0) < pc ? (np - p0) : pc;
        run_loops(E, d_tau, d_el, p0, n, scratch, d_loops);
      }
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    };
    char key[160], gkey[320];
    std::snprintf(key, sizeof key, "sigma_set nt=%d nv=%d np=%d idir=%d src0=%d step=%d nsrc=%d chunk=%d", nt, nv, np,
                  idir, src0, src_step, nsrc, pc);
    std::snprintf(gkey, sizeof gkey, "%s mx=%d tau=%p el=%p raw=%p loops=%p", key, exact ? 2 : mixed ? 1 : 0, d_tau, d_el, d_raw,
                  d_loops);
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

int lc_engine_sigma_set(const void* d_tau, const void* d_el, int nt, int nv, int np, int idir, int src0,
                        int src_step, int nsrc, long long max_scratch_bytes, void* d_raw, void* d_loops) {
  return sigma_impl(false, d_tau, d_el, nt, nv, np, idir, src0, src_step, nsrc, max_scratch_bytes, d_raw, d_loops);
}

int lc_engine_sigma_set_mixed(const void* d_tau, const void* d_el, int nt, int nv, int np, int idir, int src0,
                              int src_step, int nsrc, long long max_scratch_bytes, void* d_raw, void* d_loops) {
  return sigma_impl(true, d_tau, d_el, nt, nv, np, idir, src0, src_step, nsrc, max_scratch_bytes, d_raw, d_loops);
}

}  // extern "C"
