// C-ABI of the contraction library: device state and the sink-row launchers.
//
// Each lc_*_row function is the D-space form of one reference launcher (lib/contract_kernels_*.cu)
// for one sink time tf and every source time ti at once. Instead of one chain of Nv x Nv
// products per spin-table row, a diagram is written as a trace of a product of D x D matrices
// (D = 4 Nv) in which every gamma matrix is a block rename + exact phase (spin.h):
//
//   rho(l,k,i)      = -+ Tr[ B_rp^dag (g5 L) A_lp (K g5) ],  A_lp = tau M_lp(ti), B_rp = M_rp(tf) tau
//   rhopi(iop)      = Tr[ (G_iop Z_ip) (g5 R tau^dag) ],    Z = M12_ip(tf) tau, R = M2 D(ti) M1,
//                                                           G_iop = g5 rho_{iop+1}
//   square          = Tr[ Q (g5 tau g5) R tau^dag ],         Q = M3 D(tf) M4, R = M2 D(ti) M1
//   doubleTriangle  = Tr[ Q (g5 tau g5) R' tau^dag ],        R' = M1 D(ti) M2
//   dTriangleHor    = Tr[ M3 tau M1 W M4 tau M2 tau^dag ],   W_ab = conj(tau_ba) (no transpose of
//                                                           the eigenvector indices, as the
//                                                           reference's element-wise conj)
//   star / fish     = Tr[Ma tau M1 tau^dag] * Tr[Mb tau M2 tau^dag]  (a,b) = (4,3) / (3,4)
//   loop            = Tr[ D(t) M(t) ]
//   sigma(lp,rp)    = Tr[ (X_lp(tf) tau)^dag g5 (tau X_rp(ti)) g5 ]
//   fish_hor        = Tr[ M2 D M1 D^dag ]
//   triangle(ip)    = Tr[ (g5 X_ip(tf) tau) (g5 R tau^dag) ]
//
// with tau = tau(tf,ti), M (x) 1 in spin. These identities are checked against a row-by-row
// restatement of the reference cores in the test suite. Shared sub-products are formed once:
// e.g. rho needs 10 elemental x perambulator products per (tf,ti) where the reference forms
// 1536 block products.
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "runtime.h"
#include "spin.h"

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

void check_sizes(int prec, int nt, int nv) {
  if (prec != LC_SINGLE && prec != LC_DOUBLE) fail(LC_ERR_ARG, "precision must be LC_SINGLE or LC_DOUBLE");
  if (nt < 1 || nt > 4096) fail(LC_ERR_ARG, "Nt=%d out of range", nt);
  if (nv < 1 || nv > 1024) fail(LC_ERR_ARG, "NeV=%d out of range", nv);
}

void check_ptr(const void* p, const char* name) {
  if (!p) fail(LC_ERR_ARG, "null pointer: %s", name);
}

void check_time(int t, int nt, const char* name) {
  if (t < 0 || t >= nt) fail(LC_ERR_ARG, "%s=%d out of range [0,%d)", name, t, nt);
}

// Bump allocator over the context's persistent scratch buffer.
struct Arena {
  char* base;
  size_t off = 0, cap;
  explicit Arena(size_t bytes) {
    ctx().scratch.ensure(bytes + 4096);
    base = static_cast<char*>(ctx().scratch.p);
    cap = ctx().scratch.cap;
  }
  char* take(size_t bytes) {
    off = (off + 255) & ~size_t(255);
    if (off + bytes > cap) fail(LC_ERR_NOMEM, "scratch arena overflow");
    char* p = base + off;
    off += bytes;
    return p;
  }
};

struct Shape {
  int nt, nv;
  long long m, blk;  // complex elements of one Nv x Nv block and of one D-space matrix
  size_t esz;        // bytes per complex element
  bool dbl;
  Shape(int prec, int nt_, int nv_) : nt(nt_), nv(nv_) {
    m = (long long)nv * nv;
    blk = 16 * m;
    dbl = prec == LC_DOUBLE;
    esz = dbl ? 16 : 8;
  }
  size_t dense_bytes(long long count) const { return (size_t)count * blk * esz; }
  const char* at(const void* p, long long elems) const {
    return static_cast<const char*>(p) + (size_t)elems * esz;
  }
  char* at(char* p, long long elems) const { return p + (size_t)elems * esz; }
};

inline LcView dense(const void* p, long long s0, long long s1 = 0, long long s2 = 0, int flags = 0) {
  return make_view(p, s0, s1, s2, LC_DENSE, flags);
}
inline LcView kron(const void* p, long long s0, long long s1 = 0, long long s2 = 0, int flags = 0) {
  return make_view(p, s0, s1, s2, LC_KRON, flags);
}

void finish(Program& p, const void* dev_out, void* host_out, size_t bytes) {
  hipStream_t s = ctx().stream;
  p.run(s);
  hip_check(hipMemcpyAsync(host_out, dev_out, bytes, hipMemcpyDeviceToHost, s), "copy results");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

// ------------------------------------------------------------------------------------- rho
void rho_row(const Shape& S, const void* row, const void* mom5, int tf, void* h_corr) {
  const int nt = S.nt;
  Arena ar(S.dense_bytes(10LL * nt) + nt * 48 * S.esz + 1024);
  char* A = ar.take(S.dense_bytes(5LL * nt));
  char* B = ar.take(S.dense_bytes(5LL * nt));
  char* out = ar.take(nt * 48 * S.esz);
  Program p(S.nv, S.dbl);
  const LcView tau = dense(row, S.blk);
  // A[lp][ti] = tau(tf,ti) (M_lp(ti) x 1);  B[rp][ti] = (M_rp(tf) x 1) tau(tf,ti)
  p.gemm(tau, kron(mom5, 5 * S.m, S.m), A, S.blk, nt * S.blk, 0, nt, 5, 1);
  p.gemm(kron(S.at(mom5, tf * 5 * S.m), 0, S.m), tau, B, S.blk, nt * S.blk, 0, nt, 5, 1);
  p.barrier();
  const Spin g5 = gamma(5);
  for (int l = 0; l < 4; l++)
    for (int k = 0; k < 4; k++)
      for (int i = 0; i < 3; i++) {
        int lp, rp;
        lp_rp(l, k, i, &lp, &rp);
        const Spin Lp = g5 * rhoj(l + 1, i, false);
        const Spin Kp = rhoj(k + 1, i, true) * g5;
        const LcView X = dense(S.at(B, (long long)rp * nt * S.blk), S.blk, 0, 0, LC_DAG);
        const LcView Y = right(left(dense(S.at(A, (long long)lp * nt * S.blk), S.blk), Lp), Kp);
        const double sgn = (l * 4 + k < 12) ? -1.0 : 1.0;
        p.trace(X, Y, S.at(out, (l * 4 + k) * 3 + i), 48, 0, 0, nt, 1, 1, sgn, 0.0);
      }
  finish(p, out, h_corr, nt * 48 * S.esz);
}

// ------------------------------------------------------------------------------------ rhopi
void rhopi_row(const Shape& S, const void* row, const void* diag, const void* p1, const void* p2,
               const void* p12x5, int tf, int idir, void* h_corr) {
  const int nt = S.nt;
  Arena ar(S.dense_bytes(6LL * nt) + nt * 4 * S.esz + 2048);
  char* T1 = ar.take(S.dense_bytes(nt));
  char* R = ar.take(S.dense_bytes(nt));
  char* Y = ar.take(S.dense_bytes(nt));
  char* Z = ar.take(S.dense_bytes(3LL * nt));
  char* out = ar.take(nt * 4 * S.esz);
  Program p(S.nv, S.dbl);
  const LcView tau = dense(row, S.blk), D = dense(diag, S.blk);
  const int ips[3] = {0, 1, 2 + idir};
  p.gemm(D, kron(p1, S.m), T1, S.blk, 0, 0, nt, 1, 1);
  for (int j = 0; j < 3; j++)
    p.gemm(kron(S.at(p12x5, (tf * 5 + ips[j]) * S.m), 0), tau, S.at(Z, j * nt * S.blk), S.blk, 0, 0, nt, 1, 1);
  p.barrier();
  p.gemm(kron(p2, S.m), dense(T1, S.blk), R, S.blk, 0, 0, nt, 1, 1);
  p.barrier();
  const Spin g5 = gamma(5);
  p.gemm(left(dense(R, S.blk), g5), dense(row, S.blk, 0, 0, LC_DAG), Y, S.blk, 0, 0, nt, 1, 1);
  p.barrier();
  for (int iop = 0; iop < 4; iop++) {
    const int ip = rhopi_ip(idir, iop);
    const int j = ip < 2 ? ip : 2;
    const Spin G = g5 * rhoj(iop + 1, idir, false);
    p.trace(left(dense(S.at(Z, j * nt * S.blk), S.blk), G), dense(Y, S.blk), S.at(out, iop), 4, 0, 0,
            nt, 1, 1);
  }
  finish(p, out, h_corr, nt * 4 * S.esz);
}

// ------------------------------------------------------------------------------------- pipi
template <typename T>
void pipi_row(const Shape& S, int iso, const void* row, const void* diag, const void* m1,
              const void* m2, const void* m3, const void* m4, int tf, void* h_corr) {
  const int nt = S.nt;
  const int NTR = 7;
  Arena ar(S.dense_bytes(15LL * nt + 2) + nt * NTR * S.esz + 8192);
  char* buf[15];
  for (int i = 0; i < 15; i++) buf[i] = ar.take(S.dense_bytes(nt));
  char* Q1 = ar.take(S.dense_bytes(1));
  char* Q = ar.take(S.dense_bytes(1));
  char* out = ar.take(nt * NTR * S.esz);
  char *T1 = buf[0], *T1p = buf[1], *L1 = buf[2], *X4 = buf[3], *S1 = buf[4], *S2 = buf[5],
       *R = buf[6], *Rp = buf[7], *L = buf[8], *Rr = buf[9], *U = buf[10], *V = buf[11],
       *Vp = buf[12], *LW = buf[13], *Zh = buf[14];
  const long long b = S.blk, m = S.m;
  Program p(S.nv, S.dbl);
  const LcView tau = dense(row, b), D = dense(diag, b), tau_dag = dense(row, b, 0, 0, LC_DAG);
  const bool hor = iso == LC_I0;
  // level 1
  p.gemm(dense(S.at(diag, tf * b), 0), kron(S.at(m4, tf * m), 0), Q1, 0, 0, 0, 1, 1, 1);
  p.gemm(D, kron(m1, m), T1, b, 0, 0, nt, 1, 1);
  p.gemm(D, kron(m2, m), T1p, b, 0, 0, nt, 1, 1);
  p.gemm(kron(S.at(m3, tf * m), 0), tau, L1, b, 0, 0, nt, 1, 1);
  p.gemm(kron(S.at(m4, tf * m), 0), tau, X4, b, 0, 0, nt, 1, 1);
  p.gemm(kron(m1, m), tau_dag, S1, b, 0, 0, nt, 1, 1);
  p.gemm(kron(m2, m), tau_dag, S2, b, 0, 0, nt, 1, 1);
  p.barrier();
  // level 2
  p.gemm(kron(S.at(m3, tf * m), 0), dense(Q1, 0), Q, 0, 0, 0, 1, 1, 1);
  p.gemm(kron(m2, m), dense(T1, b), R, b, 0, 0, nt, 1, 1);
  p.gemm(kron(m1, m), dense(T1p, b), Rp, b, 0, 0, nt, 1, 1);
  if (hor) {
    p.gemm(dense(L1, b), kron(m1, m), L, b, 0, 0, nt, 1, 1);
    p.gemm(dense(X4, b), kron(m2, m), Rr, b, 0, 0, nt, 1, 1);
  }
  p.barrier();
  // level 3: the dense x dense products
  const Spin g5 = gamma(5);
  p.gemm(dense(Q, 0), right(left(tau, g5), g5), U, b, 0, 0, nt, 1, 1);
  p.gemm(dense(R, b), tau_dag, V, b, 0, 0, nt, 1, 1);
  p.gemm(dense(Rp, b), tau_dag, Vp, b, 0, 0, nt, 1, 1);
  if (hor) {
    // W_ab = conj(tau_ba), no eigenvector transpose (lib/calculate_pipi_doubleTriangleHor_core.h:126), or
    // tau^dag under lc_set_dthor_dagger
    p.gemm(dense(L, b), dense(row, b, 0, 0, ctx().dthor_dagger ? LC_DAG : (LC_CONJ | LC_TS)), LW, b, 0, 0, nt, 1, 1);
    p.gemm(dense(Rr, b), tau_dag, Zh, b, 0, 0, nt, 1, 1);
  }
  p.barrier();
  // closing traces, out[ti][7] = {square, dT, dTHor, star1, star2, fish1, fish2}
  p.trace(dense(U, b), dense(V, b), S.at(out, 0), NTR, 0, 0, nt, 1, 1);
  p.trace(dense(U, b), dense(Vp, b), S.at(out, 1), NTR, 0, 0, nt, 1, 1);
  if (hor) p.trace(dense(LW, b), dense(Zh, b), S.at(out, 2), NTR, 0, 0, nt, 1, 1);
  p.trace(dense(X4, b), dense(S1, b), S.at(out, 3), NTR, 0, 0, nt, 1, 1);
  p.trace(dense(L1, b), dense(S2, b), S.at(out, 4), NTR, 0, 0, nt, 1, 1);
  p.trace(dense(L1, b), dense(S1, b), S.at(out, 5), NTR, 0, 0, nt, 1, 1);
  p.trace(dense(X4, b), dense(S2, b), S.at(out, 6), NTR, 0, 0, nt, 1, 1);
  std::vector<T> tr((size_t)nt * NTR * 2, T(0));
  finish(p, out, tr.data(), nt * NTR * S.esz);
  // host combination, in the compute type as lib/contract_kernels_pipi.cu:261-303 / :446-478
  T* c = static_cast<T*>(h_corr);
  for (int ti = 0; ti < nt; ti++) {
    const T* t = &tr[(size_t)ti * NTR * 2];
    auto mul = [](const T* x, const T* y, T* z) {
      z[0] = x[0] * y[0] - x[1] * y[1];
      z[1] = x[0] * y[1] + x[1] * y[0];
    };
    T st[2], fi[2];
    mul(t + 6, t + 8, st);
    mul(t + 10, t + 12, fi);
    T* o = c + (size_t)ti * 10;
    if (iso == LC_I1) {
      o[0] = T(-2.) * t[0]; o[1] = T(-2.) * t[1];
      o[2] = T(2.) * t[2];  o[3] = T(2.) * t[3];
      o[4] = -st[0];        o[5] = -st[1];
      o[6] = fi[0];         o[7] = fi[1];
      o[8] = 0;             o[9] = 0;
    } else {
      o[0] = T(-1. / 3.) * t[0]; o[1] = T(-1. / 3.) * t[1];
      o[2] = -T(1. / 3.) * t[2]; o[3] = -T(1. / 3.) * t[3];
      o[4] = st[0];              o[5] = st[1];
      o[6] = fi[0];              o[7] = fi[1];
      o[8] = -T(5. / 3.) * t[4]; o[9] = -T(5. / 3.) * t[5];
    }
  }
}

// ------------------------------------------------------------------------------------- loop
void loop_unity(const Shape& S, const void* diag, const void* mom, void* h_loop) {
  Arena ar(S.nt * S.esz + 1024);
  char* out = ar.take(S.nt * S.esz);
  Program p(S.nv, S.dbl);
  p.trace(dense(diag, S.blk), kron(mom, S.m), out, 1, 0, 0, S.nt, 1, 1);
  finish(p, out, h_loop, S.nt * S.esz);
}

// ------------------------------------------------------------------------------------ sigma
void sigma_row(const Shape& S, const void* row, const void* x2, int tf, void* h_corr) {
  const int nt = S.nt;
  const long long b = S.blk, m = S.m;
  Arena ar(S.dense_bytes(4LL * nt) + nt * 4 * S.esz + 1024);
  char* B = ar.take(S.dense_bytes(2LL * nt));
  char* A = ar.take(S.dense_bytes(2LL * nt));
  char* out = ar.take(nt * 4 * S.esz);
  Program p(S.nv, S.dbl);
  const LcView tau = dense(row, b);
  p.gemm(kron(S.at(x2, tf * 2 * m), 0, m), tau, B, b, nt * b, 0, nt, 2, 1);
  p.gemm(tau, kron(x2, 2 * m, m), A, b, nt * b, 0, nt, 2, 1);
  p.barrier();
  const Spin g5 = gamma(5);
  for (int lp = 0; lp < 2; lp++)
    for (int rp = 0; rp < 2; rp++)
      p.trace(dense(S.at(B, lp * nt * b), b, 0, 0, LC_DAG),
              right(left(dense(S.at(A, rp * nt * b), b), g5), g5), S.at(out, lp * 2 + rp), 4, 0, 0,
              nt, 1, 1);
  finish(p, out, h_corr, nt * 4 * S.esz);
}

// ---------------------------------------------------------------------------------- sigmapi
void fish_hor(const Shape& S, const void* diag, const void* p1, const void* p2, void* h) {
  const int nt = S.nt;
  const long long b = S.blk, m = S.m;
  Arena ar(S.dense_bytes(2LL * nt) + nt * S.esz + 1024);
  char* T = ar.take(S.dense_bytes(nt));
  char* T2 = ar.take(S.dense_bytes(nt));
  char* out = ar.take(nt * S.esz);
  Program p(S.nv, S.dbl);
  p.gemm(kron(p2, m), dense(diag, b), T, b, 0, 0, nt, 1, 1);
  p.barrier();
  p.gemm(dense(T, b), kron(p1, m), T2, b, 0, 0, nt, 1, 1);
  p.barrier();
  p.trace(dense(T2, b), dense(diag, b, 0, 0, LC_DAG), out, 1, 0, 0, nt, 1, 1);
  finish(p, out, h, nt * S.esz);
}

void triangle_row(const Shape& S, const void* row, const void* diag, const void* p1,
                  const void* p2, const void* x2, int tf, void* h_corr) {
  const int nt = S.nt;
  const long long b = S.blk, m = S.m;
  Arena ar(S.dense_bytes(5LL * nt) + nt * 2 * S.esz + 2048);
  char* T1 = ar.take(S.dense_bytes(nt));
  char* R = ar.take(S.dense_bytes(nt));
  char* Y = ar.take(S.dense_bytes(nt));
  char* Z = ar.take(S.dense_bytes(2LL * nt));
  char* out = ar.take(nt * 2 * S.esz);
  Program p(S.nv, S.dbl);
  const LcView tau = dense(row, b);
  p.gemm(dense(diag, b), kron(p1, m), T1, b, 0, 0, nt, 1, 1);
  p.gemm(kron(S.at(x2, tf * 2 * m), 0, m), tau, Z, b, nt * b, 0, nt, 2, 1);
  p.barrier();
  p.gemm(kron(p2, m), dense(T1, b), R, b, 0, 0, nt, 1, 1);
  p.barrier();
  const Spin g5 = gamma(5);
  p.gemm(left(dense(R, b), g5), dense(row, b, 0, 0, LC_DAG), Y, b, 0, 0, nt, 1, 1);
  p.barrier();
  for (int ip = 0; ip < 2; ip++)
    p.trace(left(dense(S.at(Z, ip * nt * b), b), g5), dense(Y, b), S.at(out, ip), 2, 0, 0, nt, 1, 1);
  finish(p, out, h_corr, nt * 2 * S.esz);
}

}  // namespace

// =========================================================================== C-ABI entry points
extern "C" hipError_t lc_launch_phi(const void* V, const void* ph, int nt, int nv, long long nsite, int np, int nks,
                                    void* part, void* phi, hipStream_t s);
extern "C" long long lc_phi_partial_bytes(int nt, int nv, int np, int nks);
extern "C" hipError_t lc_launch_phi_bilinear(const void* X, const void* Y, const void* ph, int nt, int nv,
                                             long long nsite, int np, int nks, void* part, double alpha,
                                             int accumulate, void* phi, long long pstride, hipStream_t s);
extern "C" hipError_t lc_launch_covderiv(const void* U, const void* V, int nt, int nv, int nx, int ny, int nz,
                                         double c, void* dV, hipStream_t s);
extern "C" hipError_t lc_noalloc_setup(int device);
extern "C" int lc_noalloc_loop_chunks(int nv);
extern "C" long long lc_noalloc_capacity(void);
extern "C" hipError_t lc_launch_loop_unity_noalloc(int dbl, const void* diag, const void* mom, int nt, int nv,
                                                   int device, void* out, hipStream_t s);

extern "C" {

const char* lc_last_error(void) { return last_error().c_str(); }

int lc_is_init(void) { return ctx().on ? 1 : 0; }

int lc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int lc_init(int device) {
  return guarded([&] {
    Context& c = ctx();
    if (c.on) fail(LC_ERR_ALREADY, "initContract can be called once only");
    const int n = lc_device_count();
    if (n == 0) fail(LC_ERR_NODEV, "no GPU devices found");
    if (device < 0 || device >= n) fail(LC_ERR_NODEV, "device id %d is not allowed (%d devices)", device, n);
    hip_check(hipSetDevice(device), "hipSetDevice");
    hip_check(hipStreamCreate(&c.stream), "hipStreamCreate");
    hip_check(lc_noalloc_setup(device), "address of the allocation-free scratch array");
    c.device = device;
    c.on = true;
  });
}

int lc_finalize(void) {
  return guarded([&] {
    Context& c = ctx();
    if (!c.on) fail(LC_ERR_NOT_INIT, "cannot finalize: library is not initialised");
    (void)hipStreamSynchronize(c.stream);
    drop_graphs();
    c.graphs = false;
    c.groups.release();
    c.scratch.release();
    c.out.release();
    c.partials.release();
    for (hipEvent_t e : c.ev_pool) (void)hipEventDestroy(e);
    c.ev_pool.clear();
    c.tuned.clear();
    c.last_tune = TuneRecord();
    (void)hipStreamDestroy(c.stream);
    c.stream = nullptr;
    c.on = false;
    c.device = -1;
  });
}

int lc_engine_graph(int enable) {
  return guarded([&] {
    if (enable != 0 && enable != 1) fail(LC_ERR_ARG, "lc_engine_graph: enable must be 0 or 1 (got %d)", enable);
    Context& c = ctx();
    if (!enable && c.on) {
      (void)hipStreamSynchronize(c.stream);
      drop_graphs();
    }
    c.graphs = enable == 1;
  });
}

int lc_engine_graph_report(int* state) {
  return guarded([&] {
    if (!state) fail(LC_ERR_ARG, "lc_engine_graph_report: null output");
    *state = ctx().last_graph;
  });
}

int lc_set_dthor_dagger(int on) {
  return guarded([&] {
    if (on != 0 && on != 1) fail(LC_ERR_ARG, "lc_set_dthor_dagger: on must be 0 or 1 (got %d)", on);
    ctx().dthor_dagger = on == 1;
  });
}

int lc_set_tile_params(int tiles_per_wg) {
  return guarded([&] {
    if (tiles_per_wg != 0 && tiles_per_wg != 1 && tiles_per_wg != 2 && tiles_per_wg != 4)
      fail(LC_ERR_ARG, "lc_set_tile_params: tiles per workgroup must be 0, 1, 2 or 4 (got %d)", tiles_per_wg);
    ctx().tile_bpw = tiles_per_wg;
  });
}

int lc_autotune(int enable) {
  return guarded([&] {
    if (enable < -1 || enable > 1) fail(LC_ERR_ARG, "lc_autotune: enable must be -1, 0 or 1 (got %d)", enable);
    Context& c = ctx();
    if (enable < 0) c.tuned.clear();
    c.autotune = enable == 1;
  });
}

int lc_autotune_report(double* out5) {
  return guarded([&] {
    if (!out5) fail(LC_ERR_ARG, "lc_autotune_report: null output");
    const TuneRecord& r = ctx().last_tune;
    out5[0] = r.bpw;
    out5[1] = r.how;
    for (int i = 0; i < 3; i++) out5[2 + i] = r.ms[i];
  });
}

int lc_rho_row(int prec, const void* d_row, const void* d_mom5, int tf, int nt, int nv,
               void* h_corr) {
  return guarded([&] {
    require_init();
    check_sizes(prec, nt, nv);
    check_time(tf, nt, "tf");
    check_ptr(d_row, "d_row"); check_ptr(d_mom5, "d_mom5"); check_ptr(h_corr, "h_corr");
    rho_row(Shape(prec, nt, nv), d_row, d_mom5, tf, h_corr);
  });
}

int lc_rhopi_row(int prec, const void* d_row, const void* d_diag, const void* d_p1,
                 const void* d_p2, const void* d_p12x5, int tf, int nt, int nv, int idir,
                 void* h_corr) {
  return guarded([&] {
    require_init();
    check_sizes(prec, nt, nv);
    check_time(tf, nt, "tf");
    if (idir < 0 || idir > 2) fail(LC_ERR_ARG, "idir=%d out of range [0,3)", idir);
    check_ptr(d_row, "d_row"); check_ptr(d_diag, "d_diag"); check_ptr(d_p1, "d_p1");
    check_ptr(d_p2, "d_p2"); check_ptr(d_p12x5, "d_p12x5"); check_ptr(h_corr, "h_corr");
    rhopi_row(Shape(prec, nt, nv), d_row, d_diag, d_p1, d_p2, d_p12x5, tf, idir, h_corr);
  });
}

int lc_pipi_row(int prec, int iso, const void* d_row, const void* d_diag, const void* d_m1,
                const void* d_m2, const void* d_m3, const void* d_m4, int tf, int nt, int nv,
                void* h_corr) {
  return guarded([&] {
    require_init();
    check_sizes(prec, nt, nv);
    check_time(tf, nt, "tf");
    if (iso != LC_I0 && iso != LC_I1) fail(LC_ERR_ARG, "This Isospin is not implemented");
    check_ptr(d_row, "d_row"); check_ptr(d_diag, "d_diag"); check_ptr(d_m1, "d_m1");
    check_ptr(d_m2, "d_m2"); check_ptr(d_m3, "d_m3"); check_ptr(d_m4, "d_m4");
    check_ptr(h_corr, "h_corr");
    const Shape S(prec, nt, nv);
    if (S.dbl) pipi_row<double>(S, iso, d_row, d_diag, d_m1, d_m2, d_m3, d_m4, tf, h_corr);
    else pipi_row<float>(S, iso, d_row, d_diag, d_m1, d_m2, d_m3, d_m4, tf, h_corr);
  });
}

int lc_loop_unity(int prec, const void* d_diag, const void* d_mom, int nt, int nv, void* h_loop) {
  return guarded([&] {
    require_init();
    check_sizes(prec, nt, nv);
    check_ptr(d_diag, "d_diag"); check_ptr(d_mom, "d_mom"); check_ptr(h_loop, "h_loop");
    loop_unity(Shape(prec, nt, nv), d_diag, d_mom, h_loop);
  });
}

int lc_loop_unity_noalloc(int prec, const void* d_diag, const void* d_mom, int nt, int nv, void* d_loop) {
  return guarded([&] {
    require_init();
    check_sizes(prec, nt, nv);
    check_ptr(d_diag, "d_diag"); check_ptr(d_mom, "d_mom"); check_ptr(d_loop, "d_loop");
    const long long parts = (long long)nt * lc_noalloc_loop_chunks(nv);
    if (parts > lc_noalloc_capacity())
      fail(LC_ERR_ARG, "lc_loop_unity_noalloc: Nt=%d, NeV=%d need %lld partial sums, above the fixed %lld", nt, nv,
           parts, lc_noalloc_capacity());
    hip_check(lc_launch_loop_unity_noalloc(prec == LC_DOUBLE, d_diag, d_mom, nt, nv, ctx().device, d_loop,
                                           ctx().stream),
              "allocation-free loop_unity launch");
  });
}

int lc_sigma_row(int prec, const void* d_row, const void* d_x2, int tf, int nt, int nv,
                 void* h_corr) {
  return guarded([&] {
    require_init();
    check_sizes(prec, nt, nv);
    check_time(tf, nt, "tf");
    check_ptr(d_row, "d_row"); check_ptr(d_x2, "d_x2"); check_ptr(h_corr, "h_corr");
    sigma_row(Shape(prec, nt, nv), d_row, d_x2, tf, h_corr);
  });
}

int lc_sigmapi_fish_hor(int prec, const void* d_diag, const void* d_p1, const void* d_p2,
                        int nt, int nv, void* h_corr) {
  return guarded([&] {
    require_init();
    check_sizes(prec, nt, nv);
    check_ptr(d_diag, "d_diag"); check_ptr(d_p1, "d_p1"); check_ptr(d_p2, "d_p2");
    check_ptr(h_corr, "h_corr");
    fish_hor(Shape(prec, nt, nv), d_diag, d_p1, d_p2, h_corr);
  });
}

int lc_sigmapi_triangle_row(int prec, const void* d_row, const void* d_diag, const void* d_p1,
                            const void* d_p2, const void* d_x2, int tf, int nt, int nv,
                            void* h_corr) {
  return guarded([&] {
    require_init();
    check_sizes(prec, nt, nv);
    check_time(tf, nt, "tf");
    check_ptr(d_row, "d_row"); check_ptr(d_diag, "d_diag"); check_ptr(d_p1, "d_p1");
    check_ptr(d_p2, "d_p2"); check_ptr(d_x2, "d_x2"); check_ptr(h_corr, "h_corr");
    triangle_row(Shape(prec, nt, nv), d_row, d_diag, d_p1, d_p2, d_x2, tf, h_corr);
  });
}

int lc_dev_alloc(long long bytes, void** out) {
  return guarded([&] {
    require_init();
    if (bytes < 0 || !out) fail(LC_ERR_ARG, "bad allocation request");
    *out = nullptr;
    if (hipMalloc(out, bytes > 0 ? (size_t)bytes : 1) != hipSuccess) {
      (void)hipGetLastError();
      *out = nullptr;
      fail(LC_ERR_NOMEM, "hipMalloc of %lld bytes failed", bytes);
    }
  });
}

int lc_dev_free(void* p) {
  return guarded([&] {
    if (p) hip_check(hipFree(p), "hipFree");
  });
}

int lc_memcpy_h2d(void* d, const void* h, long long bytes) {
  return guarded([&] {
    require_init();
    if (bytes < 0 || (bytes > 0 && (!d || !h))) fail(LC_ERR_ARG, "bad copy");
    hip_check(hipMemcpyAsync(d, h, (size_t)bytes, hipMemcpyHostToDevice, ctx().stream), "copy h2d");
    hip_check(hipStreamSynchronize(ctx().stream), "sync");
  });
}

int lc_memcpy_d2h(void* h, const void* d, long long bytes) {
  return guarded([&] {
    require_init();
    if (bytes < 0 || (bytes > 0 && (!d || !h))) fail(LC_ERR_ARG, "bad copy");
    hip_check(hipMemcpyAsync(h, d, (size_t)bytes, hipMemcpyDeviceToHost, ctx().stream), "copy d2h");
    hip_check(hipStreamSynchronize(ctx().stream), "sync");
  });
}

int lc_memcpy_d2d(void* d_dst, const void* d_src, long long bytes) {
  return guarded([&] {
    require_init();
    if (bytes < 0 || (bytes > 0 && (!d_dst || !d_src))) fail(LC_ERR_ARG, "bad copy");
    hip_check(hipMemcpyAsync(d_dst, d_src, (size_t)bytes, hipMemcpyDeviceToDevice, ctx().stream), "copy d2d");
    hip_check(hipStreamSynchronize(ctx().stream), "sync");
  });
}

int lc_synchronize(void) {
  return guarded([&] {
    require_init();
    hip_check(hipStreamSynchronize(ctx().stream), "sync");
  });
}

// K slices of the elemental GEMMs: enough to give every CU several workgroups, at least 256
// entries per slice; the partial-tile buffer sized for them.
static int phi_slices(int nt, int nv, int np, long long nsite) {
  const long long tiles = (nv + 63) / 64;
  const long long tiles_total = (long long)np * nt * tiles * tiles;
  long long nks = (2048 + tiles_total - 1) / tiles_total;
  if (nks > nsite / 256) nks = nsite / 256;
  if (nks < 1) nks = 1;
  ctx().partials.ensure((size_t)lc_phi_partial_bytes(nt, nv, np, (int)nks));
  return (int)nks;
}

int lc_build_elementals(const void* d_evec, const void* d_phase, int nt, int nv, long long nsite, int np,
                        void* d_phi) {
  return guarded([&] {
    require_init();
    if (nt < 1 || nv < 1 || nv > 1024 || np < 1 || nsite < 1)
      fail(LC_ERR_ARG, "bad sizes nt=%d nv=%d np=%d nsite=%lld", nt, nv, np, nsite);
    check_ptr(d_evec, "d_evec"); check_ptr(d_phase, "d_phase"); check_ptr(d_phi, "d_phi");
    const int nks = phi_slices(nt, nv, np, nsite);
    hipStream_t s = ctx().stream;
    hip_check(lc_launch_phi(d_evec, d_phase, nt, nv, nsite, np, nks, ctx().partials.p, d_phi, s), "phi launch");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  });
}

int lc_covariant_derivative(const void* d_links, const void* d_evec, int nt, int nv, int nx, int ny, int nz,
                            void* d_dvec) {
  return lc_covariant_derivative_scaled(d_links, d_evec, nt, nv, nx, ny, nz, 0.5, d_dvec);
}

int lc_covariant_derivative_scaled(const void* d_links, const void* d_evec, int nt, int nv, int nx, int ny, int nz,
                                   double c, void* d_dvec) {
  return guarded([&] {
    require_init();
    if (nt < 1 || nv < 1 || nv > 1024 || nx < 1 || ny < 1 || nz < 1 ||
        3LL * nt * (((long long)nx * ny * nz + 255) / 256) > 0xffffffffLL)
      fail(LC_ERR_ARG, "bad sizes nt=%d nv=%d lattice %dx%dx%d", nt, nv, nx, ny, nz);
    check_ptr(d_links, "d_links"); check_ptr(d_evec, "d_evec"); check_ptr(d_dvec, "d_dvec");
    hipStream_t s = ctx().stream;
    hip_check(lc_launch_covderiv(d_links, d_evec, nt, nv, nx, ny, nz, c, d_dvec, s), "covariant derivative launch");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  });
}

int lc_build_elementals_types(const void* d_evec, const void* d_dvec, const void* d_phase, int nt, int nv,
                              long long nsite, int np, void* d_mom5) {
  return guarded([&] {
    require_init();
    if (nt < 1 || nv < 1 || nv > 1024 || np < 1 || nsite < 1)
      fail(LC_ERR_ARG, "bad sizes nt=%d nv=%d np=%d nsite=%lld", nt, nv, np, nsite);
    check_ptr(d_evec, "d_evec"); check_ptr(d_dvec, "d_dvec"); check_ptr(d_phase, "d_phase");
    check_ptr(d_mom5, "d_mom5");
    const int nks = phi_slices(nt, nv, np, nsite);
    hipStream_t s = ctx().stream;
    void* part = ctx().partials.p;
    const size_t vec_bytes = (size_t)nt * nv * nsite * 16;  // one vector set [nt][nv][nsite]
    const size_t type_bytes = (size_t)nt * nv * nv * 16;    // one type [nt][nv][nv]
    const long long pstride = 5LL * nt;                     // matrices per momentum in d_mom5
    auto dv = [&](int j) { return static_cast<const char*>(d_dvec) + j * vec_bytes; };
    auto type = [&](int k) { return static_cast<char*>(d_mom5) + k * type_bytes; };
    auto bil = [&](const void* X, const void* Y, double alpha, int acc, int k) {
      hip_check(lc_launch_phi_bilinear(X, Y, d_phase, nt, nv, nsite, np, nks, part, alpha, acc, type(k), pstride, s),
                "elemental launch");
    };
    // type 0: V^dag ph V; type 1: sum_j (nabla_j V)^dag ph (nabla_j V);
    // type 2 + j: V^dag ph (nabla_j V) - (nabla_j V)^dag ph V
    bil(d_evec, d_evec, 1.0, 0, 0);
    for (int j = 0; j < 3; j++) bil(dv(j), dv(j), 1.0, j > 0, 1);
    for (int j = 0; j < 3; j++) {
      bil(d_evec, dv(j), 1.0, 0, 2 + j);
      bil(dv(j), d_evec, -1.0, 1, 2 + j);
    }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  });
}

int lc_selftest_mfma_bf16(const float* h_a, const float* h_b, float* h_d) {
  return guarded([&] {
    require_init();
    check_ptr(h_a, "h_a"); check_ptr(h_b, "h_b"); check_ptr(h_d, "h_d");
    ctx().out.ensure(16384);
    char* d = static_cast<char*>(ctx().out.p);
    hipStream_t s = ctx().stream;
    hip_check(hipMemcpyAsync(d, h_a, 512 * 4, hipMemcpyHostToDevice, s), "copy a");
    hip_check(hipMemcpyAsync(d + 4096, h_b, 512 * 4, hipMemcpyHostToDevice, s), "copy b");
    hip_check(lc_launch_mfma_bf16_selftest((const float*)d, (const float*)(d + 4096), (float*)(d + 8192), s),
              "selftest launch");
    hip_check(hipMemcpyAsync(h_d, d + 8192, 512 * 4, hipMemcpyDeviceToHost, s), "copy d");
    hip_check(hipStreamSynchronize(s), "sync");
  });
}

int lc_selftest_mfma(int prec, const void* h_a, const void* h_b, void* h_d) {
  return guarded([&] {
    require_init();
    if (prec != LC_SINGLE && prec != LC_DOUBLE) fail(LC_ERR_ARG, "bad precision");
    check_ptr(h_a, "h_a"); check_ptr(h_b, "h_b"); check_ptr(h_d, "h_d");
    const size_t es = prec == LC_DOUBLE ? 8 : 4;
    ctx().out.ensure(8192);
    char* d = static_cast<char*>(ctx().out.p);
    hipStream_t s = ctx().stream;
    hip_check(hipMemcpyAsync(d, h_a, 64 * es, hipMemcpyHostToDevice, s), "copy a");
    hip_check(hipMemcpyAsync(d + 1024, h_b, 64 * es, hipMemcpyHostToDevice, s), "copy b");
    hip_check(lc_launch_mfma_selftest(prec == LC_DOUBLE, d, d + 1024, d + 2048, s), "selftest launch");
    hip_check(hipMemcpyAsync(h_d, d + 2048, 512 * es, hipMemcpyDeviceToHost, s), "copy d");
    hip_check(hipStreamSynchronize(s), "sync");
  });
}

}  // extern "C"
