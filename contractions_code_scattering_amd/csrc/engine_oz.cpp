// Exact form of the batched engines: the rho + pi-pi chains of engine.cpp and the sigma / sigma-pi
// / rho-pi chains of engine_sigma.cpp, with every product on the int8 matrix cores through digit
// storage and exact integer accumulation (gemm_oz.hip), FP64 outputs, closing factors and traces.
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

}  // namespace

extern "C" {

int lc_selftest_oz_gemm(int nv, int ka, int kb, int fa, int fb, const int* gam4, const void* h_a, const void* h_b,
                        void* h_c, void* h_cd) {
  return guarded([&] {
    require_init();
    if (nv < 1 || nv > 128) fail(LC_ERR_ARG, "nv=%d out of range", nv);
    if (ka && kb) fail(LC_ERR_ARG, "elemental x elemental product");
    if (!h_a || !h_b || !h_c || !gam4) fail(LC_ERR_ARG, "null pointer");
    if ((fa | fb) & ~(LC_CONJ | LC_TV | LC_TS)) fail(LC_ERR_ARG, "bad view flags");
    const int P = oz_pad(nv);
    const long long nv2 = (long long)nv * nv, P2 = (long long)P * P;
    const int nba = ka ? 1 : 16, nbb = kb ? 1 : 16;
    hipStream_t s = ctx().stream;
    Scoped xa(nba * nv2 * 16), xb(nbb * nv2 * 16);
    Scoped ya(2 * nba * 8 * P2), yb(2 * nbb * 8 * P2), ea(2 * P * 4), eb(2 * P * 4);
    Scoped c(16 * nv2 * 16), cd(16 * 8 * P2), ce(4 * P * 4), dec(16 * nv2 * 16);
    hip_check(hipMemcpyAsync(xa.p, h_a, nba * nv2 * 16, hipMemcpyHostToDevice, s), "copy a");
    hip_check(hipMemcpyAsync(xb.p, h_b, nbb * nv2 * 16, hipMemcpyHostToDevice, s), "copy b");
    int* exa = static_cast<int*>(ea.p);
    int* exb = static_cast<int*>(eb.p);
    hip_check(lc_launch_oz_convert(xa.p, 0, 1, nv, P, nba, 0, exa, exa + P, ya.p, ya.c() + nba * 8 * P2, s),
              "convert a");
    hip_check(lc_launch_oz_convert(xb.p, 0, 1, nv, P, nbb, 0, exb, exb + P, yb.p, yb.c() + nbb * 8 * P2, s),
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
    const bool digits = h_cd && 4 * P <= 256;
    DigitOut d;
    d.d = cd.p;
    d.e = static_cast<int*>(ce.p);
    Program pg(nv, Program::OZ);
    pg.gemm_oz(A, B, c.p, 0, 0, 0, digits ? &d : nullptr, 1, 1, 1);
    pg.run(s);
    hip_check(hipMemcpyAsync(h_c, c.p, 16 * nv2 * 16, hipMemcpyDeviceToHost, s), "copy c");
    if (digits) {
      hip_check(lc_launch_oz_decode(cd.p, d.e, nv, P, 16, 1, 1, dec.p, s), "decode");
      hip_check(hipMemcpyAsync(h_cd, dec.p, 16 * nv2 * 16, hipMemcpyDeviceToHost, s), "copy cd");
    }
    hip_check(hipStreamSynchronize(s), "sync");
  });
}

}  // extern "C"
