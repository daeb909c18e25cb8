// Drop-in public API (reference lib/interface.cpp): one call per correlator family, with the
// elemental re-layouts and casts of the reference, timing print, and the C wrappers that let
// FFI callers (ctypes, cgo, JNI) run the same flows with error codes instead of exit().
#include <interface.h>
#include <sys/time.h>

#include <cmath>
#include <string>
#include <vector>

#include "lapH_contract.h"
#include "runtime.h"

using namespace contract;

namespace lcdrop {
extern bool g_throw;
}

namespace {

double now() {
  struct timeval t;
  gettimeofday(&t, NULL);
  return t.tv_sec + t.tv_usec * 1e-6;
}

template <typename Float>
void report(double sec, double flo) {
  printf("Elapsed time is %lf sec, kernel performance is %lf GFlops in %s precision\n", sec,
         flo / sec / 1.e9, sizeof(Float) == sizeof(float) ? "single" : "double");
}

// [k][t][v][v] (driver layout, FP64) -> [t][nk][v][v] of Float, taking types ks[0..nk).
template <typename Float>
void gather_types(Float* dst, const double* src, const int* ks, int nk, int nt, int nv) {
  const size_t m = (size_t)nv * nv * 2;
  for (int t = 0; t < nt; t++)
    for (int j = 0; j < nk; j++) {
      const double* s = src + ((size_t)ks[j] * nt + t) * m;
      Float* d = dst + ((size_t)t * nk + j) * m;
      for (size_t e = 0; e < m; e++) d[e] = (Float)s[e];
    }
}

template <typename Float>
void cast_copy(Float* dst, const double* src, size_t n) {
  for (size_t e = 0; e < n; e++) dst[e] = (Float)src[e];
}

// Work definitions printed by the reference, generalised from NeV = 100 (lib/interface.cpp:33-39,
// 93-98, 150-184, 228-233, 262-269, 356-368): g = 8 NeV^3 per block product, tr = 8 NeV^2.
double g_of(int nv) { return 8.0 * nv * (double)nv * nv; }
double tr_of(int nv) { return 8.0 * nv * (double)nv; }

template <typename Float>
void rho_impl(contractInfo info, Float* prop, double* mom, std::string outfile) {
  const int nt = info.Nt, nv = info.NeV;
  LapH_prop<Float> propH(DEVICE, LAPH_PROP_CLASS, prop);
  Mom_matrices<Float> momM(BOTH, MOM_MATRICES_CLASS);
  const int ks[5] = {0, 1, 2, 3, 4};
  gather_types(momM.H_elem(), mom, ks, 5, nt, nv);
  Contract_rho_rho<Float> rho(HOST, CONTRACT_RHO_RHO_CLASS, outfile);
  const double flo = (2 * g_of(nv) + tr_of(nv)) * 768.0 * nt * nt;
  const double t0 = now();
  rho.contractRho(momM, propH);
  report<Float>(now() - t0, flo);
}

template <typename Float>
void rhopi_impl(int idir, contractInfo info, Float* prop, double* p1, double* p2, double* p12,
                std::string outfile) {
  const int nt = info.Nt, nv = info.NeV;
  LapH_prop<Float> propH(DEVICE_EXTRA, LAPH_PROP_CLASS, prop);
  Mom_matrices<Float> m12(BOTH, MOM_MATRICES_CLASS);
  const int ks[5] = {0, 1, 2, 3, 4};
  gather_types(m12.H_elem(), p12, ks, 5, nt, nv);
  Mom_matrix<Float> m1(BOTH, MOM_MATRIX_CLASS), m2(BOTH, MOM_MATRIX_CLASS);
  cast_copy(m1.H_elem(), p1, (size_t)nt * nv * nv * 2);
  cast_copy(m2.H_elem(), p2, (size_t)nt * nv * nv * 2);
  Contract_rho_pipi<Float> obj(HOST, CONTRACT_RHO_PION_PION_CLASS, outfile);
  const double flo = (4 * g_of(nv) + tr_of(nv)) * 256.0 * nt * nt;
  const double t0 = now();
  obj.contractRhoPi(idir, m1, m2, m12, propH);
  report<Float>(now() - t0, flo);
}

template <typename Float>
void pipi_impl(contractInfo info, Float* prop, double* p1, double* p2, double* p3, double* p4,
               std::string outfile, ISOSPIN iso) {
  const int nt = info.Nt, nv = info.NeV;
  if (iso != I0 && iso != I1) throw lc::Error{LC_ERR_ARG, "This Isospin is not implemeted"};
  LapH_prop<Float> propH(DEVICE_EXTRA, LAPH_PROP_CLASS, prop);
  Mom_matrix<Float> m1(BOTH, MOM_MATRIX_CLASS), m2(BOTH, MOM_MATRIX_CLASS), m3(BOTH, MOM_MATRIX_CLASS),
      m4(BOTH, MOM_MATRIX_CLASS);
  const size_t n = (size_t)nt * nv * nv * 2;
  cast_copy(m1.H_elem(), p1, n);
  cast_copy(m2.H_elem(), p2, n);
  cast_copy(m3.H_elem(), p3, n);
  cast_copy(m4.H_elem(), p4, n);
  Contract_pipi_pipi<Float> obj(HOST, CONTRACT_PION_PION_PION_PION_CLASS, outfile);
  const double six = (6 * g_of(nv) + tr_of(nv)) * 256.0, two = (2 * g_of(nv) + tr_of(nv)) * 16.0 * 2;
  const double flo = ((iso == I0 ? 3 : 2) * six + 2 * two) * nt * (double)nt;
  const double t0 = now();
  obj.contractPiPi(m1, m2, m3, m4, propH, iso);
  report<Float>(now() - t0, flo);
}

template <typename Float>
void sigma_impl(contractInfo info, Float* prop, double* mom, std::string outfile) {
  const int nt = info.Nt, nv = info.NeV;
  LapH_prop<Float> propH(DEVICE_EXTRA, LAPH_PROP_CLASS, prop);
  Mom_matrix<Float> m(BOTH, MOM_MATRIX_CLASS);
  Contract_loop_unity<Float> loop(HOST, CONTRACT_LOOP_UNITY, outfile);
  const size_t n = (size_t)nt * nv * nv * 2;
  for (int i = 0; i < 4; i++) {  // e^{+ip}, e^{-ip}, and their smeared variants
    cast_copy(m.H_elem(), mom + i * n, n);
    const double t0 = now();
    loop.contractLoopUnity(m, propH);
    report<Float>(now() - t0, 4 * tr_of(nv) * nt);
  }
  Mom_matrices_x2<Float> x2(BOTH, MOM_MATRICES_CLASS_x2);
  const int ks[2] = {1, 3};  // e^{-ip} and e^{-ip} smeared (lib/interface.cpp:253-259)
  gather_types(x2.H_elem(), mom, ks, 2, nt, nv);
  Contract_sigma_sigma<Float> obj(HOST, CONTRACT_SIGMA_SIGMA_CLASS, outfile);
  const double flo = (2 * g_of(nv) + tr_of(nv)) * 64.0 * nt * nt;
  const double t0 = now();
  obj.contractSigma(x2, propH);
  report<Float>(now() - t0, flo);
}

template <typename Float>
void sigmapi_impl(contractInfo info, Float* prop, double* p1, double* p2, double* p12, std::string outfile) {
  const int nt = info.Nt, nv = info.NeV;
  LapH_prop<Float> propH(DEVICE_EXTRA, LAPH_PROP_CLASS, prop);
  Mom_matrix<Float> m1(BOTH, MOM_MATRIX_CLASS), m2(BOTH, MOM_MATRIX_CLASS);
  Mom_matrices_x2<Float> x2(BOTH, MOM_MATRICES_CLASS_x2);
  Contract_loop_unity<Float> loop(HOST, CONTRACT_LOOP_UNITY, outfile);
  Contract_sigma_pipi<Float> obj(HOST, CONTRACT_SIGMA_PION_PION_CLASS, outfile);
  const size_t n = (size_t)nt * nv * nv * 2;
  for (int i = 0; i < 2; i++) {
    cast_copy(m1.H_elem(), p12 + i * n, n);
    const double t0 = now();
    loop.contractLoopUnity(m1, propH);
    report<Float>(now() - t0, 4 * tr_of(nv) * nt);
  }
  const int ks[2] = {0, 1};
  gather_types(x2.H_elem(), p12, ks, 2, nt, nv);
  cast_copy(m1.H_elem(), p1, n);
  cast_copy(m2.H_elem(), p2, n);
  const double flo = (2 * g_of(nv) + tr_of(nv)) * 16.0 * nt + (4 * g_of(nv) + tr_of(nv)) * 128.0 * nt * nt;
  const double t0 = now();
  obj.contractSigmaPi(m1, m2, x2, propH);
  report<Float>(now() - t0, flo);
}

// Run f with errors reported as exceptions; returns an LC_* code.
template <typename F>
int as_code(F&& f) {
  const bool prev = lcdrop::g_throw;
  lcdrop::g_throw = true;
  int rc = LC_OK;
  try {
    f();
  } catch (const lc::Error& e) {
    lc::last_error() = e.msg;
    rc = e.code;
  } catch (const std::exception& e) {
    lc::last_error() = e.what();
    rc = LC_ERR_ARG;
  }
  lcdrop::g_throw = prev;
  return rc;
}

[[noreturn]] void die_uncaught(const lc::Error& e) {
  fprintf(stderr, "%s\nError in libcontract_hip\n", e.msg.c_str());
  fflush(stderr);
  exit(EXIT_FAILURE);
}

}  // namespace

// ------------------------------------------------------------------- reference C++ signatures
#define DISPATCH(call_single, call_double)               \
  try {                                                  \
    if (prec == SINGLE) call_single; else call_double;   \
  } catch (const lc::Error& e) {                         \
    if (lcdrop::g_throw) throw;                          \
    die_uncaught(e);                                     \
  }

void calculateRho(contractInfo info, void* prop, double* mom, PRECISION prec, std::string outfile) {
  DISPATCH(rho_impl<float>(info, (float*)prop, mom, outfile), rho_impl<double>(info, (double*)prop, mom, outfile))
}

void calculateRhoPi(int idir, contractInfo info, void* prop, double* p1, double* p2, double* p12,
                    PRECISION prec, std::string outfile) {
  DISPATCH(rhopi_impl<float>(idir, info, (float*)prop, p1, p2, p12, outfile),
           rhopi_impl<double>(idir, info, (double*)prop, p1, p2, p12, outfile))
}

void calculatePiPi(contractInfo info, void* prop, double* p1, double* p2, double* p3, double* p4,
                   PRECISION prec, std::string outfile, ISOSPIN iso) {
  DISPATCH(pipi_impl<float>(info, (float*)prop, p1, p2, p3, p4, outfile, iso),
           pipi_impl<double>(info, (double*)prop, p1, p2, p3, p4, outfile, iso))
}

void calculateSigma(contractInfo info, void* prop, double* mom, PRECISION prec, std::string outfile) {
  DISPATCH(sigma_impl<float>(info, (float*)prop, mom, outfile), sigma_impl<double>(info, (double*)prop, mom, outfile))
}

void calculateSigmaPi(contractInfo info, void* prop, double* p1, double* p2, double* p12, PRECISION prec,
                      std::string outfile) {
  DISPATCH(sigmapi_impl<float>(info, (float*)prop, p1, p2, p12, outfile),
           sigmapi_impl<double>(info, (double*)prop, p1, p2, p12, outfile))
}

// ---------------------------------------------------------------------------- C wrappers
extern "C" {

int lc_contract_init(int nev, int nt, int dev) {
  return as_code([&] {
    contractInfo info;
    info.NeV = nev;
    info.Nt = nt;
    info.dev = dev;
    initContract(info);
  });
}

int lc_contract_destroy(void) { return as_code([&] { destroyContract(); }); }

int lc_set_binary_output(int on) {
  setBinaryOutput(on != 0);
  return LC_OK;
}

int lc_set_i_dir(int idir) {
  contract::i_dir = idir;
  return LC_OK;
}

int lc_read_prop(const char* path, int prec, void** out) {
  return as_code([&] {
    if (!path || !out) throw lc::Error{LC_ERR_ARG, "null pointer"};
    *out = readPropH(path, prec == LC_SINGLE ? SINGLE : DOUBLE);
  });
}

void lc_free_host(void* p) { free(p); }

static contractInfo info_of(int nev, int nt) {
  contractInfo info;
  info.NeV = nev;
  info.Nt = nt;
  info.dev = 0;
  return info;
}

int lc_calculate_rho(int nev, int nt, void* prop, double* mom, int prec, const char* outfile) {
  return as_code([&] {
    calculateRho(info_of(nev, nt), prop, mom, prec == LC_SINGLE ? SINGLE : DOUBLE, outfile);
  });
}

int lc_calculate_rhopi(int idir, int nev, int nt, void* prop, double* p1, double* p2, double* p12,
                       int prec, const char* outfile) {
  return as_code([&] {
    calculateRhoPi(idir, info_of(nev, nt), prop, p1, p2, p12, prec == LC_SINGLE ? SINGLE : DOUBLE, outfile);
  });
}

int lc_calculate_pipi(int nev, int nt, void* prop, double* p1, double* p2, double* p3, double* p4,
                      int prec, const char* outfile, int iso) {
  return as_code([&] {
    calculatePiPi(info_of(nev, nt), prop, p1, p2, p3, p4, prec == LC_SINGLE ? SINGLE : DOUBLE, outfile,
                  iso == LC_I0 ? I0 : I1);
  });
}

int lc_calculate_sigma(int nev, int nt, void* prop, double* mom, int prec, const char* outfile) {
  return as_code([&] {
    calculateSigma(info_of(nev, nt), prop, mom, prec == LC_SINGLE ? SINGLE : DOUBLE, outfile);
  });
}

int lc_calculate_sigmapi(int nev, int nt, void* prop, double* p1, double* p2, double* p12, int prec,
                         const char* outfile) {
  return as_code([&] {
    calculateSigmaPi(info_of(nev, nt), prop, p1, p2, p12, prec == LC_SINGLE ? SINGLE : DOUBLE, outfile);
  });
}

}  // extern "C"
