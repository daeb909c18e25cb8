// An allocation-free form of the loop_unity trace (lc_loop_unity_noalloc) for callers that cannot
// allocate device memory: code under HIP graph capture, tight serving loops. The caller passes the
// inputs and a device output; the partial sums live in one file-scope device array of fixed size,
// whose address on each device is taken once, when the library is initialised on that device.
//
//   loop[t] = Tr[diag(t) (mom(t) x 1_spin)] = sum_a sum_ij diag(t)_aa(i, j) mom(t)(j, i)
//
// (run_ContractLoopUnity, lib/contract_kernels_loop_unity.cu:95-107.) Pass 1: workgroup (t, chunk)
// sums a strided share of the 4 nv^2 terms of slice t in FP64 into scratch[t * nchunk + chunk];
// pass 2: one thread per slice adds its nchunk partial sums in a fixed order. Every slot that
// pass 2 reads is written by pass 1 of the same call, so nothing is cleared; calls that overlap on
// one device share the array and race.
#include "lc_device.h"

namespace lc {

constexpr long long kNoallocParts = 65536;  // complex FP64 partial sums: 1 MiB per device
static __device__ __align__(256) double2 scratch[kNoallocParts];

template <typename T>
__global__ void __launch_bounds__(256) lc_loop_partial_kernel(const Cplx<T>* __restrict__ diag,
                                                              const Cplx<T>* __restrict__ mom, int nv, int nchunk,
                                                              double2* __restrict__ part) {
  __shared__ double red[2][4];
  const int t = blockIdx.x / nchunk, ch = blockIdx.x - t * nchunk;
  const long long m = (long long)nv * nv;
  const Cplx<T>* X = diag + (long long)t * 16 * m;  // D-space matrix of slice t: 4 x 4 blocks of nv x nv
  const Cplx<T>* M = mom + (long long)t * m;
  double sr = 0, si = 0;
  for (long long e = (long long)ch * 256 + threadIdx.x; e < 4 * m; e += 256LL * nchunk) {
    const int a = (int)(e / m);
    const long long r = e - a * m;
    const int i = (int)(r / nv), j = (int)(r - (long long)i * nv);
    const Cplx<T> x = X[(long long)(5 * a) * m + r];  // block (a, a), element (i, j)
    const Cplx<T> y = M[(long long)j * nv + i];
    sr += (double)x.x * y.x - (double)x.y * y.y;
    si += (double)x.x * y.y + (double)x.y * y.x;
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = sr;
    red[1][wave] = si;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    part[blockIdx.x] = double2{red[0][0] + red[0][1] + red[0][2] + red[0][3], red[1][0] + red[1][1] + red[1][2] + red[1][3]};
}

template <typename T>
__global__ void __launch_bounds__(256) lc_loop_final_kernel(const double2* __restrict__ part, int nt, int nchunk,
                                                            Cplx<T>* __restrict__ out) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nt) return;
  double sr = 0, si = 0;
  for (int k = 0; k < nchunk; k++) {
    const double2 v = part[(long long)t * nchunk + k];
    sr += v.x;
    si += v.y;
  }
  out[t] = Cplx<T>{(T)sr, (T)si};
}

}  // namespace lc

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device (host-side copy)
}  // namespace

// With `device` current: takes the address of this device's copy of the scratch array (once).
extern "C" hipError_t lc_noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(lc::scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = p;
  return hipSuccess;
}

// Chunks per slice of the first pass, and whether nt slices fit the scratch array.
extern "C" int lc_noalloc_loop_chunks(int nv) {
  const long long n = 4LL * nv * nv;
  long long c = (n + 4095) / 4096;
  return (int)(c < 1 ? 1 : c > 64 ? 64 : c);
}
extern "C" long long lc_noalloc_capacity(void) { return lc::kNoallocParts; }

// The two passes on stream s; no allocation, no synchronisation. hipErrorInvalidValue when the
// device was not set up or the partial sums do not fit.
extern "C" hipError_t lc_launch_loop_unity_noalloc(int dbl, const void* diag, const void* mom, int nt, int nv,
                                                   int device, void* out, hipStream_t s) {
  if (device < 0 || device >= kMaxDevices || !g_scratch[device]) return hipErrorInvalidValue;
  const int nchunk = lc_noalloc_loop_chunks(nv);
  if ((long long)nt * nchunk > lc::kNoallocParts) return hipErrorInvalidValue;
  double2* part = static_cast<double2*>(g_scratch[device]);
  const unsigned g2 = (unsigned)((nt + 255) / 256);
  if (dbl) {
    hipLaunchKernelGGL(lc::lc_loop_partial_kernel<double>, dim3((unsigned)(nt * nchunk)), dim3(256), 0, s,
                       (const lc::Cplx<double>*)diag, (const lc::Cplx<double>*)mom, nv, nchunk, part);
    hipLaunchKernelGGL(lc::lc_loop_final_kernel<double>, dim3(g2), dim3(256), 0, s, (const double2*)part, nt, nchunk,
                       (lc::Cplx<double>*)out);
  } else {
    hipLaunchKernelGGL(lc::lc_loop_partial_kernel<float>, dim3((unsigned)(nt * nchunk)), dim3(256), 0, s,
                       (const lc::Cplx<float>*)diag, (const lc::Cplx<float>*)mom, nv, nchunk, part);
    hipLaunchKernelGGL(lc::lc_loop_final_kernel<float>, dim3(g2), dim3(256), 0, s, (const double2*)part, nt, nchunk,
                       (lc::Cplx<float>*)out);
  }
  return hipGetLastError();
}
