// Device-side helpers shared by the kernels: complex element access through a view, batch
// addressing, group lookup, the XCD-aware workgroup order, accumulator layouts and reductions.
#pragma once
#include <hip/hip_runtime.h>

#include "lc_internal.h"

namespace lc {

typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));

template <typename T> struct Cplx { T x, y; };

// Device memory reached through descriptors (pointers read from a group struct) is generic to
// the compiler, which would emit FLAT instructions; FLAT loads count against the LDS counter as
// well, so every wait for an LDS read would also wait for the prefetch in flight. All global
// traffic goes through these address-space-1 accessors instead (global_load / global_store).
#define LC_GLOBAL __attribute__((address_space(1)))
template <typename V> __device__ __forceinline__ V gload(const void* p) {
  return *(const LC_GLOBAL V*)(p);
}
template <typename V> __device__ __forceinline__ void gstore(void* p, const V& v) {
  *(LC_GLOBAL V*)(p) = v;
}
// complex values move as two-element vectors (a struct cannot be copied across address spaces)
template <typename T> __device__ __forceinline__ Cplx<T> gload_c(const Cplx<T>* p) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 v = *(const LC_GLOBAL v2*)(p);
  return Cplx<T>{v[0], v[1]};
}
template <typename T> __device__ __forceinline__ void gstore_c(Cplx<T>* p, const Cplx<T>& z) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  *(LC_GLOBAL v2*)(p) = v2{z.x, z.y};
}
template <typename T> struct Acc4;
template <> struct Acc4<double> { typedef dx4 type; };
template <> struct Acc4<float> { typedef fx4 type; };

__device__ __forceinline__ dx4 mfma16(double a, double b, dx4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ fx4 mfma16(float a, float b, fx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row of the 16x16 result tile held in accumulator element q of lane l (the column is l % 16),
// as measured on gfx950 with lc_mfma_selftest_kernel (raw accumulator dump):
//   v_mfma_f64_16x16x4_f64: row = l/16 + 4q       v_mfma_f32_16x16x4_f32: row = 4(l/16) + q
template <typename T> __device__ __forceinline__ int acc_row(int l, int q);
template <> __device__ __forceinline__ int acc_row<double>(int l, int q) { return (l >> 4) + 4 * q; }
template <> __device__ __forceinline__ int acc_row<float>(int l, int q) { return ((l >> 4) << 2) + q; }

// Raw block of a view for spin block (a,b): pointer, element flags, exact phase.
template <typename T> struct BlockRef {
  const Cplx<T>* p;
  int flags;
  T sr, si;
  bool zero;  // block of an elemental (x) 1 view off the spin diagonal
};

template <typename T>
__device__ __forceinline__ BlockRef<T> view_block(const LcView& v, long long off, int a, int b,
                                                  long long nv2) {
  BlockRef<T> r;
  const Cplx<T>* base = reinterpret_cast<const Cplx<T>*>(v.base) + off;
  r.flags = v.flags;
  if (v.kind == LC_KRON) {
    r.p = base;
    r.sr = T(1);
    r.si = T(0);
    r.zero = a != b;
    return r;
  }
  const int ra = v.perm_r[a], cb = v.perm_c[b];
  const int blk = (v.flags & LC_TS) ? (cb * 4 + ra) : (ra * 4 + cb);
  r.p = base + (long long)blk * nv2;
  const double pr = v.ph_r[a][0] * v.ph_c[b][0] - v.ph_r[a][1] * v.ph_c[b][1];
  const double pi = v.ph_r[a][0] * v.ph_c[b][1] + v.ph_r[a][1] * v.ph_c[b][0];
  r.sr = T(pr);
  r.si = T(pi);
  r.zero = false;
  return r;
}

// Element (i,j) of f(X) for a raw block.
template <typename T>
__device__ __forceinline__ Cplx<T> elem(const BlockRef<T>& r, int i, int j, int nv) {
  const long long idx = (r.flags & LC_TV) ? (long long)j * nv + i : (long long)i * nv + j;
  Cplx<T> z = gload_c(r.p + idx);
  if (r.flags & LC_CONJ) z.y = -z.y;
  return z;
}

__device__ __forceinline__ void decode_item(long long item, const int* n, long long* c) {
  c[0] = item % n[0];
  const long long q = item / n[0];
  c[1] = q % n[1];
  c[2] = q / n[1];
}

__device__ __forceinline__ long long voff(const long long* s, const long long* c) {
  return s[0] * c[0] + s[1] * c[1] + s[2] * c[2];
}

template <typename G>
__device__ __forceinline__ int find_group(const G* groups, int ng, long long wg) {
  int lo = 0, hi = ng - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (groups[mid].wg_begin <= wg) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8). Renumber them so that
// each XCD receives one contiguous range of logical workgroups: all tiles of one product then
// run on the same XCD and share its L2. The grid is 8 * ceil(total / 8); surplus ids return -1.
__device__ __forceinline__ long long xcd_order(long long total) {
  const long long per = gridDim.x >> 3;
  const long long wg = (long long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  return wg < total ? wg : -1;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Sum of (re, im) over the 256-thread workgroup; result valid in thread 0.
__device__ __forceinline__ void block_sum2(double& re, double& im, double* red /* [8] LDS */) {
  re = wave_sum(re);
  im = wave_sum(im);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[wave] = re;
    red[4 + wave] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    re = red[0] + red[1] + red[2] + red[3];
    im = red[4] + red[5] + red[6] + red[7];
  }
  __syncthreads();
}

}  // namespace lc
