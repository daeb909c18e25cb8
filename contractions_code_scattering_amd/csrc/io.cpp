// Perambulator reader onto the device (the reference's readPropH, lib/contract.cpp:91-160, with the
// permutation moved to the GPU). The file is read one (s1, v1) slab at a time into pinned host
// memory; each slab is copied to the device and scattered into the block layout by a kernel while
// the host reads the next slab (two staging buffers, one event each).
#include <cstdint>
#include <cstdio>
#include <exception>
#include <string>

#include "runtime.h"

extern "C" hipError_t lc_launch_prop_scatter(const void* slab, void* tau, int is_double, int nt, int nv, int s1,
                                             int v1, hipStream_t s);

using namespace lc;

namespace {

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

struct Pinned {
  void* p = nullptr;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};

struct Events {
  hipEvent_t e[2] = {nullptr, nullptr};
  ~Events() {
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

}  // namespace

extern "C" {

int lc_read_prop_device(const char* path, int nt, int nv, int prec, void* d_tau) {
  try {
    require_init();
    if (!path || !d_tau) fail(LC_ERR_ARG, "null pointer");
    if (nt < 1 || nv < 1 || (prec != LC_SINGLE && prec != LC_DOUBLE)) fail(LC_ERR_ARG, "bad nt=%d nv=%d prec=%d", nt, nv, prec);
    File file;
    file.f = fopen(path, "rb");
    if (!file.f) fail(LC_ERR_IO, "cannot open perambulator file %s", path);
    char cfg[100];
    uint32_t hdr[2] = {0, 0};
    if (fread(cfg, 1, 100, file.f) != 100 || fread(hdr, 4, 2, file.f) != 2) fail(LC_ERR_IO, "short header in %s", path);
    if ((int)hdr[0] != nt || (int)hdr[1] != nv)
      fail(LC_ERR_IO, "header nt=%u nv=%u does not match nt=%d nv=%d", hdr[0], hdr[1], nt, nv);
    const size_t slab = (size_t)nt * 4 * nv * nt * 16;
    Context& c = ctx();
    hipStream_t s = c.stream;
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    Pinned host;
    hip_check(hipHostMalloc(&host.p, 2 * slab, hipHostMallocDefault), "hipHostMalloc");
    c.out.ensure(2 * slab);
    char* dev = static_cast<char*>(c.out.p);
    Events ev;
    for (int i = 0; i < 2; i++) hip_check(hipEventCreate(&ev.e[i]), "hipEventCreate");
    int n = 0;
    for (int s1 = 0; s1 < 4; s1++)
      for (int v1 = 0; v1 < nv; v1++, n++) {
        const int b = n & 1;
        char* h = static_cast<char*>(host.p) + b * slab;
        if (n >= 2) hip_check(hipEventSynchronize(ev.e[b]), "hipEventSynchronize");  // buffer b is free again
        if (fread(h, 1, slab, file.f) != slab) {
          (void)hipStreamSynchronize(s);  // no copy may still read the staging buffers
          fail(LC_ERR_IO, "short read in %s (slab s1=%d v1=%d)", path, s1, v1);
        }
        hip_check(hipMemcpyAsync(dev + b * slab, h, slab, hipMemcpyHostToDevice, s), "slab copy");
        hip_check(lc_launch_prop_scatter(dev + b * slab, d_tau, prec == LC_DOUBLE, nt, nv, s1, v1, s), "slab scatter");
        hip_check(hipEventRecord(ev.e[b], s), "hipEventRecord");
      }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    return LC_OK;
  } catch (const Error& e) {
    last_error() = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return LC_ERR_ARG;
  }
}

}  // extern "C"
