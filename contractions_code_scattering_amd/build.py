"""Build libcontract_hip.so in-tree with hipcc for gfx950.

    python -m contractions_code_scattering_amd.build

Every HIP translation unit is compiled with packed-FP32 instructions disabled: the compiler
would otherwise form v_pk_{add,mul,fma}_f32 from ordinary float code, and those are kept out
of this library.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcontract_hip.so")
OBJDIR = os.path.join(HERE, "_build")

HIP_SOURCES = ["kernels.hip", "gemm_mx.hip", "gemm_oz.hip", "elementals.hip", "noalloc.hip"]
CXX_SOURCES = ["runtime.cpp", "capi.cpp", "engine.cpp", "engine_sigma.cpp", "engine_oz.cpp", "io.cpp", "contract.cpp", "interface.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(HERE, "..", "include"), "-I" + CSRC]
GPU_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + os.path.basename(cmd[-1]))
    return r


def build(verbose=False):
    cc = hipcc()
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJDIR, src + ".o")
        _run([cc, *COMMON, *GPU_FLAGS, "-c", "-o", o, os.path.join(CSRC, src)])
        objs.append(o)
    for src in CXX_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        o = os.path.join(OBJDIR, src + ".o")
        # host-only code: compiled as C++ against the HIP runtime headers
        _run([cc, *COMMON, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-x", "c++", "-c", "-o", o, path])
        objs.append(o)
    tmp = OUT + ".tmp"
    _run([cc, "-shared", "--offload-arch=gfx950", "-fPIC", "-o", tmp, *objs, "-lamdhip64"])
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
