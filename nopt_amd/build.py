"""Build recipe for libnoptjk.so (host C++ with g++, HIP kernels with hipcc
for gfx950), in-tree. Run as ``python -m nopt_amd.build``."""
import os
import subprocess
import sys
import hashlib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnoptjk.so")

HOST_SRCS = ["host/basis.cpp", "host/ints.cpp", "host/errors.cpp", "host/mol_capi.cpp", "host/order.cpp"]
# host code that calls the GPU entry points (left out of a --host-only build)
GLUE_SRCS = ["host/ri_shim.cpp", "host/ri_cxx.cpp"]
HIP_SRCS = ["kernels/linalg.hip", "kernels/bf16.hip", "kernels/ozaki.hip", "kernels/jk.hip"]

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

CXXFLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-march=x86-64-v3", "-Wall",
            "-Wno-unused-function", "-I" + os.path.join(ROOT, "include")]
HIPFLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950",
            # packed FP32 VALU ops stay off: hipcc would otherwise form
            # v_pk_{add,mul,fma}_f32 from float2 code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics", "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(hip=True, verbose=True):
    """Compile every source (always; no staleness check) and link the library."""
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for s in HOST_SRCS + (GLUE_SRCS if hip else []):
        o = os.path.join(BUILD, s.replace("/", "_") + ".o")
        _run(["g++"] + CXXFLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    if hip:
        for s in HIP_SRCS:
            o = os.path.join(BUILD, s.replace("/", "_") + ".o")
            _run([HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
            objs.append(o)
    tmp = LIB + ".tmp"
    if hip:
        _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs +
             ["-lgomp", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    else:
        _run(["g++", "-shared", "-fPIC", "-fopenmp", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(hip="--host-only" not in sys.argv)
