"""Builds libtsr.so (HIP kernels for gfx950 + C++ host code) in-tree with hipcc.

    python -m twoscalerich_amd.build        (or __graft_entry__.build())

Every HIP source is compiled with packed FP32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtsr.so")
HIP_SOURCES = ["gfunc.hip", "krylov.hip", "mlp.hip"]
CXX_SOURCES = ["mesh.cpp", "problem.cpp", "dense.cpp", "solver.cpp", "refabi.cpp"]
ARCH = os.environ.get("TSR_OFFLOAD_ARCH", "gfx950")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    common = ["-O3", "-fPIC", "-std=c++17", "-I" + os.path.join(HERE, "..", "include"), "-Wall",
              "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-label", "-Wno-unused-value", "-Wno-unused-result"]
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        _run([hipcc, "-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
              "-packed-fp32-ops", "-munsafe-fp-atomics", "-c", os.path.join(CSRC, src), "-o", obj] + common)
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        _run([hipcc, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", os.path.join(CSRC, src), "-o", obj] + common)
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + objs + ["-lamdhip64", "-pthread"])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
