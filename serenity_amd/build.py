"""Build the in-tree shared library ``libserenity_xc.so`` for gfx950.

hipcc compiles the HIP kernels, the C ABI and the host C++ (grid, basis,
overlap) into one library next to this file.  Called by
``__graft_entry__.build()``; ``python -m serenity_amd.build`` does the same.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libserenity_xc.so")

SOURCES = [
    "xc_api.hip",
    "kernels/xc_kernels.hip",
    "kernels/grid_kernels.hip",
    "kernels/gradient_kernels.hip",
    "kernels/coulomb_kernels.hip",
    "host/basis.cpp",
    "host/grid.cpp",
    "host/overlap.cpp",
    "host/integrals.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-fopenmp",
    "-Wall", "-Wno-unused-result",
]
DEVICE = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def build(verbose=True):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src.replace("/", "_") + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC, "-x", "hip"] + DEVICE + COMMON + ["-c", path, "-o", obj]
        else:
            cmd = [HIPCC] + COMMON + ["-march=x86-64-v3", "-c", path, "-o", obj]
        if verbose:
            print("[build]", " ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, stderr=subprocess.PIPE if not verbose else None)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fopenmp", "-o", tmp] + objs
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_cpp_driver(verbose)
    return LIB


CPP_DRIVER_SRC = os.path.join(HERE, "..", "tests", "cpp", "cpp_parity.cpp")
CPP_DRIVER = os.path.join(HERE, "..", "tests", "cpp", "cpp_parity")


def build_cpp_driver(verbose=True):
    """The C++ parity driver over include/serenity_xc.hpp, linked to the in-tree library."""
    if not os.path.exists(CPP_DRIVER_SRC):
        return None
    inc = os.path.join(HERE, "..", "include")
    cmd = ["g++", "-O2", "-std=c++17", "-I", inc, CPP_DRIVER_SRC, "-o", CPP_DRIVER + ".tmp", "-L", HERE,
           "-lserenity_xc", "-Wl,-rpath,$ORIGIN/../../serenity_amd"]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(CPP_DRIVER + ".tmp", CPP_DRIVER)
    return CPP_DRIVER


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
