"""Builds libviennals_gpu.so for gfx950 in-tree (hipcc, no JIT)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libviennals_gpu.so")
SOURCES = ["lsg_kernels.hip", "lsg_learned.hip", "lsg_host.hip", "lsg_geometry.cpp"]
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    # bitwise parity with the CPU restatement: no FMA contraction
    "-ffp-contract=off",
    # packed FP32 instructions are disabled on this platform
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def build_cpp_example(verbose=False):
    """C++ program using the include/viennals_gpu/lsAdvect.hpp facade (tests/cpp)."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "advect_trench.cpp")
    out = os.path.join(root, "tests", "cpp", "advect_trench")
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(root, "include"), src, "-L" + HERE, "-lviennals_gpu",
           "-Wl,-rpath,$ORIGIN/../../viennals_amd", "-o", out]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    return out


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [hipcc] + FLAGS + ["-I" + os.path.join(HERE, "..", "include"), "-o", LIB + ".tmp"] + srcs
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_example(verbose=True)
    print("built", LIB)
