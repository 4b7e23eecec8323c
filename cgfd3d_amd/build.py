"""Builds the gfx950 engine library in-tree: cgfd3d_amd/libcgfdh.so.

hipcc flags:
  -ffp-contract=off            no FMA contraction: the kernels evaluate the
                               reference's expressions with separate roundings
  -fno-gpu-flush-denormals-to-zero, -fhip-fp32-correctly-rounded-divide-sqrt
                               IEEE FP32 as the CPU evaluates it
  -Xclang -target-feature -Xclang -packed-fp32-ops
                               no packed FP32 instructions
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libcgfdh.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the gfx950 engine cannot be built")


def build_engine(verbose=False):
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    obj = os.path.join(CSRC, "host_setup.o")
    # host tables: plain C, no FMA contraction (the reference's x86-64 rounding)
    ccmd = ["gcc", "-O2", "-fPIC", "-std=gnu99", "-ffp-contract=off", "-c",
            os.path.join(CSRC, "host_setup.c"), "-o", obj] + inc
    if verbose:
        print(" ".join(ccmd), flush=True)
    subprocess.run(ccmd, check=True)
    srcs = [os.path.join(CSRC, "cgfdh.hip"), "-x", "none", obj]
    cmd = [hipcc()] + HIPCC_FLAGS + inc + srcs + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_oracle(verbose=False):
    cmd = ["make", "-C", os.path.join(ROOT, "oracle"), "-j4", "all"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stdout=None if verbose else subprocess.DEVNULL)


def build_all(verbose=False):
    build_engine(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
