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
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the gfx950 engine cannot be built")


# one translation unit per medium's stage kernels plus the surrogate network's
# kernels, compiled in parallel
UNITS = ["cgfdh.hip", "stage_iso.hip", "stage_vti.hip", "stage_aniso.hip", "stage_visco.hip",
         "surrogate.hip"]


def build_engine(verbose=False):
    from concurrent.futures import ThreadPoolExecutor
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    objdir = os.path.join(CSRC, "_obj")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    # host tables and output files: plain C, no FMA contraction (the
    # reference's x86-64 rounding)
    for src in ("host_setup.c", "host_io.c"):
        obj = os.path.join(objdir, src.replace(".c", ".o"))
        cmds.append(["gcc", "-O2", "-fPIC", "-std=gnu99", "-ffp-contract=off", "-Wall",
                     "-Wno-unknown-pragmas", "-Werror", "-c", os.path.join(CSRC, src), "-o",
                     obj] + inc)
        objs.append(obj)
    for u in UNITS:
        o = os.path.join(objdir, u.replace(".hip", ".o"))
        cmds.append([hipcc()] + HIPCC_FLAGS + inc + ["-c", os.path.join(CSRC, u), "-o", o])
        objs.append(o)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        return subprocess.run(cmd, capture_output=not verbose, text=True)

    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        results = list(ex.map(run, cmds))
    for cmd, r in zip(cmds, results):
        if r.returncode != 0:
            raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stderr or ''}")
    link = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", LIB + ".tmp"]
    r = run(link)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(link)}\n{r.stderr or ''}")
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_oracle(verbose=False):
    cmd = ["make", "-C", os.path.join(ROOT, "oracle"), "-j4", "all"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stdout=None if verbose else subprocess.DEVNULL)


def build_c_caller(verbose=False):
    """tests/c/abi_replay: a C program built only from include/cgfdh.h that
    drives libcgfdh.so in the drop-in shim's call order (a GPU test runs it)."""
    src = os.path.join(ROOT, "tests", "c", "abi_replay.c")
    if not os.path.exists(src):
        return
    cmd = ["gcc", "-O2", "-std=gnu99", "-Wall", "-Wextra", "-Werror",
           "-I" + os.path.join(ROOT, "include"), src, "-L" + PKG, "-l:libcgfdh.so",
           "-Wl,-rpath,$ORIGIN/../../cgfd3d_amd", "-o", os.path.join(ROOT, "tests", "c",
                                                                    "abi_replay")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_all(verbose=False):
    build_engine(verbose)
    build_c_caller(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
