"""Speed of the CPU restatement (liboracle.so) against the reference's own
RHS (oracle/_ref/libcgfd3d_ref.so, compiled from the reference's sources with
the reference Makefile's -O3), on the same inputs, one thread each.

TEST AND MEASUREMENT INFRASTRUCTURE ONLY: this decides whether the oracle's
timing is a fair stand-in for the reference's CPU path in bench.py's
cpu_baseline leg (the reference's sources cannot travel to the GPU box).

The case is bench.py's CPU sub-block shape: a 160 x 160 x 80 isotropic grid
from the benchmark's workload generator (topography, layered medium with a
random field), CFS-PML 20 layers on 5 faces, free surface, the double-couple
source, run for --steps RK4 steps from a random O(1) wavefield.  A run that
starts from rest spends its first hundreds of steps on subnormal numbers
ahead of the wave front (the stencil's reach grows faster than the wave),
where x86 arithmetic slows by an order of magnitude and the measurement
would say more about the CPU's subnormal handling than about either code;
the random start measures both in the normal regime.  Both runs go through
the same time-loop structure (the harness's refh_run restates
drv_rk_curv_col.c's combine around the reference RHS; orc_run is the
oracle's own loop) and must end bit-identical.  Writes oracle/calibration.json.

    OMP_NUM_THREADS=1 python oracle/calibrate.py [--steps 12]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import pyoracle as po  # noqa: E402
from tests import cases as cs  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--size", default="160x160x80")
    args = ap.parse_args()
    if os.environ.get("OMP_NUM_THREADS") != "1":
        sys.exit("run with OMP_NUM_THREADS=1 (one thread for both codes)")
    if not po.ref_available():
        sys.exit("oracle/_ref/libcgfd3d_ref.so not built (needs the reference tree)")
    from cgfd3d_amd import workloads as wk
    ni, nj, nk = (int(v) for v in args.size.split("x"))
    wl = wk.iso_topo_grf(ni=ni, nj=nj, nk=nk, pml=20, n_side=4, dt=0.004, check_dt=False,
                         nt_total=args.steps + 1)
    case = cs.workload_case(wl)
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    w0 = cs.random_wavefield(case, seed=7).reshape(-1)
    orc = po.Oracle(case)
    t = time.perf_counter()
    w_o, _, _ = orc.run(wl.dt, args.steps, w=w0, aux=None, it0=0, nt_total=wl.nt_total)
    t_orc = time.perf_counter() - t
    ref = po.load_ref()
    cfg, keep = cs.ref_cfg(case)
    co = np.ascontiguousarray(case["coords"])
    w_r = np.zeros(9 * nx * ny * nz, np.float32)
    w_in = np.ascontiguousarray(w0, np.float32)
    t = time.perf_counter()
    ref.refh_run(C.byref(cfg), po.ptr(co), po.ptr(case["metric"]), po.ptr(case["med"]),
                 wl.dt, args.steps, po.ptr(w_in), 0, None, None, po.ptr(w_r), None, None, None)
    t_ref = time.perf_counter() - t
    same = bool(np.array_equal(w_r, w_o))
    npts = ni * nj * nk
    out = {
        "case": f"{ni}x{nj}x{nk} iso sub-block, CFS-PML 20 x5, free surface, double-couple "
                f"source, {args.steps} RK4 steps from a random O(1) wavefield",
        "threads": 1,
        "oracle_updates_per_s": npts * args.steps / t_orc,
        "reference_updates_per_s": npts * args.steps / t_ref,
        "restatement_over_reference": t_ref / t_orc,
        "oracle_seconds": t_orc, "reference_seconds": t_ref,
        "bit_identical": same,
        "oracle_build": "gcc -O3 -march=x86-64-v3 -ffp-contract=off (oracle/Makefile)",
        "reference_build": "gcc -O3 (oracle/Makefile target ref, the reference Makefile's -O3)",
        "cpu": _cpu(),
    }
    with open(os.path.join(HERE, "calibration.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def _cpu():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
