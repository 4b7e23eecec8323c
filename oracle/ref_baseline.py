"""The reference's own CPU path as bench.py's CPU baseline: the reference's
RHS functions (oracle/_ref/libcgfd3d_ref.so, compiled from the reference's
sources, driven by oracle/ref_harness.c's restatement of the time loop) run
as one single-threaded process per host core, each on its own copy of the
same sub-block -- the way the reference's MPI build runs one rank per core on
its own block (without the halo exchange, which would only slow it).

TEST AND MEASUREMENT INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg uses it
where oracle/_ref exists; nothing in the product loads it.

    python -m oracle.ref_baseline --case CASE.npz --steps N    (one worker)
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def save_case(path, case, dt):
    """The fields a worker needs, as one npz (no pickles)."""
    import numpy as np
    src = case["src"]
    np.savez(path, ni=case["ni"], nj=case["nj"], nk=case["nk"], dt=dt,
             coords=case["coords"], metric=case["metric"], med=case["med"],
             pml_on=np.array(case["pml_on"]), pml_layers=np.array(case["pml_layers"]),
             pml_vel=case["pml_vel"], pml_alpha=case["pml_alpha"], pml_beta=case["pml_beta"],
             free_top=case["free_top"], src_ijk=np.array(src["ijk"], np.int32),
             src_itb=np.array(src["it_begin"], np.int32),
             src_ite=np.array(src["it_end"], np.int32), src_max_nt=src["max_nt"],
             src_mij=np.asarray(src["mij"], np.float32),
             medium_type=int(case.get("medium_type", 2)),
             visco_wl=np.asarray(case.get("visco_wl", []), np.float32))


def _load_case(path):
    import numpy as np
    z = np.load(path, allow_pickle=False)
    case = dict(ni=int(z["ni"]), nj=int(z["nj"]), nk=int(z["nk"]), coords=z["coords"],
                metric=np.ascontiguousarray(z["metric"]), med=np.ascontiguousarray(z["med"]),
                pml_on=[int(v) for v in z["pml_on"]],
                pml_layers=[int(v) for v in z["pml_layers"]],
                pml_vel=float(z["pml_vel"]), pml_alpha=float(z["pml_alpha"]),
                pml_beta=float(z["pml_beta"]), free_top=int(z["free_top"]),
                src=dict(nsrc=1, ijk=z["src_ijk"].tolist(), it_begin=z["src_itb"].tolist(),
                         it_end=z["src_ite"].tolist(), max_nt=int(z["src_max_nt"]),
                         mij=z["src_mij"]),
                medium_type=int(z["medium_type"]))
    if z["visco_wl"].size:
        case["visco_wl"] = z["visco_wl"]
    return case, float(z["dt"])


def worker(path, steps):
    """One reference run of `steps` RK4 steps from rest; seconds it took."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import pyoracle as po
    from tests import cases as cs
    case, dt = _load_case(path)
    ref = po.load_ref()
    cfg, keep = cs.ref_cfg(case)
    co = np.ascontiguousarray(case["coords"])
    nx, ny, nz = case["ni"] + 6, case["nj"] + 6, case["nk"] + 6
    w = np.zeros(cs.ncomp(case) * nx * ny * nz, np.float32)
    t = time.perf_counter()
    ref.refh_run(C.byref(cfg), po.ptr(co), po.ptr(case["metric"]), po.ptr(case["med"]), dt,
                 steps, None, 0, None, None, po.ptr(w), None, None, None)
    el = time.perf_counter() - t
    if not np.isfinite(w).all():
        raise RuntimeError("reference run gave non-finite values")
    return el


def run_parallel(path, nproc, steps, timeout=600):
    """nproc single-threaded workers at once; their elapsed seconds."""
    env = dict(os.environ, OMP_NUM_THREADS="1", PYTHONPATH=ROOT)
    cmd = [sys.executable, "-m", "oracle.ref_baseline", "--case", path, "--steps", str(steps)]
    procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                              cwd=ROOT, env=env) for _ in range(nproc)]
    out = []
    for p in procs:
        so, se = p.communicate(timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"reference worker failed: {se[-500:]}")
        out.append(json.loads(so.strip().splitlines()[-1])["seconds"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", required=True)
    ap.add_argument("--steps", type=int, required=True)
    args = ap.parse_args()
    print(json.dumps({"seconds": worker(args.case, args.steps)}), flush=True)


if __name__ == "__main__":
    main()
