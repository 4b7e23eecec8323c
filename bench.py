"""Benchmark of the time-stepping hot path on one MI355X.

Default workload (the roofline configuration): isotropic elastic 600x600x300
curvilinear grid (~108 M points), dx = 50 m, topography from 20 seeded
Gaussian hills, 3-layer medium times a seeded Gaussian random field, seeded
double-couple source with a 2 Hz Ricker, CFS-PML 20 layers on 5 faces,
traction-image free surface, 256 surface receivers, FP32.

A step is one full RK4 step (4 fused stage launches + receiver sampling) over
the whole grid; inputs are resident in HBM before timing starts.  Prints one
JSON line: grid-point RK4 updates/s, the HBM roofline of the stage kernel
(algorithmic bytes per launch / HIP-event launch time) and a CPU baseline (the
oracle, OpenMP over the host cores, on a bounded sub-block of the same
workload).

--workload vti | aniso | visco: the same topography and layered GRF model on
400x400x200 with a VTI medium (Thomsen eps, delta, gamma seeded per layer in
[0, 0.2]), its seeded random rotation (21 Cij), or the GMB visco-elastic
medium (3 Maxwell bodies, Qp 100, Qs 50 over 0.5-5 Hz).
--workload batch: 64 independent grids (24 at 128^3, 24 at 192x192x128, 16
at 256x256x160; own topography, medium, source position, mechanism and 1-3 Hz
dominant frequency each) advanced together by one replayed hipGraph with
autotuned kernels.
"""
import argparse
import collections
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# per variant: (profile name, the kernels of one stage launch); the shell
# kernels are stage_shell_kernel (the five PML slabs) and stage_box_kernel
# (the free-surface slab)
KERNELS = [
    ("stage_point", "stage_point_kernel"),
    ("stage_tiled", "stage_tiled_kernel"),
    ("stage_interior_ty8", "stage_interior_kernel<TY=8> + shell kernels"),
    ("stage_interior_ty16", "stage_interior_kernel<TY=16> + shell kernels"),
    ("stage_interior_point_by4", "stage_interior_point_kernel<BY=4> + shell kernels"),
    ("stage_interior_point_by2", "stage_interior_point_kernel<BY=2> + shell kernels"),
    ("stage_interior_vec2", "stage_interior_vec_kernel<NQ=2> + shell kernels"),
    ("stage_interior_vec4", "stage_interior_vec_kernel<NQ=4> + shell kernels"),
    ("stage_interior_lds8", "stage_interior_lds_kernel<TY=8> + shell kernels"),
    ("stage_interior_lds16", "stage_interior_lds_kernel<TY=16> + shell kernels"),
    ("stage_interior_pointz2", "stage_interior_pointz_kernel<KZ=2> + shell kernels"),
    ("stage_interior_pointz4", "stage_interior_pointz_kernel<KZ=4> + shell kernels"),
    ("stage_interior_phased", "stage_interior_phased_kernel<BY=4> + shell kernels"),
    ("stage_interior_phased_late", "stage_interior_phased_kernel<BY=4, combine loads last> + "
                                   "shell kernels"),
    ("stage_interior_ldsp8", "stage_interior_ldsp_kernel<TY=8> + shell kernels"),
    ("stage_interior_phased_w3", "stage_interior_phased_kernel<BY=4, 3 waves/SIMD> + "
                                 "shell kernels"),
    ("stage_rows_phased", "stage_rows_phased_kernel<BY=4> (whole rows, PML folded in) + "
                          "free-surface rows"),
    ("stage_interior_phased_shellp", "stage_interior_phased_kernel<BY=4> + "
                                     "stage_shellp_kernel (phased PML slabs) + free-surface rows"),
    ("stage_rows_lds", "stage_rows_lds_kernel<BY=4> (whole rows, PML folded in, non-stencil "
                       "operands loaded straight to LDS) + free-surface rows"),
]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel, workload_name):
    """HBM bytes per stage launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None
    when no summary for this kernel and workload is committed."""
    import re
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload_name}_traffic.md")
    try:
        with open(path) as f:
            m = re.search(r"^traffic_bytes_per_stage_launch:\s*([0-9.eE+]+)", f.read(), re.M)
        return float(m.group(1)) if m else None
    except OSError:
        return None


def cpu_baseline(wl, budget_s=15.0):
    """The CPU path on a sub-block cut from the same workload (top of the
    grid, free surface, PML on 5 faces, the source inside, the workload's
    medium), timed on the host cores: the reference's own RHS compiled from
    its sources (oracle/_ref, one process per core) where that build is
    present, else the oracle (OpenMP over the cores) with the measured
    speed ratio of the two attached."""
    import numpy as np
    from oracle import pyoracle as po
    from tests import cases as cs

    ni, nj, nk = 160, 160, 80
    i0 = (wl.ni - ni) // 2
    j0 = (wl.nj - nj) // 2
    k0 = wl.nk - nk
    sl = (slice(k0, k0 + nk + 6), slice(j0, j0 + nj + 6), slice(i0, i0 + ni + 6))
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    xyz = np.ascontiguousarray(wl.xyz.reshape(3, wl.nk + 6, wl.nj + 6, wl.ni + 6)[(slice(None),) + sl])
    shp = (wl.nk + 6, wl.nj + 6, wl.ni + 6)
    vols = [wl.lam, wl.mu, wl.slw] if wl.med_vols is None else wl.med_vols
    med = np.stack([np.asarray(v).reshape(shp)[sl] for v in vols])
    orc = po.load_oracle()
    metric = np.zeros((10, nz, ny, nx), np.float32)
    orc.orc_metric(ni, nj, nk, po.ptr(xyz), po.ptr(metric))
    layers = 10
    pml_on = [1, 1, 1, 1, 1, 0]
    A, B, D, region = cs.pml_tables(orc, ni, nj, nk, xyz, pml_on, [layers] * 5 + [0],
                                    vel=wl.pml_vel, alpha=wl.pml_alpha, beta=wl.pml_beta)
    case = dict(ni=ni, nj=nj, nk=nk, metric=metric, med=np.ascontiguousarray(med),
                free_top=1, pml_on=pml_on, pml_layers=[layers] * 5 + [0],
                pml_A=A, pml_B=B, pml_D=D, pml_region=region, src=None,
                medium_type=wl.medium_type, coords=xyz)
    if wl.wl is not None:
        case["visco_wl"] = wl.wl
    src = dict(wl.src)
    src.update(nsrc=1, ijk=[3 + ni // 2, 3 + nj // 2, 3 + nk - 1 - 20], inc=[0, 0, 0], gauss=0)
    case["src"] = src
    case.update(pml_vel=wl.pml_vel, pml_alpha=wl.pml_alpha, pml_beta=wl.pml_beta)
    if po.ref_available():
        try:
            return reference_baseline(case, wl, budget_s)
        except Exception as e:      # the port below still gives a baseline
            log(f"reference baseline unavailable ({e}); timing the oracle instead")
    o = po.Oracle(case)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    t = time.perf_counter()
    w, aux, _ = o.run(wl.dt, 1)
    t1 = time.perf_counter() - t
    n = max(1, min(1000, int(budget_s / max(t1, 1e-3))))
    t = time.perf_counter()
    o.run(wl.dt, n, w=w, aux=aux, it0=1)
    tn = time.perf_counter() - t
    value = ni * nj * nk * n / tn
    out = {"value": value, "unit": "grid-point RK4 updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle (C restatement, OpenMP {threads} threads) on a {ni}x{nj}x{nk} "
                      f"{wl.name.rstrip('0123456789x')} "
                      f"sub-block of the same workload with free surface, 10-layer PML on 5 "
                      f"faces and the source, {n} RK4 steps in {tn:.1f} s"}
    cal = os.path.join(ROOT, "oracle", "calibration.json")
    if os.path.exists(cal):
        with open(cal) as f:
            c = json.load(f)
        out["calibration"] = {"restatement_over_reference": c["restatement_over_reference"],
                              "measured": c["case"] + ", one thread each, on " + c["cpu"]}
    return out


def reference_baseline(case, wl, budget_s):
    """The reference's own RHS (oracle/_ref, compiled from its sources) as
    one single-threaded process per host core, each advancing its own copy
    of the sub-block from rest with the source -- the reference's MPI layout
    of one rank per core without its halo exchange."""
    import tempfile
    from oracle import ref_baseline as rb
    ni, nj, nk = case["ni"], case["nj"], case["nk"]
    nproc = max(1, min(16, os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "case.npz")
        rb.save_case(path, case, wl.dt)
        t1 = rb.run_parallel(path, 1, 1)[0]
        n = max(1, min(50, int(budget_s / max(t1, 1e-3))))
        times = rb.run_parallel(path, nproc, n)
    tmax = max(times)
    return {"value": nproc * ni * nj * nk * n / tmax, "unit": "grid-point RK4 updates/s",
            "cores": nproc, "kind": "reference",
            "sample": f"the reference's own RHS (forward/sv_curv_col_el*.c, bdry_t.c, fd_t.c, "
                      f"compiled -O3) in {nproc} single-threaded processes at once, one per "
                      f"core, each running its own copy of a {ni}x{nj}x{nk} "
                      f"{wl.name.rstrip('0123456789x')} sub-block of the same workload (free "
                      f"surface, 10-layer CFS-PML on 5 faces, the source) for {n} RK4 steps "
                      f"from rest; slowest process {tmax:.1f} s"}


MEDIUM_NAMES = {2: "iso elastic", 3: "VTI elastic (Thomsen eps/delta/gamma in [0, 0.2])",
                4: "general anisotropic (rotated VTI, 21 Cij)",
                5: "visco-elastic GMB (3 Maxwell bodies, Qp 100, Qs 50, 0.5-5 Hz)"}


def bench_batch(args):
    """The batched configuration: every grid built, uploaded and freed on the
    host one at a time, then all advanced together."""
    import torch
    from cgfd3d_amd import Batch, workloads
    t = time.perf_counter()
    solvers, npts, bstep, big = [], 0, 0.0, None
    nt_total = max(1001, args.steps + args.warmup + 16)
    for wl in workloads.batch_workloads(nt_total=nt_total):
        solvers.append(wl.build_solver(variant=args.variant))
        npts += wl.npoints
        bstep += wl.algorithmic_bytes_per_step()
        if big is None or wl.npoints > big.npoints:
            big = wl               # the CPU baseline's sub-block comes from the largest grid
        del wl
    log(f"{len(solvers)} grids, {npts / 1e6:.1f} M points, set up in "
        f"{time.perf_counter() - t:.1f} s")
    b = Batch(solvers)
    t = time.perf_counter()
    if args.variant is None:      # a variant given on the command line is kept for every grid
        b.autotune(0)
    choices = collections.Counter(tuple(s.kernel_choice()) for s in solvers)
    log(f"kernel choice in {time.perf_counter() - t:.1f} s: {dict(choices)}")
    b.run(0, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(args.warmup, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = b.last_ms()
    import numpy as np
    for s in solvers:
        if not np.isfinite(s.seismograms()).all():
            raise RuntimeError("non-finite seismograms")
    b.close()
    for s in solvers:
        s.close()
    achieved = bstep * args.steps / (dev_ms / 1e3) / 1e9
    out = {
        "metric": "grid-point RK4 updates/s", "value": npts * args.steps / wall,
        "unit": "grid-point RK4 updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batch of 64 iso elastic grids (24x128^3, 24x192x192x128, "
                               "16x256x256x160), own topography/GRF/source each, fc 1-3 Hz, "
                               "CFS-PML 20 x5, free surface, one replayed hipGraph per 8 steps, "
                               "autotuned kernels",
                   "grids": len(solvers), "points": npts, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_step": bstep,
                     "kernel": "all stage kernels of the batch (graph replay device time)"},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(big),
    }
    print(json.dumps(out), flush=True)


def _media_model_file(path, x0, y0, x1, y1, d=100.0):
    """A three-layer isotropic .md3lay over [x0, x1] x [y0, y1]: a flat top
    above the grid, two undulating interfaces at about -1.5 and -6 km, the
    default workload's layer velocities with depth gradients."""
    import numpy as np
    xs = np.arange(x0, x1 + d, d, dtype=np.float64)
    ys = np.arange(y0, y1 + d, d, dtype=np.float64)
    X, Y = np.meshgrid(xs, ys)
    elev = [np.full(X.shape, 2000.0), -1500.0 + 300.0 * np.sin(X / 4000.0) * np.cos(Y / 5000.0),
            -6000.0 + 800.0 * np.sin(X / 7000.0 + 1.0) * np.sin(Y / 6000.0)]
    base = [(2400.0, 4500.0, 2600.0), (2670.0, 6000.0, 3464.0), (2900.0, 6800.0, 3900.0)]
    grads = (0.02, 0.05, 0.03)
    with open(path, "w") as fp:
        fp.write(f"elastic_isotropic\n3\n{len(xs)} {len(ys)} {x0:.9g} {y0:.9g} {d:.9g} {d:.9g}\n")
        for L in range(3):
            e = elev[L].reshape(-1)
            cols = [e]
            for q in range(3):
                cols += [np.full(e.shape, base[L][q]), np.full(e.shape, grads[q]),
                         np.full(e.shape, 1.0)]
            fp.write("\n".join(" ".join(f"{v:.9g}" for v in row) for row in zip(*cols)) + "\n")
    return len(xs), len(ys)


def bench_media(args):
    """The media discretisation of the default grid (600x600x300 with its
    topography, ghosts included: 112 M points) from a three-layer .md3lay
    with the harmonic averaging, on the GPU (cgfdh_media_layer2model_el_iso,
    file read, upload, kernels, download); the reference's own
    media_layer2model_el_iso (oracle/_ref, single-threaded) on a column of
    the same grid through every depth, scaled by point count."""
    import tempfile
    import numpy as np
    from cgfd3d_amd import L2M_EL_ISO, media_layer2model, workloads
    ni, nj, nk = 600, 600, 300
    dh = 50.0
    top = workloads.topography(ni, nj, dh, 20, 2024)
    xyz = workloads.layered_coords(ni, nj, nk, dh, top, (nk - 1) * dh)
    n = xyz[0].size
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "layers.md3lay")
        _media_model_file(path, -500.0, -500.0, 30500.0, 30500.0)
        media_layer2model(L2M_EL_ISO, path, xyz[:, :8, :8, :8], args.media_method)   # warm-up
        times = []
        for _ in range(2):
            t = time.perf_counter()
            out = media_layer2model(L2M_EL_ISO, path, xyz, args.media_method)
            times.append(time.perf_counter() - t)
        gpu_s = min(times)
        if not all(np.isfinite(o).all() for o in out):
            raise RuntimeError("non-finite media values")
        del out
        cpu = None
        if not args.no_cpu:
            cpu = media_reference_baseline(path, xyz, args.media_method, n)
    log(f"media {args.media_method}: {gpu_s * 1e3:.1f} ms on the GPU for {n / 1e6:.1f} M points")
    print(json.dumps({
        "metric": "media discretisation grid points/s", "value": n / gpu_s,
        "unit": "grid points/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": gpu_s * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32/f64", "data": "synthetic",
        "config": {"workload": f"media_layer2model_el_iso '{args.media_method}' on the 600x600x300 "
                               "grid with topography (606x606x306 points with ghosts), 3-layer "
                               ".md3lay with undulating interfaces and power-law gradients; "
                               "file read + upload + kernels + download",
                   "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": cpu}), flush=True)


def media_reference_baseline(path, xyz, method, n_full, ncol=64):
    """The reference's media_layer2model_el_iso on an ncol x ncol column of
    the grid through every depth (its stdout progress discarded), one core;
    its file parse, timed on one point, is counted once, not per point."""
    import ctypes as C
    import numpy as np
    from oracle import pyoracle as po
    if not po.ref_available():
        return None
    ref = po.load_ref()
    nz, ny, nx = xyz.shape[1:]
    j0, i0 = ny // 2 - ncol // 2, nx // 2 - ncol // 2
    sub = [np.ascontiguousarray(xyz[d][:, j0:j0 + ncol, i0:i0 + ncol]) for d in range(3)]
    outs = [np.zeros(sub[0].shape, np.float32) for _ in range(3)]
    fn = ref.media_layer2model_el_iso
    fn.argtypes = [C.c_void_p] * 6 + [C.c_size_t] * 3 + [C.c_int, C.c_char_p, C.c_char_p, C.c_int]
    sys.stdout.flush()
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        one = [np.ascontiguousarray(a[:1, :1, :1]) for a in sub]
        o1 = [np.zeros(1, np.float32) for _ in range(3)]
        t = time.perf_counter()
        fn(*[o.ctypes.data for o in o1], *[a.ctypes.data for a in one], 1, 1, 1, 3,
           path.encode(), b"loc", 1)
        parse_s = time.perf_counter() - t
        t = time.perf_counter()
        fn(*[o.ctypes.data for o in outs], *[a.ctypes.data for a in sub], ncol, ncol, nz, 3,
           path.encode(), method.encode(), 1)
        cpu_s = time.perf_counter() - t
    finally:
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)
    npts = sub[0].size
    work_s = max(cpu_s - parse_s, 1e-9)
    whole = parse_s + work_s * n_full / npts
    return {"value": n_full / whole, "unit": "grid points/s", "cores": 1, "kind": "reference",
            "sample": f"the reference's media_layer2model_el_iso ('{method}', compiled -O3, "
                      f"one thread as the reference runs it) on a {ncol}x{ncol}x{nz} column of "
                      f"the same grid through every depth: {cpu_s:.2f} s for {npts} points, of "
                      f"which {parse_s:.2f} s read the file; the whole grid at that rate: "
                      f"{whole:.0f} s"}


def bench_outputs(args, wl, s):
    """The workload with output every step: K z slices (all 9 components)
    and the free surface's velocities, packed on the device and copied to
    host buffers; async overlaps packing and copies with the following
    steps (two sets of pinned buffers in turn), sync waits for each step's
    output before the next step.  Wall clock over the timed steps, with the
    host reading every output (a checksum standing in for the file write)."""
    import numpy as np
    import torch
    from cgfd3d_amd import PinnedBuffer
    nx, ny = wl.ni + 6, wl.nj + 6
    ks = np.linspace(3, 3 + wl.nk - 1, args.outputs + 2).astype(int)[1:-1]
    boxes = [((0, nx, 1, 0, ny, 1, int(k), 1, 1), list(range(9))) for k in ks]
    boxes.append(((3, wl.ni, 1, 3, wl.nj, 1, 3 + wl.nk - 1, 1, 1), [0, 1, 2]))
    shapes = [(len(c), b[7], b[4], b[1]) for b, c in boxes]
    nbytes = sum(4 * int(np.prod(sh)) for sh in shapes)
    s.run(0, args.warmup)
    torch.cuda.synchronize()
    check = 0.0
    it = args.warmup
    if args.output_mode == "async":
        sets = [[PinnedBuffer(sh) for sh in shapes] for _ in range(2)]
        pend = [None, None]
        t0 = time.perf_counter()
        for n in range(args.steps):
            q = n % 2
            if pend[q] is not None:          # this set's copies, two steps back
                for t in pend[q]:
                    s.wait(t)
                check += sum(float(b.array[0].ravel()[0]) for b in sets[q])
            s.run_async(it + n, 1)
            pend[q] = [s.subvolume_async(b, c, sets[q][m].array)
                       for m, (b, c) in enumerate(boxes)]
        for q in range(2):
            if pend[q] is not None:
                for t in pend[q]:
                    s.wait(t)
                check += sum(float(b.array[0].ravel()[0]) for b in sets[q])
        s.sync()
        wall = time.perf_counter() - t0
        for st in sets:
            for b in st:
                b.close()
    else:
        t0 = time.perf_counter()
        for n in range(args.steps):
            s.run(it + n, 1)
            for b, c in boxes:
                check += float(s.subvolume(b, c)[0].ravel()[0])
        wall = time.perf_counter() - t0
    s.close()
    if not np.isfinite(check):
        raise RuntimeError("non-finite output")
    N = wl.npoints
    bstep = wl.algorithmic_bytes_per_step()
    out = {
        "metric": "grid-point RK4 updates/s", "value": N * args.steps / wall,
        "unit": "grid-point RK4 updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{wl.name} with output every step: {args.outputs} z slices "
                               f"of 9 components + surface velocities "
                               f"({nbytes / 1e6:.1f} MB per step), {args.output_mode}",
                   "grid": [wl.ni, wl.nj, wl.nk], "dt": wl.dt, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(bstep / (wall / args.steps) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bstep / (wall / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "all kernels of a step, wall clock with output"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--variant", type=int, default=None,
                    help="stage kernels (default: the engine's choice for the medium, 16 "
                         "isotropic, VTI and visco-elastic, 18 anisotropic): 0 per-point, 1 "
                         "tiled, 2/3 z-marching interior + "
                         "shell, 4/5 per-point interior (64x4 / 64x2) + shell, 6/7 "
                         "x-vectorised interior (2 / 4 points per thread) + shell, 8/9 "
                         "LDS-tiled interior (64x8 / 64x16 plane tiles) + shell, 10/11 "
                         "per-point interior marching 2 / 4 planes per block + shell, 12 "
                         "two-phase per-point interior + shell, 13 the same with the RK "
                         "combine loads last, 14 LDS-tiled interior with phased loads + shell, "
                         "15 variant 12 built for 3 waves per SIMD, 16 whole-row two-phase "
                         "kernel with the PML faces folded in, 17 variant 12 with a two-phase "
                         "shell, 18 variant 16 with the operands that have no stencil loaded "
                         "straight to LDS (4 waves per SIMD); with --workload batch a "
                         "given variant serves every grid and the autotune is skipped")
    ap.add_argument("--workload", default="iso",
                    choices=["iso", "vti", "aniso", "visco", "batch", "media"])
    ap.add_argument("--media-method", default="har", choices=["loc", "har", "ari"],
                    help="with --workload media: the equivalent-medium method")
    ap.add_argument("--size", type=str, default=None,
                    help="grid NIxNJxNK (default 600x600x300 iso, 400x400x200 media)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device copies of MIB MiB (PMC counter calibration)")
    ap.add_argument("--outputs", type=int, default=0, metavar="K",
                    help="after every timed step also write out K z slices of all 9 "
                         "components and the velocities of the free surface (the "
                         "reference's slice and snapshot output), timed on the wall clock")
    ap.add_argument("--output-mode", default="async", choices=["async", "sync"],
                    help="with --outputs: packed and copied while the next steps run "
                         "(async, pinned buffers) or between steps (sync)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "the path shards over x/y blocks with per-stage halo exchange, which is "
              "not part of this build", file=sys.stderr)
        sys.exit(2)
    import torch
    from cgfd3d_amd import workloads

    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "media":
        return bench_media(args)
    size = args.size or ("600x600x300" if args.workload == "iso" else "400x400x200")
    ni, nj, nk = (int(v) for v in size.split("x"))
    t = time.perf_counter()
    if args.workload == "iso":
        wl = workloads.iso_topo_grf(ni=ni, nj=nj, nk=nk,
                                    nt_total=max(2001, args.steps + args.warmup))
    else:
        wl = workloads.media_topo_grf(args.workload, ni=ni, nj=nj, nk=nk,
                                      nt_total=max(1501, args.steps + args.warmup))
    log(f"workload {wl.name}: {wl.npoints / 1e6:.1f} M points, aux {wl.n_aux() / 1e6:.2f} M, "
        f"built in {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    s = wl.build_solver(variant=args.variant)
    log(f"engine set up in {time.perf_counter() - t:.1f} s")
    if args.outputs:
        return bench_outputs(args, wl, s)
    s.run(0, args.warmup)
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} steps done")
    t0 = time.perf_counter()
    s.run(args.warmup, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    run_ms, stage_ms, nlaunch = s.last_timing()
    used_variant = s.kernel_choice()[0]
    seis = s.seismograms()
    s.close()
    log(f"timed {args.steps} steps: wall {wall:.3f} s, device {run_ms:.1f} ms, "
        f"stage launches {nlaunch} totalling {stage_ms:.1f} ms")
    import numpy as np
    if not np.isfinite(seis).all():
        raise RuntimeError("non-finite seismograms")
    N = wl.npoints
    value = N * args.steps / wall
    bstep = wl.algorithmic_bytes_per_step()
    per_launch = bstep / 4.0
    achieved = per_launch / (stage_ms / nlaunch / 1e3) / 1e9
    kernel = KERNELS[used_variant]
    traffic = measured_traffic(kernel[0], wl.name)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": per_launch, "kernel": kernel[1]}
    cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        cpu = cpu_baseline(wl)
        log(f"cpu baseline {cpu['value']:.3e} updates/s ({time.perf_counter() - t:.1f} s)")
    out = {
        "metric": "grid-point RK4 updates/s", "value": value, "unit": "grid-point RK4 updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{wl.name} {MEDIUM_NAMES[wl.medium_type]}, 20 Gaussian hills, "
                               f"3-layer+GRF, DC Ricker 2 Hz, CFS-PML 20 x5, free surface, "
                               f"256 receivers",
                   "grid": [wl.ni, wl.nj, wl.nk], "dt": wl.dt, "parallelism": "single GPU",
                   "bytes_per_point_step": wl.bytes_per_point()},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
