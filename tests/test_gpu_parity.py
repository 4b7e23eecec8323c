"""GPU parity: the HIP engine against the CPU oracle on the same seeded
inputs.  Tolerance: relative L2 <= 1e-6 per RHS evaluation and <= 1e-5 over a
run (FP32, the bar of the parity target); the kernels evaluate the reference's
expressions in the same order without FMA contraction, so the tests also
record that the results are bit-identical."""
import numpy as np
import pytest

from tests import cases as cs
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CONFIGS = [
    ("inner", False, (0,) * 6),
    ("pml6", False, (1,) * 6),
    ("free", True, (0,) * 6),
    ("free_pml5", True, (1, 1, 1, 1, 1, 0)),
]


# (medium, kernel variant): every variant of the isotropic medium, the
# per-point, lean-interior and vectorised variants of the VTI and
# anisotropic media, the per-point and lean-interior ones of the
# visco-elastic medium
MEDIUM_VARIANTS = ([("iso", v) for v in range(19)] +
                   [(m, v) for m in ("vti", "aniso")
                    for v in (0, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18)] +
                   [("visco", v) for v in (0, 4, 5, 8, 9, 10, 11, 16)])


@pytest.mark.parametrize("medium,variant", MEDIUM_VARIANTS)
@pytest.mark.parametrize("name,free,pml", CONFIGS)
def test_onestage_all_pairs(name, free, pml, medium, variant):
    case = cs.make_case(18, 15, 13, seed=2, free_top=free, pml=pml, layers=4, medium=medium)
    w = cs.random_wavefield(case, seed=4).reshape(-1)
    aux = cs.random_aux(case, seed=4) if any(pml) else None
    orc = po.Oracle(case)
    s = cs.make_solver(case, 0.005, 10, variant=variant)
    worst, mism = 0.0, 0
    for ipair in range(8):
        for istage in range(4):
            r_o, a_o = orc.onestage(3, ipair, istage, w, aux)
            r_g, a_g = s.onestage(3, ipair, istage, w, aux)
            worst = max(worst, cs.rel_l2(r_g, r_o))
            mism += int(np.count_nonzero(r_g != r_o))
            if aux is not None:
                worst = max(worst, cs.rel_l2(a_g, a_o))
                mism += int(np.count_nonzero(a_g != a_o))
    s.close()
    print(f"{name} {medium} variant={variant}: worst rel L2 {worst:.3e}, differing values {mism}")
    assert worst <= 1e-6
    assert mism == 0, "expected bit-identical RHS"


RUN_VARIANTS = ([("iso", v) for v in range(19)] +
                [(m, v) for m in ("vti", "aniso") for v in (0, 4, 6, 8, 11, 12, 16, 17, 18)] +
                [("visco", v) for v in (0, 4, 8, 11, 16)])


@pytest.mark.parametrize("medium,variant", RUN_VARIANTS)
@pytest.mark.parametrize("gauss", [False, True])
def test_run_matches_oracle(medium, variant, gauss):
    ni, nj, nk, nt, dt = 26, 22, 20, 60, 0.004
    case = cs.make_case(ni, nj, nk, seed=5, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=5,
                        medium=medium)
    case["src"] = cs.ricker_point_source(case, 3 + ni // 2, 3 + nj // 2, 3 + nk // 2, nt, dt,
                                         fc=10.0, t0=0.12, gauss=gauss,
                                         inc=(0.2, -0.3, 0.1) if gauss else (0, 0, 0))
    ridx, rw = cs.surface_receivers(case, n=5, seed=1)
    bi, bw = cs.buried_receiver(case, 9, 10, 8, 0.25, 0.5, 0.75)
    ridx = np.concatenate([ridx, bi]); rw = np.concatenate([rw, bw])
    orc = po.Oracle(case)
    w_o, a_o, s_o = orc.run(dt, nt, recv_idx=ridx, recv_w=rw)
    pg_o, dis_o = po.oracle_pg(orc)
    s = cs.make_solver(case, dt, nt, variant=variant)
    s.set_receivers(ridx, rw)
    s.run(0, nt)
    w_g, a_g, s_g = s.wavefield(), s.pml_aux(), s.seismograms()
    pg_g, dis_g = s.peak_ground_motion()
    s.close()
    assert pg_o.max() > 0
    assert np.array_equal(pg_g, pg_o) and np.array_equal(dis_g, dis_o), "peak ground motion"
    e_w, e_s, e_a = cs.rel_l2(w_g, w_o), cs.rel_l2(s_g, s_o), cs.rel_l2(a_g, a_o)
    print(f"{medium} gauss={gauss} variant={variant}: rel L2 wave {e_w:.2e} seis {e_s:.2e} "
          f"aux {e_a:.2e}")
    assert np.abs(s_o).max() > 0
    assert e_w <= 1e-5 and e_s <= 1e-5 and e_a <= 1e-5
    assert np.array_equal(w_g, w_o) and np.array_equal(s_g, s_o)


@pytest.mark.parametrize("medium", ["iso", "vti", "aniso", "visco"])
def test_device_setup_matches_oracle(medium):
    """Metric (gd_curv_metric_cal) and free-surface matrices
    (sv_curv_col_el_{iso,vti,aniso}_dvh2dvz) built on the device equal the
    oracle's."""
    from cgfd3d_amd import Solver, workloads
    case = cs.make_case(21, 17, 15, seed=3, hill=0.4, medium=medium)
    L = workloads._setup_sigs()
    wl = case.get("visco_wl")
    s = Solver(21, 17, 15, 0.004, 4, free_top=True, medium_type=case["medium_type"],
               nmaxwell=0 if wl is None else len(wl))
    xyz = np.ascontiguousarray(case["coords"])
    assert L.cgfdh_compute_metric(s.h, po.ptr(xyz)) == 0
    m = np.zeros_like(case["metric"])
    assert L.cgfdh_get_metric(s.h, po.ptr(m)) == 0
    assert np.array_equal(m, case["metric"])
    s.set_medium(case["medium_type"], case["med"], wl=wl)
    assert L.cgfdh_compute_free_surface(s.h) == 0
    mats = np.zeros(3 * 27 * 23 * 9, np.float32)
    assert L.cgfdh_get_free_surface(s.h, po.ptr(mats)) == 0
    want = cs.free_surface_mats(case)
    assert np.array_equal(mats, want[: mats.size])
    if wl is not None:
        # matD, built from the coordinates by cgfdh_compute_metric
        matD = np.zeros(27 * 23 * 9, np.float32)
        from cgfd3d_amd._lib import lib
        assert lib().cgfdh_get_free_surface_matD(s.h, po.ptr(matD)) == 0
        assert np.array_equal(matD, want[mats.size:])
    s.close()


def test_engine_matches_reference_golden():
    """The committed fixture written by the reference's own code."""
    import os
    from tests.golden.make_golden import RUN_CASES, run_receivers
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "run_free_pml_src.npz"))
    spec = RUN_CASES["free_pml_src"]
    case = cs.make_case(**spec["case"])
    ni, nj, nk = case["ni"], case["nj"], case["nk"]
    case["src"] = cs.ricker_point_source(case, 3 + ni // 2, 3 + nj // 2, 3 + nk // 2, spec["nt"],
                                         spec["dt"], fc=spec["fc"], t0=spec["t0"])
    ridx, rw = run_receivers(case)
    s = cs.make_solver(case, spec["dt"], spec["nt"])
    s.set_receivers(ridx, rw)
    s.run(0, spec["nt"])
    seis, w = s.seismograms(), s.wavefield().reshape(9, nk + 6, nj + 6, ni + 6)
    s.close()
    print(f"golden: rel L2 seismograms {cs.rel_l2(seis, g['seismo']):.2e}")
    assert cs.rel_l2(seis, g["seismo"]) <= 1e-5
    assert np.array_equal(seis, g["seismo"])
    assert np.array_equal(w[:, 3 + nk - 1, 3:3 + nj, 3:3 + ni], g["surface"])


@pytest.mark.parametrize("variant", [4, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18])
def test_midsize_run_matches_oracle(variant):
    """120x110x80 grid, 12 steps of the benchmark's kind of workload (topography,
    random medium, double couple, PML 10 on 5 faces), engine vs oracle; the
    interior box spans whole and partial tiles of every interior kernel."""
    from cgfd3d_amd import workloads
    wl = workloads.iso_topo_grf(ni=120, nj=110, nk=80, nt_total=40, dt=0.004, pml=10, n_side=6)
    s = wl.build_solver(variant=variant)
    s.run(0, 12)
    w_g, seis_g = s.wavefield(), s.seismograms()
    s.close()
    orc = po.load_oracle()
    metric = np.zeros((10, 86, 116, 126), np.float32)
    orc.orc_metric(120, 110, 80, po.ptr(wl.xyz), po.ptr(metric))
    A, B, D, region = cs.pml_tables(orc, 120, 110, 80, wl.xyz, wl.pml_on, wl.pml_layers,
                                    vel=wl.pml_vel, alpha=wl.pml_alpha, beta=wl.pml_beta)
    case = dict(ni=120, nj=110, nk=80, metric=metric,
                med=np.stack([wl.lam, wl.mu, wl.slw]), free_top=1, pml_on=wl.pml_on,
                pml_layers=wl.pml_layers, pml_A=A, pml_B=B, pml_D=D, pml_region=region)
    src = dict(wl.src)
    src.update(nsrc=1, inc=[0, 0, 0], gauss=0)
    case["src"] = src
    w_o, _, seis_o = po.Oracle(case).run(wl.dt, 12, recv_idx=wl.recv_idx, recv_w=wl.recv_w,
                                          nt_total=40)
    e = cs.rel_l2(w_g, w_o)
    print(f"midsize: rel L2 wavefield {e:.2e}")
    assert np.abs(w_o).max() > 0 and e <= 1e-5
    assert np.array_equal(w_g, w_o) and np.array_equal(seis_g, seis_o)


def test_full_size_linearity():
    """600x600x300 benchmark workload: doubling the source doubles every value
    of the wavefield and every seismogram sample exactly (all operations are
    linear and scaling by 2 is exact in binary floating point) except where
    products fall into the subnormal range, whose rounding has fewer bits; a
    run without a source stays exactly zero."""
    from cgfd3d_amd import workloads
    wl = workloads.iso_topo_grf(nt_total=10)
    print("full-size workload built", flush=True)
    s = wl.build_solver()
    s.run(0, 6)
    w1, r1 = s.wavefield(), s.seismograms()
    s.close()
    print("run 1 done", flush=True)
    wl.src = dict(wl.src)
    wl.src["mij"] = (wl.src["mij"] * np.float32(2.0)).astype(np.float32)
    s = wl.build_solver()
    s.run(0, 6)
    w2, r2 = s.wavefield(), s.seismograms()
    s.close()
    print("run 2 done", flush=True)
    assert np.abs(w1).max() > 0
    # Products that fall into the subnormal range round with fewer bits, and
    # the stress-velocity coupling (lambda * metric * dt ~ 1e7 per stage)
    # carries those last-bit differences upward through the steps; they stay
    # many orders of magnitude below each component's peak.
    w1c, w2c = w1.reshape(9, -1), w2.reshape(9, -1)
    for c in range(9):
        peak = float(np.abs(w1c[c]).max())
        differ = w2c[c] != 2.0 * w1c[c]
        worst = float(np.abs(w1c[c][differ]).max()) if differ.any() else 0.0
        print(f"component {c}: peak {peak:.3e}, not exactly doubled {int(differ.sum())}, "
              f"largest among them {worst:.3e}")
        assert worst <= 1e-15 * peak
    assert cs.rel_l2(w2, 2.0 * w1) <= 1e-12
    rdiff = r2 != 2.0 * r1
    assert not rdiff.any() or np.abs(r1[rdiff]).max() <= 1e-15 * np.abs(r1).max()
    wl.src = None
    s = wl.build_solver()
    s.run(0, 3)
    assert not np.any(s.wavefield())
    s.close()


@pytest.mark.parametrize("variant", [16, 17, 18])
def test_wide_run_matches_oracle(variant):
    """A grid wide enough for every box of the whole-row (16) and phased-shell
    (17) layouts: x edge tiles, a 64-point interior, y and z PML slabs in
    whole rows, free-surface rows; 10 steps against the oracle."""
    from cgfd3d_amd import workloads
    wl = workloads.iso_topo_grf(ni=200, nj=48, nk=36, nt_total=20, dt=0.004, pml=10, n_side=4)
    s = wl.build_solver(variant=variant)
    s.run(0, 10)
    w_g, a_g, seis_g = s.wavefield(), s.pml_aux(), s.seismograms()
    s.close()
    case = cs.workload_case(wl)
    w_o, a_o, s_o = po.Oracle(case).run(wl.dt, 10, recv_idx=wl.recv_idx, recv_w=wl.recv_w,
                                        nt_total=wl.nt_total)
    print(f"wide variant {variant}: rel L2 wavefield {cs.rel_l2(w_g, w_o):.1e}", flush=True)
    assert np.abs(w_o).max() > 0
    assert np.array_equal(w_g, w_o) and np.array_equal(a_g, a_o) and np.array_equal(seis_g, s_o)


def test_full_size_values_match_oracle():
    """600x600x300 benchmark workload at its full size: from a random state
    on every physical point and every PML memory variable (so every point,
    face and byte offset of the 4 GB levels carries data), two RK4 steps on
    the GPU equal the oracle's two steps on the host bit for bit, wavefield
    and PML memory variables."""
    from cgfd3d_amd import workloads
    wl = workloads.iso_topo_grf(nt_total=10)
    print("full-size workload built, dt", wl.dt, flush=True)
    ni, nj, nk = wl.ni, wl.nj, wl.nk
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    rng = np.random.default_rng(42)
    w0 = np.zeros((9, nz, ny, nx), np.float32)
    for c in range(9):
        w0[c, 3:3 + nk, 3:3 + nj, 3:3 + ni] = rng.standard_normal((nk, nj, ni), np.float32)
    s = wl.build_solver()
    try:
        aux0 = rng.standard_normal(s.aux_len(), np.float32)
        s.set_wavefield(w0)
        s.set_pml_aux(aux0)
        s.run(0, 2)
        w_g, a_g = s.wavefield(), s.pml_aux()
    finally:
        s.close()
    print("GPU: 2 steps done", flush=True)
    case = cs.workload_case(wl)
    w_o, a_o, _ = po.Oracle(case).run(wl.dt, 2, w=w0.reshape(-1), aux=aux0,
                                      nt_total=wl.nt_total)
    print("oracle: 2 steps done", flush=True)
    e_w, e_a = cs.rel_l2(w_g, w_o), cs.rel_l2(a_g, a_o)
    print(f"full size: rel L2 wavefield {e_w:.2e}, PML memory {e_a:.2e}, bit-identical "
          f"{bool(np.array_equal(w_g, w_o))} / {bool(np.array_equal(a_g, a_o))}", flush=True)
    assert e_w <= 1e-5 and e_a <= 1e-5
    assert np.array_equal(w_g, w_o) and np.array_equal(a_g, a_o)
