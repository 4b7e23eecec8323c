"""CPU checks that pin the oracle's force sources on the free surface.

The reference's default is the strict surface force (par_t.c:821-825): a
force whose extent reaches the surface is left out of the RHS on the surface
row (sv_curv_col_el.c:388, 447) and enters through the surface traction
TxSrc..TzSrc of the traction image (sv_curv_col_el.c:147-163) and, for the
isotropic medium, through matF2Vz2 times the velocity source VxSrc..VzSrc in
vlow (sv_curv_col_el_iso.c:583-592).  The slices are set before every stage
by src_set_surface_layer_for_force (src_t.c:154-314), which lives in a file
that needs netCDF; its restatement is pinned here by its closed forms, and
everything that consumes the slices is pinned against the reference's own
functions compiled from sv_curv_col_el*.c (the harness hands them the
oracle's slices)."""
import ctypes as C

import numpy as np
import pytest

from tests import cases as cs
from oracle import pyoracle as po

needs_ref = pytest.mark.skipif(not po.ref_available(), reason="reference build absent")


def _idx(case, i, j, k):
    nx, ny = case["ni"] + 6, case["nj"] + 6
    return i + j * nx + k * nx * ny


@needs_ref
@pytest.mark.parametrize("medium", ["iso", "vti", "aniso", "visco"])
@pytest.mark.parametrize("depth", [0, 1])
def test_strict_point_force_onestage_matches_reference(medium, depth):
    """One RHS with a strict point force on the surface (depth 0: a surface
    force, TxSrc/VxSrc set, RHS term dropped) or one row below it (depth 1:
    an ordinary force), beside a moment source, every operator direction."""
    case = cs.make_case(13, 12, 11, seed=6, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3,
                        medium=medium)
    nk2 = 3 + case["nk"] - 1
    case["src"] = cs.ricker_force_source(case, [(8, 7, nk2 - depth), (6, 9, nk2)], 10, 0.004,
                                         fxyz=[(1.0e12, -0.6e12, 2.0e12), (-3e11, 5e11, 1e11)],
                                         strict=1, moment=(1e15, 2e15, -1e15, 3e14, -2e14, 1e14))
    assert len(case["src"]["surf_idx"]) == (2 if depth == 0 else 1)
    w = cs.random_wavefield(case, seed=1)
    aux = cs.random_aux(case, seed=1)
    orc = po.Oracle(case)
    ref = po.load_ref()
    co = np.ascontiguousarray(case["coords"])
    for it, ipair, istage in [(3, 0, 0), (3, 1, 1), (4, 2, 2), (4, 3, 3), (5, 5, 0), (6, 6, 3),
                              (2, 7, 1), (5, 4, 2)]:
        r_o, a_o = orc.onestage(it, ipair, istage, w.reshape(-1), aux)
        ts, vs = orc.surface_force_layer(it, istage)
        assert np.abs(ts).max() > 0
        slices = np.concatenate([ts, vs])[None, None]
        cfg, keep = cs.ref_cfg(case, surf_slices=slices)
        r_r = np.zeros_like(r_o)
        a_r = np.zeros(max(len(a_o), 1), np.float32)
        ref.refh_onestage(C.byref(cfg), po.ptr(co), po.ptr(case["metric"]),
                          po.ptr(case["med"]), it, ipair, istage, 1, po.ptr(w.reshape(-1)),
                          po.ptr(aux), po.ptr(r_r), po.ptr(a_r))
        assert np.array_equal(r_o, r_r), (ipair, istage)
        assert np.array_equal(a_o, a_r[: len(a_o)]), (ipair, istage)


@needs_ref
@pytest.mark.parametrize("medium", ["iso", "visco"])
def test_strict_surface_force_run_matches_reference(medium):
    """Whole RK4 steps with a strict surface point force, PML and the free
    surface: the reference's RHS functions in the harness's time loop, fed the
    oracle's surface slices stage by stage, equal the oracle bit for bit
    (wavefield and seismograms)."""
    case = cs.make_case(14, 13, 12, seed=3, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3,
                        medium=medium)
    nk2 = 3 + case["nk"] - 1
    nt, dt = 14, 0.004
    case["src"] = cs.ricker_force_source(case, [(9, 8, nk2)], nt, dt, strict=1, fc=12.0, t0=0.02)
    orc = po.Oracle(case)
    ridx, rw = cs.surface_receivers(case, n=3)
    w_o, _, s_o = orc.run(dt, nt, recv_idx=ridx, recv_w=rw)
    slices = orc.surface_slices(nt)
    cfg, keep = cs.ref_cfg(case, surf_slices=slices)
    ref = po.load_ref()
    nc = cs.ncomp(case)
    w_r = np.zeros_like(w_o)
    s_r = np.zeros((len(ridx), nc, nt), np.float32)
    ref.refh_run(C.byref(cfg), po.ptr(np.ascontiguousarray(case["coords"])),
                 po.ptr(case["metric"]), po.ptr(case["med"]), dt, nt, None, len(ridx),
                 po.ptr(ridx), po.ptr(rw), po.ptr(w_r), po.ptr(s_r), None, None)
    assert np.abs(w_o).max() > 0
    assert np.array_equal(w_o, w_r)
    assert np.array_equal(s_o, s_r)


def test_surface_layer_point_closed_form():
    """src_set_surface_layer_for_force, point source on the surface: TxSrc..
    hold the force of (it, stage) at (si, sj) and VxSrc.. its rate over the
    surface jacobian there (src_t.c:236-250); zero elsewhere, and zero outside
    the source's time window."""
    case = cs.make_case(12, 11, 10, seed=7, free_top=True, pml=(0,) * 6)
    nk2 = 3 + case["nk"] - 1
    nt = 8
    case["src"] = cs.ricker_force_source(case, [(7, 6, nk2), (5, 8, nk2 - 1)], nt, 0.004,
                                         strict=1)
    assert case["src"]["surf_idx"] == [0]
    orc = po.Oracle(case)
    n = nt * 4
    F = case["src"]["fxyz"].reshape(3, 2, n)
    R = case["src"]["frate"].reshape(3, 1, n)
    jac = case["metric"][0, nk2, 6, 7]
    for it, s in [(0, 0), (2, 1), (5, 3)]:
        ts, vs = orc.surface_force_layer(it, s)
        q = it * 4 + s
        assert np.array_equal(ts[:, 6, 7], F[:, 0, q])
        assert np.array_equal(vs[:, 6, 7], (R[:, 0, q] / jac).astype(np.float32))
        ts[:, 6, 7] = 0
        vs[:, 6, 7] = 0
        assert not ts.any() and not vs.any()
    ts, vs = orc.surface_force_layer(nt, 0)     # after it_end
    assert not ts.any() and not vs.any()


@pytest.mark.parametrize("shift", [(0.0, 0.0, 0.0), (0.3, -0.2, 0.4)])
@pytest.mark.parametrize("depth", [0, 2])
def test_surface_layer_gaussian_closed_form(shift, depth):
    """Gaussian surface force: TxSrc.. = F * w on the surface row of the
    source's own weight block (normalised to the rows at and below the
    surface), VxSrc.. = F_rate * w / jac (src_t.c:252-305)."""
    case = cs.make_case(14, 13, 10, seed=2, free_top=True, pml=(0,) * 6)
    nk2 = 3 + case["nk"] - 1
    si, sj, sk = 8, 7, nk2 - depth
    nt = 6
    case["src"] = cs.ricker_force_source(case, [(si, sj, sk)], nt, 0.004, strict=1,
                                         gauss=True, inc=shift)
    assert case["src"]["surf_idx"] == [0]
    orc = po.Oracle(case)
    W = cs.gauss_coefs(case).reshape(7, 7, 7)
    assert abs(W[: 3 + 1 + depth].astype(np.float64).sum() - 1) < 1e-6
    Ws = W[3 + depth]                         # the surface row of the block
    n = nt * 4
    F = case["src"]["fxyz"].reshape(3, n)
    R = case["src"]["frate"].reshape(3, n)
    jac = case["metric"][0, nk2, sj - 3:sj + 4, si - 3:si + 4]
    it, s = 3, 2
    q = it * 4 + s
    ts, vs = orc.surface_force_layer(it, s)
    for m in range(3):
        want_t = (F[m, q] * Ws).astype(np.float32)
        want_v = ((R[m, q] * Ws).astype(np.float32) / jac).astype(np.float32)
        assert np.array_equal(ts[m, sj - 3:sj + 4, si - 3:si + 4], want_t)
        assert np.array_equal(vs[m, sj - 3:sj + 4, si - 3:si + 4], want_v)
        assert abs(ts[m].astype(np.float64).sum() - F[m, q] * Ws.astype(np.float64).sum()) \
            <= 1e-5 * abs(F[m, q])


@pytest.mark.parametrize("strict", [0, 1])
def test_gaussian_force_rhs_closed_form(strict):
    """Gaussian force in a zero wavefield without boundaries: the velocity
    rates are F * w * slw / jac over the block (w the normalised weights,
    float products left to right as sv_curv_col_el.c:448-452), with the
    surface row left out when strict."""
    case = cs.make_case(14, 13, 10, seed=4, free_top=False, pml=(0,) * 6)
    nk2 = 3 + case["nk"] - 1
    si, sj, sk = 8, 7, nk2 - 2
    nt = 6
    case["src"] = cs.ricker_force_source(case, [(si, sj, sk)], nt, 0.004, strict=strict,
                                         gauss=True, inc=(0.2, -0.1, 0.3))
    orc = po.Oracle(case)
    nc = cs.ncomp(case)
    nx, ny, nz = case["ni"] + 6, case["nj"] + 6, case["nk"] + 6
    w = np.zeros(nc * nx * ny * nz, np.float32)
    it, s = 2, 3
    r, _ = orc.onestage(it, 1, s, w, None)
    r = r.reshape(nc, nz, ny, nx)
    W = cs.gauss_coefs(case).reshape(7, 7, 7)
    F = case["src"]["fxyz"].reshape(3, nt * 4)[:, it * 4 + s]
    slw = case["med"][-1]
    jac = case["metric"][0]
    want = np.zeros((3, nz, ny, nx), np.float32)
    for ke in range(-3, 3):      # kb = nk2 - sk = 2 rows above the source
        k = sk + ke
        if strict and k == nk2:
            continue
        for je in range(-3, 4):
            for ie in range(-3, 4):
                i, j = si + ie, sj + je
                coef = W[ke + 3, je + 3, ie + 3]
                V = np.float32(np.float32(coef * slw[k, j, i]) / jac[k, j, i])
                want[:, k, j, i] = (F * V).astype(np.float32)
    assert np.array_equal(r[:3], want)
    assert not r[3:].any()
    if strict:
        assert not r[:3, nk2].any()
    else:
        assert r[:3, nk2].any()
