"""CPU checks that pin the oracle's force sources (sv_curv_col_el_rhs_src,
sv_curv_col_el.c:356-364, the non-strict point force: f * slw / jac added to
the velocity rates) against the reference's own function compiled from
sv_curv_col_el.c, alone and beside a moment source.  The engine has no force
kernel yet; this oracle is its checker."""
import ctypes as C

import numpy as np
import pytest

from tests import cases as cs
from oracle import pyoracle as po

needs_ref = pytest.mark.skipif(not po.ref_available(), reason="reference build absent")


def _force_source(case, nt, seed, with_moment):
    src = cs.ricker_point_source(case, 8, 7, 6, nt, 0.004, fc=10.0, t0=0.05)
    rng = np.random.default_rng(seed)
    n = src["nsrc"] * src["max_nt"] * 4
    # a Ricker-shaped force along a random direction
    stf = src["mij"].reshape(6, n)[0] / max(np.abs(src["mij"]).max(), 1.0)
    d = rng.standard_normal(3)
    src["fxyz"] = (np.outer(d, stf) * 1.0e12).astype(np.float32).reshape(-1)
    if not with_moment:
        src["mij"] = None
    return src


@needs_ref
@pytest.mark.parametrize("medium", ["iso", "vti", "aniso", "visco"])
@pytest.mark.parametrize("with_moment", [False, True])
def test_force_onestage_matches_reference(medium, with_moment):
    case = cs.make_case(13, 12, 11, seed=6, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3,
                        medium=medium)
    case["src"] = _force_source(case, 10, 3, with_moment)
    w = cs.random_wavefield(case, seed=1)
    aux = cs.random_aux(case, seed=1)
    orc = po.Oracle(case)
    ref = po.load_ref()
    cfg, keep = cs.ref_cfg(case)
    co = np.ascontiguousarray(case["coords"])
    nc = cs.ncomp(case)
    vol = w.size // nc
    p = 8 + 7 * (case["ni"] + 6) + 6 * (case["ni"] + 6) * (case["nj"] + 6)
    moved = 0.0
    for ipair, istage in [(0, 0), (1, 1), (2, 2), (3, 3), (5, 0), (6, 3)]:
        r_o, a_o = orc.onestage(3, ipair, istage, w.reshape(-1), aux)
        r_r = np.zeros_like(r_o)
        a_r = np.zeros(max(len(a_o), 1), np.float32)
        ref.refh_onestage(C.byref(cfg), po.ptr(co), po.ptr(case["metric"]),
                          po.ptr(case["med"]), 3, ipair, istage, 1, po.ptr(w.reshape(-1)),
                          po.ptr(aux), po.ptr(r_r), po.ptr(a_r))
        assert np.array_equal(r_o, r_r), (ipair, istage)
        assert np.array_equal(a_o, a_r[: len(a_o)]), (ipair, istage)
        moved = max(moved, float(np.abs(r_o.reshape(nc, vol)[:3, p]).max()))
    assert moved > 0


def test_force_adds_slowness_over_jacobian():
    """Closed form: with a zero wavefield the velocity rates are exactly
    f * slw / jac at the source point and zero elsewhere."""
    case = cs.make_case(12, 11, 10, seed=7, free_top=False, pml=(0,) * 6)
    case["src"] = _force_source(case, 10, 4, with_moment=False)
    orc = po.Oracle(case)
    nc = cs.ncomp(case)
    w = np.zeros(nc * (case["ni"] + 6) * (case["nj"] + 6) * (case["nk"] + 6), np.float32)
    r, _ = orc.onestage(3, 0, 1, w, None)
    vol = w.size // nc
    r = r.reshape(nc, vol)
    p = 8 + 7 * (case["ni"] + 6) + 6 * (case["ni"] + 6) * (case["nj"] + 6)
    n = case["src"]["max_nt"] * 4
    f = case["src"]["fxyz"].reshape(3, n)[:, 3 * 4 + 1]       # it 3, stage 1
    slw = case["med"][-1].reshape(-1)[p]
    jac = case["metric"][0].reshape(-1)[p]
    want = f * np.float32(slw / jac)
    assert np.array_equal(r[:3, p], want.astype(np.float32))
    r[:3, p] = 0
    assert not r.any()
