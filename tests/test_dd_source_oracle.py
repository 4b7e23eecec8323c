"""CPU checks that pin the oracle's dd sources at grid points (the reference's
sv_curv_col_el_rhs_srcdd, sv_curv_col_el.c:486-555, called after the point
sources by the isotropic and GMB right-hand sides, sv_curv_col_el_iso.c:
193-199 and sv_curv_col_vis_iso.c:235-241) against the reference's own
function compiled from sv_curv_col_el.c, and by a closed form: on a zero
wavefield the rates are exactly the dd terms, duplicate points summed in
input order."""
import ctypes as C

import numpy as np
import pytest

from tests import cases as cs
from oracle import pyoracle as po

needs_ref = pytest.mark.skipif(not po.ref_available(), reason="reference build absent")


def _ref_onestage(case, w, aux, it, ipair, istage):
    ref = po.load_ref()
    cfg, keep = cs.ref_cfg(case)
    co = np.ascontiguousarray(case["coords"])
    r = np.zeros(w.size, np.float32)
    a = np.zeros(max(aux.size, 1), np.float32)
    ref.refh_onestage(C.byref(cfg), po.ptr(co), po.ptr(case["metric"]), po.ptr(case["med"]),
                      it, ipair, istage, 1, po.ptr(w), po.ptr(aux), po.ptr(r), po.ptr(a))
    return r, a[: aux.size]


@needs_ref
@pytest.mark.parametrize("medium", ["iso", "visco"])
@pytest.mark.parametrize("kinds", [(True, False), (False, True), (True, True)])
def test_dd_onestage_matches_reference(medium, kinds):
    case = cs.make_case(13, 12, 11, seed=9, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3,
                        medium=medium)
    case["src"] = cs.ricker_point_source(case, 8, 7, 6, 10, 0.004, fc=10.0, t0=0.05)
    case["dd"] = cs.dd_sources(case, 23, 6, seed=2, vi=kinds[0], mij=kinds[1])
    w = cs.random_wavefield(case, seed=3).reshape(-1)
    aux = cs.random_aux(case, seed=3)
    orc = po.Oracle(case)
    base = dict(case)
    base.pop("dd")
    orc0 = po.Oracle(base)
    moved = 0
    # steps inside and past dd_max_nt (6): nothing is added from step 6 on
    for it, ipair, istage in [(0, 0, 0), (2, 1, 1), (5, 3, 3), (5, 6, 2), (6, 2, 0), (9, 7, 3)]:
        r_o, a_o = orc.onestage(it, ipair, istage, w, aux)
        r_r, a_r = _ref_onestage(case, w, aux, it, ipair, istage)
        assert np.array_equal(r_o, r_r), (it, ipair, istage)
        assert np.array_equal(a_o, a_r), (it, ipair, istage)
        r0, _ = orc0.onestage(it, ipair, istage, w, aux)
        if it < 6:
            moved += int((r0 != r_o).sum())
        else:
            assert np.array_equal(r0, r_o)
    assert moved > 0


@needs_ref
@pytest.mark.parametrize("medium", ["iso", "visco"])
def test_dd_run_matches_reference(medium):
    ni, nj, nk, nt, dt = 16, 14, 12, 24, 0.004
    case = cs.make_case(ni, nj, nk, seed=13, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3,
                        medium=medium)
    case["dd"] = cs.dd_sources(case, 40, 17, seed=5, scale=1.0e-2)
    ridx, rw = cs.surface_receivers(case, n=3)
    orc = po.Oracle(case)
    w_o, a_o, s_o = orc.run(dt, nt, recv_idx=ridx, recv_w=rw)
    ref = po.load_ref()
    cfg, keep = cs.ref_cfg(case)
    w_r = np.zeros_like(w_o)
    s_r = np.zeros((len(ridx), cs.ncomp(case), nt), np.float32)
    ref.refh_run(C.byref(cfg), po.ptr(np.ascontiguousarray(case["coords"])),
                 po.ptr(case["metric"]), po.ptr(case["med"]), dt, nt, None, len(ridx),
                 po.ptr(ridx), po.ptr(rw), po.ptr(w_r), po.ptr(s_r), None, None)
    assert np.abs(w_r).max() > 0 and np.abs(s_r).max() > 0
    assert np.array_equal(w_o, w_r)
    assert np.array_equal(s_o, s_r)


def test_dd_closed_form_on_zero_wavefield():
    """No PML, no free surface, zero wavefield: the rates are the dd terms
    alone, vi * (slw / jac) and -mij / jac, summed over duplicates in input
    order (Mxx,Myy,Mzz,Myz,Mxz,Mxy onto Txx,Tyy,Tzz,Tyz,Txz,Txy)."""
    case = cs.make_case(12, 11, 10, seed=7, free_top=False, pml=(0,) * 6)
    dd = cs.dd_sources(case, 9, 4, seed=8, dup=3, surface=False)
    dd["indx"][7] = dd["indx"][8] = dd["indx"][2]          # three at one point
    case["dd"] = dd
    orc = po.Oracle(case)
    nc = cs.ncomp(case)
    vol = (case["ni"] + 6) * (case["nj"] + 6) * (case["nk"] + 6)
    it, st = 3, 2
    r, _ = orc.onestage(it, 0, st, np.zeros(nc * vol, np.float32), None)
    r = r.reshape(nc, vol)
    slw = case["med"][-1].reshape(-1)
    jac = case["metric"][0].reshape(-1)
    want = np.zeros((9, vol), np.float32)
    slot = [3, 4, 5, 6, 7, 8]                 # Txx Tyy Tzz Tyz Txz Txy
    for q, p in enumerate(dd["indx"]):
        v = np.float32(slw[p] / jac[p])
        rj = np.float32(1.0 / np.float64(jac[p]))
        for c in range(3):
            want[c, p] = np.float32(want[c, p] + np.float32(dd["vi"][it, st, q, c] * v))
        for c in range(6):
            want[slot[c], p] = np.float32(want[slot[c], p]
                                          - np.float32(dd["mij"][it, st, q, c] * rj))
    assert np.array_equal(r, want)
    assert np.count_nonzero(r[0]) == len(set(dd["indx"].tolist()))


def test_dd_not_added_in_anisotropic_media():
    """The VTI and general anisotropic right-hand sides add no dd sources."""
    for medium in ("vti", "aniso"):
        case = cs.make_case(12, 11, 10, seed=4, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3,
                            medium=medium)
        w = cs.random_wavefield(case, seed=2).reshape(-1)
        r0, _ = po.Oracle(case).onestage(1, 2, 1, w)
        case["dd"] = cs.dd_sources(case, 5, 3, seed=1)
        r1, _ = po.Oracle(case).onestage(1, 2, 1, w)
        assert np.array_equal(r0, r1), medium
