"""GPU: dd sources at grid points (sv_curv_col_el_rhs_srcdd, sv_curv_col_el.c:
486-555, called by the isotropic and GMB right-hand sides after the point
sources) through cgfdh_set_dd_sources / cgfdh_put_dd_values, against the
oracle, which equals the reference's own function bit for bit
(tests/test_dd_source_oracle.py).  Bit-exact: one RHS evaluation for
operator pairs x stages, whole runs through every isotropic kernel variant
with duplicate points, a surface point, PML and the free surface, values
streamed through a table smaller than the source duration as the
reference's chunked reads do (src_dd_accit_loadstf, src_t.c:1929-2005), a
captured batch, and the host-side refusals."""
import numpy as np
import pytest

from cgfd3d_amd import Batch, CgfdhError
from oracle import pyoracle as po
from tests import cases as cs

pytestmark = pytest.mark.gpu


def _case(medium, ndd=30, max_nt=17, seed=3, **kw):
    case = cs.make_case(20, 18, 14, seed=seed, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=4,
                        medium=medium)
    case["src"] = cs.ricker_point_source(case, 11, 9, 7, 24, 0.004, fc=12.0, t0=0.03)
    case["dd"] = cs.dd_sources(case, ndd, max_nt, seed=seed, scale=1.0e6, **kw)
    return case


@pytest.mark.parametrize("medium", ["iso", "visco"])
def test_dd_onestage_matches_oracle(medium):
    case = _case(medium)
    w = cs.random_wavefield(case, seed=2).reshape(-1)
    aux = cs.random_aux(case, seed=2)
    orc = po.Oracle(case)
    s = cs.make_solver(case, 0.004, 24, variant=16 if medium == "iso" else 4)
    try:
        for it, ipair, istage in [(0, 0, 0), (3, 1, 1), (9, 2, 2), (16, 3, 3), (17, 5, 1),
                                  (20, 6, 3)]:
            r_o, a_o = orc.onestage(it, ipair, istage, w, aux)
            r_g, a_g = s.onestage(it, ipair, istage, w, aux)
            assert np.array_equal(r_g, r_o), (it, ipair, istage)
            assert np.array_equal(a_g, a_o), (it, ipair, istage)
    finally:
        s.close()


RUN_VARIANTS = [("iso", v) for v in range(19)] + [("visco", v) for v in (0, 4, 5, 8, 9, 10, 11, 16)]


@pytest.mark.parametrize("medium,variant", RUN_VARIANTS)
def test_dd_run_matches_oracle(medium, variant):
    case = _case(medium)
    nt, dt = 24, 0.004
    ridx, rw = cs.surface_receivers(case, n=3)
    w_o, a_o, s_o = po.Oracle(case).run(dt, nt, recv_idx=ridx, recv_w=rw)
    s = cs.make_solver(case, dt, nt, variant=variant)
    try:
        s.set_receivers(ridx, rw)
        s.run(0, nt)
        w_g, a_g, s_g = s.wavefield(), s.pml_aux(), s.seismograms()
    finally:
        s.close()
    assert np.abs(s_o).max() > 0
    assert np.array_equal(w_g, w_o), (medium, variant)
    assert np.array_equal(a_g, a_o), (medium, variant)
    assert np.array_equal(s_g, s_o), (medium, variant)


@pytest.mark.parametrize("kinds", [(True, False), (False, True)])
def test_dd_values_streamed_in_chunks(kinds):
    """A 5-step table for a 17-step source: each chunk is put before the
    steps that read it (the reference's dd_nt_per_read reads), the run split
    at the chunk boundaries; a step whose values are not in the table is
    refused."""
    case = _case("iso", vi=kinds[0], mij=kinds[1], nt_cap=5)
    dd = case["dd"]
    nt, dt = 24, 0.004
    w_o, a_o, _ = po.Oracle(case).run(dt, nt)
    s = cs.make_solver(case, dt, nt, variant=16)
    try:
        with pytest.raises(CgfdhError, match="not loaded"):
            s.run(0, 1)
        it = 0
        while it < nt:
            if it < dd["max_nt"]:
                n = min(5, dd["max_nt"] - it)
                sl = slice(it, it + n)
                s.put_dd_values(it, None if dd["vi"] is None else dd["vi"][sl],
                                None if dd["mij"] is None else dd["mij"][sl])
                if it == 5:
                    with pytest.raises(CgfdhError, match="step 10"):
                        s.run(5, 6)
            else:
                n = nt - it
            s.run(it, n)
            it += n
        assert np.array_equal(s.wavefield(), w_o)
        assert np.array_equal(s.pml_aux(), a_o)
    finally:
        s.close()


def test_dd_in_captured_batch():
    """A batch of an isotropic and a GMB solver with dd sources, advanced in
    captured 8-step cycles; values put after the capture are read on replay."""
    nt, dt = 24, 0.004
    cases = [_case("iso", seed=3, nt_cap=8), _case("visco", seed=4, nt_cap=8)]
    want = [po.Oracle(c).run(dt, nt)[0] for c in cases]
    sv = [cs.make_solver(c, dt, nt, variant=16 if i == 0 else 4) for i, c in enumerate(cases)]
    b = Batch(sv)
    try:
        for it0 in (0, 8, 16):
            for s, c in zip(sv, cases):
                dd = c["dd"]
                if it0 < dd["max_nt"]:
                    sl = slice(it0, min(it0 + 8, dd["max_nt"]))
                    s.put_dd_values(it0, dd["vi"][sl], dd["mij"][sl])
            b.run(it0, 8)
        for s, w in zip(sv, want):
            assert np.array_equal(s.wavefield(), w)
    finally:
        b.close()
        for s in sv:
            s.close()


def test_dd_refusals():
    case = _case("iso")
    s = cs.make_solver(case, 0.004, 24, variant=16)
    try:
        nx, ny = case["ni"] + 6, case["nj"] + 6
        ghost = 2 + 5 * nx + 5 * nx * ny
        with pytest.raises(CgfdhError, match="not a physical grid point"):
            s.set_dd_sources([ghost], 4)
        s.set_dd_sources(case["dd"]["indx"], 17, nt_cap=4)
        with pytest.raises(CgfdhError, match="exceed"):
            s.put_dd_values(0, case["dd"]["vi"][:5], case["dd"]["mij"][:5])
        with pytest.raises(CgfdhError, match="missing value tables"):
            s.put_dd_values(0, case["dd"]["vi"][:4], None)
        s.set_dd_sources([], 4)                  # removed: runs need no values
        s.run(0, 2)
    finally:
        s.close()
    vti = cs.make_case(12, 11, 10, seed=1, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3,
                       medium="vti")
    s = cs.make_solver(vti, 0.004, 8, variant=12)
    try:
        with pytest.raises(CgfdhError, match="isotropic or GMB"):
            s.set_dd_sources([3 + 3 * 18 + 3 * 18 * 17], 4)
    finally:
        s.close()
