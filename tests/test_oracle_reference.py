"""The CPU oracle against the reference's own hot-path code, compiled from
the reference sources into oracle/_ref (skipped where that build is absent),
and against the golden fixtures that build wrote (always run).

Bar: bit-identical.  Both sides evaluate the same FP32 expressions in the same
order without FMA contraction."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import cases as cs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built here")


@needs_ref
def test_operator_tables_match_reference():
    ref, orc = po.load_ref(), po.load_oracle()
    out = []
    for lib, fn in ((ref, "refh_fd_ops"), (orc, "orc_fd_ops")):
        ix = np.zeros(8 * 4 * 6 * 5, np.int32)
        cf = np.zeros(8 * 4 * 6 * 5, np.float32)
        ln = np.zeros(8 * 4 * 6, np.int32)
        a, b, t = np.zeros(3, np.float32), np.zeros(4, np.float32), np.zeros(4, np.float32)
        getattr(lib, fn)(po.ptr(ix), po.ptr(cf), po.ptr(ln), po.ptr(a), po.ptr(b), po.ptr(t))
        out.append((ix, cf, ln, a, b, t))
    for x, y in zip(*out):
        assert np.array_equal(x, y)


@needs_ref
@pytest.mark.parametrize("layers", [3, 7])
def test_pml_profiles_match_reference(layers):
    case = cs.make_case(14, 13, 12, seed=9, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=layers)
    ref = po.load_ref()
    cfg, keep = cs.ref_cfg(case)
    maxlen = layers + 1
    A = np.zeros((6, maxlen), np.float32); B = np.zeros_like(A); D = np.zeros_like(A)
    region = np.zeros((6, 6), np.int32)
    ref.refh_pml(C.byref(cfg), po.ptr(np.ascontiguousarray(case["coords"])), maxlen,
                 po.ptr(A), po.ptr(B), po.ptr(D), po.ptr(region))
    assert np.array_equal(region, case["pml_region"])
    for f in range(5):
        assert np.array_equal(A[f], case["pml_A"][f])
        assert np.array_equal(B[f], case["pml_B"][f])
        assert np.array_equal(D[f], case["pml_D"][f])


MEDIA = ["iso", "vti", "aniso", "visco"]


@needs_ref
@pytest.mark.parametrize("medium", ["iso", "vti", "aniso"])
def test_free_surface_matrices_match_reference(medium):
    case = cs.make_case(13, 12, 10, seed=4, free_top=True, medium=medium)
    ref = po.load_ref()
    cfg, keep = cs.ref_cfg(case)
    nx, ny = case["ni"] + 6, case["nj"] + 6
    m_ref = np.zeros(3 * nx * ny * 9, np.float32)
    ref.refh_dvh2dvz(C.byref(cfg), po.ptr(case["metric"]), po.ptr(case["med"]), po.ptr(m_ref))
    assert np.abs(m_ref).max() > 0
    assert np.array_equal(cs.free_surface_mats(case), m_ref)


@needs_ref
def test_visco_free_surface_matrices_match_reference():
    """sv_curv_col_vis_iso_dvh2dvz, including matD from the coordinates."""
    case = cs.make_case(13, 12, 10, seed=4, free_top=True, medium="visco")
    ref = po.load_ref()
    cfg, keep = cs.ref_cfg(case)
    nx, ny = case["ni"] + 6, case["nj"] + 6
    m_ref = np.zeros(4 * nx * ny * 9, np.float32)
    ref.refh_dvh2dvz_vis(C.byref(cfg), po.ptr(np.ascontiguousarray(case["coords"])),
                         po.ptr(case["metric"]), po.ptr(case["med"]), po.ptr(m_ref))
    assert np.abs(m_ref[3 * nx * ny * 9:]).max() > 0
    assert np.array_equal(cs.free_surface_mats(case), m_ref)


@needs_ref
@pytest.mark.parametrize("medium", MEDIA)
@pytest.mark.parametrize("free,pml", [(False, (0,) * 6), (False, (1,) * 6), (True, (0,) * 6),
                                      (True, (1, 1, 1, 1, 1, 0))])
def test_onestage_matches_reference(free, pml, medium):
    case = cs.make_case(13, 12, 11, seed=6, free_top=free, pml=pml, layers=3, medium=medium)
    case["src"] = cs.ricker_point_source(case, 9, 8, 7, 10, 0.004, fc=10.0, t0=0.05)
    w = cs.random_wavefield(case, seed=1)
    aux = cs.random_aux(case, seed=1)
    orc = po.Oracle(case)
    ref = po.load_ref()
    cfg, keep = cs.ref_cfg(case)
    co = np.ascontiguousarray(case["coords"])
    for ipair in range(8):
        for istage in range(4):
            r_o, a_o = orc.onestage(3, ipair, istage, w.reshape(-1), aux)
            r_r = np.zeros_like(r_o)
            a_r = np.zeros(max(len(a_o), 1), np.float32)
            ref.refh_onestage(C.byref(cfg), po.ptr(co), po.ptr(case["metric"]),
                                  po.ptr(case["med"]), 3, ipair, istage, 1, po.ptr(w),
                                  po.ptr(aux), po.ptr(r_r), po.ptr(a_r))
            assert np.array_equal(r_o, r_r), (ipair, istage)
            assert np.array_equal(a_o, a_r[: len(a_o)]), (ipair, istage)


@needs_ref
def test_reference_patch_touches_only_the_undefined_row():
    """The harness's re-evaluation changes nothing but hV on the surface row of
    stages whose z operator is the forward one: everywhere else the unpatched
    reference output is returned as is."""
    case = cs.make_case(12, 11, 10, seed=8, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3)
    w = cs.random_wavefield(case, seed=2)
    aux = cs.random_aux(case, seed=2)
    ref = po.load_ref()
    cfg, keep = cs.ref_cfg(case)
    co = np.ascontiguousarray(case["coords"])
    nx, ny, nz = case["ni"] + 6, case["nj"] + 6, case["nk"] + 6
    nk2 = 3 + case["nk"] - 1
    for ipair, istage in [(0, 0), (2, 0), (4, 0), (1, 1)]:
        out = []
        for patch in (0, 1):
            r = np.zeros(w.size, np.float32)
            a = np.zeros(aux.size, np.float32)
            ref.refh_onestage(C.byref(cfg), po.ptr(co), po.ptr(case["metric"]),
                                  po.ptr(case["med"]), 0, ipair, istage, patch, po.ptr(w),
                                  po.ptr(aux), po.ptr(r), po.ptr(a))
            out.append((r.reshape(9, nz, ny, nx), a))
        diff = out[0][0] != out[1][0]
        diff[0:3, nk2] = False
        assert not diff.any(), (ipair, istage)
        assert np.array_equal(out[0][1], out[1][1])


@needs_ref
@pytest.mark.parametrize("medium", MEDIA)
def test_run_matches_reference(medium):
    ni, nj, nk, nt, dt = 20, 18, 16, 80, 0.004
    case = cs.make_case(ni, nj, nk, seed=12, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=4,
                        medium=medium)
    case["src"] = cs.ricker_point_source(case, 13, 12, 10, nt, dt, fc=9.0, t0=0.12)
    ridx, rw = cs.surface_receivers(case, n=4)
    bi, bw = cs.buried_receiver(case, 7, 8, 6, 0.2, 0.7, 0.35)
    ridx = np.concatenate([ridx, bi]); rw = np.concatenate([rw, bw])
    orc = po.Oracle(case)
    w_o, a_o, s_o = orc.run(dt, nt, recv_idx=ridx, recv_w=rw)
    pg_o, dis_o = po.oracle_pg(orc)
    ref = po.load_ref()
    cfg, keep = cs.ref_cfg(case)
    w_r = np.zeros_like(w_o)
    s_r = np.zeros((len(ridx), cs.ncomp(case), nt), np.float32)
    pg_r = np.zeros_like(pg_o)
    dis_r = np.zeros_like(dis_o)
    ref.refh_run(C.byref(cfg), po.ptr(np.ascontiguousarray(case["coords"])),
                     po.ptr(case["metric"]), po.ptr(case["med"]), dt, nt, None, len(ridx),
                     po.ptr(ridx), po.ptr(rw), po.ptr(w_r), po.ptr(s_r), po.ptr(pg_r),
                     po.ptr(dis_r))
    assert np.abs(s_r).max() > 0
    assert np.array_equal(w_o, w_r)
    assert np.array_equal(s_o, s_r)
    # peak ground motion: the reference's own PG_calcu (wav_t.c:380-455)
    assert pg_r.max() > 0
    assert np.array_equal(pg_o, pg_r) and np.array_equal(dis_o, dis_r)


# ---- golden fixtures (written by tests/golden/make_golden.py) ------------

def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.mark.parametrize("name", ["free_pml", "pml6"])
def test_oracle_onestage_matches_golden(name):
    from tests.golden.make_golden import ONESTAGE_CASES, checksum, input_checksum
    kw = ONESTAGE_CASES[name]
    g = _load(f"onestage_{name}.npz")
    case = cs.make_case(**kw)
    assert input_checksum(case) == str(g["input_sha256"])
    w = cs.random_wavefield(case, seed=kw["seed"])
    aux = cs.random_aux(case, seed=kw["seed"])
    assert checksum(w) == str(g["w_sha256"]) and checksum(aux) == str(g["aux_sha256"])
    orc = po.Oracle(case)
    for ipair in range(8):
        for istage in range(4):
            r, a = orc.onestage(4, ipair, istage, w.reshape(-1), aux)
            if (ipair, istage) == (0, 0):
                assert np.array_equal(r, g["rhs_p0s0"]) and np.array_equal(a, g["aux_rhs_p0s0"])
            if (ipair, istage) == (2, 0):
                assert np.array_equal(r, g["rhs_p2s0"]) and np.array_equal(a, g["aux_rhs_p2s0"])
            assert checksum(r, a) == str(g["sha256"][ipair, istage]), (ipair, istage)


def test_oracle_run_matches_golden():
    from tests.golden.make_golden import RUN_CASES, checksum, input_checksum, run_receivers
    spec = RUN_CASES["free_pml_src"]
    g = _load("run_free_pml_src.npz")
    case = cs.make_case(**spec["case"])
    assert input_checksum(case) == str(g["input_sha256"])
    ni, nj, nk = case["ni"], case["nj"], case["nk"]
    case["src"] = cs.ricker_point_source(case, 3 + ni // 2, 3 + nj // 2, 3 + nk // 2, spec["nt"],
                                         spec["dt"], fc=spec["fc"], t0=spec["t0"])
    assert checksum(case["src"]["mij"]) == str(g["src_sha256"])
    ridx, rw = run_receivers(case)
    w, _, s = po.Oracle(case).run(spec["dt"], spec["nt"], recv_idx=ridx, recv_w=rw)
    assert np.array_equal(s, g["seismo"])
    W = w.reshape(9, nk + 6, nj + 6, ni + 6)
    assert np.array_equal(W[:, 3 + nk - 1, 3:3 + nj, 3:3 + ni], g["surface"])
    assert checksum(w) == str(g["wave_sha256"])
