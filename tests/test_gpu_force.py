"""GPU: force sources (sv_curv_col_el_rhs_src's force part, sv_curv_col_el.c:
356-364, 440-452) and the strict surface force (the surface traction
TxSrc..TzSrc in the traction image, sv_curv_col_el.c:147-163, and matF2Vz2 x
VxSrc..VzSrc in the isotropic vlow, sv_curv_col_el_iso.c:583-592, from the
slices src_set_surface_layer_for_force forms before every stage, src_t.c:
154-314) through cgfdh_set_sources, against the oracle, which equals the
reference's own functions bit for bit (tests/test_force_source_oracle.py,
tests/test_surface_force_oracle.py).  Bit-exact: one RHS evaluation for
operator pairs x stages, whole runs with PML, the free surface and surface
receivers, and a run shaped like the reference's example (a surface point
force under an ablexp sponge, example/cgfd3d.example.sh:98-111, 153-187)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests import cases as cs

pytestmark = pytest.mark.gpu

STAGES = [(3, 0, 0), (3, 1, 1), (4, 2, 2), (4, 3, 3), (5, 5, 0), (6, 6, 3), (2, 7, 1),
          (5, 4, 2)]


def _onestage_case(medium, gauss, strict, depth):
    case = cs.make_case(14, 13, 12, seed=16, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3,
                        medium=medium)
    nk2 = 3 + case["nk"] - 1
    ijk = [(8, 7, nk2 - depth), (6, 9, nk2), (9, 10, 6)]
    case["src"] = cs.ricker_force_source(
        case, ijk, 10, 0.004, fxyz=[(1.0e12, -0.6e12, 2.0e12), (-3e11, 5e11, 1e11),
                                    (2e11, 1e11, -4e11)],
        strict=strict, gauss=gauss, inc=(0.2, -0.3, 0.1) * 3,
        moment=(1e15, 2e15, -1e15, 3e14, -2e14, 1e14))
    return case


@pytest.mark.parametrize("medium", ["iso", "vti", "aniso", "visco"])
@pytest.mark.parametrize("gauss,strict,depth", [(False, 0, 0), (False, 1, 0), (False, 1, 1),
                                                (True, 0, 1), (True, 1, 2)])
def test_force_onestage_matches_oracle(medium, gauss, strict, depth):
    case = _onestage_case(medium, gauss, strict, depth)
    if strict:
        assert len(case["src"]["surf_idx"]) >= 1
    w = cs.random_wavefield(case, seed=2)
    aux = cs.random_aux(case, seed=2)
    orc = po.Oracle(case)
    s = cs.make_solver(case, 0.004, 10, variant=12 if medium != "visco" else 4)
    try:
        for it, ipair, istage in STAGES:
            r_o, a_o = orc.onestage(it, ipair, istage, w.reshape(-1), aux)
            r_g, a_g = s.onestage(it, ipair, istage, w.reshape(-1), aux)
            assert np.array_equal(r_g, r_o), (medium, it, ipair, istage)
            assert np.array_equal(a_g, a_o), (medium, it, ipair, istage)
    finally:
        s.close()


@pytest.mark.parametrize("medium,variant", [("iso", 12), ("iso", 4), ("iso", 6), ("iso", 7),
                                            ("iso", 16), ("iso", 17), ("vti", 12), ("vti", 7),
                                            ("aniso", 6), ("visco", 4)])
@pytest.mark.parametrize("gauss", [False, True])
def test_strict_surface_force_run_matches_oracle(medium, variant, gauss):
    """Whole RK4 steps with strict surface forces (on the surface, and for
    the Gaussian also two rows below it, still reaching it), a buried force
    and a moment part, PML on 5 faces and the free surface."""
    case = cs.make_case(20, 18, 14, seed=5, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=4,
                        medium=medium)
    nk2 = 3 + case["nk"] - 1
    nt, dt = 24, 0.004
    ijk = [(12, 10, nk2), (8, 9, nk2 - 2 if gauss else nk2 - 1), (11, 12, 7)]
    case["src"] = cs.ricker_force_source(case, ijk, nt, dt, strict=1, gauss=gauss, fc=12.0,
                                         t0=0.03, inc=(0.1, 0.2, -0.3) * 3,
                                         moment=(1e15, -2e15, 1e15, 0.0, 3e14, -1e14))
    assert len(case["src"]["surf_idx"]) == (2 if gauss else 1)
    ridx, rw = cs.surface_receivers(case, n=3)
    w_o, a_o, s_o = po.Oracle(case).run(dt, nt, recv_idx=ridx, recv_w=rw)
    s = cs.make_solver(case, dt, nt, variant=variant)
    try:
        s.set_receivers(ridx, rw)
        s.run(0, nt)
        w_g, a_g, s_g = s.wavefield(), s.pml_aux(), s.seismograms()
    finally:
        s.close()
    print(medium, variant, "gauss" if gauss else "point", "rel L2",
          cs.rel_l2(w_g, w_o), cs.rel_l2(s_g, s_o), flush=True)
    assert np.abs(s_o).max() > 0
    assert np.array_equal(w_g, w_o) and np.array_equal(a_g, a_o) and np.array_equal(s_g, s_o)


def test_example_shaped_surface_force_with_sponge():
    """The reference example's source and boundaries in small: one strict
    surface point force (0, 0, 1e16) with a 2 Hz Ricker, ablexp sponge on 5
    faces, free surface; engine = oracle stepped with the sponge between
    steps."""
    from tests.test_ablexp_oracle import masks, sponge_blocks
    ni, nj, nk = 24, 22, 16
    case = cs.make_case(ni, nj, nk, seed=12, free_top=True, pml=(0,) * 6, layers=3)
    nk2 = 3 + nk - 1
    nt, dt = 30, 0.01
    case["src"] = cs.ricker_force_source(case, [(3 + 15, 3 + 11, nk2)], nt, dt,
                                         fxyz=(0.0, 0.0, 1.0e16), fc=2.0, t0=0.5, strict=1)
    ex, ey, ez = masks(ni + 6, nj + 6, nk + 6, 6)
    blocks = sponge_blocks(ni, nj, nk)
    ridx, rw = cs.surface_receivers(case, n=3)
    orc = po.Oracle(case)
    lib = po.load_oracle()
    nc = cs.ncomp(case)
    w = np.zeros(nc * (ni + 6) * (nj + 6) * (nk + 6), np.float32)
    seis = np.zeros((len(ridx), nc, nt), np.float32)
    for it in range(nt):
        w, _, sv = orc.run(dt, 1, w, None, it0=it, nt_total=nt)
        w = np.ascontiguousarray(w, np.float32)
        lib.orc_ablexp_apply(ni + 6, nj + 6, nk + 6, nc, po.ptr(w), po.ptr(ex), po.ptr(ey),
                             po.ptr(ez), po.ptr(blocks))
        # receivers sample after the sponge (drv_rk_curv_col.c:482-500)
        W = w.reshape(nc, -1)
        for r in range(len(ridx)):
            ix = ridx[r]
            Lx2, Ly2, Lz2 = (np.float32(x) for x in rw[r])
            Lx1, Ly1, Lz1 = (np.float32(1.0 - float(x)) for x in rw[r])
            for c in range(nc):
                v = W[c]
                seis[r, c, it] = (v[ix[0]] * Lx1 * Ly1 * Lz1 + v[ix[1]] * Lx2 * Ly1 * Lz1
                                  + v[ix[2]] * Lx1 * Ly2 * Lz1 + v[ix[3]] * Lx2 * Ly2 * Lz1
                                  + v[ix[5]] * Lx1 * Ly1 * Lz2 + v[ix[5]] * Lx2 * Ly1 * Lz2
                                  + v[ix[6]] * Lx1 * Ly2 * Lz2 + v[ix[7]] * Lx2 * Ly2 * Lz2)
    s = cs.make_solver(case, dt, nt, variant=12)
    try:
        s.set_ablexp(ex, ey, ez, blocks)
        s.set_receivers(ridx, rw)
        s.run(0, nt)
        w_g, s_g = s.wavefield(), s.seismograms()
    finally:
        s.close()
    print("example-shaped: rel L2", cs.rel_l2(w_g, w), cs.rel_l2(s_g, seis), flush=True)
    assert np.abs(w).max() > 0
    assert np.array_equal(w_g, w)
    assert cs.rel_l2(s_g, seis) <= 1e-6


def test_set_sources_rejects_bad_surface_forces():
    from cgfd3d_amd import CgfdhError
    case = cs.make_case(14, 13, 12, seed=16, free_top=True, pml=(0,) * 6)
    nk2 = 3 + case["nk"] - 1
    src = cs.ricker_force_source(case, [(8, 7, nk2 - 1)], 6, 0.004, strict=0)
    s = cs.make_solver(case, 0.004, 6)
    try:
        fr = np.zeros(3 * 6 * 4, np.float32)
        # a point force one row below the surface does not reach it
        with pytest.raises(CgfdhError):
            s.set_sources(src["ijk"], src["it_begin"], src["it_end"], 6, fxyz=src["fxyz"],
                          strict=True, surf_idx=[0], frate=fr)
        # surface forces need strict mode
        with pytest.raises(CgfdhError):
            s.set_sources([8, 7, nk2], src["it_begin"], src["it_end"], 6, fxyz=src["fxyz"],
                          strict=False, surf_idx=[0], frate=fr)
        # neither forces nor moments
        with pytest.raises(CgfdhError):
            s.set_sources([8, 7, nk2], src["it_begin"], src["it_end"], 6)
    finally:
        s.close()
