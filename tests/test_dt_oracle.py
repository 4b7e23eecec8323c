"""CPU checks that pin the stable-time-step oracle (blk_dt_esti_curv,
blk_t.c:2840-2930): the point-to-plane distance against the reference's own
fdlib_math_dist_point2plane (compiled from its source) and the estimate
against closed forms on Cartesian grids."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po

MT = {"iso": po.MEDIUM_ISO, "vti": po.MEDIUM_VTI, "aniso": po.MEDIUM_ANISO,
      "visco": po.MEDIUM_VISCO}


def _cart(ni, nj, nk, dh):
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([x * dh, y * dh, z * dh]).astype(np.float32)


def dt_estimate(xyz, mt, vols, cfl, nmaxwell=0):
    orc = po.load_oracle()
    nz, ny, nx = xyz.shape[1:]
    out = np.zeros(3, np.float32)
    ijk = np.zeros(3, np.int32)
    vols = np.ascontiguousarray(vols, np.float32)
    xyz = np.ascontiguousarray(xyz, np.float32)
    orc.orc_dt_estimate(nx - 6, ny - 6, nz - 6, po.ptr(xyz), MT[mt], po.ptr(vols), nmaxwell,
                        cfl, po.ptr(out[0:]), po.ptr(out[1:]), po.ptr(out[2:]), po.ptr(ijk))
    return float(out[0]), float(out[1]), float(out[2]), tuple(int(v) for v in ijk)


@pytest.mark.skipif(not po.ref_available(), reason="reference build absent")
def test_distance_matches_reference():
    ref = C.CDLL(po.os.path.join(po.HERE, "_ref", "libcgfd3d_ref.so"), mode=po.os.RTLD_LAZY)
    ref.fdlib_math_dist_point2plane.restype = C.c_float
    ref.fdlib_math_dist_point2plane.argtypes = [C.c_void_p] * 4
    orc = po.load_oracle()
    rng = np.random.default_rng(0)
    for n in range(2000):
        pts = (rng.standard_normal((4, 3)) * 10.0 ** rng.integers(0, 5)).astype(np.float32)
        pts += np.float32(rng.uniform(-1e4, 1e4))
        args = [po.ptr(np.ascontiguousarray(p)) for p in pts]
        a = np.float32(ref.fdlib_math_dist_point2plane(*args))
        b = np.float32(orc.orc_dist_point2plane(*args))
        assert a.view(np.uint32) == b.view(np.uint32) or (np.isnan(a) and np.isnan(b)), n


def test_distance_closed_form():
    orc = po.load_oracle()
    f = lambda *p: orc.orc_dist_point2plane(*[po.ptr(np.array(q, np.float32)) for q in p])
    # plane z = 2 through three points; the point at height 5 is 3 away
    assert f([0.3, 0.1, 5.0], [0, 0, 2], [1, 0, 2], [0, 1, 2]) == pytest.approx(3.0, rel=1e-6)
    # the plane x + y + z = 1 from the origin: 1/sqrt(3)
    assert f([0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]) == pytest.approx(3 ** -0.5, rel=1e-6)


def _iso(shape, vp, vs, rho):
    mu = np.full(shape, rho * vs * vs, np.float32)
    lam = np.full(shape, rho * vp * vp - 2 * rho * vs * vs, np.float32)
    return np.stack([lam, mu, np.full(shape, rho, np.float32)])


def test_cartesian_homogeneous_closed_form():
    dh, cfl = 50.0, 1.3
    xyz = _cart(10, 9, 8, dh)
    vols = _iso(xyz.shape[1:], 6000.0, 3464.0, 2670.0)
    dt, vp, L, ijk = dt_estimate(xyz, "iso", vols, cfl)
    assert L == pytest.approx(dh / 3 ** 0.5, rel=1e-5)
    assert vp == pytest.approx(6000.0, rel=1e-6)
    assert dt == pytest.approx(cfl * dh / 3 ** 0.5 / 6000.0, rel=1e-5)
    assert all(3 <= v for v in ijk)


def test_fastest_point_sets_the_step():
    dh, cfl = 100.0, 1.0
    xyz = _cart(10, 9, 8, dh)
    vols = _iso(xyz.shape[1:], 4000.0, 2300.0, 2500.0)
    p = (6, 7, 8)                                   # k, j, i
    vols[0][p] = 2500.0 * 7000.0 ** 2 - 2 * 2500.0 * 2300.0 ** 2
    dt, vp, L, ijk = dt_estimate(xyz, "iso", vols, cfl)
    assert ijk == (8, 7, 6)
    assert vp == pytest.approx(7000.0, rel=1e-6)
    # a fast point in the ghost layer is not looked at
    vols[0][1, 4, 4] = vols[0][p] * 4
    assert dt_estimate(xyz, "iso", vols, cfl)[3] == (8, 7, 6)


def test_vti_aniso_visco_velocities():
    dh, cfl, rho = 100.0, 1.0, 2000.0
    xyz = _cart(8, 8, 8, dh)
    shp = xyz.shape[1:]
    c11, c33 = rho * 5000.0 ** 2, rho * 4000.0 ** 2
    full = lambda v: np.full(shp, v, np.float32)
    # VTI: the larger of the horizontal and vertical P speeds
    assert dt_estimate(xyz, "vti", [full(c11), full(c33), full(rho)], cfl)[1] == \
        pytest.approx(5000.0, rel=1e-6)
    assert dt_estimate(xyz, "vti", [full(c33), full(c11), full(rho)], cfl)[1] == \
        pytest.approx(5000.0, rel=1e-6)
    # anisotropic: sqrt(c11 / rho)
    assert dt_estimate(xyz, "aniso", [full(c33), full(rho)], cfl)[1] == \
        pytest.approx(4000.0, rel=1e-6)
    # visco: vp (1 + sqrt(n Ymax)), Ymax over every Ylam and Ymu
    iso = _iso(shp, 6000.0, 3000.0, rho)
    y = [full(0.01), full(0.02), full(0.04), full(0.03)]       # Ylam0, Ymu0, Ylam1, Ymu1
    vp = dt_estimate(xyz, "visco", np.concatenate([iso, np.stack(y)]), cfl, nmaxwell=2)[1]
    assert vp == pytest.approx(6000.0 * (1 + (2 * 0.04) ** 0.5), rel=1e-6)


def _ref_dt(xyz, mt, vols, nmw, cfl):
    ref = po.load_ref()
    nz, ny, nx = xyz.shape[1:]
    out = np.zeros(3, np.float32)
    ijk = np.zeros(3, np.int32)
    ref.refh_dt_esti(nx - 6, ny - 6, nz - 6, po.ptr(xyz), MT[mt], po.ptr(vols), nmw, cfl,
                     po.ptr(out), po.ptr(ijk))
    return out, ijk


@pytest.mark.skipif(not po.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("medium", ["iso", "vti", "aniso", "visco", "cartesian"])
def test_estimate_matches_reference(medium):
    """The oracle against blk_dt_esti_curv compiled from the reference's
    blk_t.c: dtmax, Vp, L bit for bit and the same point, on curvilinear grids
    for every medium and on a Cartesian grid full of ties."""
    from tests import cases as cs
    xyz, mt, vols, nmw = cs.dt_case(medium)
    out, ijk = _ref_dt(xyz, mt, vols, nmw, 1.3)
    got = dt_estimate(xyz, mt, vols, 1.3, nmaxwell=nmw)
    assert [np.float32(g).view(np.uint32) for g in got[:3]] == list(out.view(np.uint32))
    assert got[3] == tuple(int(v) for v in ijk)
    assert 0 < out[0] < 1


@pytest.mark.parametrize("medium", ["iso", "vti", "aniso", "visco", "cartesian"])
def test_estimate_matches_golden(medium):
    """The same comparison against the reference's outputs stored in
    tests/golden/dt_estimate.npz (written by tests/golden/make_golden.py dt)."""
    from tests import cases as cs
    from tests.golden.make_golden import checksum
    g = np.load(po.os.path.join(po.os.path.dirname(po.HERE), "tests", "golden",
                                "dt_estimate.npz"), allow_pickle=False)
    xyz, mt, vols, nmw = cs.dt_case(medium)
    assert checksum(xyz, vols) == str(g[f"{medium}_input_sha256"])
    got = dt_estimate(xyz, mt, vols, 1.3, nmaxwell=nmw)
    assert np.array_equal(np.array(got[:3], np.float32).view(np.uint32),
                          g[f"{medium}_out"].view(np.uint32))
    assert got[3] == tuple(int(v) for v in g[f"{medium}_ijk"])
