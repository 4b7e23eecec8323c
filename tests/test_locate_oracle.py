"""CPU checks that pin the oracle's point location on the curvilinear grid
(oracle/locate_oracle.c): the point-in-cell test against the reference's own
isPointInHexahedron compiled from media/media_geometry3d.cpp; the cell search
of gd_curv_coord_to_local_indx (gd_t.c:1998-2173, in a file that needs
netCDF) against a brute force over every cell, in the reference's tile
order, with that compiled test; the sampled shift of
gd_curv_coord2shift_sample (gd_t.c:2304-2378) on points placed exactly on a
sample of the trilinear cell and in Cartesian cells (closed form); the
surface search of gd_curv_depth_to_axis (gd_t.c:2179-2296) against a brute
force with the reference's own fdlib_math_Point2InWhichTri and
fdlib_math_computeZfromTri (compiled from fdlib_math.c), and a planar top
(closed form)."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import cases as cs

needs_ref = pytest.mark.skipif(not po.ref_available(), reason="reference build absent")


def _ref_hexa():
    f = po.load_ref().isPointInHexahedron
    f.restype = C.c_bool
    f.argtypes = [C.c_float] * 3 + [C.c_void_p] * 3
    return f


def _corners(co, i, j, k):
    nz, ny, nx = co.shape[1:]
    v = np.empty((3, 8), np.float32)
    for n in range(8):
        v[:, n] = co[:, k + (n >> 2), j + ((n >> 1) & 1), i + (n & 1)]
    return v


@needs_ref
def test_point_in_hexa_matches_reference():
    rng = np.random.default_rng(0)
    ref = _ref_hexa()
    orc = po.load_oracle()
    cube = np.array([[n & 1, (n >> 1) & 1, n >> 2] for n in range(8)], np.float64)
    ins = 0
    for t in range(300):
        v = (cube + 0.15 * rng.standard_normal((8, 3))) * 100.0 + rng.uniform(-1e4, 1e4, 3)
        v = np.ascontiguousarray(v.T, np.float32)
        pts = [v[:, n] for n in range(8)]                                  # vertices
        pts += [0.5 * (v[:, a] + v[:, b]) for a, b in ((0, 1), (0, 7), (3, 5), (2, 6))]
        w = rng.dirichlet(np.ones(4), 6)                                   # on faces
        for face in ((0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 4, 5)):
            pts += [(v[:, face] * ww).sum(axis=1) for ww in w[:2]]
        c = v.mean(axis=1)
        pts += [c + 70.0 * rng.standard_normal(3) for _ in range(24)]      # in and out
        for p in pts:
            p = p.astype(np.float32)
            vx, vy, vz = (np.ascontiguousarray(v[d]) for d in range(3))
            a = orc.orc_point_in_hexa(float(p[0]), float(p[1]), float(p[2]), po.ptr(vx),
                                      po.ptr(vy), po.ptr(vz))
            b = ref(float(p[0]), float(p[1]), float(p[2]), vx.ctypes.data, vy.ctypes.data,
                    vz.ctypes.data)
            assert bool(a) == bool(b), (t, p)
            ins += bool(a)
    assert ins > 1000


def _tile_ranges(n):
    avg, left = n // 4, n % 4
    s, e = [], []
    for t in range(4):
        s.append(3 if t == 0 else e[-1] + 1)
        e.append(s[-1] + avg - 1 + (1 if t < left else 0))
    return s, e


@needs_ref
def test_cell_search_is_the_first_hit_in_tile_order():
    ni, nj, nk = 13, 11, 9
    co = cs.curvilinear_coords(ni, nj, nk, seed=3, hill=0.25, wiggle=0.12)
    ref = _ref_hexa()
    rng = np.random.default_rng(1)
    lo = co.reshape(3, -1).min(axis=1)
    hi = co.reshape(3, -1).max(axis=1)
    pts = rng.uniform(lo, hi, (60, 3)).astype(np.float32)
    # and grid points and cell-face points, which several cells hold
    pts = np.concatenate([pts, co[:, 6, 5, 7][None], co[:, 3, 3, 3][None],
                          (0.5 * (co[:, 5, 4, 6] + co[:, 5, 5, 6]))[None].astype(np.float32)])
    found, ijk, inc, cell = po.locate_points(co, ni, nj, nk, pts, cells=True)
    si, ei = _tile_ranges(ni)
    sj, ej = _tile_ranges(nj)
    sk, ek = _tile_ranges(nk)
    nhit = 0
    for q, p in enumerate(pts):
        want = None
        for tk in range(4):
            for tj in range(4):
                for ti in range(4):
                    for k in range(sk[tk], ek[tk] + 1):
                        for j in range(sj[tj], ej[tj] + 1):
                            for i in range(si[ti], ei[ti] + 1):
                                v = _corners(co, i, j, k)
                                if (p < v.min(axis=1)).any() or (p > v.max(axis=1)).any():
                                    continue
                                if ref(float(p[0]), float(p[1]), float(p[2]), v[0].ctypes.data,
                                       v[1].ctypes.data, v[2].ctypes.data):
                                    want = (i, j, k)
                                    break
                            if want: break
                        if want: break
                    if want: break
                if want: break
            if want: break
        assert bool(found[q]) == (want is not None), q
        if want is not None:
            nhit += 1
            assert tuple(cell[q]) == want, (q, want, cell[q])
            # the index is the cell's corner, or the next point where the
            # shift passed one half (the shift then lies in (-0.5, 0])
            moved = ijk[q] - cell[q]
            assert set(moved.tolist()) <= {0, 1}
            assert np.all(inc[q][moved == 1] <= 0) and np.all(inc[q][moved == 1] > -0.5)
            assert np.all(inc[q][moved == 0] >= 0) and np.all(inc[q][moved == 0] <= 0.5)
    assert nhit > 12


def test_shift_of_a_point_on_a_sample():
    """A point placed on sample (n1, n2, n3)/100 of a cell's trilinear map is
    that sample's distance-zero neighbour: the shift is n/100, and past one
    half the index moves to the next point and the shift to n/100 - 1."""
    ni, nj, nk = 10, 9, 8
    co = cs.curvilinear_coords(ni, nj, nk, seed=5, hill=0.2, wiggle=0.1).astype(np.float64)
    rng = np.random.default_rng(2)
    orc = po.load_oracle()
    hits = 0
    for _ in range(40):
        i, j, k = rng.integers(4, 4 + ni - 3), rng.integers(4, 4 + nj - 3), rng.integers(4, 4 + nk - 3)
        n = rng.integers(5, 96, 3)
        f = n / 100.0
        w = np.array([((1 - f[0]) if (c & 1) == 0 else f[0]) *
                      ((1 - f[1]) if ((c >> 1) & 1) == 0 else f[1]) *
                      ((1 - f[2]) if (c >> 2) == 0 else f[2]) for c in range(8)])
        v = _corners(co, i, j, k).astype(np.float64)
        p = (v * w).sum(axis=1).astype(np.float32)
        vx, vy, vz = (np.ascontiguousarray(v[d], np.float32) for d in range(3))
        sh = np.zeros(3, np.float32)
        orc.orc_sample_shift(float(p[0]), float(p[1]), float(p[2]), po.ptr(vx), po.ptr(vy),
                             po.ptr(vz), 100, 100, 100, po.ptr(sh))
        assert np.array_equal(sh, (n.astype(np.float32) / np.float32(100)).astype(np.float32))
        found, ijk, inc, cell = po.locate_points(co.astype(np.float32), ni, nj, nk, p[None],
                                                 cells=True)
        if found[0] and tuple(cell[0]) == (i, j, k):
            hits += 1
            for d in range(3):
                want = np.float32(n[d]) / np.float32(100)
                if want > 0.5:
                    assert inc[0, d] == np.float32(np.float64(want) - 1.0)
                else:
                    assert inc[0, d] == want
    assert hits > 30


def test_cartesian_closed_form():
    ni, nj, nk, h = 12, 10, 8, 50.0
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    K, J, I = np.meshgrid(np.arange(nz) - 3.0, np.arange(ny) - 3.0, np.arange(nx) - 3.0,
                          indexing="ij")
    co = np.stack([I * h, J * h, K * h - 1000.0]).astype(np.float32)
    rng = np.random.default_rng(4)
    g = rng.integers(0, [ni - 1, nj - 1, nk - 1], (30, 3))
    frac = (rng.integers(0, 100, (30, 3)) + rng.uniform(-0.2, 0.2, (30, 3))).clip(0.3, 99.7) / 100
    pts = ((g + frac) * h + np.array([0, 0, -1000.0])).astype(np.float32)
    found, ijk, inc = po.locate_points(co, ni, nj, nk, pts)
    assert found.all()
    r = np.round(frac * 100).astype(int)
    wrap = r > 50
    assert np.array_equal(ijk, 3 + g + wrap)
    assert np.allclose(inc, r / 100.0 - wrap, atol=1e-6)


@needs_ref
def test_depth_to_axis_matches_brute_force_with_reference_primitives():
    ref = po.load_ref()
    which = ref.fdlib_math_Point2InWhichTri
    which.restype = C.c_int
    which.argtypes = [C.c_float, C.c_float, C.c_void_p, C.c_void_p]
    ztri = ref.fdlib_math_computeZfromTri
    ztri.restype = C.c_float
    ztri.argtypes = [C.c_float] * 11
    ni, nj, nk = 14, 12, 8
    co = cs.curvilinear_coords(ni, nj, nk, seed=6, hill=0.3, wiggle=0.1)
    rng = np.random.default_rng(3)
    lo = co.reshape(3, -1).min(axis=1)
    hi = co.reshape(3, -1).max(axis=1)
    xyd = np.stack([rng.uniform(lo[0], hi[0], 50), rng.uniform(lo[1], hi[1], 50),
                    rng.uniform(0, 800, 50)], axis=1).astype(np.float32)
    found, z = po.depth_to_axis(co, ni, nj, nk, xyd)
    si, ei = _tile_ranges(ni)
    sj, ej = _tile_ranges(nj)
    k = 3 + nk - 1
    for q, (x, y, dep) in enumerate(xyd):
        want = None
        for tj in range(4):
            for ti in range(4):
                for j in range(sj[tj], ej[tj] + 1):
                    for i in range(si[ti], ei[ti] + 1):
                        v = _corners(co, i, j, k)
                        if x < v[0].min() or x > v[0].max() or y < v[1].min() or y > v[1].max():
                            continue
                        qx = np.ascontiguousarray([co[0, k, j, i], co[0, k, j, i + 1],
                                                   co[0, k, j + 1, i], co[0, k, j + 1, i + 1]],
                                                  np.float32)
                        qy = np.ascontiguousarray([co[1, k, j, i], co[1, k, j, i + 1],
                                                   co[1, k, j + 1, i], co[1, k, j + 1, i + 1]],
                                                  np.float32)
                        qz = [co[2, k, j, i], co[2, k, j, i + 1], co[2, k, j + 1, i],
                              co[2, k, j + 1, i + 1]]
                        t = which(float(x), float(y), qx.ctypes.data, qy.ctypes.data)
                        if t == 1:
                            want = ztri(x, y, qx[0], qy[0], qz[0], qx[1], qy[1], qz[1], qx[3],
                                        qy[3], qz[3]) - dep
                        elif t == 2:
                            want = ztri(x, y, qx[3], qy[3], qz[3], qx[2], qy[2], qz[2], qx[0],
                                        qy[0], qz[0]) - dep
                        if want is not None: break
                    if want is not None: break
                if want is not None: break
            if want is not None: break
        assert bool(found[q]) == (want is not None), q
        if want is not None:
            assert z[q] == np.float32(want), q
    assert found.sum() > 20


def test_depth_to_axis_planar_top():
    ni, nj, nk, h = 10, 9, 6, 100.0
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    K, J, I = np.meshgrid(np.arange(nz) - 3.0, np.arange(ny) - 3.0, np.arange(nx) - 3.0,
                          indexing="ij")
    top = 50.0 + 0.1 * I * h - 0.05 * J * h
    z = -3000.0 + (K / (nk - 1)) * (3000.0 + top)
    co = np.stack([I * h, J * h, z]).astype(np.float32)
    rng = np.random.default_rng(9)
    xyd = np.stack([rng.uniform(0, (ni - 1) * h, 40), rng.uniform(0, (nj - 1) * h, 40),
                    rng.uniform(0, 2000, 40)], axis=1).astype(np.float32)
    found, za = po.depth_to_axis(co, ni, nj, nk, xyd)
    assert found.all()
    want = 50.0 + 0.1 * xyd[:, 0] - 0.05 * xyd[:, 1] - xyd[:, 2]
    assert np.allclose(za, want, rtol=0, atol=2e-3)
