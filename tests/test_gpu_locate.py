"""GPU: point location on the curvilinear grid (cgfdh_locate_points, the
cell search of gd_curv_coord_to_local_indx with isPointInHexahedron and the
101^3-sample shift; cgfdh_depth_to_axis, gd_curv_depth_to_axis) against the
oracle (oracle/locate_oracle.c, pinned to the reference's compiled
isPointInHexahedron and fdlib triangle functions by
tests/test_locate_oracle.py).  Bit-exact: found flags, indices and shifts
for random points in and around a hilly grid, grid points and cell-face
points (held by several cells: the reference's tile and cell order decides),
points on trilinear samples, and a few points on the 600 x 600 x 300
benchmark grid."""
import time

import numpy as np
import pytest

from cgfd3d_amd import depth_to_axis, locate_points
from oracle import pyoracle as po
from tests import cases as cs

pytestmark = pytest.mark.gpu


def _points(co, n, seed):
    rng = np.random.default_rng(seed)
    flat = co.reshape(3, -1)
    lo, hi = flat.min(axis=1), flat.max(axis=1)
    nz, ny, nx = co.shape[1:]
    pts = [rng.uniform(lo, hi, (n, 3))]
    idx = rng.integers([3, 3, 3], [nz - 4, ny - 4, nx - 4], (n // 4, 3))
    pts.append(co[:, idx[:, 0], idx[:, 1], idx[:, 2]].T)                   # grid points
    nb = co[:, idx[:, 0], idx[:, 1], idx[:, 2] + 1].T
    pts.append(0.5 * (pts[-1] + nb))                                       # edge midpoints
    w = rng.uniform(0.05, 0.95, (n // 4, 3))
    c0 = co[:, idx[:, 0], idx[:, 1], idx[:, 2]].T
    pts.append(c0 + w * (co[:, idx[:, 0] + 1, idx[:, 1] + 1, idx[:, 2] + 1].T - c0))
    return np.concatenate(pts).astype(np.float32)


@pytest.mark.parametrize("shape,seed", [((30, 26, 20), 1), ((45, 38, 27), 2)])
def test_locate_points_match_oracle(shape, seed):
    ni, nj, nk = shape
    co = cs.curvilinear_coords(ni, nj, nk, seed=seed, hill=0.3, wiggle=0.1)
    pts = _points(co, 120, seed)
    here_o, ijk_o, inc_o = po.locate_points(co, ni, nj, nk, pts)
    here_g, ijk_g, inc_g = locate_points(co, pts)
    print(f"{shape}: {int(here_o.sum())} of {len(pts)} points located", flush=True)
    assert here_o.sum() > len(pts) // 3
    assert np.array_equal(here_g, here_o)
    assert np.array_equal(ijk_g[here_o == 1], ijk_o[here_o == 1])
    assert np.array_equal(inc_g[here_o == 1], inc_o[here_o == 1])
    assert np.all(ijk_g[here_g == 0] == -1)


def test_depth_to_axis_matches_oracle():
    ni, nj, nk = 40, 33, 12
    co = cs.curvilinear_coords(ni, nj, nk, seed=4, hill=0.35, wiggle=0.12)
    rng = np.random.default_rng(5)
    flat = co.reshape(3, -1)
    lo, hi = flat.min(axis=1), flat.max(axis=1)
    xyd = np.stack([rng.uniform(lo[0] - 50, hi[0] + 50, 300),
                    rng.uniform(lo[1] - 50, hi[1] + 50, 300),
                    rng.uniform(0, 900, 300)], axis=1).astype(np.float32)
    # surface grid points and surface-cell diagonals (both triangles touch them)
    k = 3 + nk - 1
    xyd[:20, :2] = co[:2, k, 10, 5:25].T
    xyd[20:40, :2] = 0.5 * (co[:2, k, 12, 6:26] + co[:2, k, 13, 7:27]).T
    f_o, z_o = po.depth_to_axis(co, ni, nj, nk, xyd)
    f_g, z_g = depth_to_axis(co, xyd)
    assert f_o.sum() > 150
    assert np.array_equal(f_g, f_o)
    assert np.array_equal(z_g, z_o)


def test_locate_on_the_benchmark_grid():
    from cgfd3d_amd import workloads
    wl = workloads.iso_topo_grf(ni=600, nj=600, nk=300, nt_total=10)
    co = wl.xyz
    rng = np.random.default_rng(6)
    nz, ny, nx = co.shape[1:]
    idx = rng.integers([3, 3, 3], [nz - 4, ny - 4, nx - 4], (4096, 3))
    c0 = co[:, idx[:, 0], idx[:, 1], idx[:, 2]].T
    c1 = co[:, idx[:, 0] + 1, idx[:, 1] + 1, idx[:, 2] + 1].T
    pts = (c0 + rng.uniform(0.05, 0.95, (4096, 3)) * (c1 - c0)).astype(np.float32)
    t = time.perf_counter()
    here_g, ijk_g, inc_g = locate_points(co, pts)
    el = time.perf_counter() - t
    print(f"600x600x300: {len(pts)} points located in {el:.2f} s "
          f"({int(here_g.sum())} found)", flush=True)
    sub = slice(0, 6)
    here_o, ijk_o, inc_o = po.locate_points(co, 600, 600, 300, pts[sub])
    assert np.array_equal(here_g[sub], here_o)
    assert np.array_equal(ijk_g[sub], np.where(here_o[:, None] == 1, ijk_o, -1))
    assert np.array_equal(inc_g[sub][here_o == 1], inc_o[here_o == 1])
    assert here_g.mean() > 0.95
