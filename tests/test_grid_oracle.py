"""CPU checks that pin the oracle's grid generation from a layer file
(oracle/grid_oracle.c restating gd_curv_gen_layer, gd_grid_z_interp and the
gd_SPLine/gd_seval cubic spline of gd_t.c, a file that needs netCDF and is
not compiled here): the spline against the published Forsythe-Malcolm-Moler
method (Computer Methods for Mathematical Computations, 1977, `spline` /
`seval`) computed independently in float64, and its exactness on cubics; the
vertical spacing by its closed form for equal layers and its invariants for
the graded (unequal) layers; whole grids from flat and tilted interfaces
against the Cartesian and planar grids they must give, with x/y resampling
up and down; the mirrored ghost points; the reference's parameter checks."""
import numpy as np
import pytest

from oracle import pyoracle as po


def _fmm64(x, y, u):
    """Forsythe-Malcolm-Moler cubic spline (end conditions from the cubics
    through the first and last four points) evaluated at u, in float64."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = len(x)
    h = np.diff(x)
    if n == 2:
        s = (y[1] - y[0]) / h[0]
        return y[0] + s * (np.asarray(u, np.float64) - x[0])
    # tridiagonal system for sigma (second derivative / 6 in FMM's scaling)
    A = np.zeros((n, n))
    r = np.zeros(n)
    dd = np.diff(y) / h
    for i in range(1, n - 1):
        A[i, i - 1], A[i, i], A[i, i + 1] = h[i - 1], 2 * (h[i - 1] + h[i]), h[i]
        r[i] = dd[i] - dd[i - 1]
    A[0, 0], A[0, 1] = -h[0], h[0]
    A[n - 1, n - 2], A[n - 1, n - 1] = h[n - 2], -h[n - 2]
    if n > 3:
        d3a = (dd[2] - dd[1]) / (x[3] - x[1]) - (dd[1] - dd[0]) / (x[2] - x[0])
        d3b = (dd[n - 2] - dd[n - 3]) / (x[n - 1] - x[n - 3]) - \
              (dd[n - 3] - dd[n - 4]) / (x[n - 2] - x[n - 4])
        r[0] = d3a * h[0] ** 2 / (x[3] - x[0])
        r[n - 1] = -d3b * h[n - 2] ** 2 / (x[n - 1] - x[n - 4])
    sig = np.linalg.solve(A, r)
    b = dd - h * (sig[1:] + 2 * sig[:-1])
    c = 3 * sig[:-1]
    d = (sig[1:] - sig[:-1]) / h
    u = np.asarray(u, np.float64)
    i = np.clip(np.searchsorted(x, u, side="right") - 1, 0, n - 2)
    w = u - x[i]
    return y[i] + w * (b[i] + w * (c[i] + w * d[i]))


def _spline_at(x, y, u):
    lib = po.load_oracle()
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    out = np.zeros(len(u), np.float32)
    rc = lib.orc_spline_at(len(x), po.ptr(x), po.ptr(y), len(u), po.ptr(u), po.ptr(out))
    return rc, out


@pytest.mark.parametrize("n", [2, 3, 4, 7, 30])
def test_spline_matches_fmm_in_float64(n):
    rng = np.random.default_rng(n)
    x = np.cumsum(rng.uniform(0.5, 2.0, n)).astype(np.float32)
    y = (np.sin(x) * 100 + rng.standard_normal(n)).astype(np.float32)
    u = np.sort(rng.uniform(x[0] - 0.5, x[-1] + 0.5, 200)).astype(np.float32)
    u[:n] = x                                   # the knots themselves too
    u = np.sort(u)
    rc, got = _spline_at(x, y, u)
    assert rc == 0
    want = _fmm64(x, y, u)
    scale = np.abs(y).max()
    assert np.abs(got - want).max() <= 2e-5 * scale * max(1, n / 4), np.abs(got - want).max()


def test_spline_reproduces_cubics():
    rng = np.random.default_rng(3)
    x = np.cumsum(rng.uniform(0.5, 1.5, 12)).astype(np.float32)
    p = np.poly1d([0.3, -1.2, 2.0, 5.0])
    u = np.linspace(x[0], x[-1], 300).astype(np.float32)
    rc, got = _spline_at(x, p(x.astype(np.float64)).astype(np.float32), u)
    assert rc == 0
    assert np.allclose(got, p(u.astype(np.float64)), rtol=1e-5, atol=1e-3)


def test_spline_refuses_unordered_knots():
    rc, _ = _spline_at([0, 2, 1, 3], [0, 1, 2, 3], [0.5])
    assert rc == 2


def _grid_z(zl, ncell, equal):
    lib = po.load_oracle()
    zl = np.ascontiguousarray(zl, np.float32)
    nc = np.ascontiguousarray(ncell, np.int32)
    eq = np.ascontiguousarray(equal, np.int32)
    z = np.zeros(int(nc.sum()) + 1, np.float32)
    lib.orc_grid_z(po.ptr(z), po.ptr(zl), po.ptr(nc), po.ptr(eq), len(nc))
    return z


def test_vertical_spacing_equal_layers():
    zl = np.array([-5000.0, -3000.0, -1200.0, 0.0], np.float32)
    nc = [10, 12, 8]
    z = _grid_z(zl, nc, [1, 1, 1])
    s = 0
    for i, n in enumerate(nc):
        dz = np.float32(np.float32(zl[i + 1] - zl[i]) / np.float32(n))
        for m in range(n):
            want = zl[i] if m == 0 else np.float32(zl[i] + np.float32(m) * dz)
            assert z[s + m] == want
        s += n
    assert z[-1] == zl[-1]


@pytest.mark.parametrize("zl,nc", [([-6000.0, -4000.0, -3000.0, -2500.0, 0.0], [20, 20, 20, 25]),
                                   ([-3000.0, -2500.0, -1375.0, 0.0], [10, 15, 10])])
def test_vertical_spacing_graded_layer(zl, nc):
    """A middle layer marked unequal between a coarser and a finer neighbour
    (either way round): it starts and ends on its interfaces, its points
    rise strictly inside it, its first spacing is close to the lower
    neighbour's and differs from its own equal spacing (the graded branch
    ran), and every other layer keeps equal spacing.  (The reference's
    grading can fold over for strong contrasts; these cases are in the
    range it is meant for.)"""
    zl = np.array(zl, np.float32)
    eq = [1] * len(nc)
    eq[1] = 0
    z = _grid_z(zl, nc, eq)
    dzs = np.diff(zl.astype(np.float64)) / np.array(nc)
    a, b = int(np.sum(nc[:1])), int(np.sum(nc[:2]))
    assert z[a] == zl[1] and z[b] == zl[2]
    sp = np.diff(z[a:b + 1].astype(np.float64))
    assert np.all(sp > 0)
    assert abs(sp[0] - dzs[0]) < 0.1 * dzs[0]
    assert np.abs(sp - dzs[1]).max() > 0.05 * dzs[1]
    assert np.allclose(np.diff(z[:a + 1]), dzs[0], rtol=1e-4)
    assert np.allclose(np.diff(z[b:]), np.repeat(dzs[2:], nc[2:]), rtol=1e-4)


def _layers(nxl, nyl, dx, zs, tilt=(0.0, 0.0)):
    n_if = len(zs)
    X, Y = np.meshgrid(np.arange(nxl) * dx, np.arange(nyl) * dx)
    L = np.zeros((3, n_if, nyl, nxl), np.float32)
    for k, z0 in enumerate(zs):
        L[0, k], L[1, k] = X, Y
        L[2, k] = z0 + tilt[0] * X + tilt[1] * Y
    return L


@pytest.mark.parametrize("resample", [(1, 1, 1), (2, 2, 1), (-2, -2, 1), (2, 1, 2), (1, 3, -2)])
def test_flat_layers_give_the_cartesian_grid(resample):
    ni, nj, nk = 12, 10, 8
    dx = 100.0
    f = resample
    # input sample spacing chosen so the resampled spacing is dx (a positive
    # factor splits every input interval, a negative one keeps every |f|-th)
    sx = dx * f[0] if f[0] > 0 else dx / -f[0]
    sy = dx * f[1] if f[1] > 0 else dx / -f[1]
    nxl = 40 if f[0] > 0 else 80
    nyl = 40 if f[1] > 0 else 80
    L = _layers(nxl, nyl, sx, [-2000.0, -1000.0, 0.0])
    if f[1] != f[0]:
        L[1] = L[1] * (sy / sx)
    ncell = [8, 4] if f[2] > 0 else [16, 8]
    start = (6, 6, 0)
    rc, co = po.gen_layer(ncell, [1, 1], L, resample, start, ni, nj, nk)
    assert rc == 0
    xs, ys, zs = co
    # x and y follow the resampled sample index, z the layers' equal spacing
    dxg = np.diff(xs[5, 5, :]); dyg = np.diff(ys[5, :, 5])
    assert np.allclose(dxg, dx, rtol=1e-5) and np.allclose(dyg, dx, rtol=1e-5)
    assert np.allclose(xs[:, :, 4], xs[0, 0, 4], atol=1e-2)
    top = 3 + nk - 1
    ntop = ncell[1] * f[2] if f[2] > 0 else ncell[1] // -f[2]   # top-layer cells, resampled
    assert np.allclose(zs[top], 0.0, atol=1e-3)
    col = zs[:top + 1, 5, 5].astype(np.float64)
    assert np.allclose(np.diff(col[-ntop - 1:]), 1000.0 / ntop, rtol=1e-5)
    assert np.allclose(np.diff(col[:-ntop]), 1000.0 / (2 * ntop), rtol=1e-5)
    for k in range(1, 4):                       # ghosts mirrored above the top
        assert np.array_equal(zs[top + k], zs[top] * np.float32(2) - zs[top - k])


def test_tilted_layers_give_the_planar_grid():
    ni, nj, nk = 10, 9, 6
    L = _layers(30, 30, 100.0, [-3000.0, -1500.0, 0.0], tilt=(0.05, -0.02))
    rc, co = po.gen_layer([6, 5], [1, 1], L, (1, 1, 1), (5, 5, 0), ni, nj, nk)
    assert rc == 0
    top = 3 + nk - 1
    want = 0.05 * co[0, top] - 0.02 * co[1, top]
    assert np.allclose(co[2, top], want, atol=1e-2)


def test_parameter_checks():
    L = _layers(20, 20, 100.0, [-1000.0, 0.0])
    assert po.gen_layer([8], [1], L, (1, 1, 1), (2, 5, 0), 10, 10, 6)[0] == -2   # x start
    assert po.gen_layer([8], [1], L, (1, 1, 1), (5, 5, 0), 14, 10, 6)[0] == -3   # too wide
    assert po.gen_layer([4], [1], L, (1, 1, 1), (5, 5, 0), 10, 10, 6)[0] == -4   # too deep
