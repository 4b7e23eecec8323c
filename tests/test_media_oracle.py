"""CPU checks that pin the oracle's layered-model discretisation with local
values (oracle/media_oracle.c) against the reference's own
media_layer2model_el_iso and media_layer2model_onecmp ("loc"), compiled from
media/*.cpp into oracle/_ref, on .md3lay files written here (hilly
interfaces with gradients and power laws, grid points exactly on an
interface, points above the surface), and the library's own .md3lay reader
against the file contents."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import cases as cs

needs_ref = pytest.mark.skipif(not po.ref_available(), reason="reference build absent")

GRID_CURV = 3


def md3lay_model(kind, seed=0, mx=31, my=27, dx=50.0, dy=60.0, x0=-200.0, y0=-300.0):
    """Three interfaces (top with topography), per-layer values, gradients,
    powers: fields [nf][3][my][mx], geo (minx, miny, dx, dy)."""
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(x0 + dx * np.arange(mx), y0 + dy * np.arange(my))
    top = 150.0 * np.exp(-((X - X.mean()) ** 2 + (Y - Y.mean()) ** 2) / (2 * 400.0 ** 2))
    elev = np.stack([top + 5.0 * rng.standard_normal(X.shape), -600.0 + 0.05 * X,
                     -1400.0 + 0.02 * Y])
    nprop = 1 if kind == "one_component" else 3
    f = np.zeros((1 + 3 * nprop, 3, my, mx), np.float32)
    f[0] = elev
    base = [[2200.0, 2500.0, 2800.0], [3000.0, 4500.0, 5600.0], [1700.0, 2500.0, 3200.0]]
    grad = [0.2, 0.5, 0.28]                     # rho, vp, vs per metre (vp/vs stays > sqrt 2)
    for q in range(nprop):
        for L in range(3):
            f[1 + 3 * q, L] = base[q][L] * (1 + 0.02 * rng.standard_normal(X.shape))
            f[2 + 3 * q, L] = grad[q] * (1 + 0.05 * rng.standard_normal(X.shape))
            f[3 + 3 * q, L] = 1.0 + 0.04 * L + 0.01 * rng.standard_normal(X.shape)
    return nprop, f.astype(np.float32), (np.float32(x0), np.float32(y0), np.float32(dx),
                                         np.float32(dy))


def write_md3lay(path, kind, f, geo):
    nf, ni, my, mx = f.shape
    with open(path, "w") as fp:
        fp.write("# layered model written by the tests\n")
        fp.write(f"{kind}\n{ni}\n{mx} {my} {geo[0]:.9g} {geo[1]:.9g} {geo[2]:.9g} {geo[3]:.9g}\n")
        fp.write("# samples\n")
        flat = f.reshape(nf, -1)
        for p in range(flat.shape[1]):
            fp.write(" ".join(f"{v:.9g}" for v in flat[:, p]) + "\n")


def grid_points(f, geo, ni=18, nj=15, nk=16, seed=1):
    """A curvilinear grid inside the mesh from 1700 m deep to above the top,
    with one row of points exactly on the middle interface's samples."""
    co = cs.curvilinear_coords(ni, nj, nk, dh=40.0, seed=seed, hill=0.2, wiggle=0.1)
    co[0] += float(geo[0]) + 150.0
    co[1] += float(geo[1]) + 150.0
    z = co[2]
    co[2] = -1700.0 + (z - z.min()) / (z.max() - z.min()) * 2000.0
    # grid points placed on mesh samples of interface 1 (z exactly its elevation)
    for n in range(6):
        i, j = 4 + n, 3 + n
        co[0, 5, 4, 5 + n] = geo[0] + geo[2] * i
        co[1, 5, 4, 5 + n] = geo[1] + geo[3] * j
        co[2, 5, 4, 5 + n] = f[0, 1, j, i]
    return co.astype(np.float32)


def _ref_run(kind, path, co):
    ref = po.load_ref()
    nz, ny, nx = co.shape[1:]
    x, y, z = (np.ascontiguousarray(co[d]) for d in range(3))
    if kind == "one_component":
        var = np.zeros(x.shape, np.float32)
        fn = ref.media_layer2model_onecmp
        fn.argtypes = [C.c_void_p] * 4 + [C.c_size_t] * 3 + [C.c_int, C.c_char_p, C.c_char_p,
                                                            C.c_int]
        fn(var.ctypes.data, x.ctypes.data, y.ctypes.data, z.ctypes.data, nx, ny, nz, GRID_CURV,
           path.encode(), b"loc", 0)
        return [var]
    lam, mu, rho = (np.zeros(x.shape, np.float32) for _ in range(3))
    fn = ref.media_layer2model_el_iso
    fn.argtypes = [C.c_void_p] * 6 + [C.c_size_t] * 3 + [C.c_int, C.c_char_p, C.c_char_p, C.c_int]
    fn(lam.ctypes.data, mu.ctypes.data, rho.ctypes.data, x.ctypes.data, y.ctypes.data,
       z.ctypes.data, nx, ny, nz, GRID_CURV, path.encode(), b"loc", 0)
    return [lam, mu, rho]


def oracle_run(nprop, f, geo, co):
    lib = po.load_oracle()
    n = co[0].size
    x, y, z = (np.ascontiguousarray(co[d]).reshape(-1) for d in range(3))
    outs = [np.zeros(n, np.float32) for _ in range(3)]
    dims = np.array([f.shape[1], f.shape[3], f.shape[2]], np.int32)
    g = np.array(geo, np.float32)
    rc = lib.orc_layer2model(nprop, po.ptr(dims), po.ptr(g), po.ptr(np.ascontiguousarray(f)), n,
                             po.ptr(x), po.ptr(y), po.ptr(z), po.ptr(outs[0]), po.ptr(outs[1]),
                             po.ptr(outs[2]))
    return rc, [o.reshape(co[0].shape) for o in outs[:1 if nprop == 1 else 3]]


@needs_ref
@pytest.mark.parametrize("kind", ["elastic_isotropic", "one_component"])
def test_layer2model_loc_matches_reference(kind, tmp_path):
    nprop, f, geo = md3lay_model(kind, seed=3)
    path = str(tmp_path / "model.md3lay")
    write_md3lay(path, kind, f, geo)
    co = grid_points(f, geo)
    want = _ref_run(kind, path, co)
    rc, got = oracle_run(nprop, f, geo, co)
    assert rc == 0
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    # the interface points took the two layers' mean, points above the top
    # the top layer's values
    if nprop == 3:
        assert np.all(want[2] > 1000)


def test_md3lay_reader(tmp_path):
    from cgfd3d_amd import output as out
    for kind in ("elastic_isotropic", "one_component"):
        nprop, f, geo = md3lay_model(kind, seed=4)
        path = str(tmp_path / f"{kind}.md3lay")
        write_md3lay(path, kind, f, geo)
        np_, fields, g = out.read_md3lay(path)
        assert np_ == nprop
        assert np.array_equal(g, np.array(geo, np.float32))
        assert np.array_equal(fields, f)
