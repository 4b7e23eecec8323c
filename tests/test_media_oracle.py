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
from tests import media_models as mm

needs_ref = pytest.mark.skipif(not po.ref_available(), reason="reference build absent")

GRID_CURV = 3


def md3lay_model(kind, seed=0, **kw):
    """(nprop, fields, geo) of a one-component or isotropic model
    (tests/media_models.py)."""
    code, f, geo = mm.model(kind, seed=seed, **kw)
    return (f.shape[0] - 1) // 3, f, geo


write_md3lay = mm.write_md3lay
grid_points = mm.grid_points


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
