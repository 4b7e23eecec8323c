"""CPU checks that pin the oracle's binary-volume discretisation
(oracle/media_oracle.c orc_bin_window, orc_bin2model_grid) against the
reference's own media_bin2model_el_iso and media_bin2model_vis_iso,
compiled from media/*.cpp into oracle/_ref, on float32 volumes written
here in three storage orders, with z increasing and decreasing, and a
curvilinear grid that reaches past the model (clamped to its edge).  The
oracle samples the window the library's reader (cgfdh_io_read_bin_window)
cuts, so the comparison pins the reader too.  A closed form pins the
interpolation: a field linear in x, y and z is reproduced inside the mesh."""
import ctypes as C

import numpy as np
import pytest

from cgfd3d_amd import output as out
from oracle import pyoracle as po
from tests import cases as cs
from tests import media_models as mm

needs_ref = pytest.mark.skipif(not po.ref_available(), reason="reference build absent")

ORDERS = [(2, 0, 1), (0, 1, 2), (1, 2, 0)]


def bin_grid(ni=20, nj=17, nk=15, seed=2):
    co = cs.curvilinear_coords(ni, nj, nk, dh=40.0, seed=seed, hill=0.2, wiggle=0.1)
    for d, (lo, hi) in enumerate(((-100.0, 1500.0), (-200.0, 1400.0), (-1750.0, 350.0))):
        c = co[d]
        co[d] = lo + (c - c.min()) / (c.max() - c.min()) * (hi - lo)
    return co.astype(np.float32)


def _ref_run(nvol, paths, co, order, size, spacing, origin):
    ref = po.load_ref()
    nz, ny, nx = co.shape[1:]
    xyz = [np.ascontiguousarray(co[d]) for d in range(3)]
    outs = [np.zeros(xyz[0].shape, np.float32) for _ in range(nvol)]
    o, s = (np.ascontiguousarray(a, np.int32) for a in (order, size))
    sp, org = (np.ascontiguousarray(a, np.float32) for a in (spacing, origin))
    fn = ref.media_bin2model_el_iso if nvol == 3 else ref.media_bin2model_vis_iso
    fn.argtypes = ([C.c_void_p] * (nvol + 3) + [C.c_size_t] * 3 + [C.c_float] * 4 + [C.c_int] +
                   [C.c_void_p] * 4 + [C.c_char_p] * nvol)
    fn(*[a.ctypes.data for a in outs], *[a.ctypes.data for a in xyz], nx, ny, nz, *mm.extent(co),
       3, o.ctypes.data, s.ctypes.data, sp.ctypes.data, org.ctypes.data,
       *[p.encode() for p in paths])
    return outs


def _window_reader(paths, order, size):
    def read(start, end):
        return np.stack([out.read_bin_window(p, order, size, start, end) for p in paths])
    return read


@needs_ref
@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("dz", [-70.0, 70.0])
def test_bin2model_matches_reference(order, dz, tmp_path):
    origin = (-200.0, -300.0, 300.0 if dz < 0 else -1660.0)
    vols, stored, size, sp, org = mm.bin_model(order, seed=3, spacing=(50.0, 60.0, dz),
                                               origin=origin)
    paths = mm.write_bins(tmp_path, stored)
    co = bin_grid()
    for nvol, ps in ((3, paths[:3]), (2, paths[3:])):
        want = _ref_run(nvol, ps, co, order, size, sp, org)
        rc, got, win = mm.oracle_bin(order, size, sp, org, mm.extent(co),
                                     _window_reader(ps, order, size), co, nvol)
        assert rc == 0
        for q, (a, b) in enumerate(zip(got, want)):
            assert np.array_equal(a, b), (order, dz, nvol, q)
    # the window is narrower than the model in x and y
    start, end = win[0], win[1]
    assert (end - start + 1 < np.array(size)).sum() >= 2


def test_linear_field_reproduced(tmp_path):
    order = (2, 0, 1)
    vols, stored, size, sp, org = mm.bin_model(order, seed=4)
    nx, ny, nz = 40, 33, 29
    zz, yy, xx = np.meshgrid(org[order.index(2)] + sp[order.index(2)] * np.arange(nz),
                             org[order.index(1)] + sp[order.index(1)] * np.arange(ny),
                             org[order.index(0)] + sp[order.index(0)] * np.arange(nx),
                             indexing="ij")
    lin = (1000.0 + 0.3 * xx - 0.2 * yy + 0.5 * zz).astype(np.float32)
    vol = np.stack([lin, lin]).astype(np.float32)
    co = bin_grid()
    inside = ((co[0] >= -200.0) & (co[0] <= -200.0 + 50.0 * 39) & (co[1] >= -300.0) &
              (co[1] <= -300.0 + 60.0 * 32) & (co[2] <= 300.0) & (co[2] >= 300.0 - 70.0 * 28))
    rc, res, win = mm.oracle_bin(order, size, sp, org, mm.extent(co),
                                 lambda s, e: vol[:, :, s[2]:e[2] + 1, s[1]:e[1] + 1], co, 2)
    assert rc == 0
    want = 1000.0 + 0.3 * co[0] - 0.2 * co[1] + 0.5 * co[2]
    # float weights and differences: a few 1e-7 of the field's range (~3000)
    assert np.allclose(res[0][inside], want[inside], rtol=0, atol=1e-3)
    assert not np.allclose(res[0][inside], want[inside] + 0.05, rtol=0, atol=1e-3)
    # outside the mesh the value is the nearest edge sample's plane: still
    # linear in the coordinates that stay inside
    assert inside.sum() > 0 and (~inside).sum() > 0


def test_reader_window_layout(tmp_path):
    """The window reader keeps start..end of each stored axis and lays the
    samples out x fastest, whatever the storage order."""
    for order in ORDERS:
        vols, stored, size, sp, org = mm.bin_model(order, seed=5, n=(9, 7, 5))
        paths = mm.write_bins(tmp_path, stored[:1])
        start = [1, 2, 0]
        end = [size[0] - 2, size[1] - 1, size[2] - 2]
        got = out.read_bin_window(paths[0], order, size, start, end)
        s3 = [0, 0, 0]; e3 = [0, 0, 0]
        for a in range(3):
            s3[order[a]] = start[a]; e3[order[a]] = end[a]
        want = vols[0][s3[2]:e3[2] + 1, s3[1]:e3[1] + 1, s3[0]:e3[0] + 1]
        assert np.array_equal(got, want), order
