"""GPU: binary volumes sampled at the grid points (cgfdh_bin2model_grid and
the reference-shaped cgfdh_media_bin2model_el_iso / _vis_iso) against the
oracle (oracle/media_oracle.c, equal to the reference's media_bin2model_*
bit for bit by tests/test_bin_model_oracle.py): three storage orders, z
increasing and decreasing, a grid reaching past the model.  Trilinear
interpolation and the lambda / mu conversion are the same float operations
on both sides: bit-identical."""
import numpy as np
import pytest

from cgfd3d_amd import CgfdhError, bin2model_grid, media_bin2model
from tests import media_models as mm
from tests.test_bin_model_oracle import ORDERS, _window_reader, bin_grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("dz", [-70.0, 70.0])
def test_bin2model_matches_oracle(order, dz, tmp_path):
    origin = (-200.0, -300.0, 300.0 if dz < 0 else -1660.0)
    vols, stored, size, sp, org = mm.bin_model(order, seed=6, spacing=(50.0, 60.0, dz),
                                               origin=origin)
    paths = mm.write_bins(tmp_path, stored)
    co = bin_grid(ni=34, nj=29, nk=26, seed=3)
    ext = mm.extent(co)
    for nvol, ps in ((3, paths[:3]), (2, paths[3:])):
        reader = _window_reader(ps, order, size)
        rc, want, win = mm.oracle_bin(order, size, sp, org, ext, reader, co, nvol)
        assert rc == 0
        got = media_bin2model(ps, co, ext, order, size, sp, org)
        for q, (a, b) in enumerate(zip(got, want)):
            assert np.array_equal(a, b), (order, dz, nvol, q)
        start, end, xv, yv, zv = win
        got2 = bin2model_grid(reader(start, end), xv, yv, zv, co)
        for a, b in zip(got2, want):
            assert np.array_equal(a, b)
    print(f"order {order} dz {dz}: bit-identical", flush=True)


def test_bin2model_refusals(tmp_path):
    order = (2, 0, 1)
    vols, stored, size, sp, org = mm.bin_model(order, seed=7)
    paths = mm.write_bins(tmp_path, stored)
    co = bin_grid()
    ext = mm.extent(co)
    with pytest.raises(CgfdhError, match="bin_order"):
        media_bin2model(paths[:3], co, ext, (2, 0, 0), size, sp, org)
    with pytest.raises(CgfdhError, match="in x"):
        media_bin2model(paths[:3], co, (5000.0, 6000.0, ext[2], ext[3]), order, size, sp, org)
    far = co.copy()
    far[2] -= 10000.0
    with pytest.raises(CgfdhError, match="in z"):
        media_bin2model(paths[:3], far, ext, order, size, sp, org)
    with pytest.raises(CgfdhError, match="insufficient data"):
        media_bin2model(paths[:3], co, ext, order, [size[0], size[1], size[2] + 5], sp, org)
    with pytest.raises(CgfdhError, match="grid type"):
        media_bin2model(paths[:3], co, ext, order, size, sp, org, grid_type=1)
