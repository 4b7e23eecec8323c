"""GPU: the stable-time-step estimate (blk_dt_esti_curv) against the oracle
on the same curvilinear grid and media; bit-exact, including the point."""
import numpy as np
import pytest

from tests import cases as cs
from tests.test_dt_oracle import dt_estimate as oracle_dt
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

NI, NJ, NK = 24, 21, 17


def _vols(medium):
    if medium == "iso":
        lam, mu, slw = cs.iso_medium(NI, NJ, NK, seed=2)
        return po.MEDIUM_ISO, np.stack([lam, mu, (1.0 / slw).astype(np.float32)]), 0
    if medium == "vti":
        v = cs.vti_medium(NI, NJ, NK, seed=2)
        return po.MEDIUM_VTI, np.stack([v[0], v[2], (1.0 / v[5]).astype(np.float32)]), 0
    if medium == "aniso":
        v = cs.aniso_medium(NI, NJ, NK, seed=2)
        return po.MEDIUM_ANISO, np.stack([v[0], (1.0 / v[21]).astype(np.float32)]), 0
    v, wl = cs.visco_medium(NI, NJ, NK, seed=2)
    rho = (1.0 / v[2]).astype(np.float32)
    return po.MEDIUM_VISCO, np.concatenate([v[:2], rho[None], v[3:]]), len(wl)


def _bits(x):
    return np.float32(x).view(np.uint32)


@pytest.mark.parametrize("medium", ["iso", "vti", "aniso", "visco"])
def test_dt_matches_oracle(medium):
    from cgfd3d_amd import estimate_dt
    xyz = cs.curvilinear_coords(NI, NJ, NK, seed=3, hill=0.4)
    mt, vols, nmw = _vols(medium)
    got = estimate_dt(xyz, mt, vols, 1.3, nmaxwell=nmw)
    want = oracle_dt(xyz, medium, vols, 1.3, nmaxwell=nmw)
    print(medium, got, want, flush=True)
    assert 0 < want[0] < 1.0
    assert [_bits(v) for v in got[:3]] == [_bits(v) for v in want[:3]]
    assert got[3] == want[3]


def test_dt_ties_pick_the_first_point():
    """A Cartesian homogeneous grid: many points share the minimum; the GPU
    keeps the one the reference's k-j-i loop keeps."""
    from cgfd3d_amd import estimate_dt
    nx, ny, nz = NI + 6, NJ + 6, NK + 6
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    xyz = np.stack([x, y, z]).astype(np.float32) * np.float32(64.0)   # exact coordinates
    shp = xyz.shape[1:]
    vols = np.stack([np.full(shp, 3.0e10, np.float32), np.full(shp, 2.5e10, np.float32),
                     np.full(shp, 2500.0, np.float32)])
    got = estimate_dt(xyz, po.MEDIUM_ISO, vols, 1.0)
    want = oracle_dt(xyz, "iso", vols, 1.0)
    assert want[3] == (3, 3, 3)
    assert got[3] == want[3] and _bits(got[0]) == _bits(want[0])


def test_dt_rejects_bad_arguments():
    from cgfd3d_amd import CgfdhError, estimate_dt
    xyz = cs.curvilinear_coords(8, 8, 8)
    mt, vols, _ = po.MEDIUM_ISO, cs.iso_medium(8, 8, 8), 0
    with pytest.raises(CgfdhError):
        estimate_dt(xyz, mt, vols, 0.0)
    with pytest.raises(CgfdhError):
        estimate_dt(xyz, 7, vols, 1.0)
    with pytest.raises(CgfdhError):
        estimate_dt(xyz, po.MEDIUM_VISCO, vols, 1.0, nmaxwell=0)
