"""GPU: the grid model (.md3grd) on a curvilinear grid for every media type
and equivalent-medium method (cgfdh_grid2model_grid and the
reference-shaped file entry points cgfdh_media_grid2model_*) against the
oracle (oracle/media_oracle.c, equal to the reference's
media_grid2model_* bit for bit by tests/test_grid_model_oracle.py), on the
mesh window the library's reader cuts.  Values interpolate linearly, and
only the TTI types take cos and sin (in double on both sides), so the bar
is 1e-6 of each value, 1e-5 of the array's largest magnitude for the
rotated types; the runs print how many values differ at all.  A cell cut
by a layer boundary, whose boundary sub-points the averages leave out, is
checked on its own."""
import numpy as np
import pytest

from cgfd3d_amd import CgfdhError, grid2model_grid, media_grid2model
from cgfd3d_amd import output as out
from tests import media_models as mm

pytestmark = pytest.mark.gpu

CASES = ([("one_component", mm.ONECMP), ("elastic_isotropic", mm.EL_ISO),
          ("elastic_isotropic", mm.EL_ANISO)] +
         [(k, mm.EL_VTI) for k in ("elastic_vti_prem", "elastic_vti_thomsen", "elastic_vti_cij")] +
         [(k, mm.EL_ANISO) for k in ("elastic_vti_cij", "elastic_tti_thomsen", "elastic_tti_bond",
                                     "elastic_aniso_cij")])
ROTATED = ("elastic_tti_thomsen", "elastic_tti_bond")


def _close(kind, got, want):
    nd = 0
    for q, (a, b) in enumerate(zip(got, want)):
        if kind in ROTATED:
            scale = float(np.max(np.abs(b))) or 1.0
            assert np.all(np.abs(a - b) <= 1e-5 * scale), q
        else:
            assert np.allclose(a, b, rtol=1e-6, atol=0), q
        nd += int((a != b).sum())
    return nd


@pytest.mark.parametrize("kind,target", CASES)
@pytest.mark.parametrize("method", ["loc", "har", "ari"])
def test_grid2model_matches_oracle(kind, target, method, tmp_path):
    code, ngz, f, geo = mm.grid_model(kind, seed=7)
    path = str(tmp_path / "m.md3grd")
    mm.write_md3grd(path, kind, ngz, f, geo)
    co = mm.grid_points_on(f[0, ngz[0]], geo, ni=30, nj=26, nk=28, seed=4)
    ext = mm.extent(co)
    mt, nprop, gz, fc, gc = out.read_md3grd(path, *ext)
    rc, want = mm.oracle_grid2model(code, target, method, gz, fc, gc, co)
    assert rc == 0
    got = grid2model_grid(mt, target, method, gz, fc, gc, co)
    nd = _close(kind, got, want)
    print(f"{kind} target {target} {method}: {nd} of {sum(a.size for a in got)} values differ",
          flush=True)
    got2 = media_grid2model(target, path, co, ext, method)
    for a, b in zip(got, got2):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("method", ["har", "ari"])
def test_boundary_subpoints_left_out(method):
    """Two layers of one flat surface each, the boundary at z = 0 through
    a row of cells: GPU and oracle agree where a ninth of the sub-points lie
    on the boundary, and one layer alone averages nothing."""
    mx = my = 8
    f = np.zeros((2, 2, my, mx), np.float32)
    f[0, 0], f[0, 1], f[1, 0], f[1, 1] = 1000.0, 0.0, 2.0, 5.0
    geo = (np.float32(0.0), np.float32(0.0), np.float32(100.0), np.float32(100.0))
    n = 6
    k = np.arange(n, dtype=np.float32)[:, None, None]
    z = k * 10.0 - 20.0 + np.zeros((n, n, n), np.float32)
    x = np.arange(n, dtype=np.float32)[None, None, :] * 10.0 + 200.0 + np.zeros((n, n, n), np.float32)
    y = np.arange(n, dtype=np.float32)[None, :, None] * 10.0 + 200.0 + np.zeros((n, n, n), np.float32)
    co = np.stack([x, y, z]).astype(np.float32)
    rc, want = mm.oracle_grid2model(0, mm.ONECMP, method, [1, 1], f, geo, co)
    assert rc == 0
    got = grid2model_grid(0, mm.ONECMP, method, [1, 1], f, geo, co)
    assert np.array_equal(got[0], want[0])
    loc = grid2model_grid(0, mm.ONECMP, "loc", [1, 1], f, geo, co)
    assert not np.array_equal(got[0], loc[0])
    one = grid2model_grid(0, mm.ONECMP, method, [2], f, geo, co)
    assert np.array_equal(one[0], grid2model_grid(0, mm.ONECMP, "loc", [2], f, geo, co)[0])


def test_grid2model_refusals(tmp_path):
    code, ngz, f, geo = mm.grid_model("elastic_vti_prem", seed=9)
    path = str(tmp_path / "m.md3grd")
    mm.write_md3grd(path, "elastic_vti_prem", ngz, f, geo)
    co = mm.grid_points_on(f[0, ngz[0]], geo)
    ext = mm.extent(co)
    with pytest.raises(CgfdhError, match="not supported"):
        media_grid2model(mm.EL_ISO, path, co, ext, "loc")
    with pytest.raises(CgfdhError, match="not provided"):
        media_grid2model(mm.EL_VTI, path, co, ext, "tti")
    with pytest.raises(CgfdhError, match="grid type"):
        media_grid2model(mm.EL_VTI, path, co, ext, "loc", grid_type=2)
    with pytest.raises(CgfdhError, match="does not cover"):
        media_grid2model(mm.EL_VTI, path, co, (ext[0], ext[1] + 5000.0, ext[2], ext[3]), "loc")
    off = co.copy()
    off[0, 3, 3, 3] = float(geo[0]) - 10.0
    with pytest.raises(CgfdhError, match="outside the interface mesh"):
        media_grid2model(mm.EL_VTI, path, off, mm.extent(co), "har")
    bad = f.copy()
    bad[0, ngz[0]] += 1.0                           # layers no longer meet
    mm.write_md3grd(path, "elastic_vti_prem", ngz, bad, geo)
    with pytest.raises(CgfdhError, match="not the first of layer"):
        media_grid2model(mm.EL_VTI, path, co, ext, "loc")
