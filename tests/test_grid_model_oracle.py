"""CPU checks that pin the oracle's grid-model discretisation
(oracle/media_oracle.c orc_grid2model_grid) against the reference's own
media_grid2model_onecmp / _el_iso / _el_vti / _el_aniso, compiled from
media/*.cpp into oracle/_ref, on .md3grd files written here: three layers
of 3, 4 and 3 surfaces for every media type, "loc", "har" and "ari" on a
curvilinear grid with points exactly on a layer boundary.  The oracle
reads the mesh window the library's reader (cgfdh_io_read_md3grd) cuts
with the reference's float arithmetic, so the comparison pins the reader
too.  A closed form pins the averaging where a layer boundary cuts a cell:
the sub-points on the boundary are left out of the count."""
import ctypes as C

import numpy as np
import pytest

from cgfd3d_amd import output as out
from oracle import pyoracle as po
from tests import media_models as mm

needs_ref = pytest.mark.skipif(not po.ref_available(), reason="reference build absent")

GRID_FN = {mm.ONECMP: "media_grid2model_onecmp", mm.EL_ISO: "media_grid2model_el_iso",
           mm.EL_VTI: "media_grid2model_el_vti", mm.EL_ANISO: "media_grid2model_el_aniso"}
CASES = ([("one_component", mm.ONECMP), ("elastic_isotropic", mm.EL_ISO),
          ("elastic_isotropic", mm.EL_ANISO)] +
         [(k, mm.EL_VTI) for k in ("elastic_vti_prem", "elastic_vti_thomsen", "elastic_vti_cij")] +
         [(k, mm.EL_ANISO) for k in ("elastic_vti_cij", "elastic_tti_thomsen", "elastic_tti_bond",
                                     "elastic_aniso_cij")])


def _ref_run(target, path, co, method):
    ref = po.load_ref()
    nz, ny, nx = co.shape[1:]
    xyz = [np.ascontiguousarray(co[d]) for d in range(3)]
    nout = mm.NOUT[target]
    outs = [np.zeros(xyz[0].shape, np.float32) for _ in range(nout)]
    fn = getattr(ref, GRID_FN[target])
    fn.argtypes = ([C.c_void_p] * (nout + 3) + [C.c_int] * 3 + [C.c_float] * 4 +
                   [C.c_int, C.c_char_p, C.c_char_p, C.c_int])
    fn(*[o.ctypes.data for o in outs], *[a.ctypes.data for a in xyz], nx, ny, nz, *mm.extent(co),
       3, path.encode(), method.encode(), 1)
    return outs


def _case(kind, tmp_path, seed=3):
    code, ngz, f, geo = mm.grid_model(kind, seed=seed)
    path = str(tmp_path / "m.md3grd")
    mm.write_md3grd(path, kind, ngz, f, geo)
    co = mm.grid_points_on(f[0, ngz[0]], geo)        # a row on the second layer's top
    return code, ngz, f, geo, path, co


@needs_ref
@pytest.mark.parametrize("kind,target", CASES)
@pytest.mark.parametrize("method", ["loc", "har", "ari"])
def test_grid2model_matches_reference(kind, target, method, tmp_path):
    code, ngz, f, geo, path, co = _case(kind, tmp_path)
    want = _ref_run(target, path, co, method)
    mt, nprop, gz, fc, gc = out.read_md3grd(path, *mm.extent(co))
    assert mt == code and gz == ngz
    rc, got = mm.oracle_grid2model(code, target, method, gz, fc, gc, co)
    assert rc == 0
    for q, (a, b) in enumerate(zip(got, want)):
        assert np.array_equal(a, b), (kind, target, method, q)
    if method != "loc":
        _, loc = mm.oracle_grid2model(code, target, "loc", gz, fc, gc, co)
        assert sum(int((a != b).sum()) for a, b in zip(got, loc)) > 0


def test_reader_window(tmp_path):
    """The kept window and its origin follow the reference's arithmetic:
    ix0 = (xmin - MINX) / DX truncated, ix1 = ceil((xmax - MINX) / DX)."""
    code, ngz, f, geo = mm.grid_model("elastic_isotropic", seed=4)
    path = str(tmp_path / "m.md3grd")
    mm.write_md3grd(path, "elastic_isotropic", ngz, f, geo)
    xmin, xmax, ymin, ymax = -130.0, 610.0, -250.0, 900.0
    mt, nprop, gz, fc, gc = out.read_md3grd(path, xmin, xmax, ymin, ymax)
    ix0 = int(np.float32(xmin - geo[0]) / geo[2])
    iy0 = int(np.float32(ymin - geo[1]) / geo[3])
    ix1 = int(np.ceil(np.float32(np.float32(xmax - geo[0]) / geo[2])))
    iy1 = int(np.ceil(np.float32(np.float32(ymax - geo[1]) / geo[3])))
    assert fc.shape == (4, sum(ngz), iy1 - iy0 + 1, ix1 - ix0 + 1)
    assert np.array_equal(fc, f[:, :, iy0:iy1 + 1, ix0:ix1 + 1])
    assert gc[0] == np.float32(np.float32(ix0) * geo[2] + geo[0])
    assert gc[1] == np.float32(np.float32(iy0) * geo[3] + geo[1])
    with pytest.raises(out.OutputError, match="does not cover"):
        out.read_md3grd(path, xmin, geo[0] + geo[2] * 40, ymin, ymax)


def test_boundary_subpoints_left_out():
    """Two layers of one flat surface each, the boundary at z = 0 cutting a
    cell in half: the sub-points on the boundary (a ninth of them) are left
    out, so har and ari are the harmonic and arithmetic means over the
    eight remaining planes of sub-points, four below and four above."""
    mx = my = 8
    f = np.zeros((2, 2, my, mx), np.float32)
    f[0, 0] = 1000.0                  # layer 0: one surface at z = 1000
    f[0, 1] = 0.0                     # layer 1: one surface at z = 0, the boundary
    f[1, 0] = 2.0
    f[1, 1] = 5.0
    geo = (np.float32(0.0), np.float32(0.0), np.float32(100.0), np.float32(100.0))
    nx = ny = nz = 5
    z = np.arange(nz, dtype=np.float32)[:, None, None] * 10.0 - 20.0 + np.zeros((nz, ny, nx), np.float32)
    x = np.arange(nx, dtype=np.float32)[None, None, :] * 10.0 + 200.0 + np.zeros((nz, ny, nx), np.float32)
    y = np.arange(ny, dtype=np.float32)[None, :, None] * 10.0 + 200.0 + np.zeros((nz, ny, nx), np.float32)
    co = np.stack([x, y, z]).astype(np.float32)
    # between the surfaces values interpolate linearly from 2 (z = 1000) to
    # 5 (z = 0); below the last surface 5.  The point at z = 0: sub-points at
    # z = -5 + 10 k / 8; k < 4 below (5), k == 4 on the boundary (left out),
    # k > 4 above: 2 + 3 (1000 - z) / 1000
    above = [2.0 + 3.0 * (1000.0 - (-5.0 + 10.0 * k / 8)) / 1000.0 for k in range(5, 9)]
    vals = [5.0] * 4 + above
    har_want = 8 / sum(1.0 / v for v in vals)
    ari_want = sum(vals) / 8
    for method, want in (("har", har_want), ("ari", ari_want)):
        rc, res = mm.oracle_grid2model(0, mm.ONECMP, method, [1, 1], f, geo, co)
        assert rc == 0
        got = float(res[0][2, 2, 2])
        assert abs(got - want) < 1e-5 * want, (method, got, want)
