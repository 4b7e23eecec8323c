"""CPU checks that pin the oracle's layered-model discretisation for every
media type of the .md3lay file and every equivalent-medium method
(oracle/media_oracle.c orc_layer2model_grid) against the reference's own
media_layer2model_onecmp / _el_iso / _el_vti / _el_aniso, compiled from
media/*.cpp into oracle/_ref, on files written here: one component,
isotropic, the three VTI forms, both TTI forms (rotated by azimuth and dip)
and the full 21-coefficient medium, with "loc", "har" and "ari" on a
curvilinear grid that crosses two interfaces and has points on one.  The
library's .md3lay reader is checked against the file contents for every
type.  A planted mistake (the lower layer's dz for the upper value in the
VTI-PREM type, where the reference recomputes it) must fail the check."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import media_models as mm

needs_ref = pytest.mark.skipif(not po.ref_available(), reason="reference build absent")

CASES = ([("one_component", mm.ONECMP), ("elastic_isotropic", mm.EL_ISO),
          ("elastic_isotropic", mm.EL_ANISO)] +
         [(k, mm.EL_VTI) for k in ("elastic_vti_prem", "elastic_vti_thomsen", "elastic_vti_cij")] +
         [(k, mm.EL_ANISO) for k in ("elastic_vti_prem", "elastic_vti_thomsen",
                                     "elastic_tti_thomsen", "elastic_tti_bond",
                                     "elastic_aniso_cij")])


def _ref_run(target, path, co, method):
    ref = po.load_ref()
    nz, ny, nx = co.shape[1:]
    xyz = [np.ascontiguousarray(co[d]) for d in range(3)]
    nout = mm.NOUT[target]
    outs = [np.zeros(xyz[0].shape, np.float32) for _ in range(nout)]
    fn = getattr(ref, mm.REF_FN[target])
    fn.argtypes = ([C.c_void_p] * (nout + 3) + [C.c_size_t] * 3 +
                   [C.c_int, C.c_char_p, C.c_char_p, C.c_int])
    fn(*[o.ctypes.data for o in outs], *[a.ctypes.data for a in xyz], nx, ny, nz, 3,
       path.encode(), method.encode(), 1)
    return outs


@needs_ref
@pytest.mark.parametrize("kind,target", CASES)
@pytest.mark.parametrize("method", ["loc", "har", "ari"])
def test_layer2model_grid_matches_reference(kind, target, method, tmp_path):
    code, f, geo = mm.model(kind, seed=3)
    path = str(tmp_path / "m.md3lay")
    mm.write_md3lay(path, kind, f, geo)
    co = mm.grid_points(f, geo)
    want = _ref_run(target, path, co, method)
    rc, got = mm.oracle_grid(code, target, method, f, geo, co)
    assert rc == 0
    for q, (a, b) in enumerate(zip(got, want)):
        assert np.array_equal(a, b), (kind, target, method, q)
    if method != "loc":
        # the averaging changed some points and left the rest at their local values
        _, loc = mm.oracle_grid(code, target, "loc", f, geo, co)
        changed = sum(int((a != b).sum()) for a, b in zip(got, loc))
        assert changed > 0


def test_planted_dz_mistake_fails(tmp_path):
    """The oracle's stale-dz rule applied to a type the reference recomputes
    dz for changes the interface points: the check above would catch it."""
    code, f, geo = mm.model("elastic_vti_prem", seed=3)
    co = mm.grid_points(f, geo)
    rc, good = mm.oracle_grid(code, mm.EL_VTI, "loc", f, geo, co)
    # the same samples read as the CIJ form (stale dz) give different interface values
    rc2, bad = mm.oracle_grid(24, mm.EL_VTI, "loc", f, geo, co)
    assert rc == 0 and rc2 == 0
    on_if = np.zeros(co[0].shape, bool)
    on_if[5, 4, 5:11] = True
    d = np.zeros(co[0].shape, bool)
    for a, b in zip(good, bad):
        d |= a != b
    assert d[on_if].any()


def test_averaging_closed_forms():
    """har and ari of a cell that one flat interface cuts in two: the
    one-component result equals the harmonic / arithmetic mean of the two
    constant values weighted by the share of the 9^3 sub-points on each
    side (sub-points on the interface take the mean of both)."""
    code = 0
    mx, my = 8, 8
    f = np.zeros((4, 2, my, mx), np.float32)
    f[0, 0] = 1000.0
    f[0, 1] = 0.0                      # interface at z = 0
    f[1, 0] = 2.0
    f[1, 1] = 5.0                      # constant values, no gradient
    f[3] = 1.0
    geo = (np.float32(0.0), np.float32(0.0), np.float32(100.0), np.float32(100.0))
    nx, ny, nz = 5, 5, 5
    z = np.arange(nz, dtype=np.float32)[:, None, None] * 10.0 - 20.0 + np.zeros((nz, ny, nx), np.float32)
    x = np.arange(nx, dtype=np.float32)[None, None, :] * 10.0 + 200.0 + np.zeros((nz, ny, nx), np.float32)
    y = np.arange(ny, dtype=np.float32)[None, :, None] * 10.0 + 200.0 + np.zeros((nz, ny, nx), np.float32)
    co = np.stack([x, y, z]).astype(np.float32)
    # the point at z = 0 (k = 2): its half cell spans z in [-5, 5]; sub-points
    # at -5 + 10 k / 8: k < 4 below (5.0), k == 4 on the interface (mean 3.5),
    # k > 4 above (2.0)
    nb, non, na = 4 * 81, 81, 4 * 81
    har_want = 729.0 / (nb / 5.0 + non / 3.5 + na / 2.0)
    ari_want = (nb * 5.0 + non * 3.5 + na * 2.0) / 729.0
    for method, want in (("har", har_want), ("ari", ari_want)):
        rc, out = mm.oracle_grid(code, mm.ONECMP, method, f, geo, co)
        assert rc == 0
        got = float(out[0][2, 2, 2])
        assert abs(got - want) < 1e-5 * want, (method, got, want)
        # the points away from the interface keep their local values
        assert out[0][0, 2, 2] == 5.0 and out[0][4, 2, 2] == 2.0


@pytest.mark.parametrize("kind", sorted(mm.KINDS))
def test_md3lay_reader_every_type(kind, tmp_path):
    from cgfd3d_amd import output as out
    code, f, geo = mm.model(kind, seed=4)
    path = str(tmp_path / f"{kind}.md3lay")
    mm.write_md3lay(path, kind, f, geo)
    mt, nprop, fields, g = out.read_md3lay(path, with_type=True)
    assert mt == code and nprop == (f.shape[0] - 1) // 3
    assert np.array_equal(g, np.array(geo, np.float32))
    assert np.array_equal(fields, f)
