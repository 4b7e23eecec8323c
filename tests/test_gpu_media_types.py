"""GPU: the layered model on a curvilinear grid for every media type and
equivalent-medium method (cgfdh_layer2model_grid, and the reference-shaped
file entry points cgfdh_media_layer2model_*) against the oracle
(oracle/media_oracle.c, equal to the reference's media library bit for bit
by tests/test_media_types_oracle.py).  The power law, and the TTI types'
cos and sin, are taken in double on both sides; two math libraries may
differ in their last place, so the bar is 1e-6 of each value for the types
without a rotation and 1e-5 of the array's largest magnitude for the
rotated ones (whose harmonic averages of near-zero coefficients magnify a
last-place difference); the runs print how many values differ at all."""
import numpy as np
import pytest

from cgfd3d_amd import CgfdhError, layer2model_grid, media_layer2model
from cgfd3d_amd import output as out
from tests import media_models as mm

pytestmark = pytest.mark.gpu

CASES = ([("one_component", mm.ONECMP), ("elastic_isotropic", mm.EL_ISO),
          ("elastic_isotropic", mm.EL_ANISO)] +
         [(k, mm.EL_VTI) for k in ("elastic_vti_prem", "elastic_vti_thomsen", "elastic_vti_cij")] +
         [(k, mm.EL_ANISO) for k in ("elastic_vti_prem", "elastic_tti_thomsen",
                                     "elastic_tti_bond", "elastic_aniso_cij")])
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
def test_layer2model_grid_matches_oracle(kind, target, method, tmp_path):
    code, f, geo = mm.model(kind, seed=7)
    path = str(tmp_path / "m.md3lay")
    mm.write_md3lay(path, kind, f, geo)
    mt, nprop, fields, g = out.read_md3lay(path, with_type=True)
    co = mm.grid_points(f, geo, ni=30, nj=26, nk=28, seed=4)
    rc, want = mm.oracle_grid(code, target, method, f, geo, co)
    assert rc == 0
    got = layer2model_grid(mt, target, method, fields, g, co)
    nd = _close(kind, got, want)
    print(f"{kind} target {target} {method}: {nd} of {sum(a.size for a in got)} values differ",
          flush=True)
    # the reference-shaped entry reading the file gives the same arrays
    got2 = media_layer2model(target, path, co, method)
    for a, b in zip(got, got2):
        assert np.array_equal(a, b)


def test_layer2model_grid_larger_isotropic_har(tmp_path):
    """A 96 x 80 x 64 grid (about 20 thousand averaged points) against the oracle."""
    code, f, geo = mm.model("elastic_isotropic", seed=8, mx=61, my=51, dx=40.0, dy=40.0)
    co = mm.grid_points(f, geo, ni=96, nj=80, nk=64, seed=5)
    co[0] = geo[0] + 100.0 + (co[0] - co[0].min()) * (2200.0 / (co[0].max() - co[0].min()))
    co[1] = geo[1] + 100.0 + (co[1] - co[1].min()) * (1800.0 / (co[1].max() - co[1].min()))
    rc, want = mm.oracle_grid(code, mm.EL_ISO, "har", f, geo, co)
    assert rc == 0
    got = layer2model_grid(code, mm.EL_ISO, "har", f, geo, co)
    nd = _close("elastic_isotropic", got, want)
    print(f"96x80x64 har: {nd} values differ", flush=True)


def test_layer2model_refusals(tmp_path):
    code, f, geo = mm.model("elastic_vti_prem", seed=9)
    path = str(tmp_path / "m.md3lay")
    mm.write_md3lay(path, "elastic_vti_prem", f, geo)
    co = mm.grid_points(f, geo)
    with pytest.raises(CgfdhError, match="not supported"):
        media_layer2model(mm.EL_ISO, path, co, "loc")
    with pytest.raises(CgfdhError, match="not provided"):
        media_layer2model(mm.EL_VTI, path, co, "tti")
    with pytest.raises(CgfdhError, match="grid type"):
        media_layer2model(mm.EL_VTI, path, co, "loc", grid_type=1)
    off = co.copy()
    off[0, 3, 3, 3] = geo[0] - 10.0
    with pytest.raises(CgfdhError, match="outside the interface mesh"):
        media_layer2model(mm.EL_VTI, path, off, "har")
    code, f, geo = mm.model("elastic_isotropic", seed=9)
    bad = f.copy()
    bad[7] = bad[4]                                  # vs = vp
    with pytest.raises(CgfdhError, match="sqrt"):
        layer2model_grid(code, mm.EL_ISO, "ari", bad, geo, co)
