"""GPU: the layered model discretised onto the grid with local values
(cgfdh_layer2model: media_layer2model_el_iso / _onecmp, "loc") against the
oracle (oracle/media_oracle.c, equal to the reference's own media library by
tests/test_media_oracle.py), through the .md3lay reader: lambda, mu, rho (and
the one-component value) on a curvilinear grid from 1.7 km deep to above the
topography, with points on an interface.  The power law is taken in double
on both sides; two math libraries' pow may differ in the last place, so the
bar is relative 1e-6 per value (the results were bit-identical in the runs
recorded), and the reference's refusals (a point off the mesh, vp <= sqrt(2)
vs) are errors."""
import numpy as np
import pytest

from cgfd3d_amd import CgfdhError, layer2model
from cgfd3d_amd import output as out
from tests.test_media_oracle import grid_points, md3lay_model, oracle_run, write_md3lay

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["elastic_isotropic", "one_component"])
def test_layer2model_matches_oracle(kind, tmp_path):
    nprop, f, geo = md3lay_model(kind, seed=5)
    path = str(tmp_path / "model.md3lay")
    write_md3lay(path, kind, f, geo)
    nprop_r, fields, g = out.read_md3lay(path)
    co = grid_points(fields, g, ni=28, nj=24, nk=30, seed=2)
    rc, want = oracle_run(nprop, fields, g, co)
    assert rc == 0
    got = layer2model(nprop_r, fields, g, co)
    nbit = 0
    for a, b in zip(got, want):
        assert np.allclose(a, b, rtol=1e-6, atol=0)
        nbit += int((a != b).sum())
    print(f"{kind}: {nbit} values differ in the last place of {sum(a.size for a in got)}",
          flush=True)


def test_layer2model_refusals(tmp_path):
    nprop, f, geo = md3lay_model("elastic_isotropic", seed=6)
    co = grid_points(f, geo, seed=3)
    off = co.copy()
    off[0, 2, 2, 2] = geo[0] - 10.0                  # left of the mesh
    with pytest.raises(CgfdhError, match="outside the interface mesh"):
        layer2model(nprop, f, geo, off)
    bad = f.copy()
    bad[7] = bad[4]                                  # vs = vp
    with pytest.raises(CgfdhError, match="sqrt"):
        layer2model(nprop, bad, geo, co)
