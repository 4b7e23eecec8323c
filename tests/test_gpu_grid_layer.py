"""GPU: the curvilinear grid generated from interface samples
(cgfdh_gen_grid_layer, gd_curv_gen_layer with gd_grid_z_interp and the
reference's cubic spline) against the oracle (oracle/grid_oracle.c, pinned
by tests/test_grid_oracle.py), bit for bit: a hilly four-interface model
with a graded middle layer, under every resampling branch (every |f|-th
sample, a spline along x, along y, or both, and the vertical refine factor
either way), through the layer-file reader; and the reference's parameter
checks as errors."""
import numpy as np
import pytest

from cgfd3d_amd import CgfdhError, gen_grid_layer
from cgfd3d_amd import output as out
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _model(nxl, nyl, dx, seed):
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.arange(nxl) * dx, np.arange(nyl) * dx)
    xc, yc = X.mean(), Y.mean()
    hill = 600.0 * np.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / (2 * (0.2 * X.max()) ** 2))
    zs = [-6000.0, -4000.0, -3000.0, 0.0]
    L = np.zeros((3, 4, nyl, nxl), np.float32)
    for k, z0 in enumerate(zs):
        w = 0.3 * dx * np.sin(0.21 * X / dx + k) * np.cos(0.17 * Y / dx)
        L[0, k] = X + w
        L[1, k] = Y - w
        L[2, k] = z0 + (k / 3.0) * hill + 40.0 * rng.standard_normal((nyl, nxl)) * (k > 0)
    return L


# (resample, start, samples nx x ny, cells per layer)
CASES = [((1, 1, 1), (6, 6, 0), (60, 50), [20, 8, 20]),
         ((-2, -2, 1), (8, 8, 2), (100, 90), [20, 8, 20]),
         ((2, 1, 1), (5, 7, 0), (60, 50), [20, 8, 20]),
         ((1, 3, 1), (7, 4, 0), (60, 50), [20, 8, 20]),
         ((2, 2, 2), (5, 5, 3), (60, 50), [20, 8, 20]),
         ((3, 2, -2), (4, 5, 0), (60, 50), [40, 16, 40])]


@pytest.mark.parametrize("resample,start,samples,ncell", CASES)
def test_grid_matches_oracle(resample, start, samples, ncell):
    ni, nj, nk = 34, 28, 30
    L = _model(samples[0], samples[1], 100.0, seed=1)
    equal = [1, 0, 1]
    rc, want = po.gen_layer(ncell, equal, L, resample, start, ni, nj, nk)
    assert rc == 0
    got = gen_grid_layer(ncell, equal, L, resample, start, ni, nj, nk)
    print(f"resample {resample}: top z {got[2, 3 + nk - 1].min():.1f}..{got[2, 3 + nk - 1].max():.1f}",
          flush=True)
    assert np.array_equal(got, want)


def test_grid_from_a_layer_file(tmp_path):
    L = _model(40, 36, 100.0, seed=2)
    p = str(tmp_path / "model.gdlay")
    with open(p, "w") as f:
        f.write("4\n20 8 20\n1 0 1\n40 36\n")
        flat = L.reshape(3, -1)
        for q in range(flat.shape[1]):
            f.write(f"{flat[0, q]:.9g} {flat[1, q]:.9g} {flat[2, q]:.9g}\n")
    ncell, equal, xyz = out.read_grid_layer(p)
    got = gen_grid_layer(ncell, equal, xyz, (1, 1, 1), (5, 5, 0), 28, 24, 30)
    rc, want = po.gen_layer(ncell, equal, xyz, (1, 1, 1), (5, 5, 0), 28, 24, 30)
    assert rc == 0 and np.array_equal(got, want)


def test_grid_parameter_checks():
    L = _model(20, 20, 100.0, seed=3)
    with pytest.raises(CgfdhError, match="horizontal start"):
        gen_grid_layer([8, 6, 6], [1, 1, 1], L, (1, 1, 1), (2, 5, 0), 10, 10, 6)
    with pytest.raises(CgfdhError, match="exceed"):
        gen_grid_layer([8, 6, 6], [1, 1, 1], L, (1, 1, 1), (5, 5, 0), 16, 10, 6)
    with pytest.raises(CgfdhError, match="too deep"):
        gen_grid_layer([2, 1, 1], [1, 1, 1], L, (1, 1, 1), (5, 5, 0), 10, 10, 6)
    bad = L.copy()
    bad[2, 2] = bad[2, 1] - 10.0                # interfaces out of order
    with pytest.raises(CgfdhError, match="increasing"):
        gen_grid_layer([8, 6, 6], [1, 1, 1], bad, (1, 1, 1), (5, 5, 0), 10, 10, 6)
