"""CPU: the reader of the grid layer file gd_curv_gen_layer takes
(in_grid_layer_file, gd_t.c:1031-1060: interface count, cells and spacing
flags per layer, sample counts, then x y z per sample) against the file as
written here; bad files are refused."""
import numpy as np
import pytest

from cgfd3d_amd import output as out


def _write(path, ncell, equal, L):
    n_if, ny, nx = L.shape[1:]
    with open(path, "w") as f:
        f.write(f"{n_if}\n")
        f.write(" ".join(str(v) for v in ncell) + "\n")
        f.write(" ".join(str(v) for v in equal) + "\n")
        f.write(f"{nx} {ny}\n")
        flat = L.reshape(3, -1)
        for q in range(flat.shape[1]):
            f.write(f"{flat[0, q]:.9g} {flat[1, q]:.9g} {flat[2, q]:.9g}\n")


def test_layer_file_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    L = rng.uniform(-5000, 5000, (3, 4, 7, 9)).astype(np.float32)
    p = str(tmp_path / "grid.gdlay")
    _write(p, [10, 12, 8], [1, 0, 1], L)
    ncell, equal, xyz = out.read_grid_layer(p)
    assert ncell.tolist() == [10, 12, 8] and equal.tolist() == [1, 0, 1]
    assert np.array_equal(xyz, L)


def test_layer_file_refusals(tmp_path):
    p = str(tmp_path / "short.gdlay")
    L = np.zeros((3, 2, 3, 3), np.float32)
    _write(p, [5], [1], L)
    text = open(p).read().splitlines()
    open(p, "w").write("\n".join(text[:-2]) + "\n")           # two samples missing
    with pytest.raises(out.OutputError, match="sample"):
        out.read_grid_layer(p)
    with pytest.raises(out.OutputError):
        out.read_grid_layer(str(tmp_path / "absent.gdlay"))
