"""CPU checks that pin the ablexp-sponge oracle (bdry_ablexp_apply,
bdry_t.c:840-890) against the reference's own function compiled from
bdry_t.c, and against closed forms."""
import numpy as np
import pytest

from oracle import pyoracle as po


def sponge_blocks(ni, nj, nk, lay=(3, 4, 2, 3, 0, 4)):
    """Six blocks laid out as bdry_ablexp_set lays them (bdry_t.c:560-800):
    x faces whole, y faces inside the x layers, z faces inside both;
    {enable, ni1, ni2, nj1, nj2, nk1, nk2} with 3 ghost points."""
    i1, i2, j1, j2, k1, k2 = 3, 2 + ni, 3, 2 + nj, 3, 2 + nk
    x1, x2, y1, y2, z1, z2 = lay
    b = [[x1 > 0, i1, i1 + x1 - 1, j1, j2, k1, k2],
         [x2 > 0, i2 - x2 + 1, i2, j1, j2, k1, k2],
         [y1 > 0, i1 + x1, i2 - x2, j1, j1 + y1 - 1, k1, k2],
         [y2 > 0, i1 + x1, i2 - x2, j2 - y2 + 1, j2, k1, k2],
         [z1 > 0, i1 + x1, i2 - x2, j1 + y1, j2 - y2, k1, k1 + z1 - 1],
         [z2 > 0, i1 + x1, i2 - x2, j1 + y1, j2 - y2, k2 - z2 + 1, k2]]
    return np.array(b, np.int32)


def masks(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.5, 1.0, n).astype(np.float32) for n in (nx, ny, nz)]


def apply(lib_fn, w, ex, ey, ez, blocks, ncmp):
    out = np.array(w, np.float32, copy=True)
    nz, ny, nx = w.shape[1:]
    lib_fn(nx, ny, nz, ncmp, po.ptr(out), po.ptr(ex), po.ptr(ey), po.ptr(ez),
           po.ptr(np.ascontiguousarray(blocks, np.int32)))
    return out


@pytest.mark.skipif(not po.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("overlap", [False, True])
def test_apply_matches_reference(overlap):
    ni, nj, nk, nc = 12, 11, 10, 15
    rng = np.random.default_rng(1)
    w = rng.standard_normal((nc, nk + 6, nj + 6, ni + 6)).astype(np.float32)
    ex, ey, ez = masks(ni + 6, nj + 6, nk + 6, 2)
    b = sponge_blocks(ni, nj, nk)
    if overlap:                       # blocks that share points are applied in order
        b[2, 1], b[4, 3] = 3, 3
    got = apply(po.load_oracle().orc_ablexp_apply, w, ex, ey, ez, b, nc)
    want = apply(po.load_ref().refh_ablexp_apply, w, ex, ey, ez, b, nc)
    assert np.array_equal(got, want)
    assert not np.array_equal(got, w)


def test_apply_closed_forms():
    ni, nj, nk = 8, 8, 8
    orc = po.load_oracle()
    w = np.ones((9, nk + 6, nj + 6, ni + 6), np.float32)
    one = [np.ones(n, np.float32) for n in (ni + 6, nj + 6, nk + 6)]
    b = sponge_blocks(ni, nj, nk)
    assert np.array_equal(apply(orc.orc_ablexp_apply, w, *one, b, 9), w)
    # one block, Ez = 0.5 on its rows only: exactly that block halves
    ez = np.ones(nk + 6, np.float32)
    ez[3:5] = 0.5
    only = np.zeros_like(b)
    only[4] = b[4]
    only[4, 0] = 1                    # the default layout has no z1 layers
    only[4, 5:7] = (3, 4)
    got = apply(orc.orc_ablexp_apply, w, one[0], one[1], ez, only, 9)
    want = w.copy()
    want[:, 3:5, b[4, 3]:b[4, 4] + 1, b[4, 1]:b[4, 2] + 1] = 0.5
    assert np.array_equal(got, want)
    # the mask is the smallest of the three 1-D masks
    ex = np.full(ni + 6, 0.9, np.float32)
    ey = np.full(nj + 6, 0.7, np.float32)
    ez = np.full(nk + 6, 0.8, np.float32)
    got = apply(orc.orc_ablexp_apply, w, ex, ey, ez, b, 9)
    inside = np.zeros(w.shape[1:], bool)
    for e, i1, i2, j1, j2, k1, k2 in b:
        if e:
            inside[k1:k2 + 1, j1:j2 + 1, i1:i2 + 1] = True
    assert np.all(got[:, inside] == np.float32(0.7)) and np.all(got[:, ~inside] == 1)
