"""GPU: the ablexp sponge applied after every step inside the engine's time
loop, against the oracle stepped one step at a time with the sponge applied
between steps (drv_rk_curv_col.c:482-486); bit-exact."""
import numpy as np
import pytest

from tests import cases as cs
from tests.test_ablexp_oracle import masks, sponge_blocks
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

NSTEP = 6


def oracle_run_with_sponge(case, w0, ex, ey, ez, blocks, dt):
    orc = po.Oracle(case)
    lib = po.load_oracle()
    nc = cs.ncomp(case)
    nx, ny, nz = case["ni"] + 6, case["nj"] + 6, case["nk"] + 6
    w = np.array(w0, np.float32).reshape(-1)
    for it in range(NSTEP):
        w, _, _ = orc.run(dt, 1, w, None, it0=it, nt_total=NSTEP)
        w = np.ascontiguousarray(w, np.float32)
        lib.orc_ablexp_apply(nx, ny, nz, nc, po.ptr(w), po.ptr(ex), po.ptr(ey), po.ptr(ez),
                             po.ptr(blocks))
    return w


@pytest.mark.parametrize("medium,variant", [("iso", 12), ("visco", 4)])
def test_sponge_run_matches_oracle(medium, variant):
    ni, nj, nk = 16, 14, 12
    case = cs.make_case(ni, nj, nk, seed=8, pml=(0,) * 6, medium=medium)
    w0 = cs.random_wavefield(case, seed=9)
    ex, ey, ez = masks(ni + 6, nj + 6, nk + 6, 10)
    blocks = sponge_blocks(ni, nj, nk)
    dt = 0.004
    s = cs.make_solver(case, dt, NSTEP, variant=variant)
    try:
        s.set_wavefield(w0)
        s.set_ablexp(ex, ey, ez, blocks)
        s.run(0, NSTEP)
        w_g = s.wavefield()
    finally:
        s.close()
    w_o = oracle_run_with_sponge(case, w0, ex, ey, ez, blocks, dt)
    w_plain, _, _ = po.Oracle(case).run(dt, NSTEP, w0.reshape(-1), None, nt_total=NSTEP)
    assert not np.array_equal(w_o, w_plain)          # the sponge acted
    print(medium, "rel L2", cs.rel_l2(w_g, w_o), flush=True)
    assert np.array_equal(w_g, w_o)


def test_sponge_rejects_bad_blocks():
    from cgfd3d_amd import CgfdhError
    ni, nj, nk = 16, 14, 12
    case = cs.make_case(ni, nj, nk, seed=8, pml=(0,) * 6)
    s = cs.make_solver(case, 0.004, 4, variant=12)
    try:
        ex, ey, ez = masks(ni + 6, nj + 6, nk + 6, 10)
        b = sponge_blocks(ni, nj, nk)
        b[1, 2] = ni + 6                       # one past the x end
        with pytest.raises(CgfdhError):
            s.set_ablexp(ex, ey, ez, b)
        b = sponge_blocks(ni, nj, nk)
        b[0, 1], b[0, 2] = 5, 4                # empty range
        with pytest.raises(CgfdhError):
            s.set_ablexp(ex, ey, ez, b)
    finally:
        s.close()
