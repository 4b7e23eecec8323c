"""GPU: Graves-Qs attenuation applied after every step inside the engine's
time loop, against the oracle stepped one step at a time with
sv_curv_col_el_graves_Qs's factor applied between steps.  Tolerance: the
engine's exponential is the double one rounded to float and the reference's
is libm expf, which differ by one ulp on a few inputs in 10^5, so the bar is
relative L2 <= 1e-6 over the run (the floating-point parity bar), with the
bit-identity of the result reported."""
import numpy as np
import pytest

from tests import cases as cs
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

NSTEP = 6


@pytest.mark.parametrize("medium", ["iso", "vti", "aniso"])
def test_graves_run_matches_oracle(medium):
    """Every elastic medium takes Graves-Qs (the reference applies it to
    all ncmp components whenever md->visco_type is Graves-Qs,
    drv_rk_curv_col.c:414-416)."""
    ni, nj, nk = 16, 14, 12
    case = cs.make_case(ni, nj, nk, seed=14, pml=(1, 1, 1, 1, 1, 0), layers=3, medium=medium)
    w0 = cs.random_wavefield(case, seed=15)
    rng = np.random.default_rng(16)
    qs = rng.uniform(5.0, 200.0, (nk + 6, nj + 6, ni + 6)).astype(np.float32)
    dt, freq = 0.004, 2.0
    s = cs.make_solver(case, dt, NSTEP, variant=12)
    try:
        s.set_wavefield(w0)
        s.set_graves_qs(qs, freq)
        s.run(0, NSTEP)
        w_g, a_g = s.wavefield(), s.pml_aux()
    finally:
        s.close()
    orc = po.Oracle(case)
    lib = po.load_oracle()
    nc = cs.ncomp(case)
    w, aux = w0.reshape(-1).copy(), None
    for it in range(NSTEP):
        w, aux, _ = orc.run(dt, 1, w, aux, it0=it, nt_total=NSTEP)
        w = np.ascontiguousarray(w, np.float32)
        lib.orc_graves_qs(ni, nj, nk, nc, po.ptr(w), po.ptr(qs), freq, dt)
    w_plain, _, _ = po.Oracle(case).run(dt, NSTEP, w0.reshape(-1), None, nt_total=NSTEP)
    assert cs.rel_l2(w, w_plain) > 1e-4                 # the attenuation acted
    e_w, e_a = cs.rel_l2(w_g, w), cs.rel_l2(a_g, aux)
    print(f"graves {medium}: rel L2 wave {e_w:.2e} aux {e_a:.2e} bit-identical "
          f"{np.array_equal(w_g, w)}", flush=True)
    assert e_w <= 1e-6 and e_a <= 1e-6


def test_graves_rejects_visco_medium():
    """The GMB visco-elastic medium carries its own attenuation."""
    from cgfd3d_amd import CgfdhError
    case = cs.make_case(16, 14, 12, seed=14, medium="visco")
    s = cs.make_solver(case, 0.004, 4, variant=4)
    try:
        with pytest.raises(CgfdhError):
            s.set_graves_qs(np.full(s.vol, 50.0, np.float32), 2.0)
    finally:
        s.close()
