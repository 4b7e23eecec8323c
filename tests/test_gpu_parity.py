"""GPU parity: the HIP engine against the CPU oracle on the same seeded
inputs.  Tolerance: relative L2 <= 1e-6 per RHS evaluation and <= 1e-5 over a
run (FP32, the bar of the parity target); the kernels evaluate the reference's
expressions in the same order without FMA contraction, so the tests also
record that the results are bit-identical."""
import numpy as np
import pytest

from tests import cases as cs
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CONFIGS = [
    ("inner", False, (0,) * 6),
    ("pml6", False, (1,) * 6),
    ("free", True, (0,) * 6),
    ("free_pml5", True, (1, 1, 1, 1, 1, 0)),
]


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("name,free,pml", CONFIGS)
def test_onestage_all_pairs(name, free, pml, variant):
    case = cs.make_case(18, 15, 13, seed=2, free_top=free, pml=pml, layers=4)
    w = cs.random_wavefield(case, seed=4).reshape(-1)
    aux = cs.random_aux(case, seed=4) if any(pml) else None
    orc = po.Oracle(case)
    s = cs.make_solver(case, 0.005, 10, variant=variant)
    worst, mism = 0.0, 0
    for ipair in range(8):
        for istage in range(4):
            r_o, a_o = orc.onestage(3, ipair, istage, w, aux)
            r_g, a_g = s.onestage(3, ipair, istage, w, aux)
            worst = max(worst, cs.rel_l2(r_g, r_o))
            mism += int(np.count_nonzero(r_g != r_o))
            if aux is not None:
                worst = max(worst, cs.rel_l2(a_g, a_o))
                mism += int(np.count_nonzero(a_g != a_o))
    s.close()
    print(f"{name} variant={variant}: worst rel L2 {worst:.3e}, differing values {mism}")
    assert worst <= 1e-6
    assert mism == 0, "expected bit-identical RHS"


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("gauss", [False, True])
def test_run_matches_oracle(variant, gauss):
    ni, nj, nk, nt, dt = 26, 22, 20, 60, 0.004
    case = cs.make_case(ni, nj, nk, seed=5, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=5)
    case["src"] = cs.ricker_point_source(case, 3 + ni // 2, 3 + nj // 2, 3 + nk // 2, nt, dt,
                                         fc=10.0, t0=0.12, gauss=gauss,
                                         inc=(0.2, -0.3, 0.1) if gauss else (0, 0, 0))
    ridx, rw = cs.surface_receivers(case, n=5, seed=1)
    bi, bw = cs.buried_receiver(case, 9, 10, 8, 0.25, 0.5, 0.75)
    ridx = np.concatenate([ridx, bi]); rw = np.concatenate([rw, bw])
    orc = po.Oracle(case)
    w_o, a_o, s_o = orc.run(dt, nt, recv_idx=ridx, recv_w=rw)
    s = cs.make_solver(case, dt, nt, variant=variant)
    s.set_receivers(ridx, rw)
    s.run(0, nt)
    w_g, a_g, s_g = s.wavefield(), s.pml_aux(), s.seismograms()
    s.close()
    e_w, e_s, e_a = cs.rel_l2(w_g, w_o), cs.rel_l2(s_g, s_o), cs.rel_l2(a_g, a_o)
    print(f"gauss={gauss} variant={variant}: rel L2 wave {e_w:.2e} seis {e_s:.2e} aux {e_a:.2e}")
    assert np.abs(s_o).max() > 0
    assert e_w <= 1e-5 and e_s <= 1e-5 and e_a <= 1e-5
    assert np.array_equal(w_g, w_o) and np.array_equal(s_g, s_o)
