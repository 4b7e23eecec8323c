"""GPU: the value check of check_nan_every_nummber_of_steps (wav_check_value,
wav_t.c:253-273) through cgfdh_check_nan: the first NaN in the reference's
scan order (component by component, array index within each, ghosts
included) equals numpy's first NaN of the same flat array; none on a
finite state; a NaN planted in the state is found again after it has spread
through a few steps."""
import numpy as np
import pytest

from tests import cases as cs

pytestmark = pytest.mark.gpu


def _first_nan(w, vol):
    q = np.flatnonzero(np.isnan(w))
    return None if q.size == 0 else (int(q[0] // vol), int(q[0] % vol))


@pytest.mark.parametrize("medium,shape", [("iso", (100, 100, 60)), ("visco", (20, 18, 14))])
def test_first_nan_in_scan_order(medium, shape):
    case = cs.make_case(*shape, seed=1, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=4,
                        medium=medium)
    s = cs.make_solver(case, 0.004, 8, variant=16 if medium == "iso" else 4)
    try:
        w = cs.random_wavefield(case, seed=4).reshape(-1)
        vol = s.vol
        s.set_wavefield(w)
        assert s.check_nan() is None
        rng = np.random.default_rng(7)
        # NaNs late in the array, in a later grid-stride round, and in ghosts
        for picks in ([w.size - 1], [w.size - 1, 5 * vol + 12345], rng.integers(0, w.size, 6),
                      [3 * vol + 2, 7 * vol + 1]):
            x = w.copy()
            x[np.asarray(picks)] = np.nan
            s.set_wavefield(x)
            assert s.check_nan() == _first_nan(x, vol), picks
    finally:
        s.close()


def test_nan_found_after_steps():
    case = cs.make_case(20, 18, 14, seed=2, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=4)
    s = cs.make_solver(case, 0.004, 8, variant=16)
    try:
        w = cs.random_wavefield(case, seed=5).reshape(-1)
        p = 3 * s.vol + 10 + 11 * 26 + 8 * 26 * 24          # Txx at (10, 11, 8)
        w[p] = np.nan
        s.set_wavefield(w)
        s.run(0, 3)
        got = s.check_nan()
        x = s.wavefield()
        assert got is not None and got == _first_nan(x, s.vol)
        assert got[0] == 0 and got[1] < p - 3 * s.vol      # spread to Vx, nearer the origin
    finally:
        s.close()
