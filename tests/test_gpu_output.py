"""GPU: output sub-volumes (snapshots and slices) packed on the device, and
the strain conversion, against the oracle on the same wavefield; bit-exact."""
import numpy as np
import pytest

from tests import cases as cs
from oracle import output_oracle as oo

pytestmark = pytest.mark.gpu

BOXES = [
    (3, 16, 1, 3, 14, 1, 3, 12, 1),     # every physical point
    (0, 22, 1, 0, 20, 1, 0, 18, 1),     # the whole array, ghosts included
    (4, 5, 3, 3, 4, 4, 5, 3, 4),        # strided snapshot
    (9, 1, 1, 3, 14, 1, 3, 12, 1),      # x slice (io_slice_nc_put)
    (3, 16, 1, 7, 1, 1, 3, 12, 1),      # y slice
    (3, 16, 1, 3, 14, 1, 14, 1, 1),     # z slice (the free surface row)
    (21, 1, 1, 19, 1, 1, 17, 1, 1),     # the last point
]


def _run_case(medium):
    case = cs.make_case(16, 14, 12, seed=4, medium=medium)
    s = cs.make_solver(case, 0.004, 8, variant=4 if medium == "visco" else 12)
    s.set_wavefield(cs.random_wavefield(case, seed=5, scale=1.0e5))
    s.run(0, 3)
    w = s.wavefield().reshape(cs.ncomp(case), 18, 20, 22)
    return case, s, w


@pytest.mark.parametrize("medium", ["iso", "visco"])
def test_subvolume_and_strain_match_oracle(medium):
    case, s, w = _run_case(medium)
    try:
        nc = cs.ncomp(case)
        emax = 0.0
        for box in BOXES:
            print("box", box, flush=True)
            comps = list(range(nc)) if box[4] == 1 or box[1] == 1 else [0, 1, 2, 3, 4, 5, 7, 6, 8]
            got = s.subvolume(box, comps)
            assert np.array_equal(got, oo.pack_subvolume(w, comps, box))
            lam = case["med"][0].reshape(18, 20, 22)
            mu = case["med"][1].reshape(18, 20, 22)
            T = [w[c] for c in (3, 4, 5, 6, 7, 8)]
            e_got = s.strain_subvolume(box)
            e_want = oo.stress_to_strain_eliso(lam, mu, T, box)
            emax = max(emax, float(np.abs(e_want).max()))   # the ghost corner is all zero
            assert np.array_equal(e_got, e_want)
        assert emax > 0
        # a smaller request after larger ones reuses the staging buffer
        small = (5, 2, 1, 5, 2, 1, 5, 2, 1)
        assert np.array_equal(s.subvolume(small, [2]), oo.pack_subvolume(w, [2], small))
    finally:
        s.close()


def test_subvolume_rejects_bad_requests():
    from cgfd3d_amd import CgfdhError
    case, s, w = _run_case("iso")
    try:
        for box in [(0, 23, 1, 0, 1, 1, 0, 1, 1),      # past the x end
                    (20, 2, 2, 0, 1, 1, 0, 1, 1),      # stride past the end
                    (0, 1, 1, 0, 0, 1, 0, 1, 1),       # empty axis
                    (0, 1, 1, 0, 1, 0, 0, 1, 1),       # zero stride
                    (-1, 1, 1, 0, 1, 1, 0, 1, 1)]:     # negative start
            with pytest.raises(CgfdhError):
                s.subvolume(box, [0])
            with pytest.raises(CgfdhError):
                s.strain_subvolume(box)
        for comps in ([9], [-1], list(range(10))):
            with pytest.raises(CgfdhError):
                s.subvolume(BOXES[0], comps)
        # the solver still works after the rejected calls
        assert np.array_equal(s.subvolume(BOXES[2], [0]), oo.pack_subvolume(w, [0], BOXES[2]))
    finally:
        s.close()
    vti = cs.make_case(16, 14, 12, seed=4, medium="vti")
    s = cs.make_solver(vti, 0.004, 8, variant=12)
    try:
        with pytest.raises(CgfdhError):
            s.strain_subvolume(BOXES[0])
    finally:
        s.close()
