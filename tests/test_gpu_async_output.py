"""GPU: asynchronous steps and output (cgfdh_run_async, cgfdh_subvolume_async,
cgfdh_wait, cgfdh_sync, pinned host memory from cgfdh_host_alloc) against the
synchronous cgfdh_run / cgfdh_get_subvolume: after every step, three boxes
(a z slice of all components, a strided snapshot box, a single point) are
bit-identical; more boxes in flight than staging slots (the oldest copies
are waited for and their slots reused); pinned and ordinary destinations;
the final state unchanged by the outputs."""
import numpy as np
import pytest

from cgfd3d_amd import PinnedBuffer
from tests import cases as cs

pytestmark = pytest.mark.gpu

BOXES = [((0, 32, 1, 0, 28, 1, 20, 1, 1), list(range(9))),
         ((5, 9, 3, 4, 7, 2, 6, 5, 3), [8, 0, 4]),
         ((12, 1, 1, 11, 1, 1, 9, 1, 1), [1])]


def _solver():
    case = cs.make_case(26, 22, 18, seed=31, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=4)
    case["src"] = cs.ricker_point_source(case, 13, 11, 9, 40, 0.004, fc=12.0, t0=0.03)
    return case, cs.make_solver(case, 0.004, 40, variant=16)


def _shape(box, comps):
    return (len(comps), box[7], box[4], box[1])


def test_async_output_matches_sync():
    nsteps = 9
    case, s = _solver()
    try:
        want = []
        for it in range(nsteps):
            s.run(it, 1)
            want.append([s.subvolume(b, c) for b, c in BOXES])
        w_sync = s.wavefield()
    finally:
        s.close()
    case, s = _solver()
    bufs = [[PinnedBuffer(_shape(b, c)) if q % 2 == 0 else None for q, (b, c) in enumerate(BOXES)]
            for _ in range(nsteps)]
    plain = [[np.empty(_shape(b, c), np.float32) for b, c in BOXES] for _ in range(nsteps)]
    try:
        tickets = []
        for it in range(nsteps):
            s.run_async(it, 1)
            for q, (b, c) in enumerate(BOXES):
                out = bufs[it][q].array if bufs[it][q] is not None else plain[it][q]
                tickets.append((it, q, out, s.subvolume_async(b, c, out)))
        # 27 boxes through 16 staging slots: the later requests waited for
        # earlier copies; every ticket still resolves to its own values
        for it, q, out, t in reversed(tickets):
            s.wait(t)
            assert np.array_equal(out, want[it][q]), (it, q)
        s.sync()
        assert np.array_equal(s.wavefield(), w_sync)
    finally:
        s.close()
        for row in bufs:
            for b in row:
                if b is not None:
                    b.close()


def test_async_steps_alone_match_run():
    case, s = _solver()
    try:
        s.run(0, 12)
        w0 = s.wavefield()
    finally:
        s.close()
    case, s = _solver()
    try:
        s.run_async(0, 5)
        s.run_async(5, 7)
        s.sync()
        assert np.array_equal(s.wavefield(), w0)
    finally:
        s.close()
