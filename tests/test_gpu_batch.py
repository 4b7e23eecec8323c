"""GPU: batched runs of independent solvers (cgfdh_batch_*).  The batch
captures 8-step cycles of all solvers into one hipGraph and replays them; its
results must equal each solver's own cgfdh_run bit for bit, and through that
the oracle's."""
import numpy as np
import pytest

from tests import cases as cs
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

NT, DT = 40, 0.004


def _cases():
    """Three uneven grids with different media, boundaries and sources."""
    out = []
    for n, (dims, medium, pml, seed) in enumerate([((22, 19, 17), "iso", (1, 1, 1, 1, 1, 0), 1),
                                                   ((18, 20, 16), "vti", (1, 1, 0, 1, 1, 0), 2),
                                                   ((17, 16, 18), "visco", (1, 0, 1, 1, 1, 0), 3)]):
        ni, nj, nk = dims
        case = cs.make_case(ni, nj, nk, seed=seed, free_top=True, pml=pml, layers=4,
                            medium=medium)
        case["src"] = cs.ricker_point_source(case, 3 + ni // 2, 3 + nj // 2, 3 + nk // 2, NT, DT,
                                             fc=10.0, t0=0.1)
        ridx, rw = cs.surface_receivers(case, n=3, seed=seed)
        out.append((case, ridx, rw))
    return out


def _solver(case, ridx, rw, variant=4):
    s = cs.make_solver(case, DT, NT, variant=variant)
    s.set_receivers(ridx, rw)
    return s


def _state(s):
    pg, dis = s.peak_ground_motion()
    return s.wavefield(), s.pml_aux(), s.seismograms(), pg, dis


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_batch_equals_individual_runs():
    """19 steps (2 graph cycles + 3 eager steps) then 16 more from step 19
    (a cycle captured at another start step): identical to cgfdh_run."""
    from cgfd3d_amd import Batch
    cases = _cases()
    ref = []
    for case, ridx, rw in cases:
        s = _solver(case, ridx, rw)
        s.run(0, 19)
        s.run(19, 16)
        ref.append(_state(s))
        s.close()
    sv = [_solver(case, ridx, rw) for case, ridx, rw in cases]
    b = Batch(sv)
    b.run(0, 19)
    b.run(19, 16)
    print(f"batch: last call {b.last_ms():.2f} ms")
    for n, s in enumerate(sv):
        assert _same(_state(s), ref[n]), f"solver {n}"
    b.close()
    for s in sv:
        s.close()


def test_batch_matches_oracle():
    """One batch member against the oracle (wavefield, aux, seismograms)."""
    from cgfd3d_amd import Batch
    case, ridx, rw = _cases()[0]
    s = _solver(case, ridx, rw)
    b = Batch([s])
    b.run(0, 24)
    w_g, a_g, s_g = s.wavefield(), s.pml_aux(), s.seismograms()
    b.close()
    s.close()
    w_o, a_o, s_o = po.Oracle(case).run(DT, 24, recv_idx=ridx, recv_w=rw, nt_total=NT)
    print(f"batch vs oracle: rel L2 {cs.rel_l2(w_g, w_o):.1e}")
    assert np.abs(s_o).max() > 0
    assert np.array_equal(w_g, w_o) and np.array_equal(a_g, a_o) and np.array_equal(s_g, s_o)


def test_batch_autotune_leaves_the_state():
    """Autotuning runs trial steps but restores the state: the run after it
    equals a run without it, whatever kernels were picked."""
    from cgfd3d_amd import Batch
    cases = _cases()
    ref = []
    for case, ridx, rw in cases:
        s = _solver(case, ridx, rw)
        s.run(0, 5)
        s.run(5, 11)
        ref.append(_state(s))
        s.close()
    sv = [_solver(case, ridx, rw) for case, ridx, rw in cases]
    b = Batch(sv)
    b.run(0, 5)
    b.autotune(5)
    print("autotuned:", [s.kernel_choice() for s in sv])
    b.run(5, 11)
    for n, s in enumerate(sv):
        assert _same(_state(s), ref[n]), f"solver {n}"
    b.close()
    for s in sv:
        s.close()


def test_batch_recaptures_after_reconfiguration():
    """Setters between two batch runs (new sources and receivers, Graves-Qs
    switched on, another kernel variant) and eager steps of one member in
    between (which swap its levels): the batch's next replay follows the new
    configuration, equal to the members' own cgfdh_run bit for bit."""
    from cgfd3d_amd import Batch
    cases = _cases()[:2]
    rng = np.random.default_rng(5)
    qs = [rng.uniform(10.0, 100.0, (c["nk"] + 6, c["nj"] + 6, c["ni"] + 6)).astype(np.float32)
          for c, _, _ in cases]

    def reconfigure(s, n, case):
        ni, nj, nk = case["ni"], case["nj"], case["nk"]
        src = cs.ricker_point_source(case, 3 + ni // 3, 3 + nj // 3, 3 + nk // 3, NT, DT,
                                     fc=8.0, t0=0.12, mech=(0, 0, 0, 1, 1, 0))
        s.set_moment_sources(src["ijk"], src["it_begin"], src["it_end"], src["max_nt"],
                             src["mij"])
        ridx, rw = cs.surface_receivers(case, n=2, seed=11 + n)
        s.set_receivers(ridx, rw)
        s.set_graves_qs(qs[n], 2.0)
        s.set_kernel_variant(12)

    ref = []
    for n, (case, ridx, rw) in enumerate(cases):
        s = _solver(case, ridx, rw)
        s.run(0, 16)
        reconfigure(s, n, case)
        s.run(16, 3)
        s.run(19, 16)
        ref.append((s.wavefield(), s.pml_aux(), s.seismograms()))
        s.close()
    sv = [_solver(case, ridx, rw) for case, ridx, rw in cases]
    b = Batch(sv)
    b.run(0, 16)
    for n, (case, _, _) in enumerate(cases):
        reconfigure(sv[n], n, case)
    sv[0].run(16, 3)           # eager steps of one member swap its levels
    sv[1].run(16, 3)
    b.run(19, 16)
    for n, s in enumerate(sv):
        assert np.array_equal(s.wavefield(), ref[n][0]), n
        assert np.array_equal(s.pml_aux(), ref[n][1]), n
        assert np.array_equal(s.seismograms(), ref[n][2]), n
    b.close()
    for s in sv:
        s.close()


@pytest.mark.parametrize("medium", ["iso", "vti", "visco"])
def test_fused_batch_equals_individual_runs(medium, monkeypatch):
    """Members all on the whole-row kernel (variant 16) and one medium: the
    batch's cycle runs each stage as one launch over every grid's rows (the
    cross-grid tile queue).  Three uneven grids, one without a free surface,
    one with Graves-Qs (elastic media): 19 steps with the fused cycle
    (CGFDH_BATCH_FUSED=1), 8 more with the per-member cycle, 8 fused again
    from another start step; bit-identical to each member's own cgfdh_run."""
    from cgfd3d_amd import Batch
    rng = np.random.default_rng(9)
    cases = []
    for n, (dims, pml, top) in enumerate([((70, 23, 19), (1, 1, 1, 1, 1, 0), True),
                                          ((19, 22, 17), (1, 0, 1, 1, 1, 1), False),
                                          ((24, 17, 21), (0, 1, 1, 0, 1, 0), True)]):
        ni, nj, nk = dims
        case = cs.make_case(ni, nj, nk, seed=40 + n, free_top=top, pml=pml, layers=3,
                            medium=medium)
        case["src"] = cs.ricker_point_source(case, 3 + ni // 2, 3 + nj // 2, 3 + nk // 2, NT,
                                             DT, fc=10.0, t0=0.1)
        ridx, rw = cs.surface_receivers(case, n=3, seed=n)
        cases.append((case, ridx, rw))
    qs = rng.uniform(10.0, 100.0, (cases[2][0]["nk"] + 6, cases[2][0]["nj"] + 6,
                                   cases[2][0]["ni"] + 6)).astype(np.float32)

    def state(s, n):
        # peak ground motion exists on free-surface grids only
        if cases[n][0]["free_top"]:
            return _state(s)
        return s.wavefield(), s.pml_aux(), s.seismograms()

    def make(n):
        case, ridx, rw = cases[n]
        s = _solver(case, ridx, rw, variant=16)
        if n == 2 and medium != "visco":     # Graves-Qs: elastic media only
            s.set_graves_qs(qs, 2.0)
        return s

    ref = []
    for n in range(3):
        s = make(n)
        s.run(0, 35)
        ref.append(state(s, n))
        s.close()
    sv = [make(n) for n in range(3)]
    b = Batch(sv)
    monkeypatch.setenv("CGFDH_BATCH_FUSED", "1")
    b.run(0, 19)
    monkeypatch.setenv("CGFDH_BATCH_FUSED", "0")
    b.run(19, 8)
    monkeypatch.setenv("CGFDH_BATCH_FUSED", "1")
    b.run(27, 8)
    print(f"fused batch ({medium}): last call {b.last_ms():.2f} ms")
    for n, s in enumerate(sv):
        assert _same(state(s, n), ref[n]), f"solver {n}"
    b.close()
    for s in sv:
        s.close()
