"""GPU: the benchmark workloads, built through their own solver set-up path
(device metric, medium setters, PML tables, device free-surface matrices),
advance bit-identically to the oracle on small grids of every medium, and a
small batch of the batched configuration equals its members' own runs and,
for one member, the oracle."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests import cases as cs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["iso", "vti", "aniso", "visco"])
def test_workload_matches_oracle(kind):
    from cgfd3d_amd import workloads as w
    if kind == "iso":
        wl = w.iso_topo_grf(ni=30, nj=26, nk=22, pml=5, n_side=3, nt_total=40, dt=0.004)
    else:
        wl = w.media_topo_grf(kind, ni=30, nj=26, nk=22, pml=5, n_side=3, nt_total=40)
    s = wl.build_solver(variant=4)
    s.run(0, 14)
    w_g, a_g, s_g = s.wavefield(), s.pml_aux(), s.seismograms()
    s.close()
    case = cs.workload_case(wl)
    w_o, a_o, s_o = po.Oracle(case).run(wl.dt, 14, recv_idx=wl.recv_idx, recv_w=wl.recv_w,
                                        nt_total=wl.nt_total)
    print(f"{kind}: rel L2 wavefield {cs.rel_l2(w_g, w_o):.1e}, seismograms "
          f"{cs.rel_l2(s_g, s_o):.1e}")
    assert np.abs(w_o).max() > 0
    assert np.array_equal(w_g, w_o) and np.array_equal(a_g, a_o) and np.array_equal(s_g, s_o)


def test_small_batch_configuration_equals_member_runs():
    from cgfd3d_amd import Batch, workloads as w
    sizes = ((2, (28, 24, 20)), (1, (32, 28, 24)))
    wls = list(w.batch_workloads(sizes=sizes, pml=5, n_side=2, nt_total=40))
    ref = []
    for wl in wls:
        s = wl.build_solver(variant=4)
        s.run(0, 20)
        ref.append((s.wavefield(), s.seismograms()))
        s.close()
    sv = [wl.build_solver(variant=4) for wl in wls]
    b = Batch(sv)
    b.autotune(0)
    b.run(0, 20)
    for n, s in enumerate(sv):
        assert np.array_equal(s.wavefield(), ref[n][0]), n
        assert np.array_equal(s.seismograms(), ref[n][1]), n
    # and one batched member against the oracle directly
    case = cs.workload_case(wls[2])
    w_o, _, s_o = po.Oracle(case).run(wls[2].dt, 20, recv_idx=wls[2].recv_idx,
                                      recv_w=wls[2].recv_w, nt_total=wls[2].nt_total)
    assert np.abs(w_o).max() > 0
    assert np.array_equal(sv[2].wavefield(), w_o)
    assert np.array_equal(sv[2].seismograms(), s_o)
    b.close()
    for s in sv:
        s.close()
