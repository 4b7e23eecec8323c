"""GPU: the learned-surrogate configuration at its own shapes.

(a) The FD data generator on the configuration's fixed 160 x 160 x 100 grid
(seeded topography and medium, CFS-PML 10, free surface, 8 x 8 surface
receivers), through the batched solver the surrogate's training data comes
from, against the oracle: 16 steps of two sources, seismograms bit-identical.
(b) One forward and backward pass of the configuration's network (7 source
parameters -> MLP 512 x 4 -> 1-D transposed-convolution decoder to 3
components x 64 receivers x 1024 samples, moment-tensor output) on the
library's BF16 MFMA kernels against the CPU oracle (oracle/surrogate_oracle.py,
bf16 storage emulated at the device's rounding points).  Tolerance: the
forward output rel L2 <= 2e-3 as for the other network checks; the
gradients rel L2 <= 2^-8, the unit roundoff of bf16 (an accumulation-order
difference can flip an intermediate bf16 rounding, which moves that value by
at most one bf16 step, and the backward pass through six layers carries such
flips into every gradient; a wrong term moves them by O(1))."""
import numpy as np
import pytest

from oracle import pyoracle as po
from oracle import surrogate_oracle as so
from tests import cases as cs

pytestmark = pytest.mark.gpu


def test_fd_generator_on_the_configuration_grid_matches_oracle():
    from cgfd3d_amd import surrogate_data as sd
    nt = 16
    grid = sd.SurrogateGrid(nt=nt)
    assert (grid.ni, grid.nj, grid.nk, grid.nrecv) == (160, 160, 100, 64)
    params, ijk = grid.sample_sources(2, 9)
    # shallow enough that the surface receivers see the source within 16 steps
    ijk[:, 2] = 3 + grid.nk - 1 - np.array([3, 5])
    gen = sd.FdGenerator(grid, 2, autotune=False)
    try:
        seis = gen.run(params, ijk)
    finally:
        gen.close()
    case = cs.workload_case(grid.wl)
    for q in range(2):
        src = dict(grid.source_tables(params[q], ijk[q]))
        src.update(nsrc=1, inc=[0, 0, 0], gauss=0)
        case["src"] = src
        _, _, s_o = po.Oracle(case).run(grid.dt, nt, recv_idx=grid.wl.recv_idx,
                                        recv_w=grid.wl.recv_w, nt_total=nt)
        want = s_o[:, :3, :nt]
        print(f"source {q}: rel L2 {cs.rel_l2(seis[q], want):.1e}", flush=True)
        assert np.abs(want).max() > 0
        assert np.array_equal(seis[q], want)


def test_configuration_network_forward_backward_match_oracle():
    import torch
    from cgfd3d_amd import surrogate as sg
    layers = sg.default_layers(64, 1024, 6)
    nt, nc = 1024, 3 * 64
    ns = 8     # the network runs multiples of 8 samples per call
    net = sg.SurrogateNet(layers, seed=5, mt_terms=6, out_shape=(nt, nc))
    rng = np.random.default_rng(13)
    x = rng.standard_normal((ns, layers[0][0])).astype(np.float32)
    m = rng.standard_normal((ns, 6)).astype(np.float32)
    y = net.forward(x, m=m).float().cpu().numpy()
    params = [(so.bf16_round(w), b) for w, b in net.params_numpy()]
    hs, g_o = so.forward(params, x, layers)
    y_o = so.mt_contract(m, g_o, nt, nc)
    e = cs.rel_l2(y.reshape(ns, nt, nc), y_o)
    print(f"configuration network forward rel L2 {e:.1e}", flush=True)
    assert e <= 2e-3
    t = rng.standard_normal((ns, nt, nc)).astype(np.float32)
    tn2 = (t.reshape(ns, -1).astype(np.float64) ** 2).sum(axis=1).astype(np.float32)
    _, dg = so.mt_loss_grad(m, g_o, t, tn2, 1e-12, nt, nc)
    grads = so.backward(params, hs, dg, layers)
    b = net._bufs[ns]
    net.forward(x, m=m)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda:0")  # noqa
    sg.mt_loss_grad(ns, nt, nc, 6, b["m"], b["y"], dev(t), dev(tn2), 1e-12, b["g"][-1],
                    b["loss"])
    net.backward(ns)
    for li in range(len(layers)):
        ew = cs.rel_l2(net.gw[li].float().cpu().numpy(), grads[li][0])
        eb = cs.rel_l2(net.gb[li].float().cpu().numpy(), grads[li][1])
        print(f"layer {li} {layers[li]}: dW rel L2 {ew:.1e}, db rel L2 {eb:.1e}", flush=True)
        assert ew <= 2.0 ** -8 and eb <= 2.0 ** -8
