"""GPU: the surrogate's BF16 matrix-core kernels against the CPU oracle.

GEMM inputs are bf16-representable, so the device and the oracle multiply the
same numbers; they differ only in the FP32 accumulation order (the oracle
accumulates in float64).  Tolerances: FP32 outputs rel-L2 <= 1e-5; bf16
outputs within one bf16 rounding step; whole-network forward/backward rel-L2
<= 2e-3 (an order difference can flip an intermediate bf16 rounding)."""
import numpy as np
import pytest

from oracle import surrogate_oracle as so
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _bf(a):
    """Host bf16-representable float32 array -> device bf16 tensor."""
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(so.bf16_round(a))).to("cuda:0").to(torch.bfloat16)


def _dev(a):
    return _torch().from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda:0")


def _host(t):
    return t.float().cpu().numpy()


@pytest.mark.parametrize("transA", [0, 1])
@pytest.mark.parametrize("transB", [0, 1])
@pytest.mark.parametrize("M,N,K", [(136, 200, 264), (64, 64, 32), (8, 24, 1032)])
def test_gemm_matches_numpy(transA, transB, M, N, K):
    from cgfd3d_amd import surrogate as sg
    torch = _torch()
    rng = np.random.default_rng(M + N + K + 2 * transA + transB)
    a = so.bf16_round(rng.standard_normal((M, K)).astype(np.float32))
    b = so.bf16_round(rng.standard_normal((K, N)).astype(np.float32))
    bias = rng.standard_normal(N).astype(np.float32)
    ref = a.astype(np.float64) @ b.astype(np.float64)
    A = _bf(a.T.copy() if transA else a)
    B = _bf(b.T.copy() if transB else b)
    lda = M if transA else K
    ldb = K if transB else N
    c32 = torch.full((M, N), np.nan, device="cuda:0")
    c16 = torch.empty((M, N), dtype=torch.bfloat16, device="cuda:0")
    sg.gemm(M, N, K, A, lda, transA, B, ldb, transB, bias=_dev(bias), epi=sg.EPI_RELU,
            C32=c32, ldc32=N, C16=c16, ldc16=N)
    want = np.maximum(ref + bias, 0)
    got = _host(c32)
    err = rel_l2(got, want)
    print(f"gemm tA={transA} tB={transB} {M}x{N}x{K}: rel L2 {err:.1e}")
    assert err <= 1e-5
    assert np.array_equal(_host(c16), so.bf16_round(got))
    # split-K into a work buffer, no epilogue
    work = torch.empty(4 * M * N, device="cuda:0")
    c32.fill_(np.nan)
    sg.gemm(M, N, K, A, lda, transA, B, ldb, transB, C32=c32, ldc32=N, splitk=4, work=work)
    assert rel_l2(_host(c32), ref) <= 1e-5
    # ReLU-backward mask epilogue
    mask = so.bf16_round(rng.standard_normal((M, N)).astype(np.float32))
    sg.gemm(M, N, K, A, lda, transA, B, ldb, transB, epi=sg.EPI_MASK, mask=_bf(mask), ldm=N,
            C16=c16, ldc16=N)
    got16 = _host(c16)
    want16 = so.bf16_round(np.where(mask > 0, ref, 0).astype(np.float32))
    tol = 2.0 ** -7 * np.abs(want16) + 1e-30
    assert (np.abs(got16 - want16) <= tol).all()
    assert (got16[mask <= 0] == 0).all()


def test_gemm_rejects_bad_shapes():
    from cgfd3d_amd import CgfdhError
    from cgfd3d_amd import surrogate as sg
    torch = _torch()
    A = torch.zeros((64, 64), dtype=torch.bfloat16, device="cuda:0")
    C = torch.zeros((64, 64), device="cuda:0")
    with pytest.raises(CgfdhError, match="multiples of 8"):
        sg.gemm(16, 16, 12, A, 64, 0, A, 64, 0, C32=C, ldc32=64)
    with pytest.raises(CgfdhError, match="split-K"):     # K = 64 in two 32-wide layers
        sg.gemm(16, 16, 64, A, 64, 0, A, 64, 0, C32=C, ldc32=64, epi=sg.EPI_RELU, splitk=2,
                work=C)
    with pytest.raises(CgfdhError, match="leading dimension"):
        sg.gemm(16, 16, 64, A, 32, 0, A, 64, 0, C32=C, ldc32=64)


def test_colsum_loss_adam_match_oracle():
    from cgfd3d_amd import surrogate as sg
    torch = _torch()
    rng = np.random.default_rng(7)
    g = so.bf16_round(rng.standard_normal((1000, 72)).astype(np.float32))
    out = torch.empty(72, device="cuda:0")
    work = torch.empty(16 * 72, device="cuda:0")
    sg.colsum(1000, 72, _bf(g), 72, out, 16, work)
    assert rel_l2(_host(out), g.astype(np.float64).sum(axis=0)) <= 1e-6

    ns, n = 8, 3000
    y = rng.standard_normal((ns, n)).astype(np.float32)
    t = rng.standard_normal((ns, n)).astype(np.float32)
    tn2 = (t.astype(np.float64) ** 2).sum(axis=1).astype(np.float32)
    loss = torch.empty(ns, device="cuda:0")
    g16 = torch.empty((ns, n), dtype=torch.bfloat16, device="cuda:0")
    sg.loss_grad(ns, n, _dev(y), _dev(t), _dev(tn2), 1e-12, g16, loss)
    l_o, g_o = so.loss_grad(y, t, tn2, 1e-12)
    assert np.allclose(_host(loss), l_o, rtol=1e-5)
    assert np.array_equal(_host(g16), so.bf16_round(g_o))     # same FP32 operations

    w = rng.standard_normal(5000).astype(np.float32)
    m = (0.01 * rng.standard_normal(5000)).astype(np.float32)
    v = (0.01 * rng.random(5000)).astype(np.float32)
    gr = rng.standard_normal(5000).astype(np.float32)
    W, Mm, V = _dev(w), _dev(m), _dev(v)
    w16 = torch.empty(5000, dtype=torch.bfloat16, device="cuda:0")
    sg.adam(W, Mm, V, _dev(gr), w16, 1e-3, 3)
    w_o, m_o, v_o = so.adam(w, m, v, gr, 3, 1e-3)
    assert np.array_equal(_host(W), w_o) and np.array_equal(_host(Mm), m_o)
    assert np.array_equal(_host(V), v_o) and np.array_equal(_host(w16), so.bf16_round(w_o))


SMALL = [(32, 64, 1), (64, 64, 1), (64, 128, 1), (32, 64, 1), (16, 48, 0)]


def test_network_forward_backward_match_oracle():
    from cgfd3d_amd.surrogate import SurrogateNet
    torch = _torch()
    net = SurrogateNet(SMALL, seed=3)
    rng = np.random.default_rng(8)
    for li, b in enumerate(net.b32):
        b.copy_(_dev(0.05 * rng.standard_normal(b.numel())))
    ns = 16
    x = rng.standard_normal((ns, 32)).astype(np.float32)
    y = _host(net.forward(x))
    params = [(so.bf16_round(w), b) for w, b in net.params_numpy()]
    hs, y_o = so.forward(params, x, SMALL)
    e = rel_l2(y, y_o)
    print(f"forward rel L2 {e:.1e}")
    assert e <= 2e-3
    t = rng.standard_normal(y.shape).astype(np.float32)
    tn2 = (t.reshape(ns, -1).astype(np.float64) ** 2).sum(axis=1).astype(np.float32)
    loss = net.loss(x, _dev(t), _dev(tn2))
    _, gy = so.loss_grad(y_o, t, tn2, 1e-12)
    grads = so.backward(params, hs, gy, SMALL)
    # gradients of this batch, then compare layer by layer
    b = net._bufs[ns]
    from cgfd3d_amd import surrogate as sg
    sg.loss_grad(ns, y.size // ns, net.forward(x), _dev(t), _dev(tn2), 1e-12, b["g"][-1],
                 b["loss"])
    net.backward(ns)
    for li in range(len(SMALL)):
        ew = rel_l2(_host(net.gw[li]), grads[li][0])
        eb = rel_l2(_host(net.gb[li]), grads[li][1])
        print(f"layer {li}: dW rel L2 {ew:.1e}, db rel L2 {eb:.1e}")
        assert ew <= 2e-3 and eb <= 2e-3
    assert np.isfinite(_host(loss)).all()


def test_training_reduces_the_misfit():
    """Fit a smooth family of sequences: the held-out misfit must fall well
    below that of the untrained network."""
    from cgfd3d_amd.surrogate import SurrogateNet
    from cgfd3d_amd.surrogate_experiment import train
    layers = [(32, 128, 1), (128, 128, 1), (128, 256, 1), (32, 64, 1), (16, 32, 0)]
    rng = np.random.default_rng(9)

    def data(n):
        p = rng.uniform(-1, 1, (n, 3))
        x = np.zeros((n, 32), np.float32)
        x[:, :3] = p
        tt = np.linspace(0, 1, 8 * 4 * 4)[None, :, None]
        ch = np.arange(8)[None, None, :]
        y = np.sin(6 * tt * (1 + 0.5 * p[:, None, None, 0]) + ch * p[:, None, None, 1]) \
            * np.exp(-p[:, None, None, 2] ** 2)
        return x, y.astype(np.float32).reshape(n, -1)

    xtr, ytr = data(256)
    xte, yte = data(64)
    net = SurrogateNet(layers, seed=1)
    tn2_tr = _dev((ytr.astype(np.float64) ** 2).sum(axis=1))
    tn2_te = _dev((yte.astype(np.float64) ** 2).sum(axis=1))
    before = float(net.loss(xte, _dev(yte), tn2_te).mean())
    hist = train(net, xtr, _dev(ytr), tn2_tr, 150, 32, 2e-3, seed=2)
    after = float(net.loss(xte, _dev(yte), tn2_te).mean())
    print(f"held-out relative squared misfit {before:.3f} -> {after:.4f} "
          f"(training loss {hist[0]:.3f} -> {hist[-1]:.4f})")
    # predicting zeros scores exactly 1
    assert after < 0.06 and after < 0.1 * before


def test_mt_contract_and_loss_grad_match_oracle():
    from cgfd3d_amd import surrogate as sg
    torch = _torch()
    rng = np.random.default_rng(11)
    ns, nt, nk, nc = 8, 40, 6, 24
    m = rng.standard_normal((ns, nk)).astype(np.float32)
    g = rng.standard_normal((ns, nt, nk, nc)).astype(np.float32)
    t = rng.standard_normal((ns, nt, nc)).astype(np.float32)
    tn2 = (t.reshape(ns, -1).astype(np.float64) ** 2).sum(axis=1).astype(np.float32)
    y = torch.empty((ns, nt, nc), device="cuda:0")
    sg.mt_contract(ns, nt, nc, nk, _dev(m), _dev(g), y)
    assert np.array_equal(_host(y), so.mt_contract(m, g, nt, nc))    # same FP32 operations
    dg = torch.empty((ns, nt, nk, nc), dtype=torch.bfloat16, device="cuda:0")
    loss = torch.empty(ns, device="cuda:0")
    sg.mt_loss_grad(ns, nt, nc, nk, _dev(m), _dev(g), _dev(t), _dev(tn2), 1e-12, dg, loss)
    l_o, dg_o = so.mt_loss_grad(m, g, t, tn2, 1e-12, nt, nc)
    assert np.array_equal(_host(dg), dg_o)
    assert np.allclose(_host(loss), l_o, rtol=1e-5)


def test_mt_network_forward_backward_match_oracle():
    from cgfd3d_amd.surrogate import SurrogateNet
    layers = [(32, 64, 1), (64, 64, 1), (64, 128, 1), (32, 64, 1), (16, 6 * 48, 0)]
    ns, nt, nc = 16, 64, 12    # last layer rows 16 per sample x (4 taps x 6 terms x 12 channels)
    net = SurrogateNet(layers, seed=4, mt_terms=6, out_shape=(nt, nc))
    rng = np.random.default_rng(12)
    x = rng.standard_normal((ns, 32)).astype(np.float32)
    m = rng.standard_normal((ns, 6)).astype(np.float32)
    y = _host(net.forward(x, m=m))
    params = [(so.bf16_round(w), b) for w, b in net.params_numpy()]
    hs, g_o = so.forward(params, x, layers)
    y_o = so.mt_contract(m, g_o, nt, nc)
    e = rel_l2(y.reshape(ns, nt, nc), y_o)
    print(f"mt forward rel L2 {e:.1e}")
    assert e <= 2e-3
    t = rng.standard_normal((ns, nt, nc)).astype(np.float32)
    tn2 = (t.reshape(ns, -1).astype(np.float64) ** 2).sum(axis=1).astype(np.float32)
    _, dg = so.mt_loss_grad(m, g_o, t, tn2, 1e-12, nt, nc)
    grads = so.backward(params, hs, dg, layers)
    from cgfd3d_amd import surrogate as sg
    b = net._bufs[ns]
    net.forward(x, m=m)
    sg.mt_loss_grad(ns, nt, nc, 6, b["m"], b["y"], _dev(t), _dev(tn2), 1e-12, b["g"][-1],
                    b["loss"])
    net.backward(ns)
    for li in range(len(layers)):
        ew = rel_l2(_host(net.gw[li]), grads[li][0])
        eb = rel_l2(_host(net.gb[li]), grads[li][1])
        print(f"layer {li}: dW rel L2 {ew:.1e}, db rel L2 {eb:.1e}")
        assert ew <= 2e-3 and eb <= 2e-3


def test_gemm_with_more_than_65535_row_tiles():
    """Row tiles run on the grid's x dimension: 4.2 M rows (65,600 tiles of
    64) work, checked on sampled rows."""
    from cgfd3d_amd import surrogate as sg
    torch = _torch()
    M, N, K = 65600 * 64, 8, 32
    rng = np.random.default_rng(13)
    a = so.bf16_round(rng.standard_normal((M, K)).astype(np.float32))
    b = so.bf16_round(rng.standard_normal((K, N)).astype(np.float32))
    c = torch.empty((M, N), device="cuda:0")
    sg.gemm(M, N, K, _bf(a), K, 0, _bf(b), N, 0, C32=c, ldc32=N)
    rows = np.concatenate([np.arange(64), rng.integers(0, M, 512), np.arange(M - 64, M)])
    got = _host(c)[rows]
    want = a[rows].astype(np.float64) @ b.astype(np.float64)
    assert rel_l2(got, want) <= 1e-5
