"""CPU checks that pin the surrogate oracle: bf16 rounding against torch's
conversion, the backward pass against finite differences of the forward pass
(float64), Adam against its closed form, and the layer-row bookkeeping of the
channels-last transposed convolutions against an explicit convolution."""
import numpy as np
import pytest

from oracle import surrogate_oracle as so

SMALL = [(32, 64, 1), (64, 64, 1), (64, 128, 1), (32, 64, 1), (16, 48, 0)]


def _params(layers, seed):
    rng = np.random.default_rng(seed)
    return [((rng.standard_normal((k, n)) * np.sqrt(2.0 / k)).astype(np.float32),
             (0.1 * rng.standard_normal(n)).astype(np.float32)) for k, n, _ in layers]


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 20000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, -3.0e38, 1e-40,
                                  np.inf, -np.inf], np.float32)]).astype(np.float32)
    # exact ties: 1 + 2^-8 (rounds to even 1.0) and 1 + 3*2^-8 (rounds up)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = so.bf16_round(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.isnan(so.bf16_round(np.array([np.nan], np.float32)))[0]
    assert np.array_equal(so.from_bf16_bits(so.bf16_bits(x)), got)


def test_layer_rows():
    assert so.layer_rows(SMALL, 16) == [16, 16, 16, 64, 256]
    # the full network: 7 parameters (padded to 32) -> 3 x 64 receivers x 1024 samples
    from cgfd3d_amd.surrogate import default_layers, layer_rows
    layers = default_layers(64, 1024)
    rows = layer_rows(layers, 4)
    assert rows == so.layer_rows(layers, 4)
    assert rows[-1] * layers[-1][1] == 4 * 3 * 64 * 1024
    assert rows[-1] // 4 * 4 == 4 * 1024 // 4 and layers[-1][1] == 4 * 192


def test_channels_last_gemm_is_a_transposed_convolution():
    """A GEMM of [T, Ci] x [Ci, 4*Co] reread as [4T, Co] equals the 1-D
    transposed convolution with kernel = stride = 4 written as a loop."""
    rng = np.random.default_rng(1)
    T, ci, co, s = 5, 6, 3, 4
    x = rng.standard_normal((T, ci))            # [time, channel]
    w = rng.standard_normal((ci, co, s))        # [in, out, tap]
    b = rng.standard_normal(co)
    y = np.zeros((T * s, co))
    for t in range(T):
        for r in range(s):
            for o in range(co):
                y[t * s + r, o] = b[o] + sum(x[t, c] * w[c, o, r] for c in range(ci))
    wg = w.transpose(0, 2, 1).reshape(ci, s * co)       # column index r*co + o
    yg = (x @ wg + np.tile(b, s)).reshape(T * s, co)
    assert np.allclose(y, yg, rtol=1e-12, atol=1e-12)


def test_backward_matches_finite_differences():
    layers = SMALL
    params = _params(layers, 2)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((16, 32))
    t = rng.standard_normal((16, 256 // 16 * 48))
    tn2 = (t ** 2).sum(axis=1)

    def loss_of(p):
        _, y = so.forward(p, x, layers, exact=True)
        return so.loss_grad(y, t, tn2, 0.0, exact=True)[0].mean()

    hs, y = so.forward(params, x, layers, exact=True)
    _, gy = so.loss_grad(y, t, tn2, 0.0, exact=True)
    grads = so.backward(params, hs, gy, layers, exact=True)
    h = 1e-6
    for li in range(len(layers)):
        for which in (0, 1):
            d = rng.standard_normal(params[li][which].shape)
            pp = [list(q) for q in params]
            pm = [list(q) for q in params]
            pp[li][which] = params[li][which] + h * d
            pm[li][which] = params[li][which] - h * d
            fd = (loss_of(pp) - loss_of(pm)) / (2 * h)
            an = (grads[li][which] * d).sum()
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (li, which, fd, an)


def test_planted_mistake_is_caught():
    """A backward pass with a dropped ReLU mask fails the same check."""
    layers = SMALL[:3] + [(32, 64, 0)]
    params = _params(layers, 4)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((8, 32))
    hs, y = so.forward(params, x, layers, exact=True)
    t = rng.standard_normal(y.shape)
    tn2 = (t.reshape(8, -1) ** 2).sum(axis=1)
    _, gy = so.loss_grad(y, t, tn2, 0.0, exact=True)
    good = so.backward(params, hs, gy, layers, exact=True)
    bad_hs = [np.abs(h) + 1.0 for h in hs]          # every unit looks active
    bad = so.backward(params, bad_hs, gy, layers, exact=True)
    assert not np.allclose(good[0][0], bad[0][0])


def test_adam_closed_form_first_step():
    """After one step from zero moments the update is lr * g / (|g| + eps')."""
    rng = np.random.default_rng(6)
    w = rng.standard_normal(1000).astype(np.float32)
    g = rng.standard_normal(1000).astype(np.float32)
    z = np.zeros_like(w)
    w1, m1, v1 = so.adam(w, z, z, g, 1, 1e-3)
    assert np.allclose(w - w1, 1e-3 * np.sign(g), rtol=1e-4, atol=1e-9)
    assert np.allclose(m1, 0.1 * g, rtol=1e-6) and np.allclose(v1, 0.001 * g * g, rtol=1e-5)


def test_mt_contraction_is_the_linear_combination():
    rng = np.random.default_rng(10)
    ns, nt, nk, nc = 3, 5, 6, 4
    m = rng.standard_normal((ns, nk)).astype(np.float32)
    g = rng.standard_normal((ns, nt, nk, nc)).astype(np.float32)
    y = so.mt_contract(m, g, nt, nc)
    assert np.allclose(y, np.einsum("bk,btkc->btc", m.astype(np.float64), g), rtol=1e-5,
                       atol=1e-6)
    # gradient of the contracted misfit with respect to g: finite differences (float64
    # through the exact path: d loss / d g = m * d loss / d y)
    t = rng.standard_normal((ns, nt, nc))
    tn2 = (t.reshape(ns, -1) ** 2).sum(axis=1)
    loss, dg = so.mt_loss_grad(m, g, t, tn2, 0.0, nt, nc)
    _, gy = so.loss_grad(np.einsum("bk,btkc->btc", m.astype(np.float64), g), t, tn2, 0.0,
                         exact=True)
    want = m[:, None, :, None] * gy.reshape(ns, nt, 1, nc)
    assert np.allclose(dg, want, rtol=1e-2, atol=1e-6)   # dg is bf16
