"""CPU oracle of the surrogate network (numpy): forward, backward, loss and
Adam, with bf16 storage emulated at exactly the points where the device
stores bf16 (weights, activations, gradients fed to GEMMs) and float64
accumulation of every product sum.

TEST INFRASTRUCTURE ONLY: imported by tests/ and the smoke check, never by
the product package.

The network has no counterpart in the reference (it is a learned stand-in
for the forward solver's surface seismograms).  Each function is pinned by
the tests against mathematics rather than a reference file: the backward
pass against central finite differences of the forward pass in float64, the
bf16 rounding against torch's float32 -> bfloat16 conversion, Adam against
its closed form after one step.
"""
import numpy as np


def bf16_round(x):
    """float32 -> nearest bf16 (ties to even), returned as float32 values."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).copy()
    out[nan] = np.float32(np.nan)
    return out


def bf16_bits(x):
    """bf16 words (uint16) of float32 values after rounding."""
    return (bf16_round(x).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(h):
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def layer_rows(layers, nsamp):
    """Rows of each layer's input matrix: the output [R, N] of a layer is
    reread as [R * N / K_next, K_next] (channels-last sequence upsampling)."""
    rows = [nsamp]
    for (k, n, _), nxt in zip(layers[:-1], layers[1:]):
        total = rows[-1] * n
        assert total % nxt[0] == 0
        rows.append(total // nxt[0])
    return rows


def forward(params, x0, layers, exact=False):
    """x0: [nsamp, K0] float32 input.  params: list of (W [K, N], b [N]).
    Returns (H list: the bf16-rounded inputs of each layer, Y: fp32 output).
    exact: float64 throughout, no bf16 rounding (for finite differences)."""
    rnd = (lambda a: np.asarray(a, np.float64)) if exact else bf16_round
    rows = layer_rows(layers, x0.shape[0])
    h = rnd(x0).reshape(rows[0], layers[0][0])
    hs = []
    y = None
    for li, ((k, n, act), (w, b)) in enumerate(zip(layers, params)):
        hs.append(h)
        z = h.astype(np.float64) @ np.asarray(w, np.float64)
        if exact:
            z = z + np.asarray(b, np.float64)
        else:
            z = z.astype(np.float32) + np.asarray(b, np.float32)
        if act:
            z = np.maximum(z, 0)
        if li + 1 < len(layers):
            h = rnd(z).reshape(rows[li + 1], layers[li + 1][0])
        else:
            y = z if exact else z.astype(np.float32)
    return hs, y


def loss_grad(y, t, tn2, eps, exact=False):
    """Per-sample relative squared misfit and its gradient dL/dY for
    L = mean_b |Y_b - T_b|^2 / (tn2_b + eps)."""
    ns = len(tn2)
    yb = y.reshape(ns, -1)
    tb = t.reshape(ns, -1)
    f = np.float64 if exact else np.float32
    d = yb.astype(f) - tb.astype(f)
    den = (np.asarray(tn2, f) + f(eps))
    loss = (d.astype(np.float64) ** 2).sum(axis=1) / den
    g = d * (f(2.0) / (f(ns) * den))[:, None]
    return loss, g.reshape(y.shape)


def backward(params, hs, gy, layers, exact=False):
    """gy: dL/dY of the last layer.  Returns [(dW, db)] per layer.  The
    gradient entering each layer's GEMMs is bf16-rounded (unless exact) and
    the bias gradient is the column sum of that rounded gradient."""
    rnd = (lambda a: np.asarray(a, np.float64)) if exact else bf16_round
    grads = [None] * len(layers)
    g = rnd(gy)
    for li in range(len(layers) - 1, -1, -1):
        k, n, _ = layers[li]
        w = np.asarray(params[li][0], np.float64)
        g = g.reshape(-1, n)
        h = hs[li]
        dw = h.astype(np.float64).T @ g.astype(np.float64)
        db = g.astype(np.float64).sum(axis=0)
        if not exact:
            dw, db = dw.astype(np.float32), db.astype(np.float32)
        grads[li] = (dw, db)
        if li > 0:
            dh = g.astype(np.float64) @ w.T
            if not exact:
                dh = dh.astype(np.float32)
            mask = h > 0            # ReLU of the previous layer (its output is h)
            g = rnd(np.where(mask, dh, 0))
    return grads


def adam(w, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in float32 with the device's evaluation order."""
    f = np.float32
    c1, c2 = f(1.0 - b1), f(1.0 - b2)
    bc1, bc2 = f(1.0 - b1 ** t), f(1.0 - b2 ** t)
    m = f(b1) * m + c1 * g
    v = f(b2) * v + c2 * (g * g)
    mh = m / bc1
    vh = v / bc2
    w = w - f(lr) * (mh / (np.sqrt(vh) + f(eps)))
    return w.astype(np.float32), m.astype(np.float32), v.astype(np.float32)


def mt_contract(m, g, nt, nc):
    """y[b][t][c] = sum_k m[b][k] g[b][t][k][c] in float32, k summed in
    increasing order (the device's order)."""
    m = np.asarray(m, np.float32)
    ns, nk = m.shape
    g = np.asarray(g, np.float32).reshape(ns, nt, nk, nc)
    y = m[:, None, 0, None] * g[:, :, 0, :]
    for k in range(1, nk):
        y = y + m[:, None, k, None] * g[:, :, k, :]
    return y


def mt_loss_grad(m, g, t, tn2, eps, nt, nc):
    """Contraction, per-sample relative misfit and the bf16 gradient with
    respect to the basis waveforms g ([ns, nt, nk, nc])."""
    m = np.asarray(m, np.float32)
    ns, nk = m.shape
    f = np.float32
    y = mt_contract(m, g, nt, nc)
    d = y - np.asarray(t, np.float32).reshape(ns, nt, nc)
    den = np.asarray(tn2, f) + f(eps)
    loss = (d.astype(np.float64) ** 2).sum(axis=(1, 2)) / den
    scale = f(2.0) / (f(ns) * den)
    dy = d * scale[:, None, None]
    dg = bf16_round(m[:, None, :, None] * dy[:, :, None, :])
    return loss, dg
