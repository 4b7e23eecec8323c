"""Learned surrogate of surface seismograms.

Source parameters (x, y, depth, strike, dip, rake, f0) -> 3-component
velocity seismograms at the surface receivers of one fixed grid and medium.
Network: a 4-layer MLP of width 512 (ReLU) and a 1-D transposed-convolution
decoder: a dense layer to a 16-sample sequence of 256 channels, then three
transposed convolutions with kernel = stride = 4 (16 -> 64 -> 256 -> 1024
samples; 256 -> 128 -> 128 -> 3*nrecv channels).  Optionally (mt_terms = 6)
the last layer emits 6 basis waveform sets and a fused kernel contracts them
with the source's moment tensor: the seismograms are linear in it, so the
network only has to learn the nonlinear dependence on position, depth and
frequency.  In the channels-last layout
[sample, time, channel] every layer is one GEMM whose output rows are reread
as the next layer's input rows, so the whole forward and backward pass runs
on the BF16 matrix-core GEMM of libcgfdh (cgfds_gemm_bf16): BF16 weights and
activations, FP32 accumulation, FP32 master weights and Adam moments, FP32
biases.

Device memory and streams come from torch (plumbing); every arithmetic step
of training and inference is a kernel of this library.  There is no CPU
fallback: without the library every call raises.
"""
import ctypes as C
import math

import numpy as np

from ._lib import check, lib

EPI_NONE, EPI_RELU, EPI_MASK = 0, 1, 2
IN_FEATURES = 32        # input row width (7 parameters and derived features, zero padded)


def default_layers(nrecv=64, nt=1024, mt_terms=0, width=1):
    """(K_in, N_out, relu) per layer.  The decoder upsamples by 4 per
    transposed convolution from nt/64 samples (channels 256, 128, 128 times
    width).  With mt_terms > 0 the last layer emits that many basis waveform
    sets (one per moment-tensor component), contracted with the source's
    moment tensor."""
    t0 = nt // 64
    assert t0 * 64 == nt
    c0, c1 = 256 * width, 128 * width
    return [(IN_FEATURES, 512, 1), (512, 512, 1), (512, 512, 1), (512, 512, 1),
            (512, t0 * c0, 1),
            (c0, 4 * c1, 1), (c1, 4 * c1, 1), (c1, 4 * 3 * nrecv * max(1, mt_terms), 0)]


def layer_rows(layers, nsamp):
    """Rows of each layer's input: the output [R, N] of a layer is reread as
    [R * N / K_next, K_next]."""
    rows = [nsamp]
    for (k, n, _), nxt in zip(layers[:-1], layers[1:]):
        total = rows[-1] * n
        if total % nxt[0]:
            raise ValueError(f"layer output {n} does not regroup into rows of {nxt[0]}")
        rows.append(total // nxt[0])
    return rows


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the surrogate runs on the GPU only (no HIP device visible)")
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def gemm(M, N, K, A, lda, transA, B, ldb, transB, bias=None, epi=EPI_NONE, mask=None, ldm=0,
         C32=None, ldc32=0, C16=None, ldc16=0, splitk=1, work=None):
    """cgfds_gemm_bf16 on torch device tensors (bf16 storage for A, B, C16, mask)."""
    check(lib().cgfds_gemm_bf16(M, N, K, _ptr(A), lda, int(transA), _ptr(B), ldb, int(transB),
                                _ptr(bias), epi, _ptr(mask), ldm, _ptr(C32), ldc32, _ptr(C16),
                                ldc16, splitk, _ptr(work), _stream()), "gemm_bf16")


def colsum(M, N, G, ldg, out, splits, work):
    check(lib().cgfds_colsum_bf16(M, N, _ptr(G), ldg, _ptr(out), splits, _ptr(work), _stream()),
          "colsum_bf16")


def loss_grad(nsamp, n, Y, T, tn2, eps, G16, loss):
    check(lib().cgfds_loss_grad(nsamp, n, _ptr(Y), _ptr(T), _ptr(tn2), eps, _ptr(G16),
                                _ptr(loss), _stream()), "loss_grad")


def adam(w, m, v, g, w16, lr, t, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias corrections for step t computed here in double)."""
    check(lib().cgfds_adam(w.numel(), _ptr(w), _ptr(m), _ptr(v), _ptr(g), _ptr(w16), lr, b1, b2,
                           1.0 - b1, 1.0 - b2, 1.0 - b1 ** t, 1.0 - b2 ** t, eps, _stream()),
          "adam")


def mt_contract(nsamp, nt, nc, nk, m, g, y):
    check(lib().cgfds_mt_contract(nsamp, nt, nc, nk, _ptr(m), _ptr(g), _ptr(y), _stream()),
          "mt_contract")


def mt_loss_grad(nsamp, nt, nc, nk, m, g, T, tn2, eps, dg16, loss):
    check(lib().cgfds_mt_loss_grad(nsamp, nt, nc, nk, _ptr(m), _ptr(g), _ptr(T), _ptr(tn2), eps,
                                   _ptr(dg16), _ptr(loss), _stream()), "mt_loss_grad")


def to_bf16(src, dst):
    check(lib().cgfds_f32_to_bf16(src.numel(), _ptr(src), _ptr(dst), _stream()), "f32_to_bf16")


def _splitk(M, N, K):
    """K splits for a GEMM with few output tiles and a long reduction."""
    tiles = ((M + 63) // 64) * ((N + 63) // 64)
    if tiles >= 256 or K < 512:
        return 1
    return max(1, min((512 + tiles - 1) // tiles, K // 256))


class SurrogateNet:
    """Parameters, Adam state and activation buffers of one network on cuda:0."""

    def __init__(self, layers=None, seed=0, init_params=None, mt_terms=0, out_shape=None):
        """mt_terms > 0: the last layer's output per sample is [nt, mt_terms,
        nc] basis waveforms (out_shape = (nt, nc)), contracted with each
        sample's moment tensor into the [nt, nc] prediction."""
        torch = _torch()
        self.torch = torch
        self.dev = torch.device("cuda:0")
        self.layers = list(layers or default_layers())
        self.mt_terms = int(mt_terms)
        self.out_shape = out_shape
        if self.mt_terms and out_shape is None:
            raise ValueError("a moment-tensor output needs out_shape = (nt, nc)")
        self.w32, self.w16, self.b32 = [], [], []
        rng = np.random.default_rng(seed)
        for li, (k, n, act) in enumerate(self.layers):
            if init_params is not None:
                w = np.asarray(init_params[li][0], np.float32)
                b = np.asarray(init_params[li][1], np.float32)
            else:
                # He initialisation (ReLU), zero biases
                w = (rng.standard_normal((k, n)) * math.sqrt(2.0 / k)).astype(np.float32)
                b = np.zeros(n, np.float32)
            wt = torch.from_numpy(w).to(self.dev)
            w16 = torch.empty((k, n), dtype=torch.bfloat16, device=self.dev)
            to_bf16(wt, w16)
            self.w32.append(wt)
            self.w16.append(w16)
            self.b32.append(torch.from_numpy(b).to(self.dev))
        self.gw = [torch.zeros_like(w) for w in self.w32]
        self.gb = [torch.zeros_like(b) for b in self.b32]
        self.mw = [torch.zeros_like(w) for w in self.w32]
        self.vw = [torch.zeros_like(w) for w in self.w32]
        self.mb = [torch.zeros_like(b) for b in self.b32]
        self.vb = [torch.zeros_like(b) for b in self.b32]
        self.t = 0
        self._bufs = {}

    def n_params(self):
        return sum(w.numel() + b.numel() for w, b in zip(self.w32, self.b32))

    def params_numpy(self):
        """FP32 master weights and biases as [(W [K, N], b [N])]."""
        return [(w.cpu().numpy(), b.cpu().numpy()) for w, b in zip(self.w32, self.b32)]

    def _buffers(self, nsamp):
        if nsamp in self._bufs:
            return self._bufs[nsamp]
        torch = self.torch
        if nsamp % 8:
            raise ValueError("the number of samples per call must be a multiple of 8")
        rows = layer_rows(self.layers, nsamp)
        bf = torch.bfloat16
        h = [torch.empty((r, k), dtype=bf, device=self.dev)
             for r, (k, _, _) in zip(rows, self.layers)]
        kl, nl, _ = self.layers[-1]
        y = torch.empty((rows[-1], nl), dtype=torch.float32, device=self.dev)
        g = [torch.empty((r, n), dtype=bf, device=self.dev)
             for r, (_, n, _) in zip(rows, self.layers)]
        wk = 0
        for r, (k, n, _) in zip(rows, self.layers):
            s = _splitk(k, n, r)
            wk = max(wk, s * k * n if s > 1 else 0, 256 * n)
        work = torch.empty(wk, dtype=torch.float32, device=self.dev)
        x32 = torch.empty((nsamp, self.layers[0][0]), dtype=torch.float32, device=self.dev)
        loss = torch.empty(nsamp, dtype=torch.float32, device=self.dev)
        yc = None
        if self.mt_terms:
            nt, nc = self.out_shape
            if y.numel() != nsamp * nt * nc * self.mt_terms:
                raise ValueError("the last layer does not match out_shape and mt_terms")
            yc = torch.empty((nsamp * nt, nc), dtype=torch.float32, device=self.dev)
        m = torch.zeros((nsamp, max(1, self.mt_terms)), dtype=torch.float32, device=self.dev)
        self._bufs[nsamp] = dict(rows=rows, h=h, y=y, g=g, work=work, x32=x32, loss=loss,
                                 yc=yc, m=m)
        return self._bufs[nsamp]

    def _set_m(self, b, m):
        if not self.mt_terms:
            return
        if m is None:
            raise ValueError("this network needs each sample's moment tensor")
        if isinstance(m, np.ndarray):
            b["m"].copy_(self.torch.from_numpy(np.ascontiguousarray(m, np.float32)))
        else:
            b["m"].copy_(m)

    def forward(self, x, nsamp=None, m=None):
        """x: [nsamp, IN_FEATURES] float32 (host numpy or device tensor); m:
        [nsamp, mt_terms] moment tensors when the network has that output.
        Returns the FP32 prediction (a view of an internal buffer, valid
        until the next call with the same nsamp): [rows_last, N_last], or
        [nsamp * nt, nc] after the moment-tensor contraction."""
        torch = self.torch
        nsamp = nsamp or x.shape[0]
        b = self._buffers(nsamp)
        self._set_m(b, m)
        if isinstance(x, np.ndarray):
            b["x32"].copy_(torch.from_numpy(np.ascontiguousarray(x, np.float32)))
        else:
            b["x32"].copy_(x)
        to_bf16(b["x32"], b["h"][0])
        rows = b["rows"]
        last = len(self.layers) - 1
        for li, (k, n, act) in enumerate(self.layers):
            out16 = b["h"][li + 1] if li < last else None
            gemm(rows[li], n, k, b["h"][li], k, 0, self.w16[li], n, 0, bias=self.b32[li],
                 epi=EPI_RELU if act else EPI_NONE,
                 C32=b["y"] if li == last else None, ldc32=n,
                 C16=out16, ldc16=n)
        if self.mt_terms:
            nt, nc = self.out_shape
            mt_contract(nsamp, nt, nc, self.mt_terms, b["m"], b["y"], b["yc"])
            return b["yc"]
        return b["y"]

    def backward(self, nsamp):
        """Gradients of every parameter from the bf16 output gradient already
        in the last layer's gradient buffer (g[-1])."""
        b = self._bufs[nsamp]
        rows = b["rows"]
        work = b["work"]
        for li in range(len(self.layers) - 1, -1, -1):
            k, n, _ = self.layers[li]
            r = rows[li]
            g = b["g"][li]
            # dW = H^T G  (K x N, reduction over the r rows)
            s = _splitk(k, n, r)
            gemm(k, n, r, b["h"][li], k, 1, g, n, 0, C32=self.gw[li], ldc32=n, splitk=s,
                 work=work)
            colsum(r, n, g, n, self.gb[li], min(256, max(1, r // 64)), work)
            if li > 0:
                # dH = G W^T, masked by the ReLU of the previous layer, to bf16
                gemm(r, k, n, g, n, 0, self.w16[li], n, 1, epi=EPI_MASK, mask=b["h"][li], ldm=k,
                     C16=b["g"][li - 1].view(-1)[: r * k].view(r, k), ldc16=k)

    def adam_step(self, lr):
        self.t += 1
        for li in range(len(self.layers)):
            adam(self.w32[li], self.mw[li], self.vw[li], self.gw[li], self.w16[li], lr, self.t)
            adam(self.b32[li], self.mb[li], self.vb[li], self.gb[li], None, lr, self.t)

    def _loss(self, x, t, tn2, eps, m, grad):
        nsamp = x.shape[0]
        y = self.forward(x, nsamp, m)
        b = self._bufs[nsamp]
        g16 = b["g"][-1] if grad else None
        if self.mt_terms:
            nt, nc = self.out_shape
            mt_loss_grad(nsamp, nt, nc, self.mt_terms, b["m"], b["y"], t, tn2, eps, g16, b["loss"])
        else:
            loss_grad(nsamp, y.numel() // nsamp, y, t, tn2, eps, g16, b["loss"])
        return nsamp, b

    def train_step(self, x, t, tn2, lr, eps=1e-12, m=None):
        """One step on a batch: x [nsamp, IN_FEATURES], t [nsamp, ...] FP32
        target with the prediction's layout, tn2 [nsamp] squared target
        norms, m the moment tensors (moment-tensor output only).  Returns the
        device tensor of per-sample losses (before the update)."""
        nsamp, b = self._loss(x, t, tn2, eps, m, True)
        self.backward(nsamp)
        self.adam_step(lr)
        return b["loss"]

    def loss(self, x, t, tn2, eps=1e-12, m=None):
        """Per-sample relative squared misfit without an update (device tensor)."""
        return self._loss(x, t, tn2, eps, m, False)[1]["loss"]
