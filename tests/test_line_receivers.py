"""Receiver-line points (io_line_keep, io_funcs.c:1626-1645: the wavefield
value at a grid point after every step) ride the receiver path with all 8
corner indices on the point and zero weights.  The oracle's receiver
sampling, which the engine's equals bit for bit, then records exactly the
point's value; checked here step by step against the wavefield."""
import numpy as np

from tests import cases as cs
from oracle import pyoracle as po


def test_zero_weight_receiver_records_the_grid_point():
    case = cs.make_case(12, 11, 10, seed=12, pml=(1, 1, 1, 1, 1, 0), layers=3)
    nx, ny = case["ni"] + 6, case["nj"] + 6
    pts = [(5, 6, 7), (3, 3, 12), (14, 13, 4)]           # (i, j, k) physical points
    idx = np.array([[i + j * nx + k * nx * ny] * 8 for i, j, k in pts], np.int32)
    wts = np.zeros((len(pts), 3), np.float32)
    orc = po.Oracle(case)
    nt = 5
    w = cs.random_wavefield(case, seed=13).reshape(-1)
    aux = None
    vol = w.size // cs.ncomp(case)
    for it in range(nt):
        w, aux, seis = orc.run(0.004, 1, w, aux, it0=it, recv_idx=idx, recv_w=wts, nt_total=nt)
        W = w.reshape(cs.ncomp(case), vol)
        want = W[:, idx[:, 0]].T                          # (receiver, component)
        assert np.abs(want).max() > 0
        assert np.array_equal(seis[:, :, it], want), it
