"""GPU: a C program built only from include/cgfdh.h (tests/c/abi_replay.c)
drives the engine in the drop-in shim's call order -- device metric, medium,
PML tables from the host builder, device free-surface matrices, moment and
strict surface-force sources through cgfdh_set_sources, receivers plus
receiver-line points, Graves-Qs, runs chunked at output steps with a packed
sub-volume read after each chunk, then seismograms, peak ground motion and
the wavefield.  Its outputs equal the oracle stepped with Graves-Qs between
steps (wavefield, seismograms, every sub-volume) and the same calls made
through the Python binding (peak ground motion)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import cases as cs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "c", "abi_replay")


def _inputs():
    from cgfd3d_amd import workloads as wk
    ni, nj, nk, nlay = 26, 24, 18, 4
    wl = wk.iso_topo_grf(ni=ni, nj=nj, nk=nk, pml=nlay, n_side=3, nt_total=30, dt=0.004)
    case = cs.workload_case(wl)
    nk2 = 3 + nk - 1
    nt = wl.nt_total
    src = cs.ricker_force_source(case, [(3 + 13, 3 + 11, nk2), (3 + 9, 3 + 14, 3 + 8)], nt, wl.dt,
                                 fxyz=[(0.0, 0.0, 1.0e15), (2e14, -1e14, 3e14)], fc=10.0,
                                 t0=0.08, strict=1, moment=(1e15, 1e15, 1e15, 0.0, 0.0, 0.0))
    case["src"] = src
    rng = np.random.default_rng(3)
    qs = rng.uniform(20.0, 150.0, (nk + 6, nj + 6, ni + 6)).astype(np.float32)
    nx, ny = ni + 6, nj + 6
    lines = np.array([(3 + 5) + (3 + 6) * nx + (3 + 10) * nx * ny,
                      (3 + 20) + (3 + 3) * nx + nk2 * nx * ny], np.int32)
    out_steps = np.array([7, 8, 19], np.int32)
    box = np.array([3 + 4, 9, 2, 3 + 2, 7, 3, 3 + 5, 6, 2], np.int32)
    return wl, case, qs, lines, out_steps, box


def _write_input(path, wl, case, qs, lines, out_steps, box):
    src = case["src"]
    nsurf = len(src["surf_idx"])
    ridx = np.ascontiguousarray(wl.recv_idx, np.int32)
    with open(path, "wb") as f:
        np.array([wl.ni, wl.nj, wl.nk, wl.nt_total, wl.pml_layers[0], src["nsrc"],
                  src["max_nt"], nsurf, len(ridx), len(lines), len(out_steps), src["strict"], 1],
                 np.int32).tofile(f)
        np.array([wl.dt, wl.pml_vel, wl.pml_alpha, wl.pml_beta, 2.0], np.float32).tofile(f)
        for a in (wl.xyz, wl.lam, wl.mu, wl.slw):
            np.ascontiguousarray(a, np.float32).tofile(f)
        np.array(src["ijk"], np.int32).tofile(f)
        np.array(src["it_begin"], np.int32).tofile(f)
        np.array(src["it_end"], np.int32).tofile(f)
        np.asarray(src["mij"], np.float32).tofile(f)
        np.asarray(src["fxyz"], np.float32).tofile(f)
        np.array(src["surf_idx"], np.int32).tofile(f)
        np.asarray(src["frate"], np.float32).tofile(f)
        ridx.tofile(f)
        np.ascontiguousarray(wl.recv_w, np.float32).tofile(f)
        lines.tofile(f)
        qs.tofile(f)
        out_steps.tofile(f)
        box.tofile(f)


def _oracle(case, wl, qs, recv_idx, recv_w, out_steps, box):
    """The oracle stepped one step at a time with Graves-Qs after each
    (drv_rk_curv_col.c:414-416), receivers sampled after it."""
    orc = po.Oracle(case)
    lib = po.load_oracle()
    ni, nj, nk = wl.ni, wl.nj, wl.nk
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    nt = wl.nt_total
    w = np.zeros(9 * nx * ny * nz, np.float32)
    aux = None
    seis = np.zeros((len(recv_idx), 9, nt), np.float32)
    boxes = []
    for it in range(nt):
        w, aux, _ = orc.run(wl.dt, 1, w, aux, it0=it, nt_total=nt)
        w = np.ascontiguousarray(w, np.float32)
        lib.orc_graves_qs(ni, nj, nk, 9, po.ptr(w), po.ptr(qs), 2.0, wl.dt)
        W = w.reshape(9, -1)
        for r in range(len(recv_idx)):
            ix = recv_idx[r]
            Lx2, Ly2, Lz2 = (np.float32(x) for x in recv_w[r])
            Lx1, Ly1, Lz1 = (np.float32(1.0 - float(x)) for x in recv_w[r])
            for c in range(9):
                v = W[c]
                seis[r, c, it] = (v[ix[0]] * Lx1 * Ly1 * Lz1 + v[ix[1]] * Lx2 * Ly1 * Lz1
                                  + v[ix[2]] * Lx1 * Ly2 * Lz1 + v[ix[3]] * Lx2 * Ly2 * Lz1
                                  + v[ix[5]] * Lx1 * Ly1 * Lz2 + v[ix[5]] * Lx2 * Ly1 * Lz2
                                  + v[ix[6]] * Lx1 * Ly2 * Lz2 + v[ix[7]] * Lx2 * Ly2 * Lz2)
        if it in out_steps:
            i1, n1, d1, j1, n2, d2, k1, n3, d3 = (int(v) for v in box)
            W4 = w.reshape(9, nz, ny, nx)
            boxes.append(W4[:, k1:k1 + n3 * d3:d3, j1:j1 + n2 * d2:d2, i1:i1 + n1 * d1:d1].copy())
    return w, seis, boxes


def test_c_caller_replays_the_shim_call_order():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run the package build first)")
    wl, case, qs, lines, out_steps, box = _inputs()
    assert len(case["src"]["surf_idx"]) == 1
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        _write_input(fin, wl, case, qs, lines, out_steps, box)
        print("abi_replay:", wl.ni, wl.nj, wl.nk, wl.nt_total, flush=True)
        r = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr, flush=True)
        assert r.returncode == 0
        out = np.fromfile(fout, np.float32)
    ni, nj, nk, nt = wl.ni, wl.nj, wl.nk, wl.nt_total
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    nall = len(wl.recv_idx) + len(lines)
    o = 0
    seis = out[o:o + nall * 9 * nt].reshape(nall, 9, nt); o += nall * 9 * nt
    pg = out[o:o + 15 * nx * ny].reshape(15, ny, nx); o += 15 * nx * ny
    w = out[o:o + 9 * nx * ny * nz]; o += 9 * nx * ny * nz
    nb = 9 * int(box[1] * box[4] * box[7])
    boxes = [out[o + n * nb:o + (n + 1) * nb] for n in range(len(out_steps))]
    assert o + len(out_steps) * nb == out.size
    # oracle: line points as receivers with 8 equal corners and zero weights
    ridx = np.concatenate([np.asarray(wl.recv_idx, np.int32),
                           np.repeat(lines[:, None], 8, axis=1)])
    rw = np.concatenate([np.asarray(wl.recv_w, np.float32), np.zeros((len(lines), 3),
                                                                     np.float32)])
    w_o, s_o, b_o = _oracle(case, wl, qs, ridx, rw, out_steps, box)
    print("C caller vs oracle: rel L2 wavefield", cs.rel_l2(w, w_o), "seismograms",
          cs.rel_l2(seis, s_o), flush=True)
    assert np.abs(s_o).max() > 0
    # Graves-Qs: the engine's exponential is within one ulp of libm's expf
    assert cs.rel_l2(w, w_o) <= 1e-6 and cs.rel_l2(seis, s_o) <= 1e-6
    for n in range(len(out_steps)):
        assert cs.rel_l2(boxes[n], b_o[n].reshape(-1)) <= 1e-6, n
    # the same calls through the Python binding give the same peak ground motion
    s = wl.build_solver()
    try:
        src = case["src"]
        s.set_sources(src["ijk"], src["it_begin"], src["it_end"], src["max_nt"], mij=src["mij"],
                      fxyz=src["fxyz"], strict=True, surf_idx=src["surf_idx"],
                      frate=src["frate"])
        s.set_receivers(ridx, rw)
        s.set_graves_qs(qs, 2.0)
        s.run(0, nt)
        pg_py, _ = s.peak_ground_motion()
    finally:
        s.close()
    assert np.abs(pg).max() > 0
    assert np.array_equal(pg, pg_py)
