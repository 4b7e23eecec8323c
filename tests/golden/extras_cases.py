"""Inputs and drivers for the extras.npz golden fixture: the ablexp sponge,
Graves-Qs attenuation and a force-source RHS evaluation, each run through a
library that exposes the same entry points (the reference harness
refh_ablexp_apply / refh_graves_qs / refh_onestage, or the oracle)."""
import ctypes as C

import numpy as np

from oracle import pyoracle as po
from tests import cases as cs
from tests.test_ablexp_oracle import masks, sponge_blocks
from tests.test_force_source_oracle import _force_source


def sponge_case():
    ni, nj, nk, nc = 12, 11, 10, 15
    rng = np.random.default_rng(1)
    w = rng.standard_normal((nc, nk + 6, nj + 6, ni + 6)).astype(np.float32)
    return w, masks(ni + 6, nj + 6, nk + 6, 2), sponge_blocks(ni, nj, nk), nc


def graves_case():
    rng = np.random.default_rng(1)
    shp = (9 + 6, 10 + 6, 11 + 6)
    w = rng.standard_normal((9,) + shp).astype(np.float32)
    qs = rng.uniform(20.0, 200.0, shp).astype(np.float32)
    return w, qs


def force_case():
    case = cs.make_case(13, 12, 11, seed=6, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3)
    case["src"] = _force_source(case, 10, 3, with_moment=False)
    return case, cs.random_wavefield(case, seed=1), cs.random_aux(case, seed=1)


def sponge(fn):
    w, (ex, ey, ez), b, nc = sponge_case()
    out = np.array(w, copy=True)
    nz, ny, nx = w.shape[1:]
    fn(nx, ny, nz, nc, po.ptr(out), po.ptr(ex), po.ptr(ey), po.ptr(ez), po.ptr(b))
    return out


def graves(fn):
    w, qs = graves_case()
    out = np.array(w, copy=True)
    fn(11, 10, 9, 9, po.ptr(out), po.ptr(qs), 1.0, 0.005)
    return out


def force_ref(ref):
    case, w, aux = force_case()
    cfg, keep = cs.ref_cfg(case)
    co = np.ascontiguousarray(case["coords"])
    r = np.zeros(w.size, np.float32)
    a = np.zeros(max(aux.size, 1), np.float32)
    ref.refh_onestage(C.byref(cfg), po.ptr(co), po.ptr(case["metric"]), po.ptr(case["med"]),
                      3, 1, 1, 1, po.ptr(w.reshape(-1)), po.ptr(aux), po.ptr(r), po.ptr(a))
    return r


def force_oracle():
    case, w, aux = force_case()
    r, _ = po.Oracle(case).onestage(3, 1, 1, w.reshape(-1), aux)
    return r


def run_all(ref):
    return {"ablexp": sponge(ref.refh_ablexp_apply), "graves_qs": graves(ref.refh_graves_qs),
            "force_rhs": force_ref(ref)}
