"""CPU checks that pin the Graves-Qs attenuation oracle
(sv_curv_col_el_graves_Qs, sv_curv_col_el.c:638-667) against the reference's
own function compiled from sv_curv_col_el.c, and against closed forms.  The
engine has no kernel for it yet; this oracle is its checker."""
import numpy as np
import pytest

from oracle import pyoracle as po

NI, NJ, NK, NC = 11, 10, 9, 9


def _fields(seed):
    rng = np.random.default_rng(seed)
    shp = (NK + 6, NJ + 6, NI + 6)
    w = rng.standard_normal((NC,) + shp).astype(np.float32)
    qs = rng.uniform(20.0, 200.0, shp).astype(np.float32)
    return w, qs


def _apply(fn, w, qs, freq, dt):
    out = np.array(w, np.float32, copy=True)
    fn(NI, NJ, NK, NC, po.ptr(out), po.ptr(qs), freq, dt)
    return out


@pytest.mark.skipif(not po.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("freq,dt", [(1.0, 0.005), (2.5, 0.0031)])
def test_graves_matches_reference(freq, dt):
    w, qs = _fields(1)
    got = _apply(po.load_oracle().orc_graves_qs, w, qs, freq, dt)
    want = _apply(po.load_ref().refh_graves_qs, w, qs, freq, dt)
    assert np.array_equal(got, want)
    assert not np.array_equal(got, w)


def test_graves_closed_forms():
    orc = po.load_oracle()
    w, qs = _fields(2)
    # physical points scale by exp(-pi f dt / Qs); ghosts are left alone
    got = _apply(orc.orc_graves_qs, w, np.full_like(qs, 50.0), 2.0, 0.01)
    inner = (slice(None), slice(3, 3 + NK), slice(3, 3 + NJ), slice(3, 3 + NI))
    assert np.allclose(got[inner], w[inner] * np.exp(-np.pi * 2.0 * 0.01 / 50.0), rtol=2e-7)
    mask = np.ones(w.shape, bool)
    mask[inner] = False
    assert np.array_equal(got[mask], w[mask])
    # no attenuation for an infinite quality factor
    assert np.array_equal(_apply(orc.orc_graves_qs, w, np.full_like(qs, np.inf), 2.0, 0.01), w)
