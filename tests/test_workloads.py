"""CPU checks of the benchmark workload builders: algorithmic bytes per point
against a hand count of the stage traffic, medium volumes (VTI
from Thomsen parameters, a rotated VTI tensor keeps its eigenvalues, GMB
coefficients from the host builder), and the batched configuration's
sizes."""
import numpy as np
import pytest

from cgfd3d_amd import workloads as w
from oracle import pyoracle as po

# the stable-step estimate runs on the GPU: CPU checks pass a step and skip it
FIXED_DT = dict(dt=0.004, check_dt=False)


def _bytes(nw, ncoef):
    # stage 1: R pre, W tmp+end; stages 2-3: R tmp,pre,end, W tmp+end; stage 4: R tmp,end, W end
    per_stage = [nw + 2 * nw, 3 * nw + 2 * nw, 3 * nw + 2 * nw, 2 * nw + nw]
    return 4 * (sum(per_stage) + 4 * (9 + ncoef))


@pytest.mark.parametrize("kind,nw,ncoef,total", [("iso", 9, 3, 768), ("vti", 9, 6, 816),
                                                 ("aniso", 9, 22, 1072),
                                                 ("visco", 27, 9, 2016)])
def test_bytes_per_point(kind, nw, ncoef, total):
    wl = (w.iso_topo_grf(ni=24, nj=20, nk=16, pml=4, n_side=2, **FIXED_DT) if kind == "iso"
          else w.media_topo_grf(kind, ni=24, nj=20, nk=16, pml=4, n_side=2, **FIXED_DT))
    assert wl.bytes_per_point() == _bytes(nw, ncoef) == total
    nsurf = 24 * 20
    n_aux = 2 * 5 * 20 * 16 + 2 * 24 * 5 * 16 + 24 * 20 * 5
    assert wl.n_aux() == n_aux
    surf = 352 + (36 if kind == "visco" else 0)
    assert wl.algorithmic_bytes_per_step() == total * 24 * 20 * 16 + 576 * n_aux + surf * nsurf


def test_vti_from_thomsen_and_rotation_invariants():
    v = w.media_topo_grf("vti", ni=20, nj=18, nk=16, pml=4, n_side=2, **FIXED_DT)
    a = w.media_topo_grf("aniso", ni=20, nj=18, nk=16, pml=4, n_side=2, **FIXED_DT)
    c11, c13, c33, c55, c66, slw = [np.asarray(x, np.float64) for x in v.med_vols]
    assert (c11 >= c33 * (1 - 1e-6)).all() and (c66 >= c55 * (1 - 1e-6)).all()   # eps, gamma >= 0
    assert np.array_equal(v.med_vols[5], a.med_vols[21])
    ut = [(p, q) for p in range(6) for q in range(p, 6)]
    rng = np.random.default_rng(0)
    for _ in range(5):
        k, j, i = (int(rng.integers(3, n + 3)) for n in (16, 18, 20))
        Cv = np.zeros((6, 6))
        Cv[0, 0] = Cv[1, 1] = c11[k, j, i]; Cv[2, 2] = c33[k, j, i]
        Cv[0, 1] = Cv[1, 0] = c11[k, j, i] - 2 * c66[k, j, i]
        Cv[0, 2] = Cv[2, 0] = Cv[1, 2] = Cv[2, 1] = c13[k, j, i]
        Cv[3, 3] = Cv[4, 4] = c55[k, j, i]; Cv[5, 5] = c66[k, j, i]
        Ca = np.zeros((6, 6))
        for q, (p, r) in enumerate(ut):
            Ca[p, r] = Ca[r, p] = a.med_vols[q][k, j, i]
        # rotation invariants of the stiffness tensor: C_iijj and C_ijij in Voigt form
        def inv(C):
            return (C[0, 0] + C[1, 1] + C[2, 2] + 2 * (C[0, 1] + C[0, 2] + C[1, 2]),
                    C[0, 0] + C[1, 1] + C[2, 2] + 2 * (C[3, 3] + C[4, 4] + C[5, 5]))
        assert np.allclose(inv(Ca), inv(Cv), rtol=2e-6)
        assert np.linalg.eigvalsh(Ca).min() > 0


def test_visco_workload_uses_the_host_gmb_builder():
    wl = w.media_topo_grf("visco", ni=20, nj=18, nk=16, pml=4, n_side=2, **FIXED_DT)
    assert wl.medium_type == w.MEDIUM_VISCO_ISO and wl.nmaxwell == 3
    assert len(wl.med_vols) == 3 + 6 and len(wl.wl) == 3
    # relaxation frequencies log-spaced over 0.5-5 Hz (2 pi f)
    assert np.allclose(wl.wl, 2 * np.pi * np.array([0.5, 0.5 * 10 ** 0.5, 5.0]), rtol=1e-5)


def test_batch_configuration_sizes():
    sizes = ((2, (24, 24, 20)), (1, (28, 24, 20)))
    wls = list(w.batch_workloads(sizes=sizes, pml=4, n_side=2, **FIXED_DT))
    assert [x.name.split("_")[1] for x in wls] == ["24x24x20", "24x24x20", "28x24x20"]
    assert len({x.src["ijk"][0] * 1000 + x.src["ijk"][1] for x in wls}) > 1
    total = 24 * 128 ** 3 + 24 * 192 * 192 * 128 + 16 * 256 * 256 * 160
    assert total == 331_350_016     # 331.35 M points in the default batch


def test_band_limited_store_round_trip():
    """Stored FD runs keep everything below 8 Hz: a trace made of sinusoids
    under 8 Hz comes back to float16 precision, one at 12 Hz is dropped."""
    from cgfd3d_amd import surrogate_data as sd
    dt, nt = 0.008, 1024
    t = np.arange(nt) * dt
    rng = np.random.default_rng(3)
    seis = np.zeros((2, 4, 3, nt), np.float32)
    for n in range(2):
        for r in range(4):
            for c in range(3):
                # a whole number of periods over the window: one FFT bin
                f = np.floor(rng.uniform(0.5, 7.0) * nt * dt) / (nt * dt)
                seis[n, r, c] = np.sin(2 * np.pi * f * t + r)
    z, nt2 = sd.compress_runs(seis, dt)
    assert z.dtype == np.float16 and z.shape[3] == int(8.0 * nt * dt) + 1
    back = sd.decompress_runs(z, nt2)
    assert np.max(sd.relative_misfit(back.reshape(2, -1), seis.reshape(2, -1))) < 2e-3
    f12 = np.floor(12.0 * nt * dt) / (nt * dt)
    hi = np.sin(2 * np.pi * f12 * t).astype(np.float32).reshape(1, 1, 1, nt)
    z, _ = sd.compress_runs(hi, dt)
    assert np.abs(sd.decompress_runs(z, nt)).max() < 1e-2


def test_receiver_rows_line_up_with_traces():
    """Per-receiver rows: features and traces of (source, receiver) pairs in
    the same order (receivers fastest), offsets = receiver minus source."""
    from cgfd3d_amd import surrogate_data as sd
    g = sd.SurrogateGrid(ni=48, nj=46, nk=30, nt=128, n_side=2, pml=5)
    p, _ = g.sample_sources(3, 2)
    x = sd.receiver_features(p, g)
    rxy = sd.receiver_xy(g)
    assert x.shape == (3 * 4, 32)
    Lx, Ly = (g.ni - 1) * g.dh, (g.nj - 1) * g.dh
    for s in range(3):
        for r in range(4):
            row = x[s * 4 + r]
            assert np.isclose(row[6], (rxy[r, 0] - p[s, 0]) / Lx, atol=1e-6)
            assert np.isclose(row[7], (rxy[r, 1] - p[s, 1]) / Ly, atol=1e-6)
    seis = np.random.default_rng(0).random((3, 4, 3, 128)).astype(np.float32)
    rows = sd.to_receiver_layout(seis)
    assert np.array_equal(rows[2 * 4 + 1], seis[2, 1].T)


def test_keep_two_digi_closed_forms():
    """blk_keep_two_digi keeps the two leading significant digits of the
    float it is given, truncated: 0.005 as a float prints 4.9999999e-03 and
    becomes 0.0049."""
    for x, want in [(0.0048372918, 0.0048), (0.012345, 0.012), (0.0099999, 0.0099),
                    (0.005, 0.0049), (0.0047, 0.0046), (0.00625, 0.0062), (3.14159, 3.1)]:
        assert w.keep_two_digi(x) == float(np.float32(want)), x


@pytest.mark.skipif(not po.ref_available(), reason="reference build absent")
def test_keep_two_digi_matches_reference():
    """Against the reference's own blk_keep_two_digi (blk_t.c:3004-3017,
    compiled into oracle/_ref) on 2000 log-spaced steps."""
    import ctypes as C
    ref = po.load_ref()
    f = ref.blk_keep_two_digi
    f.argtypes = [C.c_float]
    f.restype = C.c_float
    for x in np.geomspace(1e-5, 0.5, 2000).astype(np.float32):
        assert w.keep_two_digi(float(x)) == f(float(x)), x
