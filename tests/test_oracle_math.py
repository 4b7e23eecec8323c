"""Checks of the oracle that do not depend on the reference build: closed
forms, invariants and special cases the mathematics fixes.  Each one is
chosen so that a plausible slip (dropped term, wrong sign, swapped operand,
wrong table entry) fails it."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import cases as cs


def _ops():
    orc = po.load_oracle()
    ix = np.zeros(8 * 4 * 6 * 5, np.int32)
    cf = np.zeros(8 * 4 * 6 * 5, np.float32)
    ln = np.zeros(8 * 4 * 6, np.int32)
    a, b, t = np.zeros(3, np.float32), np.zeros(4, np.float32), np.zeros(4, np.float32)
    orc.orc_fd_ops(po.ptr(ix), po.ptr(cf), po.ptr(ln), po.ptr(a), po.ptr(b), po.ptr(t))
    return ix.reshape(8, 4, 6, 5), cf.reshape(8, 4, 6, 5), ln.reshape(8, 4, 6), a, b, t


def test_operator_moments():
    """Interior DRP ops: first moment 0.99998 and coefficient sum -/+6e-5 (the
    truncated table of fd_t.c:109-112 / manual eq. 10.13); the 2- and 3-point
    near-surface ops are exact first derivatives (moment 1, sum 0)."""
    ix, cf, ln, _, _, _ = _ops()
    for p in range(8):
        for s in range(4):
            for slot in range(6):
                n = ln[p, s, slot]
                idx = ix[p, s, slot, :n].astype(np.float64)
                c = cf[p, s, slot, :n].astype(np.float64)
                m1 = float((c * idx).sum())
                if n == 5:
                    assert abs(m1 - 0.99998) < 1e-6
                    assert abs(abs(c.sum()) - 6e-5) < 1e-7
                    # forward op {-1..3}, backward {-3..1}: mirror images with opposite sign
                    assert list(idx) in ([-1, 0, 1, 2, 3], [-3, -2, -1, 0, 1])
                else:
                    assert abs(m1 - 1.0) < 1e-6 and abs(c.sum()) < 1e-6


def test_pair_schedule_alternates():
    """Each pair alternates forward/backward per stage on every axis
    (dir = (flag + stage) % 2, fd_t.c:234-236); the 8 pairs cover all 8
    direction triples at stage 0."""
    ix, _, _, _, _, _ = _ops()
    triples = set()
    for p in range(8):
        d = [tuple(ix[p, s, slot, 0] == -1 for slot in (0, 1, 5)) for s in range(4)]
        for s in range(3):
            assert all(a != b for a, b in zip(d[s], d[s + 1]))
        triples.add(d[0])
    assert len(triples) == 8


def test_rk4_tables():
    _, _, _, a, b, t = _ops()
    assert np.allclose(a, [0.5, 0.5, 1.0]) and np.allclose(t, [0.0, 0.5, 0.5, 1.0])
    assert np.allclose(b, [1 / 6, 1 / 3, 1 / 3, 1 / 6]) and abs(b.astype(np.float64).sum() - 1) < 1e-7


def test_rk4_step_is_the_classical_combination():
    """One oracle step (FP64) equals w + dt/6 (k1 + 2 k2 + 2 k3 + k4) with
    k1 = L0 w, k2 = L1 (w + dt/2 k1), k3 = L2 (w + dt/2 k2), k4 = L3 (w + dt k3),
    L_s the stage-s RHS of the pair: the textbook RK4 weights, not the tables."""
    case = cs.make_case(12, 11, 10, seed=3, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3)
    o = po.Oracle(case, fp64=True)
    w = cs.random_wavefield(case, seed=5).reshape(-1).astype(np.float64)
    aux = cs.random_aux(case, seed=5).astype(np.float64)
    dt, it = float(np.float32(0.004)), 5   # the run takes dt as float
    ipair = it % 8
    k1, a1 = o.onestage(it, ipair, 0, w, aux)
    k2, a2 = o.onestage(it, ipair, 1, w + dt / 2 * k1, aux + dt / 2 * a1)
    k3, a3 = o.onestage(it, ipair, 2, w + dt / 2 * k2, aux + dt / 2 * a2)
    k4, a4 = o.onestage(it, ipair, 3, w + dt * k3, aux + dt * a3)
    w_ref = w + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    aux_ref = aux + dt / 6 * (a1 + 2 * a2 + 2 * a3 + a4)
    w1, aux1, _ = o.run(dt, 1, w=w, aux=aux, it0=it)
    assert np.abs(w1 - w_ref).max() <= 1e-12 * np.abs(w_ref).max()
    assert np.abs(aux1 - aux_ref).max() <= 1e-12 * np.abs(aux_ref).max()


def test_metric_cartesian_closed_form():
    """On a Cartesian grid the 7-point centred op gives x_xi = 1.00004 dx (its
    first moment), so jac = 1.00004^3 dx dy dz and the metric is diagonal."""
    ni, nj, nk = 10, 9, 8
    dx, dy, dz = 100.0, 80.0, 50.0
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    K, J, I = np.meshgrid(np.arange(nz) - 3.0, np.arange(ny) - 3.0, np.arange(nx) - 3.0,
                          indexing="ij")
    xyz = np.stack([I * dx, J * dy, K * dz]).astype(np.float32)
    m = np.zeros((10, nz, ny, nx), np.float32)
    po.load_oracle().orc_metric(ni, nj, nk, po.ptr(xyz), po.ptr(m))
    c = 1.00004
    ph = (slice(3, 3 + nk), slice(3, 3 + nj), slice(3, 3 + ni))
    assert np.allclose(m[0][ph], c ** 3 * dx * dy * dz, rtol=5e-6)
    for q, val in ((1, 1 / (c * dx)), (5, 1 / (c * dy)), (9, 1 / (c * dz))):
        assert np.allclose(m[q][ph], val, rtol=5e-6)
    for q in (2, 3, 4, 6, 7, 8):
        assert np.abs(m[q][ph]).max() < 1e-6 / dx   # rounding of sums that cancel
    # ghost layers mirror the first/last physical layers (gd_t.c:291-401)
    assert np.array_equal(m[:, 2], m[:, 3]) and np.array_equal(m[:, 0], m[:, 5])
    assert np.array_equal(m[:, :, :, nx - 1], m[:, :, :, nx - 6])


def test_metric_curvilinear_inverse():
    """xi/eta/zeta rows are the inverse of the coordinate Jacobian: for the
    metric M and the derivatives G computed with the same centred op,
    M G = I (to rounding), and jac = det G."""
    case = cs.make_case(12, 11, 10, seed=7, hill=0.4)
    ni, nj, nk = 12, 11, 10
    xyz = case["coords"].astype(np.float64)
    c = np.array([-0.02084, 0.1667, -0.7709, 0.0, 0.7709, -0.1667, 0.02084])
    m = case["metric"].astype(np.float64)
    for (k, j, i) in [(5, 6, 7), (3 + nk - 1, 4, 9), (3, 3, 3)]:
        G = np.zeros((3, 3))
        for a, (dk, dj, di) in enumerate(((0, 0, 1), (0, 1, 0), (1, 0, 0))):
            for q in range(3):
                G[q, a] = sum(c[n] * xyz[q, k + (n - 3) * dk, j + (n - 3) * dj, i + (n - 3) * di]
                              for n in range(7))
        Mi = np.array([[m[1 + 3 * r + q, k, j, i] for q in range(3)] for r in range(3)])
        assert np.allclose(Mi @ G, np.eye(3), atol=2e-5)
        assert abs(m[0, k, j, i] - np.linalg.det(G)) <= 2e-5 * abs(np.linalg.det(G))


def test_free_surface_matrices_flat_closed_form():
    """Flat surface, diagonal metric: traction-free rows give dVx/dz = -dVz/dx,
    dVy/dz = -dVz/dy and dVz/dz = -lam/(lam+2mu) (dVx/dx + dVy/dy), in index
    space scaled by the metric (sv_curv_col_el_iso.c:1327-1370)."""
    ni, nj, nk = 9, 8, 8
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    vol = nx * ny * nz
    ex, ey, ez = 1 / 100.0, 1 / 90.0, 1 / 50.0
    m = np.zeros((10, vol), np.float32)
    m[0] = 1 / (ex * ey * ez); m[1] = ex; m[5] = ey; m[9] = ez
    lam, mu = 3.2e10, 2.9e10
    med = np.stack([np.full(vol, lam), np.full(vol, mu), np.full(vol, 1 / 2600.0)]).astype(np.float32)
    mats = np.zeros(3 * nx * ny * 9, np.float32)
    po.load_oracle().orc_dvh2dvz_iso(ni, nj, nk, po.ptr(m), po.ptr(med), po.ptr(mats))
    mats = mats.reshape(3, ny, nx, 3, 3)[:, 4, 5]
    mx = np.zeros((3, 3)); mx[0, 2] = -ex / ez; mx[2, 0] = -lam / (lam + 2 * mu) * ex / ez
    my = np.zeros((3, 3)); my[1, 2] = -ey / ez; my[2, 1] = -lam / (lam + 2 * mu) * ey / ez
    mf = np.diag([1 / (mu * ez * ez), 1 / (mu * ez * ez), 1 / ((lam + 2 * mu) * ez * ez)])
    assert np.allclose(mats[0], mx, rtol=1e-6, atol=1e-9)
    assert np.allclose(mats[1], my, rtol=1e-6, atol=1e-9)
    assert np.allclose(mats[2], mf, rtol=1e-6, atol=1e-20)


def test_ricker_closed_form():
    orc = po.load_oracle()
    fc, t0 = 2.0, 0.5
    assert orc.orc_ricker(t0, fc, t0) == 1.0
    tz = t0 + np.sqrt(2) / (2 * np.pi * fc)       # zero crossing of (1 - u^2/2)
    assert abs(orc.orc_ricker(tz, fc, t0)) < 1e-6
    for t in np.linspace(0, 1.2, 17):
        u = 2 * np.pi * fc * (np.float32(t) - t0)
        assert abs(orc.orc_ricker(t, fc, t0) - (1 - u * u / 2) * np.exp(-u * u / 4)) < 2e-6


def test_gaussian_weights_normalised():
    """Weights of the 7^3 smoothing sum to 1 over the rows below the surface
    (src_t.c:2127-2150); symmetric for a centred source; a shifted source
    moves the centroid by the shift."""
    orc = po.load_oracle()
    w = np.zeros(343, np.float32)
    orc.orc_gauss_weights(po.ptr(w), 0.0, 0.0, 0.0, 1.5, 3, 3)
    W = w.reshape(7, 7, 7).astype(np.float64)
    assert abs(W.sum() - 1) < 1e-6
    assert np.allclose(W, W[::-1, ::-1, ::-1], rtol=1e-6)
    orc.orc_gauss_weights(po.ptr(w), 0.0, 0.0, 0.0, 1.5, 3, 1)   # 1 row to the surface
    W = w.reshape(7, 7, 7).astype(np.float64)
    assert abs(W[:5].sum() - 1) < 1e-6
    orc.orc_gauss_weights(po.ptr(w), 0.3, -0.2, 0.1, 1.5, 3, 3)
    W = w.reshape(7, 7, 7).astype(np.float64)
    r = np.arange(-3, 4)
    assert abs((W.sum(axis=(0, 1)) * r).sum() - 0.3) < 0.02
    assert abs((W.sum(axis=(0, 2)) * r).sum() + 0.2) < 0.02


def _gauss_width_error(W, x0, y0, z0, a, half):
    """Largest deviation of a (2h+1)^3 weight block from the separable form
    exp(-((i-x0)^2 + (j-y0)^2 + (k-z0)^2) / a^2) that fun_gauss
    (src_t.c:2179-2184) fixes: along each axis the log-ratio of adjacent
    weights equals -(2 i + 1 - 2 x0) / a^2, whatever the normalisation."""
    W = np.asarray(W, np.float64).reshape(2 * half + 1, 2 * half + 1, 2 * half + 1)
    r = np.arange(-half, half)
    err = 0.0
    for axis, c0 in ((2, x0), (1, y0), (0, z0)):
        lo = np.take(W, range(0, 2 * half), axis=axis)
        hi = np.take(W, range(1, 2 * half + 1), axis=axis)
        got = np.log(hi / lo)
        shape = [1, 1, 1]
        shape[axis] = 2 * half
        want = (-(2 * r + 1 - 2 * c0) / (a * a)).reshape(shape)
        err = max(err, float(np.abs(got - want).max()))
    return err


def _numpy_gauss_block(x0, y0, z0, a, half, half_below, form):
    """float64 weight blocks for the width check's self-test: the right form
    and the plausible mistakes it must reject."""
    r = np.arange(-half, half + 1, dtype=np.float64)
    f = {"right": lambda x: np.exp(-x * x / (a * a)),
         "no_square": lambda x: np.exp(-x * x / a),
         "half_width": lambda x: np.exp(-x * x / (a * a / 2)),
         "double_width": lambda x: np.exp(-x * x / (2 * a) ** 2)}[form]
    W = f(r - z0)[:, None, None] * f(r - y0)[None, :, None] * f(r - x0)[None, None, :]
    return W / W[: half + 1 + half_below].sum()


@pytest.mark.parametrize("shift", [(0.0, 0.0, 0.0), (0.3, -0.2, 0.1), (-0.45, 0.4, -0.35)])
@pytest.mark.parametrize("half_below", [3, 1, 0])
def test_gaussian_weights_width(shift, half_below):
    """src_cal_norm_delt3d_z2fre with ext_func_coef 1.5 (src_t.c:41-46,
    2111-2151): the adjacent-weight log-ratio pins the width a^2 and the
    centre along all three axes, the surface truncation leaves the weights
    above it in place (same normalisation) and the physical rows sum to 1."""
    orc = po.load_oracle()
    a, half = 1.5, 3
    w = np.zeros(343, np.float32)
    orc.orc_gauss_weights(po.ptr(w), *shift, a, half, half_below)
    assert _gauss_width_error(w, *shift, a, half) < 2e-5
    W = w.reshape(7, 7, 7).astype(np.float64)
    assert abs(W[: half + 1 + half_below].sum() - 1) < 1e-6
    ref = _numpy_gauss_block(*shift, a, half, half_below, "right")
    assert np.allclose(W, ref, rtol=2e-6, atol=1e-12)


@pytest.mark.parametrize("form", ["no_square", "half_width", "double_width"])
def test_gaussian_width_check_rejects_mistakes(form):
    """The width check fails on the plausible mistakes (exp(-x^2/a), a^2/2,
    2a) that the sum, symmetry and centroid checks all let through."""
    shift = (0.3, -0.2, 0.1)
    bad = _numpy_gauss_block(*shift, 1.5, 3, 3, form)
    assert _gauss_width_error(bad, *shift, 1.5, 3) > 0.05
    good = _numpy_gauss_block(*shift, 1.5, 3, 3, "right")
    assert _gauss_width_error(good, *shift, 1.5, 3) < 1e-9


def _fault_moment_enu(strike, dip, rake, flip_xz_yz=False, swap_xy=False):
    """Unit double couple M = n d^T + d n^T from the fault geometry of Aki &
    Richards (Box 4.4: strike clockwise from north, dip, rake; fault normal n
    and slip d in x north, y east, z down), rotated to the reference's frame
    x east, y north, z up (src_t.c:2340-2342).  Returns Mxx, Myy, Mzz, Myz,
    Mxz, Mxy in float64.  The flags build the plausible mistakes the check
    must reject."""
    f, dl, lm = np.radians([strike, dip, rake])
    n = np.array([-np.sin(dl) * np.sin(f), np.sin(dl) * np.cos(f), -np.cos(dl)])
    d = np.array([np.cos(lm) * np.cos(f) + np.cos(dl) * np.sin(lm) * np.sin(f),
                  np.cos(lm) * np.sin(f) - np.cos(dl) * np.sin(lm) * np.cos(f),
                  -np.sin(lm) * np.sin(dl)])
    M = np.outer(n, d) + np.outer(d, n)
    P = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, -1.0]])   # NED -> ENU
    if swap_xy:
        P = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, -1.0]])
    E = P @ M @ P.T
    if flip_xz_yz:
        D = np.diag([1.0, 1.0, -1.0])
        E = D @ E @ D
    return np.array([E[0, 0], E[1, 1], E[2, 2], E[1, 2], E[0, 2], E[0, 1]])


_SDR = [(0.0, 45.0, 90.0), (90.0, 45.0, 90.0), (30.0, 60.0, -90.0), (0.0, 90.0, 0.0),
        (211.0, 33.0, -47.0), (300.0, 75.0, 150.0), (45.0, 20.0, 60.0)]


@pytest.mark.parametrize("sdr", _SDR)
def test_angle2moment_matches_fault_geometry(sdr):
    """angle2moment (src_t.c:2344-2383), every component, against the moment
    tensor built from the fault normal and slip vectors and rotated into
    x east, y north, z up."""
    m = np.zeros(6, np.float32)
    po.load_oracle().orc_angle2moment(*sdr, po.ptr(m))
    want = _fault_moment_enu(*sdr)
    assert np.allclose(m.astype(np.float64), want, atol=2e-6), (m, want)


def test_angle2moment_named_mechanisms():
    """Textbook mechanisms in x east, y north, z up: a 45-degree thrust has
    its T axis vertical (Mzz = +1) and its P axis across strike (Mxx = -1 for
    strike 0, Myy = -1 for strike 90); a 60-degree normal fault on strike 0
    has Mzz < 0 and Mxx > 0; a vertical strike-slip fault striking north
    with rake 0 is a pure Mxy couple."""
    orc = po.load_oracle()
    m = np.zeros(6, np.float32)

    def mt(s, d, r):
        orc.orc_angle2moment(s, d, r, po.ptr(m))
        return m.astype(np.float64).copy()

    mxx, myy, mzz, myz, mxz, mxy = mt(0.0, 45.0, 90.0)
    assert abs(mzz - 1) < 1e-6 and abs(mxx + 1) < 1e-6 and abs(myy) < 1e-6
    assert max(abs(myz), abs(mxz), abs(mxy)) < 1e-6
    mxx, myy, mzz, myz, mxz, mxy = mt(90.0, 45.0, 90.0)
    assert abs(mzz - 1) < 1e-6 and abs(myy + 1) < 1e-6 and abs(mxx) < 1e-6
    mxx, myy, mzz, myz, mxz, mxy = mt(0.0, 60.0, -90.0)
    s120 = np.sin(np.radians(120.0))
    assert abs(mzz + s120) < 1e-6 and abs(mxx - s120) < 1e-6 and abs(myy) < 1e-6
    # the fault dips east (to the right of north), so the east-up couple is
    # cos(2 dip) = -0.5 with z up
    assert abs(mxz - np.cos(np.radians(120.0))) < 1e-6
    mxx, myy, mzz, myz, mxz, mxy = mt(0.0, 90.0, 0.0)
    assert abs(mxy - 1) < 1e-6 and max(abs(mxx), abs(myy), abs(mzz), abs(mxz), abs(myz)) < 1e-6


@pytest.mark.parametrize("mistake", ["flip_xz_yz", "swap_xy"])
def test_fault_geometry_check_rejects_mistakes(mistake):
    """A z-sign flip of Mxz/Myz, or x and y swapped, keeps the eigenvalues
    (1, 0, -1) but fails the component check on these mechanisms."""
    orc = po.load_oracle()
    m = np.zeros(6, np.float32)
    worst = 0.0
    for sdr in _SDR:
        orc.orc_angle2moment(*sdr, po.ptr(m))
        bad = _fault_moment_enu(*sdr, **{mistake: True})
        worst = max(worst, float(np.abs(m.astype(np.float64) - bad).max()))
    assert worst > 0.3


@pytest.mark.parametrize("sdr", [(30.0, 60.0, 90.0), (0.0, 90.0, 0.0), (211.0, 33.0, -47.0)])
def test_double_couple_invariants(sdr):
    """Unit double couple: symmetric, traceless, eigenvalues (1, 0, -1)."""
    m = np.zeros(6, np.float32)
    po.load_oracle().orc_angle2moment(*sdr, po.ptr(m))
    mxx, myy, mzz, myz, mxz, mxy = m.astype(np.float64)
    M = np.array([[mxx, mxy, mxz], [mxy, myy, myz], [mxz, myz, mzz]])
    ev = np.sort(np.linalg.eigvalsh(M))
    assert np.allclose(ev, [-1, 0, 1], atol=1e-6)
    # vertical strike-slip on a north-striking fault, x east / y north / z up
    if sdr == (0.0, 90.0, 0.0):
        assert abs(mxy) > 0.999 and abs(mxx) < 1e-6 and abs(myy) < 1e-6


def test_receiver_sampling_quirk():
    """io_recv_keep weights corner 5 twice and corner 4 not at all
    (io_funcs.c:1608-1609): exact on the surface (dk = 0), biased below it."""
    case = cs.make_case(10, 9, 8, seed=1, free_top=False, pml=(0,) * 6)
    nx, ny, nz = 16, 15, 14
    V = nx * ny * nz
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = (1.0 + 0.5 * I + 0.25 * J + 2.0 * K).astype(np.float32)
    w = np.zeros((9, V), np.float32)
    w[0] = f.reshape(-1)
    o = po.Oracle(case, fp64=True)
    di, dj, dk = 0.25, 0.5, 0.75
    idx, rw = cs.buried_receiver(case, 6, 5, 4, di, dj, dk)
    s_idx, s_rw = cs.buried_receiver(case, 6, 5, 4, di, dj, 0.0)
    F = f.astype(np.float64)
    c = [F[4, 5, 6], F[4, 5, 7], F[4, 6, 6], F[4, 6, 7], F[5, 5, 6], F[5, 5, 7], F[5, 6, 6],
         F[5, 6, 7]]
    Lx1, Ly1, Lz1 = 1 - di, 1 - dj, 1 - dk
    quirk = (c[0] * Lx1 * Ly1 * Lz1 + c[1] * di * Ly1 * Lz1 + c[2] * Lx1 * dj * Lz1
             + c[3] * di * dj * Lz1 + c[5] * Lx1 * Ly1 * dk + c[5] * di * Ly1 * dk
             + c[6] * Lx1 * dj * dk + c[7] * di * dj * dk)
    exact = 1.0 + 0.5 * (6 + di) + 0.25 * (5 + dj) + 2.0 * (4 + dk)
    assert abs(quirk - exact) > 0.05          # the quirk is visible below the surface
    got = _sample_once(o, w, idx, rw)
    assert abs(got - quirk) < 1e-9
    got_s = _sample_once(o, w, s_idx, s_rw)
    assert abs(got_s - (1.0 + 0.5 * (6 + di) + 0.25 * (5 + dj) + 2.0 * 4)) < 1e-9


def _sample_once(o, w, idx, rw):
    """Receiver value of component 0 after a step whose RHS is zero: a run with
    dt = 0 leaves the field unchanged, so the sample is the sampler itself."""
    _, _, s = o.run(0.0, 1, w=w.reshape(-1), recv_idx=idx, recv_w=rw)
    return float(s[0, 0, 0])
