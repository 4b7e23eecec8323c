"""Checks of the oracle's VTI, general anisotropic and GMB visco-elastic media
that do not depend on the reference build: Hooke's law against the stiffness
tensor contracted in float64, the traction-free condition the surface
matrices must satisfy, the reference's own way of writing a VTI medium as a
general anisotropic one (md_t.c:544-590), the isotropic limits, the GMB
model's closed forms (flat Q over the band, the relaxed phase velocity at the
reference frequency) and the traction-free surface rotation.  A transposed or
mistyped entry of the Cij index tables fails the first two."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests import cases as cs

VOIGT = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


def stiffness(case, p):
    """6x6 Voigt stiffness of the case's medium at flat point index p."""
    mt, med = case["medium_type"], case["med"].reshape(case["med"].shape[0], -1)
    C = np.zeros((6, 6))
    if mt == po.MEDIUM_ISO:
        lam, mu = float(med[0, p]), float(med[1, p])
        C[:3, :3] = lam
        C[np.arange(3), np.arange(3)] = lam + 2 * mu
        C[3, 3] = C[4, 4] = C[5, 5] = mu
    elif mt == po.MEDIUM_VTI:
        c11, c13, c33, c55, c66 = (float(med[q, p]) for q in range(5))
        C[0, 0] = C[1, 1] = c11
        C[0, 1] = C[1, 0] = c11 - 2 * c66
        C[0, 2] = C[2, 0] = C[1, 2] = C[2, 1] = c13
        C[2, 2] = c33
        C[3, 3] = C[4, 4] = c55
        C[5, 5] = c66
    else:
        m = 0
        for a in range(6):
            for b in range(a, 6):
                C[a, b] = C[b, a] = float(med[m, p])
                m += 1
    return C


def stress_rate(C, G):
    """sigma_dot (Voigt) = C : sym(G), G[b][j] = dV_b/dx_j."""
    e = np.array([G[0, 0], G[1, 1], G[2, 2], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0],
                  G[0, 1] + G[1, 0]])
    return C @ e


def affine_case(medium, ni=14, nj=13, nk=12, seed=0):
    """Grid x = A (i,j,k) with a sheared, rotated A: the coordinate derivatives
    are exact for every operator, and all nine metric terms are non-zero."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    A = np.array([[100.0, 14.0, -9.0], [-11.0, 95.0, 12.0], [7.0, -13.0, 60.0]])
    A += rng.uniform(-3, 3, (3, 3))
    K, J, I = np.meshgrid(np.arange(nz) - 3.0, np.arange(ny) - 3.0, np.arange(nx) - 3.0,
                          indexing="ij")
    coords = np.einsum("ab,bkji->akji", A, np.stack([I, J, K])).astype(np.float32)
    metric = np.zeros((10, nz, ny, nx), np.float32)
    po.load_oracle().orc_metric(ni, nj, nk, po.ptr(np.ascontiguousarray(coords)),
                                po.ptr(metric))
    mtype, med = cs.make_medium(medium, ni, nj, nk, seed=seed)
    return dict(ni=ni, nj=nj, nk=nk, coords=coords, metric=metric, med=med, medium_type=mtype,
                free_top=0, pml_on=[0] * 6, pml_layers=[0] * 6, pml_A=[None] * 6,
                pml_B=[None] * 6, pml_D=[None] * 6, src=None)


@pytest.mark.parametrize("medium", ["iso", "vti", "aniso"])
def test_stress_rate_is_hookes_law(medium):
    """A velocity field with a uniform gradient G (zero at the test point):
    the oracle's stress rates there equal C : sym(G) to the operators'
    truncation (their coefficients sum to 1 +- 6e-5)."""
    case = affine_case(medium, seed=3)
    ni, nj, nk = case["ni"], case["nj"], case["nk"]
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    vol = nx * ny * nz
    orc = po.Oracle(case, fp64=True)
    rng = np.random.default_rng(7)
    X = case["coords"].reshape(3, -1).astype(np.float64)
    worst = 0.0
    for (i, j, k) in [(8, 8, 8), (10, 7, 11), (6, 12, 7)]:
        p = i + j * nx + k * nx * ny
        G = rng.standard_normal((3, 3)) * 1e-3
        w = np.zeros((9, vol))
        w[:3] = G @ (X - X[:, p:p + 1])
        C = stiffness(case, p)
        want = stress_rate(C, G)
        for ipair, istage in [(0, 0), (3, 1), (6, 2)]:
            r, _ = orc.onestage(0, ipair, istage, w.reshape(-1))
            got = r.reshape(9, vol)[3:, p]
            err = np.abs(got - want).max() / np.abs(want).max()
            worst = max(worst, err)
    print(f"{medium}: worst relative deviation from C:sym(G) {worst:.2e}")
    assert worst < 1e-3


@pytest.mark.parametrize("medium", ["iso", "vti", "aniso"])
def test_free_surface_matrices_are_traction_free(medium):
    """dV/dzeta = matVx2Vz dV/dxi + matVy2Vz dV/deta makes the traction rate
    on the zeta surface vanish: sum_j sigma_dot_ij zeta_j = 0."""
    case = cs.make_case(12, 11, 9, seed=5, free_top=True, medium=medium, hill=0.4)
    nx, ny = case["ni"] + 6, case["nj"] + 6
    mats = cs.free_surface_mats(case).reshape(3, ny, nx, 3, 3)
    M = case["metric"].reshape(10, -1).astype(np.float64)
    k = 3 + case["nk"] - 1
    rng = np.random.default_rng(11)
    worst = 0.0
    for j in range(3, 3 + case["nj"], 3):
        for i in range(3, 3 + case["ni"], 2):
            p = i + j * nx + k * nx * ny
            xi, et, zt = M[1:4, p], M[4:7, p], M[7:10, p]
            Dx, Dy = rng.standard_normal(3), rng.standard_normal(3)
            Dz = mats[0, j, i].astype(np.float64) @ Dx + mats[1, j, i].astype(np.float64) @ Dy
            G = np.outer(Dx, xi) + np.outer(Dy, et) + np.outer(Dz, zt)
            s = stress_rate(stiffness(case, p), G)
            S = np.array([[s[0], s[5], s[4]], [s[5], s[1], s[3]], [s[4], s[3], s[2]]])
            t = S @ zt
            ref = np.abs(stress_rate(stiffness(case, p), np.outer(Dx, xi) + np.outer(Dy, et)))
            worst = max(worst, np.abs(t).max() / (ref.max() * np.linalg.norm(zt)))
    print(f"{medium}: worst relative traction {worst:.2e}")
    assert worst < 1e-5


def test_aniso_with_vti_constants_equals_vti():
    """The reference's test medium writes VTI as 21 Cij (c12 = c11 - 2 c66,
    c22 = c11, c23 = c13, c44 = c55, rest zero); every RHS of the general
    anisotropic path then equals the VTI path's bit for bit (the extra terms
    are exact zeros), free surface and PML included."""
    case_v = cs.make_case(13, 12, 11, seed=6, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3,
                          medium="vti")
    case_a = dict(case_v, medium_type=po.MEDIUM_ANISO, med=cs.vti_as_aniso(case_v["med"]))
    assert np.array_equal(cs.free_surface_mats(case_v)[: 2 * 19 * 18 * 9],
                          cs.free_surface_mats(case_a)[: 2 * 19 * 18 * 9])
    w = cs.random_wavefield(case_v, seed=1).reshape(-1)
    aux = cs.random_aux(case_v, seed=1)
    ov, oa = po.Oracle(case_v), po.Oracle(case_a)
    for ipair in range(8):
        for istage in range(4):
            rv, av = ov.onestage(2, ipair, istage, w, aux)
            ra, aa = oa.onestage(2, ipair, istage, w, aux)
            assert np.array_equal(rv, ra) and np.array_equal(av, aa), (ipair, istage)


@pytest.mark.parametrize("medium", ["vti", "aniso"])
def test_isotropic_limit(medium):
    """c11 = c33 = lambda + 2 mu, c13 = lambda, c55 = c66 = mu (and the matching
    21 Cij): the VTI and anisotropic paths give the isotropic path's RHS up
    to FP32 rounding of their differently associated sums."""
    case_i = cs.make_case(13, 12, 11, seed=6, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3)
    lam, mu, slw = case_i["med"].astype(np.float64)
    vti = np.stack([lam + 2 * mu, lam, lam + 2 * mu, mu, mu, slw]).astype(np.float32)
    if medium == "vti":
        case_x = dict(case_i, medium_type=po.MEDIUM_VTI, med=vti)
    else:
        case_x = dict(case_i, medium_type=po.MEDIUM_ANISO, med=cs.vti_as_aniso(vti))
    w = cs.random_wavefield(case_i, seed=2).reshape(-1)
    aux = cs.random_aux(case_i, seed=2)
    oi, ox = po.Oracle(case_i), po.Oracle(case_x)
    worst = 0.0
    for ipair, istage in [(0, 0), (1, 1), (2, 2), (5, 3), (7, 0)]:
        ri, ai = oi.onestage(2, ipair, istage, w, aux)
        rx, ax = ox.onestage(2, ipair, istage, w, aux)
        worst = max(worst, cs.rel_l2(rx, ri), cs.rel_l2(ax, ai))
    print(f"{medium}: isotropic limit rel L2 {worst:.2e}")
    assert 0 < worst < 1e-5 or worst == 0.0
    assert worst < 1e-5


@pytest.mark.parametrize("medium", ["vti", "aniso"])
def test_fp32_run_error_against_fp64(medium):
    """State the FP32 error of a free-surface + PML + source run of each
    medium against the FP64 build of the same scheme."""
    ni, nj, nk, nt, dt = 18, 16, 14, 60, 0.004
    case = cs.make_case(ni, nj, nk, seed=13, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=4,
                        medium=medium)
    case["src"] = cs.ricker_point_source(case, 12, 10, 9, nt, dt, fc=9.0, t0=0.12)
    ridx, rw = cs.surface_receivers(case, n=4)
    _, _, s32 = po.Oracle(case).run(dt, nt, recv_idx=ridx, recv_w=rw)
    _, _, s64 = po.Oracle(case, fp64=True).run(dt, nt, recv_idx=ridx, recv_w=rw)
    e = cs.rel_l2(s32, s64)
    print(f"{medium}: FP32 vs FP64 seismograms rel L2 {e:.2e}")
    assert np.abs(s64).max() > 0 and e < 1e-4


# ---- visco-elastic GMB (sv_curv_col_vis_iso.c, md_t.c:643-914) -----------

def _gmb(n=3, npts=40, seed=0, qp=100.0, qs=50.0, freq=1.0, fmin=0.5, fmax=5.0):
    rng = np.random.default_rng(seed)
    lam = rng.uniform(2e10, 4e10, npts).astype(np.float32)
    mu = rng.uniform(1e10, 3e10, npts).astype(np.float32)
    Qp = (qp * rng.uniform(0.8, 1.2, npts)).astype(np.float32)
    Qs = (qs * rng.uniform(0.8, 1.2, npts)).astype(np.float32)
    return lam, mu, Qp, Qs


def test_gmb_host_builder_matches_oracle():
    """The engine's host builder and the oracle's restatement agree bit for
    bit (unrelaxed moduli, Ylam, Ymu, wl)."""
    import ctypes as C
    from cgfd3d_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    L.cgfdh_visco_gmb.argtypes = [C.c_long, C.c_int, C.c_float, C.c_float, C.c_float] + \
        [C.c_void_p] * 6
    lam, mu, Qp, Qs = _gmb()
    n, npts = 3, len(lam)
    l1, m1 = lam.copy(), mu.copy()
    ym = np.zeros((2 * n, npts), np.float32)
    wl1 = np.zeros(n, np.float32)
    assert L.cgfdh_visco_gmb(npts, n, 1.0, 0.5, 5.0, po.ptr(l1), po.ptr(m1), po.ptr(Qp),
                             po.ptr(Qs), po.ptr(ym), po.ptr(wl1)) == 0
    l2, m2 = lam.copy(), mu.copy()
    Yl = np.zeros((npts, n), np.float32); Ym = np.zeros((npts, n), np.float32)
    wl2 = np.zeros(n, np.float32)
    po.load_oracle().orc_visco_gmb(npts, n, 1.0, 0.5, 5.0, po.ptr(l2), po.ptr(m2), po.ptr(Qp),
                                   po.ptr(Qs), po.ptr(Yl), po.ptr(Ym), po.ptr(wl2))
    assert np.array_equal(l1, l2) and np.array_equal(m1, m2) and np.array_equal(wl1, wl2)
    assert np.array_equal(ym[0::2].T, Yl) and np.array_equal(ym[1::2].T, Ym)


def _modulus(Y, wl, w):
    """GMB modulus over its unrelaxed value: 1 - sum Y_n wl_n / (wl_n + i w)."""
    return 1 - np.sum(Y[None, :] * wl[None, :] / (wl[None, :] + 1j * w[:, None]), axis=1)


@pytest.mark.parametrize("n", [2, 3, 4])
def test_gmb_fit_gives_flat_q_and_keeps_the_reference_velocity(n):
    """Closed forms of the generalized Maxwell body: Q^-1 = Im M / Re M stays
    near the target over [fmin, fmax], and the unrelaxed moduli give the
    relaxed (input) phase velocity at the reference frequency, for P
    (Y_P = (lam Ylam + 2 mu Ymu)/(lam + 2 mu)) and S (Y_S = Ymu)."""
    lam, mu, Qp, Qs = _gmb(n=n, npts=12, seed=n)
    rho = 2500.0
    l2, m2 = lam.copy(), mu.copy()
    Yl = np.zeros((len(lam), n), np.float32); Ym = np.zeros((len(lam), n), np.float32)
    wl = np.zeros(n, np.float32)
    po.load_oracle().orc_visco_gmb(len(lam), n, 1.0, 0.5, 5.0, po.ptr(l2), po.ptr(m2),
                                   po.ptr(Qp), po.ptr(Qs), po.ptr(Yl), po.ptr(Ym), po.ptr(wl))
    wl = wl.astype(np.float64)
    band = 2 * np.pi * np.geomspace(0.5, 5.0, 25)
    wr = np.array([2 * np.pi * 1.0])
    worst_q, worst_v = 0.0, 0.0
    for p in range(len(lam)):
        lu, mu_u = float(l2[p]), float(m2[p])
        YP = (lu * Yl[p].astype(np.float64) + 2 * mu_u * Ym[p]) / (lu + 2 * mu_u)
        YS = Ym[p].astype(np.float64)
        for Y, Q, Mu, Mr in ((YP, Qp[p], lu + 2 * mu_u, float(lam[p]) + 2 * float(mu[p])),
                             (YS, Qs[p], mu_u, float(mu[p]))):
            M = _modulus(Y, wl, band)
            qinv = M.imag / M.real
            worst_q = max(worst_q, float(np.max(np.abs(qinv * round(float(Q)) - 1))))
            k = 1 / np.sqrt(Mu * _modulus(Y, wl, wr) / rho)
            v = 1 / k.real[0]
            worst_v = max(worst_v, abs(v / np.sqrt(Mr / rho) - 1))
    print(f"n={n}: worst |Q/Q0 - 1| over the band {worst_q:.3f}, velocity at fr {worst_v:.1e}")
    assert worst_q < (0.25 if n == 2 else 0.1)
    assert worst_v < 1e-5


def test_visco_surface_rotation_leaves_no_traction():
    """sv_curv_col_vis_iso_free rotates the surface-row stresses into the
    frame (tangent x normal, tangent, normal) of matD, zeroes the rows of the
    normal and rotates back: after every step the traction T.n on the
    surface vanishes, and matD is orthonormal to FP32 precision."""
    ni, nj, nk, nt, dt = 14, 13, 12, 6, 0.004
    case = cs.make_case(ni, nj, nk, seed=9, free_top=True, pml=(0,) * 6, medium="visco",
                        hill=0.4)
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    mats = cs.free_surface_mats(case).reshape(4, ny, nx, 3, 3)
    D = mats[3, 3:3 + nj, 3:3 + ni].astype(np.float64)
    eye = np.einsum("jiab,jicb->jiac", D, D)
    assert np.abs(eye - np.eye(3)).max() < 1e-5
    w = cs.random_wavefield(case, seed=3).reshape(-1)
    orc = po.Oracle(case, fp64=False)
    w1, _, _ = orc.run(dt, nt, w=w)
    W = w1.reshape(cs.ncomp(case), nz, ny, nx)[:, 3 + nk - 1, 3:3 + nj, 3:3 + ni].astype(np.float64)
    T = np.stack([np.stack([W[3], W[8], W[7]]), np.stack([W[8], W[4], W[6]]),
                  np.stack([W[7], W[6], W[5]])])           # [a][b][j][i]
    n = D[:, :, 2, :]                                        # unit normal rows
    t = np.einsum("abji,jib->jia", T, n)
    rel = np.abs(t).max() / np.abs(T).max()
    print(f"surface traction / stress after {nt} steps: {rel:.1e}")
    assert rel < 1e-5


def test_visco_high_q_limit():
    """Q -> 1e8: the memory variables' couplings scale with 1/Q, so the
    visco-elastic RHS of the elastic fields (and the PML auxiliaries) matches
    the isotropic one to ~1e-6 (its unrelaxed moduli equal the relaxed ones to
    FP32 rounding); the attenuation terms themselves are pinned bit for bit
    by the reference build."""
    case_i = cs.make_case(13, 12, 11, seed=6, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3)
    lam, mu, slw = (v.copy() for v in case_i["med"])
    shp = lam.shape
    n, npts = 3, lam.size
    l2, m2 = np.ascontiguousarray(lam.reshape(-1)), np.ascontiguousarray(mu.reshape(-1))
    Q = np.full(npts, 1e8, np.float32)
    Yl = np.zeros((npts, n), np.float32); Ym = np.zeros((npts, n), np.float32)
    wl = np.zeros(n, np.float32)
    po.load_oracle().orc_visco_gmb(npts, n, 1.0, 0.5, 5.0, po.ptr(l2), po.ptr(m2), po.ptr(Q),
                                   po.ptr(Q), po.ptr(Yl), po.ptr(Ym), po.ptr(wl))
    vols = [l2.reshape(shp), m2.reshape(shp), slw]
    for b in range(n):
        vols += [Yl[:, b].reshape(shp), Ym[:, b].reshape(shp)]
    case_v = dict(case_i, medium_type=po.MEDIUM_VISCO, med=np.stack(vols).astype(np.float32),
                  visco_wl=wl)
    wv = cs.random_wavefield(case_v, seed=2).reshape(-1)
    wi = wv[: 9 * lam.size]
    aux = cs.random_aux(case_i, seed=2)
    oi, ov = po.Oracle(case_i), po.Oracle(case_v)
    worst = 0.0
    for ipair, istage in [(0, 0), (3, 1), (6, 2)]:
        ri, ai = oi.onestage(2, ipair, istage, wi, aux)
        rv, av = ov.onestage(2, ipair, istage, wv, aux)
        worst = max(worst, cs.rel_l2(rv[: ri.size], ri), cs.rel_l2(av, ai))
    print(f"high-Q visco vs isotropic RHS rel L2 {worst:.1e}")
    assert worst < 1e-5
