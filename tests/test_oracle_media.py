"""Checks of the oracle's VTI and general anisotropic media that do not depend
on the reference build: Hooke's law against the stiffness tensor contracted
in float64, the traction-free condition the surface matrices must satisfy,
the reference's own way of writing a VTI medium as a general anisotropic one
(md_t.c:544-590), and the isotropic limit.  A transposed or mistyped entry of
the Cij index tables fails the first two."""
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
