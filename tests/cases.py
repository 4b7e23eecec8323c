"""Seeded synthetic inputs shared by the parity tests: small curvilinear grids
with topography, heterogeneous isotropic, VTI and general anisotropic media,
PML tables and sources.

Layout follows the reference (x fastest, 3 ghost points per side).  The metric
and the PML profile tables are produced by the oracle's restatement of the
reference setup code; the product's own setup code is checked against the
same restatement in test_setup.py.
"""
import numpy as np

from oracle import pyoracle


def dims(ni, nj, nk):
    return ni + 6, nj + 6, nk + 6


def curvilinear_coords(ni, nj, nk, dh=100.0, seed=0, hill=0.3, wiggle=0.08):
    """Body-fitted grid: flat bottom, Gaussian-hill top, mildly skewed columns.
    Ghost coordinates extend the same analytic map, as a generated grid does."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims(ni, nj, nk)
    i = np.arange(nx) - 3.0
    j = np.arange(ny) - 3.0
    k = np.arange(nz) - 3.0
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 4)
    x = I * dh + wiggle * dh * np.sin(0.7 * J + ph[0]) * np.sin(0.5 * K + ph[1])
    y = J * dh + wiggle * dh * np.sin(0.6 * I + ph[2]) * np.cos(0.4 * K + ph[3])
    depth = (nk - 1) * dh
    xc, yc = 0.5 * (ni - 1) * dh, 0.5 * (nj - 1) * dh
    sig = 0.25 * min(ni, nj) * dh
    top = hill * depth * np.exp(-((I * dh - xc) ** 2 + (J * dh - yc) ** 2) / (2 * sig ** 2))
    frac = K / (nk - 1)
    z = -depth + frac * (depth + top)
    return np.stack([x, y, z]).astype(np.float32)


def iso_medium(ni, nj, nk, seed=0, vp=6000.0, vs=3464.0, rho=2670.0, pert=0.05):
    rng = np.random.default_rng(seed + 1)
    nx, ny, nz = dims(ni, nj, nk)
    f = 1.0 + pert * rng.standard_normal((3, nz, ny, nx))
    vp_, vs_, rho_ = vp * f[0], vs * f[1], rho * f[2]
    mu = (rho_ * vs_ ** 2).astype(np.float32)
    lam = (rho_ * vp_ ** 2 - 2 * rho_ * vs_ ** 2).astype(np.float32)
    slw = (1.0 / rho_).astype(np.float32)
    return np.stack([lam, mu, slw]).astype(np.float32)


def vti_medium(ni, nj, nk, seed=0, vp=6000.0, vs=3464.0, rho=2670.0, eps=0.12,
               delta=0.06, gamma=0.09, pert=0.05):
    """c11,c13,c33,c55,c66,slw from Thomsen parameters with seeded per-point
    perturbations (vertical axis of symmetry)."""
    rng = np.random.default_rng(seed + 2)
    nx, ny, nz = dims(ni, nj, nk)
    f = 1.0 + pert * rng.standard_normal((6, nz, ny, nx))
    vp_, vs_, rho_ = vp * f[0], vs * f[1], rho * f[2]
    e, d, g = eps * f[3], delta * f[4], gamma * f[5]
    c33 = rho_ * vp_ ** 2
    c55 = rho_ * vs_ ** 2
    c11 = c33 * (1 + 2 * e)
    c66 = c55 * (1 + 2 * g)
    c13 = np.sqrt((c33 - c55) * (2 * d * c33 + c33 - c55)) - c55
    return np.stack([c11, c13, c33, c55, c66, 1.0 / rho_]).astype(np.float32)


# Voigt index pairs and the 21 upper-triangle entries c11,c12,..,c66
_VOIGT = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
_UPPER = [(a, b) for a in range(6) for b in range(a, 6)]


def aniso_medium(ni, nj, nk, seed=0, pert=0.03):
    """The 21 Cij (c11,c12,..,c66) and slw of a triclinic medium: a VTI
    stiffness rotated by a seeded per-point rotation, then perturbed
    symmetrically (small enough to stay positive definite)."""
    rng = np.random.default_rng(seed + 3)
    nx, ny, nz = dims(ni, nj, nk)
    n = nx * ny * nz
    v = vti_medium(ni, nj, nk, seed=seed).reshape(6, n).astype(np.float64)
    c11, c13, c33, c55, c66, slw = v
    C = np.zeros((n, 6, 6))
    C[:, 0, 0] = C[:, 1, 1] = c11
    C[:, 0, 1] = C[:, 1, 0] = c11 - 2 * c66
    C[:, 0, 2] = C[:, 2, 0] = C[:, 1, 2] = C[:, 2, 1] = c13
    C[:, 2, 2] = c33
    C[:, 3, 3] = C[:, 4, 4] = c55
    C[:, 5, 5] = c66
    # 4th-order tensor, rotate, back to Voigt
    T = np.zeros((n, 3, 3, 3, 3))
    for a, (i, j) in enumerate(_VOIGT):
        for b, (k, l) in enumerate(_VOIGT):
            for (p, q) in {(i, j), (j, i)}:
                for (r, s) in {(k, l), (l, k)}:
                    T[:, p, q, r, s] = C[:, a, b]
    ang = rng.uniform(-0.6, 0.6, (3, n))
    ca, sa = np.cos(ang), np.sin(ang)
    Rz = np.zeros((n, 3, 3)); Rz[:, 0, 0] = ca[0]; Rz[:, 0, 1] = -sa[0]; Rz[:, 1, 0] = sa[0]
    Rz[:, 1, 1] = ca[0]; Rz[:, 2, 2] = 1
    Ry = np.zeros((n, 3, 3)); Ry[:, 0, 0] = ca[1]; Ry[:, 0, 2] = sa[1]; Ry[:, 2, 0] = -sa[1]
    Ry[:, 2, 2] = ca[1]; Ry[:, 1, 1] = 1
    Rx = np.zeros((n, 3, 3)); Rx[:, 1, 1] = ca[2]; Rx[:, 1, 2] = -sa[2]; Rx[:, 2, 1] = sa[2]
    Rx[:, 2, 2] = ca[2]; Rx[:, 0, 0] = 1
    R = Rz @ Ry @ Rx
    T = np.einsum("nia,njb,nkc,nld,nabcd->nijkl", R, R, R, R, T, optimize=True)
    out = np.zeros((22, n))
    scale = np.abs(C).max(axis=(1, 2))
    noise = pert * rng.standard_normal((21, n)) * scale
    for m, (a, b) in enumerate(_UPPER):
        (i, j), (k, l) = _VOIGT[a], _VOIGT[b]
        out[m] = T[:, i, j, k, l] + (noise[m] if a != b else abs(noise[m]))
    out[21] = slw
    return out.reshape(22, nz, ny, nx).astype(np.float32)


def vti_as_aniso(vti):
    """The 21 Cij of a VTI medium written the way the reference's own test
    medium does (md_t.c:544-590): c12 = c11 - 2 c66, c22 = c11, c23 = c13,
    c44 = c55, the rest zero."""
    c11, c13, c33, c55, c66, slw = vti.astype(np.float32)
    z = np.zeros_like(c11)
    c12 = (c11.astype(np.float64) - 2.0 * c66.astype(np.float64)).astype(np.float32)
    comp = dict(c11=c11, c12=c12, c13=c13, c22=c11, c23=c13, c33=c33, c44=c55, c55=c55,
                c66=c66)
    names = ["c%d%d" % (a + 1, b + 1) for a, b in _UPPER]
    return np.stack([comp.get(nm, z) for nm in names] + [slw]).astype(np.float32)


def visco_medium(ni, nj, nk, seed=0, nmaxwell=3, qp=100.0, qs=50.0, freq=1.0, fmin=0.5,
                 fmax=5.0, pert=0.1):
    """GMB visco-elastic medium: the isotropic medium with seeded Qp, Qs around
    qp, qs, converted by the oracle's restatement of md_vis_GMB_cal_Y.
    Returns (volumes lambda, mu, slw, Ylam_0, Ymu_0, ..., wl)."""
    rng = np.random.default_rng(seed + 4)
    lam, mu, slw = iso_medium(ni, nj, nk, seed=seed)
    shp = lam.shape
    Qp = (qp * (1 + pert * rng.standard_normal(shp))).astype(np.float32)
    Qs = (qs * (1 + pert * rng.standard_normal(shp))).astype(np.float32)
    n = lam.size
    lam = np.ascontiguousarray(lam.reshape(-1)); mu = np.ascontiguousarray(mu.reshape(-1))
    Yl = np.zeros((n, nmaxwell), np.float32); Ym = np.zeros((n, nmaxwell), np.float32)
    wl = np.zeros(nmaxwell, np.float32)
    pyoracle.load_oracle().orc_visco_gmb(n, nmaxwell, freq, fmin, fmax, pyoracle.ptr(lam),
                                         pyoracle.ptr(mu), pyoracle.ptr(Qp.reshape(-1)),
                                         pyoracle.ptr(Qs.reshape(-1)), pyoracle.ptr(Yl),
                                         pyoracle.ptr(Ym), pyoracle.ptr(wl))
    vols = [lam.reshape(shp), mu.reshape(shp), slw]
    for b in range(nmaxwell):
        vols += [Yl[:, b].reshape(shp), Ym[:, b].reshape(shp)]
    return np.stack(vols).astype(np.float32), wl


def make_medium(kind, ni, nj, nk, seed=0):
    """(medium_type, volumes) for kind 'iso', 'vti' or 'aniso'."""
    if kind == "vti":
        return pyoracle.MEDIUM_VTI, vti_medium(ni, nj, nk, seed=seed)
    if kind == "aniso":
        return pyoracle.MEDIUM_ANISO, aniso_medium(ni, nj, nk, seed=seed)
    return pyoracle.MEDIUM_ISO, iso_medium(ni, nj, nk, seed=seed)


def ncomp(case):
    """Wavefield components: 9, plus 6 memory variables per Maxwell body."""
    return 9 + 6 * len(case.get("visco_wl", ()))


def pml_tables(orc, ni, nj, nk, coords, pml_on, pml_layers, vel=6000.0,
               alpha=None, beta=3.0):
    """A/B/D per face with the reference's profile (bdry_t.c:215-272)."""
    region = np.zeros((6, 6), np.int32)
    on = np.array(pml_on, np.int32)
    lay = np.array(pml_layers, np.int32)
    orc.orc_pml_region(ni, nj, nk, pyoracle.ptr(lay), pyoracle.ptr(on), pyoracle.ptr(region))
    nx, ny, nz = dims(ni, nj, nk)
    vol = nx * ny * nz
    c = np.ascontiguousarray(coords.reshape(3, vol))
    A, B, D = [], [], []
    for f in range(6):
        if not pml_on[f]:
            A.append(None); B.append(None); D.append(None)
            continue
        L0 = np.zeros(1, np.float32)
        r = np.ascontiguousarray(region[f])
        import ctypes
        dh = orc.orc_pml_avg_dh(ni, nj, nk, pyoracle.ptr(c[0]), pyoracle.ptr(c[1]),
                                pyoracle.ptr(c[2]), pyoracle.ptr(r), f // 2,
                                pyoracle.ptr(L0))
        n = pml_layers[f] + 1
        a = np.zeros(n, np.float32); b = np.zeros(n, np.float32); d = np.zeros(n, np.float32)
        am = alpha if alpha is not None else np.pi * 1.0  # ~ pi * fc
        orc.orc_pml_profile(pml_layers[f], float(L0[0]), float(dh), vel, am, beta, f % 2,
                            pyoracle.ptr(a), pyoracle.ptr(b), pyoracle.ptr(d))
        A.append(a); B.append(b); D.append(d)
    return A, B, D, region


def make_case(ni=16, nj=14, nk=12, seed=0, free_top=True, pml=(1, 1, 1, 1, 1, 0),
              layers=4, src=None, dh=100.0, hill=0.3, medium="iso"):
    orc = pyoracle.load_oracle()
    nx, ny, nz = dims(ni, nj, nk)
    coords = curvilinear_coords(ni, nj, nk, dh=dh, seed=seed, hill=hill)
    metric = np.zeros((10, nz, ny, nx), np.float32)
    orc.orc_metric(ni, nj, nk, pyoracle.ptr(np.ascontiguousarray(coords)), pyoracle.ptr(metric))
    visco_wl = None
    if medium == "visco":
        mtype = pyoracle.MEDIUM_VISCO
        med, visco_wl = visco_medium(ni, nj, nk, seed=seed)
    else:
        mtype, med = make_medium(medium, ni, nj, nk, seed=seed)
    pml_on = list(pml)
    pml_layers = [layers if p else 0 for p in pml_on]
    A, B, D, region = pml_tables(orc, ni, nj, nk, coords, pml_on, pml_layers)
    case = dict(ni=ni, nj=nj, nk=nk, coords=coords, metric=metric, med=med, medium_type=mtype,
                free_top=int(free_top), pml_on=pml_on, pml_layers=pml_layers,
                pml_A=A, pml_B=B, pml_D=D, pml_region=region, src=src,
                pml_vel=6000.0, pml_alpha=np.pi * 1.0, pml_beta=3.0)
    if visco_wl is not None:
        case["visco_wl"] = visco_wl
    return case


def random_wavefield(case, seed=0, scale=1.0):
    """Random values on physical points, zero ghosts (the reference keeps its
    ghost layers at zero: wav_t.c:276-378 and the calloc'ed levels)."""
    rng = np.random.default_rng(seed + 7)
    ni, nj, nk = case["ni"], case["nj"], case["nk"]
    nx, ny, nz = dims(ni, nj, nk)
    nc = ncomp(case)
    w = np.zeros((nc, nz, ny, nx), np.float32)
    w[:, 3:3 + nk, 3:3 + nj, 3:3 + ni] = scale * rng.standard_normal((nc, nk, nj, ni))
    return w


def aux_shapes(case):
    """(nz_f, ny_f, nx_f) of each enabled PML face's aux block, face order."""
    out = []
    for f in range(6):
        if not case["pml_on"][f]:
            out.append(None)
            continue
        r = case["pml_region"][f]
        out.append((r[5] - r[4] + 1, r[3] - r[2] + 1, r[1] - r[0] + 1))
    return out


def random_aux(case, seed=0, scale=1.0):
    rng = np.random.default_rng(seed + 11)
    n = sum(9 * int(np.prod(s)) for s in aux_shapes(case) if s is not None)
    return (scale * rng.standard_normal(n)).astype(np.float32)


def ricker_point_source(case, i, j, k, nt, dt, fc=2.0, t0=0.5, m0=1e16,
                        mech=(1, 1, 1, 0, 0, 0), gauss=False, inc=(0.0, 0.0, 0.0)):
    """Moment source with a Ricker STF table laid out as src_t.c:1047-1091:
    M??[is][it][stage] at t = it*dt + rk_t[stage]*dt (wavelet_tstart = t0 of run = 0)."""
    orc = pyoracle.load_oracle()
    rk_t = [0.0, 0.5, 0.5, 1.0]
    stf = np.zeros((nt, 4), np.float32)
    for it in range(nt):
        for s in range(4):
            t = np.float32(np.float32(it * np.float32(dt)) + np.float32(np.float32(rk_t[s]) * np.float32(dt)))
            stf[it, s] = orc.orc_ricker(float(t), fc, t0)
    mij = np.zeros((6, 1, nt, 4), np.float32)
    for m in range(6):
        mij[m, 0] = (stf * np.float32(mech[m] * m0)).astype(np.float32)
    return dict(nsrc=1, ijk=[i, j, k], inc=list(inc), it_begin=[0], it_end=[nt - 1],
                max_nt=nt, mij=mij.reshape(-1), gauss=int(gauss), ext_half=3, ext_coef=1.5)


def ricker_force_source(case, ijk, nt, dt, fxyz=(1.0e12, -0.6e12, 2.0e12), fc=10.0, t0=0.05,
                        gauss=False, inc=(0.0, 0.0, 0.0), strict=0, moment=None):
    """Force sources with Ricker time functions, laid out as src_t.c:1047-1091:
    F?[is][it][stage] = ricker(t) * f at t = it*dt + rk_t[stage]*dt.  ijk: one
    (i, j, k) or a list of them.  strict: is_surface_force_strict; the sources
    whose extent reaches the surface (sk + half >= nk2, src_t.c:361-371, half
    0 for a point source, src_t.c:418) are listed in surf_idx with their rate
    tables F?_rate = ricker'(t) * f (src_t.c:1082-1089).  moment: optional
    (Mxx,..,Mxy) of a moment part with the same time function."""
    orc = pyoracle.load_oracle()
    ijk = np.array(ijk, np.int32).reshape(-1, 3)
    ns = len(ijk)
    rk_t = [0.0, 0.5, 0.5, 1.0]
    stf = np.zeros((nt, 4), np.float32)
    rate = np.zeros((nt, 4), np.float32)
    for it in range(nt):
        for s in range(4):
            t = np.float32(np.float32(it * np.float32(dt)) + np.float32(np.float32(rk_t[s]) * np.float32(dt)))
            stf[it, s] = orc.orc_ricker(float(t), fc, t0)
            rate[it, s] = orc.orc_ricker_deriv(float(t), fc, t0)
    f = np.array(fxyz, np.float32).reshape(-1, 3)
    if len(f) == 1:
        f = np.repeat(f, ns, axis=0)
    fx = np.zeros((3, ns, nt, 4), np.float32)
    for q in range(ns):
        for m in range(3):
            fx[m, q] = (stf * f[q, m]).astype(np.float32)
    half = 3 if gauss else 0
    nk2 = 3 + case["nk"] - 1
    surf = [q for q in range(ns) if strict and ijk[q, 2] + half >= nk2]
    fr = np.zeros((3, len(surf), nt, 4), np.float32)
    for n, q in enumerate(surf):
        for m in range(3):
            fr[m, n] = (rate * f[q, m]).astype(np.float32)
    src = dict(nsrc=ns, ijk=ijk.reshape(-1).tolist(), inc=list(inc) * ns if len(inc) == 3 else list(inc),
               it_begin=[0] * ns, it_end=[nt - 1] * ns, max_nt=nt, mij=None,
               fxyz=fx.reshape(-1), gauss=int(gauss), ext_half=3, ext_coef=1.5,
               strict=int(strict), surf_idx=surf, frate=fr.reshape(-1))
    if moment is not None:
        mij = np.zeros((6, ns, nt, 4), np.float32)
        for m in range(6):
            mij[m] = (stf * np.float32(moment[m])).astype(np.float32)
        src["mij"] = mij.reshape(-1)
    return src


def dd_sources(case, ndd, max_nt, seed=0, vi=True, mij=True, dup=2, scale=1.0e3,
               nt_cap=None, surface=True):
    """dd sources at ndd random physical points (the last `dup` repeat
    earlier ones, as overlapping input points do) with random value tables
    vi [max_nt][4][ndd][3] and mij [max_nt][4][ndd][6]; with surface, one
    point sits on the top row."""
    rng = np.random.default_rng(seed + 101)
    ni, nj, nk = case["ni"], case["nj"], case["nk"]
    nx, ny, _ = dims(ni, nj, nk)
    n0 = ndd - dup
    i = rng.integers(3, 3 + ni, n0)
    j = rng.integers(3, 3 + nj, n0)
    k = rng.integers(3, 3 + nk, n0)
    if surface and n0 > 0:
        k[0] = 3 + nk - 1
    idx = i + j * nx + k * nx * ny
    idx = np.concatenate([idx, idx[rng.integers(0, n0, dup)]]).astype(np.int64)
    dd = dict(indx=idx, max_nt=max_nt, nt_cap=nt_cap, vi=None, mij=None)
    if vi:
        dd["vi"] = (scale * rng.standard_normal((max_nt, 4, ndd, 3))).astype(np.float32)
    if mij:
        dd["mij"] = (scale * 1.0e9 * rng.standard_normal((max_nt, 4, ndd, 6))).astype(np.float32)
    return dd


def surface_receivers(case, n=4, seed=0):
    """Receivers on the free surface at grid nodes (dk forced to 0 there,
    io_funcs.c:366-368) with fractional di/dj."""
    rng = np.random.default_rng(seed + 5)
    ni, nj, nk = case["ni"], case["nj"], case["nk"]
    nx, ny, _ = dims(ni, nj, nk)
    idx, w = [], []
    for _ in range(n):
        i = int(rng.integers(3, 3 + ni - 1)); j = int(rng.integers(3, 3 + nj - 1))
        k = 3 + nk - 1
        di, dj, dk = float(rng.uniform(0, 1)), float(rng.uniform(0, 1)), 0.0
        idx.append(corner_indices(i, j, k, nx, ny))
        w.append((di, dj, dk))
    return np.array(idx, np.int32), np.array(w, np.float32)


def buried_receiver(case, i, j, k, di, dj, dk):
    nx, ny, _ = dims(case["ni"], case["nj"], case["nk"])
    return (np.array([corner_indices(i, j, k, nx, ny)], np.int32),
            np.array([(di, dj, dk)], np.float32))


def corner_indices(i, j, k, nx, ny):
    """indx1d[0..7] of io_funcs.c:382-389."""
    s = nx * ny
    return [i + j * nx + k * s, i + 1 + j * nx + k * s,
            i + (j + 1) * nx + k * s, i + 1 + (j + 1) * nx + k * s,
            i + j * nx + (k + 1) * s, i + 1 + j * nx + (k + 1) * s,
            i + (j + 1) * nx + (k + 1) * s, i + 1 + (j + 1) * nx + (k + 1) * s]


def ref_cfg(case, src_arrays=None, surf_slices=None):
    """The reference harness's config for the same case; surf_slices: the
    surface-force slices [nt][4][6][ny][nx] the harness hands to the
    reference's traction image and vlow (Oracle.surface_slices)."""
    c = pyoracle.RefCfg()
    c.ni, c.nj, c.nk = case["ni"], case["nj"], case["nk"]
    c.free_top = int(case["free_top"])
    c.medium_type = int(case.get("medium_type", pyoracle.MEDIUM_ISO))
    keep = []
    if "visco_wl" in case:
        wl = np.ascontiguousarray(case["visco_wl"], np.float32)
        keep.append(wl)
        c.nmaxwell = len(wl)
        c.visco_wl = wl.ctypes.data
    for f in range(6):
        c.pml_on[f] = int(case["pml_on"][f])
        c.pml_layers[f] = int(case["pml_layers"][f])
        c.pml_alpha[f] = case["pml_alpha"]
        c.pml_beta[f] = case["pml_beta"]
        c.pml_vel[f] = case["pml_vel"]
    src = case.get("src")
    if src is not None and src["nsrc"] > 0:
        assert not src.get("gauss"), "the reference harness runs point sources only"
        a = [np.ascontiguousarray(src["ijk"], np.int32),
             np.ascontiguousarray(src["it_begin"], np.int32),
             np.ascontiguousarray(src["it_end"], np.int32)]
        keep += a
        c.nsrc = src["nsrc"]
        c.src_max_nt = src["max_nt"]
        c.src_ijk, c.src_it_begin, c.src_it_end = [x.ctypes.data for x in a]
        for key, fld in (("mij", "src_mij"), ("fxyz", "src_fxyz")):
            if src.get(key) is not None:
                t = np.ascontiguousarray(src[key], np.float32)
                keep.append(t)
                setattr(c, fld, t.ctypes.data)
        c.src_force_strict = int(src.get("strict", 0))
    if surf_slices is not None:
        t = np.ascontiguousarray(surf_slices, np.float32)
        keep.append(t)
        c.surf_slices = t.ctypes.data
        c.surf_nt = t.shape[0]
    pyoracle.set_dd_fields(c, case.get("dd"), keep)
    return c, keep


def make_solver(case, dt, nt_total, variant=4):
    """The product engine loaded with a case's static fields."""
    from cgfd3d_amd import Solver, SRC_GAUSSIAN, SRC_POINT
    mtype = int(case.get("medium_type", pyoracle.MEDIUM_ISO))
    wl = case.get("visco_wl")
    s = Solver(case["ni"], case["nj"], case["nk"], dt, nt_total, free_top=bool(case["free_top"]),
               pml_on=case["pml_on"], pml_layers=case["pml_layers"], medium_type=mtype,
               nmaxwell=0 if wl is None else len(wl))
    s.set_kernel_variant(variant)
    s.set_metric(case["metric"])
    s.set_medium(mtype, case["med"], wl=wl)
    for f in range(6):
        if case["pml_on"][f]:
            s.set_pml_face(f, case["pml_A"][f], case["pml_B"][f], case["pml_D"][f])
    if case["free_top"]:
        s.set_free_surface(free_surface_mats(case))
    src = case.get("src")
    if src is not None and src["nsrc"] > 0 and src.get("fxyz") is not None:
        # force sources (with or without a moment part): cgfdh_set_sources
        gauss = bool(src.get("gauss"))
        surf = src.get("surf_idx")
        surf = None if surf is None or len(surf) == 0 else surf
        s.set_sources(src["ijk"], src["it_begin"], src["it_end"], src["max_nt"],
                      mij=src.get("mij"), fxyz=src["fxyz"],
                      spatial=SRC_GAUSSIAN if gauss else SRC_POINT,
                      ext_half=src.get("ext_half", 3) if gauss else 0,
                      ext_coefs=gauss_coefs(case) if gauss else None,
                      strict=bool(src.get("strict", 0)), surf_idx=surf,
                      frate=None if surf is None else src["frate"])
    elif src is not None and src["nsrc"] > 0:
        if src.get("gauss"):
            coefs = gauss_coefs(case)
            s.set_moment_sources(src["ijk"], src["it_begin"], src["it_end"], src["max_nt"],
                                 src["mij"], SRC_GAUSSIAN, src.get("ext_half", 3), coefs)
        else:
            s.set_moment_sources(src["ijk"], src["it_begin"], src["it_end"], src["max_nt"],
                                 src["mij"], SRC_POINT)
    dd = case.get("dd")
    if dd is not None and len(dd["indx"]) > 0:
        s.set_dd_sources(dd["indx"], dd["max_nt"], nt_cap=dd.get("nt_cap"),
                         vi=dd.get("vi") is not None, mij=dd.get("mij") is not None)
        if dd.get("nt_cap") is None:
            s.put_dd_values(0, dd.get("vi"), dd.get("mij"))
    return s


def free_surface_mats(case):
    """matVx2Vz2, matVy2Vz2, matF2Vz2 (and matD for the visco-elastic medium)."""
    orc = pyoracle.load_oracle()
    nx, ny, _ = dims(case["ni"], case["nj"], case["nk"])
    if case.get("medium_type") == pyoracle.MEDIUM_VISCO:
        mats = np.zeros(4 * nx * ny * 9, np.float32)
        orc.orc_dvh2dvz_vis(case["ni"], case["nj"], case["nk"],
                            pyoracle.ptr(np.ascontiguousarray(case["metric"])),
                            pyoracle.ptr(np.ascontiguousarray(case["med"])),
                            pyoracle.ptr(np.ascontiguousarray(case["coords"])), pyoracle.ptr(mats))
        return mats
    mats = np.zeros(3 * nx * ny * 9, np.float32)
    orc.orc_dvh2dvz(int(case.get("medium_type", pyoracle.MEDIUM_ISO)),
                    case["ni"], case["nj"], case["nk"],
                    pyoracle.ptr(np.ascontiguousarray(case["metric"])),
                    pyoracle.ptr(np.ascontiguousarray(case["med"])), pyoracle.ptr(mats))
    return mats


def gauss_coefs(case):
    """Normalised Gaussian weights per source (src_t.c:2111-2151)."""
    orc = pyoracle.load_oracle()
    src = case["src"]
    half = src.get("ext_half", 3)
    L = 2 * half + 1
    nk2 = 3 + case["nk"] - 1
    out = np.zeros((src["nsrc"], L ** 3), np.float32)
    ijk = np.array(src["ijk"]).reshape(-1, 3)
    inc = np.array(src["inc"], np.float32).reshape(-1, 3)
    for s in range(src["nsrc"]):
        sk = ijk[s, 2]
        kb = half if sk + half < nk2 else nk2 - sk
        w = np.zeros(L ** 3, np.float32)
        orc.orc_gauss_weights(pyoracle.ptr(w), float(inc[s, 0]), float(inc[s, 1]),
                              float(inc[s, 2]), src.get("ext_coef", 1.5), half, kb)
        out[s] = w
    return out.reshape(-1)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def workload_case(wl):
    """The oracle case of a whole Workload of cgfd3d_amd.workloads (metric from
    the oracle's own restatement, PML tables, medium volumes, point source)."""
    ni, nj, nk = wl.ni, wl.nj, wl.nk
    orc = pyoracle.load_oracle()
    nx, ny, nz = dims(ni, nj, nk)
    xyz = np.ascontiguousarray(wl.xyz.reshape(3, nz, ny, nx), np.float32)
    metric = np.zeros((10, nz, ny, nx), np.float32)
    orc.orc_metric(ni, nj, nk, pyoracle.ptr(xyz), pyoracle.ptr(metric))
    A, B, D, region = pml_tables(orc, ni, nj, nk, xyz, wl.pml_on, wl.pml_layers,
                                 vel=wl.pml_vel, alpha=wl.pml_alpha, beta=wl.pml_beta)
    vols = [wl.lam, wl.mu, wl.slw] if wl.med_vols is None else wl.med_vols
    med = np.ascontiguousarray(np.stack([np.asarray(v).reshape(nz, ny, nx) for v in vols]),
                               np.float32)
    case = dict(ni=ni, nj=nj, nk=nk, metric=metric, med=med, medium_type=wl.medium_type,
                free_top=int(wl.free_top), pml_on=list(wl.pml_on), pml_layers=list(wl.pml_layers),
                pml_A=A, pml_B=B, pml_D=D, pml_region=region, coords=xyz,
                pml_vel=wl.pml_vel, pml_alpha=wl.pml_alpha, pml_beta=wl.pml_beta)
    if wl.wl is not None:
        case["visco_wl"] = np.asarray(wl.wl, np.float32)
    if wl.src is not None:
        src = dict(wl.src)
        src.update(nsrc=1, inc=[0, 0, 0], gauss=0)
        case["src"] = src
    return case


DT_MEDIA = ("iso", "vti", "aniso", "visco", "cartesian")


def dt_case(medium, ni=14, nj=12, nk=10):
    """Inputs of the stable-time-step estimate (blk_dt_esti_curv): coordinates
    (3, nz, ny, nx), the medium name the estimate takes, its volumes in the
    order cgfdh_estimate_dt takes them, and nmaxwell.  'cartesian' is a
    homogeneous isotropic medium on an exact 64 m Cartesian grid, where every
    point ties."""
    xyz = curvilinear_coords(ni, nj, nk, seed=5, hill=0.4)
    nmw = 0
    if medium in ("iso", "cartesian"):
        lam, mu, slw = iso_medium(ni, nj, nk, seed=6)
        vols = np.stack([lam, mu, (1.0 / slw).astype(np.float32)])
        if medium == "cartesian":
            nx, ny, nz = dims(ni, nj, nk)
            z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
            xyz = np.stack([x, y, z]).astype(np.float32) * np.float32(64.0)
            vols = np.stack([np.full_like(lam, 3.0e10), np.full_like(lam, 2.5e10),
                             np.full_like(lam, 2500.0)])
        medium = "iso"
    elif medium == "vti":
        v = vti_medium(ni, nj, nk, seed=6)
        vols = np.stack([v[0], v[2], (1.0 / v[5]).astype(np.float32)])
    elif medium == "aniso":
        v = aniso_medium(ni, nj, nk, seed=6)
        vols = np.stack([v[0], (1.0 / v[21]).astype(np.float32)])
    else:
        v, wl = visco_medium(ni, nj, nk, seed=6)
        vols = np.concatenate([v[:2], (1.0 / v[2]).astype(np.float32)[None], v[3:]])
        nmw = len(wl)
    return (np.ascontiguousarray(xyz, np.float32), medium,
            np.ascontiguousarray(vols, np.float32), nmw)
