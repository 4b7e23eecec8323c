"""Seeded synthetic inputs of the benchmark configurations, built for the
engine: curvilinear coordinates over Gaussian-hill topography, layered media
with a Gaussian random field, moment sources and surface receivers.

Layouts are the reference's (3 ghost points per side, x fastest).  Large
fields are generated with torch on the GPU when one is present (plumbing
only: the arrays are handed to the engine as host buffers).
"""
import ctypes as C
import math

import numpy as np

from ._lib import check, lib
from .solver import (MEDIUM_ELASTIC_ANISO, MEDIUM_ELASTIC_ISO, MEDIUM_ELASTIC_VTI,
                     MEDIUM_VISCO_ISO, SRC_POINT, Solver)

# medium coefficient volumes per point (besides the 9 metric volumes)
_NCOEF = {MEDIUM_ELASTIC_ISO: 3, MEDIUM_ELASTIC_VTI: 6, MEDIUM_ELASTIC_ANISO: 22}

_P = C.c_void_p


def _p(a):
    return a.ctypes.data_as(_P)


def _setup_sigs():
    L = lib()
    L.cgfdh_compute_metric.argtypes = [_P, _P]
    L.cgfdh_compute_metric.restype = C.c_int
    L.cgfdh_compute_free_surface.argtypes = [_P]
    L.cgfdh_compute_free_surface.restype = C.c_int
    L.cgfdh_get_metric.argtypes = [_P, _P]
    L.cgfdh_get_metric.restype = C.c_int
    L.cgfdh_get_free_surface.argtypes = [_P, _P]
    L.cgfdh_get_free_surface.restype = C.c_int
    L.cgfdh_pml_face_coefs.argtypes = [C.c_int, C.c_int, C.c_int, _P, C.c_int, C.c_int,
                                       C.c_float, C.c_float, C.c_float, _P, _P, _P]
    L.cgfdh_pml_face_coefs.restype = C.c_int
    L.cgfdh_ricker.argtypes = [C.c_float, C.c_float, C.c_float]
    L.cgfdh_ricker.restype = C.c_float
    L.cgfdh_moment_from_angles.argtypes = [C.c_float] * 6 + [_P]
    L.cgfdh_moment_from_angles.restype = None
    L.cgfdh_ricker_moment_table.argtypes = [C.c_float] * 6 + [_P, _P, _P, _P, _P, C.c_long]
    L.cgfdh_ricker_moment_table.restype = C.c_int
    L.cgfdh_gauss_weights.argtypes = [_P] + [C.c_float] * 4 + [C.c_int, C.c_int]
    L.cgfdh_gauss_weights.restype = C.c_int
    L.cgfdh_visco_gmb.argtypes = [C.c_long, C.c_int, C.c_float, C.c_float, C.c_float,
                                  _P, _P, _P, _P, _P, _P]
    L.cgfdh_visco_gmb.restype = C.c_int
    return L


def _torch_device():
    try:
        import torch
        if torch.cuda.is_available():
            return torch, torch.device("cuda:0")
        return torch, torch.device("cpu")
    except Exception:
        return None, None


def topography(ni, nj, dh, nhills, seed, amp_max=1000.0, sig_min=1000.0, sig_max=5000.0):
    """Surface elevation (ny, nx) with ghosts: a sum of seeded Gaussian hills."""
    rng = np.random.default_rng(seed)
    nx, ny = ni + 6, nj + 6
    x = (np.arange(nx) - 3.0) * dh
    y = (np.arange(ny) - 3.0) * dh
    X, Y = np.meshgrid(x, y)
    top = np.zeros((ny, nx))
    Lx, Ly = (ni - 1) * dh, (nj - 1) * dh
    for _ in range(nhills):
        cx, cy = rng.uniform(0, Lx), rng.uniform(0, Ly)
        amp = rng.uniform(0.2, 1.0) * amp_max
        sig = rng.uniform(sig_min, sig_max)
        top += amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sig * sig))
    return top


def layered_coords(ni, nj, nk, dh, top, depth):
    """x3d,y3d,z3d: vertical columns from a flat bottom at -depth up to the
    surface elevation, equally divided (ghost layers extend the same map)."""
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    xyz = np.empty((3, nz, ny, nx), np.float32)
    xyz[0] = ((np.arange(nx) - 3.0) * dh).astype(np.float32)[None, None, :]
    xyz[1] = ((np.arange(ny) - 3.0) * dh).astype(np.float32)[None, :, None]
    frac = ((np.arange(nz) - 3.0) / (nk - 1)).astype(np.float64)
    for k in range(nz):
        xyz[2, k] = (-depth + frac[k] * (top + depth)).astype(np.float32)
    return xyz


def gaussian_random_field(shape, dh, corr_len, std, seed):
    """Zero-mean field with Gaussian correlation exp(-r^2/corr_len^2)."""
    torch, dev = _torch_device()
    nz, ny, nx = shape
    if torch is not None:
        g = torch.Generator(device=dev).manual_seed(seed)
        f = torch.randn(shape, generator=g, device=dev, dtype=torch.float32)
        F = torch.fft.rfftn(f)
        kz = torch.fft.fftfreq(nz, d=dh, device=dev)
        ky = torch.fft.fftfreq(ny, d=dh, device=dev)
        kx = torch.fft.rfftfreq(nx, d=dh, device=dev)
        k2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
        F *= torch.exp(-k2 * (math.pi * corr_len) ** 2 / 2.0)
        f = torch.fft.irfftn(F, s=shape)
        f *= std / f.std()
        out = f.cpu().numpy().astype(np.float32)
        del f, F, k2
        if dev.type == "cuda":
            torch.cuda.empty_cache()
        return out
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(shape).astype(np.float32)
    F = np.fft.rfftn(f)
    kz = np.fft.fftfreq(nz, d=dh); ky = np.fft.fftfreq(ny, d=dh); kx = np.fft.rfftfreq(nx, d=dh)
    k2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
    F *= np.exp(-k2 * (math.pi * corr_len) ** 2 / 2.0)
    f = np.fft.irfftn(F, s=shape).astype(np.float32)
    return f * (std / f.std())


def layered_velocity(z3d, layers, grf):
    """vp, vs, rho from a layer table [(z_top_of_layer, vp, vs, rho)] ordered
    from the top down, scaled by (1 + grf)."""
    vp = np.empty(z3d.shape, np.float32)
    vs = np.empty_like(vp)
    rho = np.empty_like(vp)
    vp[...] = layers[0][1]; vs[...] = layers[0][2]; rho[...] = layers[0][3]
    for ztop, a, b, r in layers[1:]:
        m = z3d < ztop
        vp[m] = a; vs[m] = b; rho[m] = r
    if grf is not None:
        s = 1.0 + grf
        vp *= s; vs *= s; rho *= s
    return vp, vs, rho


def layered_medium(z3d, layers, grf):
    """lambda, mu, slowness (and the largest vp) of the layered model."""
    vp, vs, rho = layered_velocity(z3d, layers, grf)
    mu = (rho * vs * vs).astype(np.float32)
    lam = (rho * vp * vp - 2.0 * mu).astype(np.float32)
    slw = np.where(rho > 1e-10, 1.0 / rho.astype(np.float64), 0.0).astype(np.float32)
    return lam, mu, slw, float(vp.max())


def corner_indices(i, j, k, nx, ny):
    s = nx * ny
    return [i + j * nx + k * s, i + 1 + j * nx + k * s, i + (j + 1) * nx + k * s,
            i + 1 + (j + 1) * nx + k * s, i + j * nx + (k + 1) * s,
            i + 1 + j * nx + (k + 1) * s, i + (j + 1) * nx + (k + 1) * s,
            i + 1 + (j + 1) * nx + (k + 1) * s]


def surface_receiver_grid(ni, nj, nk, n_side):
    """n_side x n_side receivers at surface nodes (dk = 0 there, io_funcs.c:366-368)."""
    nx, ny = ni + 6, nj + 6
    idx, w = [], []
    for a in range(n_side):
        for b in range(n_side):
            i = 3 + int((a + 0.5) * (ni - 1) / n_side)
            j = 3 + int((b + 0.5) * (nj - 1) / n_side)
            idx.append(corner_indices(i, j, 3 + nk - 1, nx, ny))
            w.append((0.0, 0.0, 0.0))
    return np.array(idx, np.int32), np.array(w, np.float32)


def keep_two_digi(dt):
    """blk_keep_two_digi (blk_t.c:3004-3017): the time step cut to its two
    leading significant digits, as the reference prints it with %9.7e, zeroes
    the mantissa digits after the first decimal and reads it back."""
    s = list("%9.7e" % np.float32(dt))
    for i in range(3, 9):
        s[i] = "0"
    return float(np.float32("".join(s)))


def stable_dt(xyz, medium_type, vols, nmaxwell=0, cfl=1.3):
    """The reference's check_stability path (main_curv_col_el_3d.c:617-668):
    dtmax from blk_dt_esti_curv, computed on the GPU by cgfdh_estimate_dt
    (cfl = fd->CFL, fd_t.c:55), and the step it sets when none is given,
    blk_keep_two_digi(dtmax).  vols: the estimate's medium volumes (lambda,
    mu, rho for the isotropic medium).  Returns (dtmax, dt)."""
    from .solver import estimate_dt
    dtmax = estimate_dt(xyz, medium_type, vols, cfl, nmaxwell)[0]
    if not dtmax > 0.0:
        raise ValueError(f"maximum dt {dtmax} <= 0")
    return dtmax, keep_two_digi(dtmax)


def _check_or_set_dt(dt, xyz, medium_type, vols, nmaxwell=0):
    """dt None: the stable step blk_keep_two_digi(dtmax); a given dt above
    dtmax stops, as the reference does (main_curv_col_el_3d.c:663-667)."""
    dtmax, auto = stable_dt(xyz, medium_type, np.stack([np.asarray(v, np.float32)
                                                         for v in vols]), nmaxwell)
    if dt is None:
        return auto, dtmax
    if dtmax < dt:
        raise ValueError(f"dt={dt} > dtmax={dtmax}: the run would be unstable")
    return dt, dtmax


class Workload:
    """Everything one run of the engine needs, as host arrays."""

    def __init__(self, name, ni, nj, nk, dt, nt_total, xyz, lam, mu, slw, pml_layers,
                 pml_vel, pml_alpha, pml_beta, free_top, src, recv_idx, recv_w,
                 medium_type=MEDIUM_ELASTIC_ISO, med_vols=None, wl=None, nmaxwell=0):
        """lam, mu, slw: the isotropic medium; other media pass medium_type
        and med_vols (the volumes in cgfdh_set_medium_* order; visco with wl)."""
        self.name = name
        self.medium_type, self.med_vols, self.wl, self.nmaxwell = medium_type, med_vols, wl, nmaxwell
        self.ni, self.nj, self.nk = ni, nj, nk
        self.dt, self.nt_total = dt, nt_total
        self.xyz, self.lam, self.mu, self.slw = xyz, lam, mu, slw
        self.pml_layers = pml_layers
        self.pml_on = [1 if n > 0 else 0 for n in pml_layers]
        self.pml_vel, self.pml_alpha, self.pml_beta = pml_vel, pml_alpha, pml_beta
        self.free_top = free_top
        self.src = src
        self.recv_idx, self.recv_w = recv_idx, recv_w
        self.dtmax = None          # the stable-step estimate, when it was made

    @property
    def npoints(self):
        return self.ni * self.nj * self.nk

    def n_aux(self):
        n = 0
        for f, L in enumerate(self.pml_layers):
            if L <= 0:
                continue
            dims = [self.ni, self.nj, self.nk]
            dims[f // 2] = L + 1
            n += dims[0] * dims[1] * dims[2]
        return n

    def bytes_per_point(self):
        """Minimum FP32 traffic per point and RK4 step: 16 wavefield
        components' worth per component (stage 1 reads pre and writes tmp,
        end; stages 2-3 read tmp, pre, end and write tmp, end; stage 4 reads
        tmp, end and writes end) plus the 9 metric and the medium volumes
        once per stage: 4 (16 nw + 4 ncoef) bytes.  Iso 768, VTI 816, aniso
        1072, visco with 3 Maxwell bodies 2016."""
        nw = 9 + 6 * self.nmaxwell
        ncoef = 9 + (3 + 2 * self.nmaxwell if self.medium_type == MEDIUM_VISCO_ISO
                     else _NCOEF[self.medium_type])
        return 4.0 * (16 * nw + 4 * ncoef)

    def algorithmic_bytes_per_step(self):
        """Minimum HBM traffic of one RK4 step (FP32): bytes_per_point per
        point, 576 B per PML aux point (9 components), 352 B per free-surface
        column (18 matrix floats and 4 jacobian values per stage; +36 B of
        the visco surface rotation)."""
        nsurf = self.ni * self.nj if self.free_top else 0
        surf = 352.0 + (36.0 if self.medium_type == MEDIUM_VISCO_ISO else 0.0)
        return self.bytes_per_point() * self.npoints + 576.0 * self.n_aux() + surf * nsurf

    def build_solver(self, variant=None):
        """A configured solver; variant None keeps the engine's default
        interior kernel for the medium."""
        L = _setup_sigs()
        s = Solver(self.ni, self.nj, self.nk, self.dt, self.nt_total, free_top=self.free_top,
                   pml_on=self.pml_on, pml_layers=self.pml_layers,
                   medium_type=self.medium_type, nmaxwell=self.nmaxwell)
        if variant is not None:
            s.set_kernel_variant(variant)
        check(L.cgfdh_compute_metric(s.h, _p(self.xyz)), "compute_metric")
        if self.medium_type == MEDIUM_ELASTIC_ISO:
            s.set_medium_iso(self.lam, self.mu, self.slw)
        else:
            s.set_medium(self.medium_type, self.med_vols, self.wl)
        for f in range(6):
            n = self.pml_layers[f]
            if n <= 0:
                continue
            A = np.zeros(n + 1, np.float32); B = np.zeros_like(A); D = np.zeros_like(A)
            check(L.cgfdh_pml_face_coefs(self.ni, self.nj, self.nk, _p(self.xyz), f, n,
                                         self.pml_vel, self.pml_alpha, self.pml_beta,
                                         _p(A), _p(B), _p(D)), "pml_face_coefs")
            s.set_pml_face(f, A, B, D)
        if self.free_top:
            check(L.cgfdh_compute_free_surface(s.h), "compute_free_surface")
        if self.src is not None:
            s.set_moment_sources(**self.src)
        if self.recv_idx is not None and len(self.recv_idx):
            s.set_receivers(self.recv_idx, self.recv_w)
        return s


def moment_source(ni, nj, nk, i, j, k, dt, t0, mech6, fc, tpeak, stf_length, tstart=0.0):
    """Point moment source with a Ricker rate table (src_t.c:985-1174)."""
    L = _setup_sigs()
    m6 = np.ascontiguousarray(mech6, np.float32)
    itb, ite, mnt = C.c_int(), C.c_int(), C.c_int()
    check(L.cgfdh_ricker_moment_table(t0, dt, stf_length, tstart, fc, tpeak, _p(m6),
                                      C.byref(itb), C.byref(ite), C.byref(mnt), None, 0),
          "ricker_moment_table")
    n = mnt.value
    table = np.zeros((6, n * 4), np.float32)
    check(L.cgfdh_ricker_moment_table(t0, dt, stf_length, tstart, fc, tpeak, _p(m6),
                                      C.byref(itb), C.byref(ite), C.byref(mnt), _p(table),
                                      n * 4), "ricker_moment_table")
    return dict(ijk=[i, j, k], it_begin=[itb.value], it_end=[ite.value], max_nt=n,
                mij=table.reshape(-1), spatial=SRC_POINT)


def iso_hill_explosive(ni=100, nj=100, nk=60, dh=100.0, dt=0.005, nt_total=1001,
                       hill_height=500.0, hill_sigma=1500.0, bottom=-5900.0, src_depth=2000.0,
                       m0=1.0e16, fc=2.0, tpeak=0.5, stf_length=2.0, pml=10, n_side=4,
                       check_dt=False):
    """Isotropic elastic homogeneous half-space under one Gaussian hill: an
    ni x nj x nk curvilinear grid, dx = dy = dh, columns from a flat bottom at
    z = bottom up to the surface (cartesian origin -5900 m with 100 m spacing
    in the reference's example, example/cgfd3d.example.sh:235-236), Vp 6000,
    Vs 3464 m/s, rho 2670 kg/m^3, an explosive point source (Mxx = Myy = Mzz
    = m0) src_depth below the surface of the hill's centre column at the
    nearest grid node, a Ricker rate with dominant frequency fc and peak time
    tpeak, CFS-PML of pml layers on 5 faces (alpha_max 3.14, beta_max 2,
    ref_vel 7000, as in the example's cfspml blocks), the traction-image free
    surface and n_side^2 surface receivers.  check_dt: stop if dt exceeds the
    stable-step estimate (GPU), as the reference does."""
    nx, ny = ni + 6, nj + 6
    x = (np.arange(nx) - 3.0) * dh
    y = (np.arange(ny) - 3.0) * dh
    ci, cj = ni // 2, nj // 2
    cx, cy = ci * dh, cj * dh
    X, Y = np.meshgrid(x, y)
    top = hill_height * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * hill_sigma ** 2))
    xyz = layered_coords(ni, nj, nk, dh, top, -bottom)
    layers = [(1e9, 6000.0, 3464.0, 2670.0)]
    lam, mu, slw, _ = layered_medium(xyz[2], layers, None)
    dtmax = None
    if check_dt:
        rho = np.full(lam.shape, 2670.0, np.float32)
        dt, dtmax = _check_or_set_dt(dt, xyz, MEDIUM_ELASTIC_ISO, (lam, mu, rho))
    si, sj = 3 + ci, 3 + cj
    zcol = xyz[2, 3:3 + nk, sj, si].astype(np.float64)
    sk = 3 + int(np.argmin(np.abs(zcol - (zcol[-1] - src_depth))))
    m6 = np.array([m0, m0, m0, 0.0, 0.0, 0.0], np.float32)
    src = moment_source(ni, nj, nk, si, sj, sk, dt, 0.0, m6, fc, tpeak, stf_length)
    ridx, rw = surface_receiver_grid(ni, nj, nk, n_side)
    wlo = Workload(f"iso_hill{ni}x{nj}x{nk}", ni, nj, nk, dt, nt_total, xyz, lam, mu, slw,
                   [pml] * 5 + [0], 7000.0, 3.14, 2.0, True, src, ridx, rw)
    wlo.dtmax = dtmax
    return wlo


def iso_topo_grf(ni=600, nj=600, nk=300, dh=50.0, nt_total=2001, dt=None, seed=2024,
                 nhills=20, pml=20, n_side=16, check_dt=True):
    """Isotropic elastic, ni x nj x nk curvilinear grid over a sum of seeded
    Gaussian hills (amplitude <= 1 km, sigma 1-5 km), 3-layer model times a
    Gaussian random field (5 % std, 500 m correlation), double-couple source
    (seeded strike/dip/rake) with a 2 Hz Ricker, CFS-PML on 5 faces, traction
    free surface, n_side^2 surface receivers.  dt None: the reference's
    stable step blk_keep_two_digi(dtmax) (estimated on the GPU); a given dt
    is checked against dtmax unless check_dt is False."""
    top = topography(ni, nj, dh, nhills, seed)
    depth = (nk - 1) * dh
    xyz = layered_coords(ni, nj, nk, dh, top, depth)
    grf = gaussian_random_field((nk + 6, nj + 6, ni + 6), dh, 500.0, 0.05, seed + 1)
    layers = [(1e9, 4500.0, 2600.0, 2400.0), (-1500.0, 6000.0, 3464.0, 2670.0),
              (-6000.0, 6800.0, 3900.0, 2900.0)]
    vp, vs, rho = layered_velocity(xyz[2], layers, grf)
    del grf
    mu = (rho * vs * vs).astype(np.float32)
    lam = (rho * vp * vp - 2.0 * mu).astype(np.float32)
    slw = np.where(rho > 1e-10, 1.0 / rho.astype(np.float64), 0.0).astype(np.float32)
    del vp, vs
    dtmax = None
    if dt is None or check_dt:
        dt, dtmax = _check_or_set_dt(dt, xyz, MEDIUM_ELASTIC_ISO, (lam, mu, rho))
    del rho
    rng = np.random.default_rng(seed + 2)
    strike, dip, rake = rng.uniform(0, 360), rng.uniform(10, 90), rng.uniform(-180, 180)
    L = _setup_sigs()
    m6 = np.zeros(6, np.float32)
    L.cgfdh_moment_from_angles(strike, dip, rake, 3.0e10, 1.0, 1.0e6, _p(m6))
    si, sj = 3 + ni // 2, 3 + nj // 2
    sk = 3 + nk - 1 - min(int(round(5000.0 / dh)), int(0.4 * nk))
    src = moment_source(ni, nj, nk, si, sj, sk, dt, 0.0, m6, 2.0, 0.5, 2.0)
    ridx, rw = surface_receiver_grid(ni, nj, nk, n_side)
    wlo = Workload(f"iso{ni}x{nj}x{nk}", ni, nj, nk, dt, nt_total, xyz, lam, mu, slw,
                   [pml, pml, pml, pml, pml, 0], 7000.0, 3.14, 2.0, True, src, ridx, rw)
    wlo.dtmax = dtmax
    return wlo


def _layer_index(z3d, layers):
    """Layer number (0 = top) of every point for a layer table ordered from the top."""
    idx = np.zeros(z3d.shape, np.int8)
    for n, (ztop, *_r) in enumerate(layers[1:], start=1):
        idx[z3d < ztop] = n
    return idx


def _bond_matrix(R):
    """6x6 Bond matrix M of a rotation R (Voigt order 11,22,33,23,13,12) with
    C' = M C M^T for the stiffness in Voigt notation."""
    M = np.zeros((6, 6))
    pairs = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    for a, (i, j) in enumerate(pairs):
        for b, (k, l) in enumerate(pairs):
            if b < 3:
                M[a, b] = R[i, k] * R[j, l]
            else:
                M[a, b] = R[i, k] * R[j, l] + R[i, l] * R[j, k]
    return M


def _random_rotation(rng):
    """Uniformly distributed rotation matrix (QR of a Gaussian matrix)."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q *= np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def media_topo_grf(kind, ni=400, nj=400, nk=200, dh=50.0, nt_total=1501, dt=None, seed=2025,
                   nhills=20, pml=20, n_side=16, nmaxwell=3, qp=100.0, qs=50.0,
                   fmin=0.5, fmax=5.0, fref=1.0, check_dt=True):
    """The anisotropy / attenuation workloads on the topography + layered GRF
    model of iso_topo_grf: kind 'vti' (Thomsen eps, delta, gamma seeded per
    layer in [0, 0.2]), 'aniso' (the VTI stiffness under one seeded random
    rotation: 21 Cij) or 'visco' (GMB with nmaxwell bodies, constant Qp, Qs
    over fmin..fmax Hz).  dt as for iso_topo_grf, from each medium's own
    stable-step estimate."""
    top = topography(ni, nj, dh, nhills, seed)
    depth = (nk - 1) * dh
    xyz = layered_coords(ni, nj, nk, dh, top, depth)
    grf = gaussian_random_field((nk + 6, nj + 6, ni + 6), dh, 500.0, 0.05, seed + 1)
    layers = [(1e9, 4500.0, 2600.0, 2400.0), (-1500.0, 6000.0, 3464.0, 2670.0),
              (-6000.0, 6800.0, 3900.0, 2900.0)]
    vp, vs, rho = layered_velocity(xyz[2], layers, grf)
    lid = _layer_index(xyz[2], layers)
    del grf
    rng = np.random.default_rng(seed + 5)
    slw = np.where(rho > 1e-10, 1.0 / rho.astype(np.float64), 0.0).astype(np.float32)
    wl = None
    if kind in ("vti", "aniso"):
        th = rng.uniform(0.0, 0.2, (len(layers), 3))        # eps, delta, gamma per layer
        e, d, g = th[lid, 0], th[lid, 1], th[lid, 2]
        c33 = rho.astype(np.float64) * vp * vp
        c55 = rho.astype(np.float64) * vs * vs
        c11 = c33 * (1 + 2 * e)
        c66 = c55 * (1 + 2 * g)
        c13 = np.sqrt((c33 - c55) * (2 * d * c33 + c33 - c55)) - c55
        del e, d, g
        if kind == "vti":
            vols = [c.astype(np.float32) for c in (c11, c13, c33, c55, c66)] + [slw]
            mtype = MEDIUM_ELASTIC_VTI
        else:
            M = _bond_matrix(_random_rotation(rng))
            n = c11.size
            cij = np.empty((21, n), np.float32)
            ut = [(a, b) for a in range(6) for b in range(a, 6)]
            chunk = 1 << 20
            f = [c.reshape(-1) for c in (c11, c13, c33, c55, c66)]
            for s0 in range(0, n, chunk):
                sl = slice(s0, min(n, s0 + chunk))
                m = sl.stop - sl.start
                C = np.zeros((m, 6, 6))
                C[:, 0, 0] = C[:, 1, 1] = f[0][sl]
                C[:, 0, 1] = C[:, 1, 0] = f[0][sl] - 2.0 * f[4][sl]
                C[:, 0, 2] = C[:, 2, 0] = C[:, 1, 2] = C[:, 2, 1] = f[1][sl]
                C[:, 2, 2] = f[2][sl]
                C[:, 3, 3] = C[:, 4, 4] = f[3][sl]
                C[:, 5, 5] = f[4][sl]
                Cr = M @ C @ M.T
                for q, (a, b) in enumerate(ut):
                    cij[q, sl] = Cr[:, a, b]
            vols = [cij.reshape((21,) + c11.shape)[q] for q in range(21)] + [slw]
            mtype = MEDIUM_ELASTIC_ANISO
        del c11, c13, c33, c55, c66
    elif kind == "visco":
        mu = (rho * vs * vs).astype(np.float32).reshape(-1)
        lam = (rho * vp * vp - 2.0 * mu.reshape(rho.shape)).astype(np.float32).reshape(-1)
        n = lam.size
        Qp = np.full(n, qp, np.float32)
        Qs = np.full(n, qs, np.float32)
        y = np.zeros((2 * nmaxwell, n), np.float32)
        wl = np.zeros(nmaxwell, np.float32)
        L = _setup_sigs()
        check(L.cgfdh_visco_gmb(n, nmaxwell, fref, fmin, fmax, _p(lam), _p(mu), _p(Qp), _p(Qs),
                                _p(y), _p(wl)), "visco_gmb")
        shp = rho.shape
        vols = [lam.reshape(shp), mu.reshape(shp), slw] + [y[q].reshape(shp)
                                                           for q in range(2 * nmaxwell)]
        mtype = MEDIUM_VISCO_ISO
    else:
        raise ValueError(f"unknown medium kind {kind!r}")
    dtmax = None
    if dt is None or check_dt:
        # the estimate's volumes per medium (cgfdh_estimate_dt): c11, c33, rho
        # (VTI), c11, rho (anisotropic), lambda, mu, rho, Ylam_n, Ymu_n (GMB)
        if mtype == MEDIUM_ELASTIC_VTI:
            ev = (vols[0], vols[2], rho)
        elif mtype == MEDIUM_ELASTIC_ANISO:
            ev = (vols[0], rho)
        else:
            ev = tuple(vols[:2]) + (rho,) + tuple(vols[3:])
        dt, dtmax = _check_or_set_dt(dt, xyz, mtype, ev,
                                     nmaxwell if mtype == MEDIUM_VISCO_ISO else 0)
        del ev
    del vp, vs, rho
    strike, dip, rake = rng.uniform(0, 360), rng.uniform(10, 90), rng.uniform(-180, 180)
    L = _setup_sigs()
    m6 = np.zeros(6, np.float32)
    L.cgfdh_moment_from_angles(strike, dip, rake, 3.0e10, 1.0, 1.0e6, _p(m6))
    si, sj = 3 + ni // 2, 3 + nj // 2
    sk = 3 + nk - 1 - min(int(round(5000.0 / dh)), int(0.4 * nk))
    src = moment_source(ni, nj, nk, si, sj, sk, dt, 0.0, m6, 2.0, 0.5, 2.0)
    ridx, rw = surface_receiver_grid(ni, nj, nk, n_side)
    wlo = Workload(f"{kind}{ni}x{nj}x{nk}", ni, nj, nk, dt, nt_total, xyz, None, None, None,
                   [pml, pml, pml, pml, pml, 0], 7000.0, 3.14, 2.0, True, src, ridx, rw,
                   medium_type=mtype, med_vols=vols, wl=wl,
                   nmaxwell=nmaxwell if mtype == MEDIUM_VISCO_ISO else 0)
    wlo.dtmax = dtmax
    return wlo


def batch_workloads(seed=4242, dh=50.0, dt=None, nt_total=1001, pml=20, n_side=4,
                    sizes=((24, (128, 128, 128)), (24, (192, 192, 128)), (16, (256, 256, 160))),
                    check_dt=True):
    """Yields the independent runs of the batched configuration one at a time
    (64 by default: 24 at 128^3, 24 at 192x192x128, 16 at 256x256x160), each
    with its own seeded topography, GRF medium, source position, mechanism
    and dominant frequency in 1-3 Hz, and its own stable time step (dt None)
    or the given dt checked against its dtmax."""
    rng = np.random.default_rng(seed)
    n = 0
    for count, (ni, nj, nk) in sizes:
        for _ in range(count):
            s = int(rng.integers(1 << 30))
            r = np.random.default_rng(s)
            top = topography(ni, nj, dh, int(r.integers(3, 9)), s, amp_max=600.0,
                             sig_min=600.0, sig_max=2500.0)
            depth = (nk - 1) * dh
            xyz = layered_coords(ni, nj, nk, dh, top, depth)
            grf = gaussian_random_field((nk + 6, nj + 6, ni + 6), dh, 500.0, 0.05, s + 1)
            layers = [(1e9, 4500.0, 2600.0, 2400.0), (-1500.0, 6000.0, 3464.0, 2670.0),
                      (-4000.0, 6800.0, 3900.0, 2900.0)]
            vp, vs, rho = layered_velocity(xyz[2], layers, grf)
            del grf
            mu = (rho * vs * vs).astype(np.float32)
            lam = (rho * vp * vp - 2.0 * mu).astype(np.float32)
            slw = np.where(rho > 1e-10, 1.0 / rho.astype(np.float64), 0.0).astype(np.float32)
            del vp, vs
            dti, dtmax = dt, None
            if dt is None or check_dt:
                dti, dtmax = _check_or_set_dt(dt, xyz, MEDIUM_ELASTIC_ISO, (lam, mu, rho))
            del rho
            strike, dip, rake = r.uniform(0, 360), r.uniform(10, 90), r.uniform(-180, 180)
            fc = r.uniform(1.0, 3.0)
            L = _setup_sigs()
            m6 = np.zeros(6, np.float32)
            L.cgfdh_moment_from_angles(strike, dip, rake, 3.0e10, 1.0, 1.0e6, _p(m6))
            # inside the PML frame, at least 3 nodes below the surface
            si = 3 + int(r.integers(pml + 3, max(pml + 4, ni - pml - 3)))
            sj = 3 + int(r.integers(pml + 3, max(pml + 4, nj - pml - 3)))
            sk = 3 + int(r.integers(pml + 3, max(pml + 4, nk - 4)))
            tpeak = 1.2 / fc
            src = moment_source(ni, nj, nk, si, sj, sk, dti, 0.0, m6, fc, tpeak, 2.0 * tpeak)
            ridx, rw = surface_receiver_grid(ni, nj, nk, n_side)
            wlo = Workload(f"batch{n}_{ni}x{nj}x{nk}", ni, nj, nk, dti, nt_total, xyz, lam, mu,
                           slw, [pml] * 5 + [0], 7000.0, 3.14, 2.0, True, src, ridx, rw)
            wlo.dtmax = dtmax
            yield wlo
            n += 1
