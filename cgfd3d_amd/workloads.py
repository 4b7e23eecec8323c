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
from .solver import SRC_POINT, Solver

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


def layered_medium(z3d, layers, grf):
    """lambda, mu, slowness from a layer table [(z_top_of_layer, vp, vs, rho)]
    ordered from the top down, scaled by (1 + grf)."""
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


class Workload:
    """Everything one run of the engine needs, as host arrays."""

    def __init__(self, name, ni, nj, nk, dt, nt_total, xyz, lam, mu, slw, pml_layers,
                 pml_vel, pml_alpha, pml_beta, free_top, src, recv_idx, recv_w):
        self.name = name
        self.ni, self.nj, self.nk = ni, nj, nk
        self.dt, self.nt_total = dt, nt_total
        self.xyz, self.lam, self.mu, self.slw = xyz, lam, mu, slw
        self.pml_layers = pml_layers
        self.pml_on = [1 if n > 0 else 0 for n in pml_layers]
        self.pml_vel, self.pml_alpha, self.pml_beta = pml_vel, pml_alpha, pml_beta
        self.free_top = free_top
        self.src = src
        self.recv_idx, self.recv_w = recv_idx, recv_w

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

    def algorithmic_bytes_per_step(self):
        """Minimum HBM traffic of one RK4 step of the isotropic path (FP32):
        768 B per point, 576 B per PML aux point, 352 B per surface column."""
        nsurf = self.ni * self.nj if self.free_top else 0
        return 768.0 * self.npoints + 576.0 * self.n_aux() + 352.0 * nsurf

    def build_solver(self, variant=4):
        L = _setup_sigs()
        s = Solver(self.ni, self.nj, self.nk, self.dt, self.nt_total, free_top=self.free_top,
                   pml_on=self.pml_on, pml_layers=self.pml_layers)
        s.set_kernel_variant(variant)
        check(L.cgfdh_compute_metric(s.h, _p(self.xyz)), "compute_metric")
        s.set_medium_iso(self.lam, self.mu, self.slw)
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


def iso_topo_grf(ni=600, nj=600, nk=300, dh=50.0, nt_total=2001, dt=0.004, seed=2024,
                 nhills=20, pml=20, n_side=16):
    """Isotropic elastic, ni x nj x nk curvilinear grid over a sum of seeded
    Gaussian hills (amplitude <= 1 km, sigma 1-5 km), 3-layer model times a
    Gaussian random field (5 % std, 500 m correlation), double-couple source
    (seeded strike/dip/rake) with a 2 Hz Ricker, CFS-PML on 5 faces, traction
    free surface, n_side^2 surface receivers."""
    top = topography(ni, nj, dh, nhills, seed)
    depth = (nk - 1) * dh
    xyz = layered_coords(ni, nj, nk, dh, top, depth)
    grf = gaussian_random_field((nk + 6, nj + 6, ni + 6), dh, 500.0, 0.05, seed + 1)
    layers = [(1e9, 4500.0, 2600.0, 2400.0), (-1500.0, 6000.0, 3464.0, 2670.0),
              (-6000.0, 6800.0, 3900.0, 2900.0)]
    lam, mu, slw, vpmax = layered_medium(xyz[2], layers, grf)
    del grf
    rng = np.random.default_rng(seed + 2)
    strike, dip, rake = rng.uniform(0, 360), rng.uniform(10, 90), rng.uniform(-180, 180)
    L = _setup_sigs()
    m6 = np.zeros(6, np.float32)
    L.cgfdh_moment_from_angles(strike, dip, rake, 3.0e10, 1.0, 1.0e6, _p(m6))
    si, sj = 3 + ni // 2, 3 + nj // 2
    sk = 3 + nk - 1 - min(int(round(5000.0 / dh)), int(0.4 * nk))
    src = moment_source(ni, nj, nk, si, sj, sk, dt, 0.0, m6, 2.0, 0.5, 2.0)
    ridx, rw = surface_receiver_grid(ni, nj, nk, n_side)
    return Workload(f"iso{ni}x{nj}x{nk}", ni, nj, nk, dt, nt_total, xyz, lam, mu, slw,
                    [pml, pml, pml, pml, pml, 0], 7000.0, 3.14, 2.0, True, src, ridx, rw)
