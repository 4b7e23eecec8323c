"""Training and evaluation data of the seismogram surrogate, produced by the
batched forward solver on one fixed grid and medium.

Grid: 160 x 160 x 100 curvilinear, 100 m horizontal spacing, a flat bottom
9.9 km below datum, seeded Gaussian-hill topography (6 hills, <= 600 m), the
3-layer model times a seeded Gaussian random field (5 % std, 500 m
correlation), CFS-PML 10 layers on 5 faces, traction-image free surface,
dt = 8 ms, 1024 recorded steps.  Receivers: an 8 x 8 grid of surface nodes.

Sources: point moment tensors at grid nodes of the interior (x, y at least
20 nodes from the sides), depth 1-8 km below the local surface, strike /
dip / rake uniform in [0, 360) / [0, 90] / [-180, 180), Ricker f0 uniform in
1-3 Hz with its peak at 1.2 / f0.  Target of one source: the 3 velocity
components at the 64 receivers over 1024 steps.
"""
import sys
import time

import numpy as np

from . import workloads as wk
from .solver import Batch

NI, NJ, NK, DH = 160, 160, 100, 100.0
NT, DT = 1024, 0.008
N_SIDE = 8
PML = 10
MARGIN = 20
RANGES = dict(depth=(1000.0, 8000.0), strike=(0.0, 360.0), dip=(0.0, 90.0),
              rake=(-180.0, 180.0), f0=(1.0, 3.0))


class SurrogateGrid:
    """The fixed grid, medium and receivers, built once."""

    def __init__(self, seed=77, ni=NI, nj=NJ, nk=NK, dh=DH, nt=NT, dt=DT, n_side=N_SIDE,
                 pml=PML):
        self.ni, self.nj, self.nk, self.dh, self.nt, self.dt = ni, nj, nk, dh, nt, dt
        top = wk.topography(ni, nj, dh, 6, seed, amp_max=600.0, sig_min=1500.0, sig_max=4000.0)
        self.top = top
        self.depth = (nk - 1) * dh
        xyz = wk.layered_coords(ni, nj, nk, dh, top, self.depth)
        grf = wk.gaussian_random_field((nk + 6, nj + 6, ni + 6), dh, 500.0, 0.05, seed + 1)
        layers = [(1e9, 4500.0, 2600.0, 2400.0), (-1500.0, 6000.0, 3464.0, 2670.0),
                  (-6000.0, 6800.0, 3900.0, 2900.0)]
        lam, mu, slw, vpmax = wk.layered_medium(xyz[2], layers, grf)
        ridx, rw = wk.surface_receiver_grid(ni, nj, nk, n_side)
        self.nrecv = n_side * n_side
        self.wl = wk.Workload(f"surrogate{ni}x{nj}x{nk}", ni, nj, nk, dt, nt, xyz, lam, mu, slw,
                              [pml] * 5 + [0], 7000.0, 3.14, 2.0, True, None, ridx, rw)
        self.x_range = ((3 + MARGIN - 3) * dh, (ni - 1 - MARGIN) * dh)
        self.y_range = ((3 + MARGIN - 3) * dh, (nj - 1 - MARGIN) * dh)
        # sources stay 3 nodes above the bottom PML
        self.kmin = 3 + pml + 3
        self.depth_range = (RANGES["depth"][0],
                            min(RANGES["depth"][1], (nk - 1 - (self.kmin - 3)) * dh * 0.95))

    def sample_sources(self, n, seed):
        """n seeded sources: rows (x, y, depth, strike, dip, rake, f0) with x,
        y, depth those of the grid node the source sits on, and its (i,j,k)."""
        rng = np.random.default_rng(seed)
        xyz = self.wl.xyz.reshape(3, self.nk + 6, self.nj + 6, self.ni + 6)
        params = np.zeros((n, 7), np.float64)
        ijk = np.zeros((n, 3), np.int32)
        ksurf = 3 + self.nk - 1
        for s in range(n):
            i = int(rng.integers(3 + MARGIN, 3 + self.ni - MARGIN))
            j = int(rng.integers(3 + MARGIN, 3 + self.nj - MARGIN))
            want = rng.uniform(*self.depth_range)
            col = xyz[2, :, j, i]
            k = int(np.argmin(np.abs((col[ksurf] - col[self.kmin:ksurf + 1]) - want))) + self.kmin
            params[s] = (xyz[0, k, j, i], xyz[1, k, j, i], col[ksurf] - col[k],
                         rng.uniform(*RANGES["strike"]), rng.uniform(*RANGES["dip"]),
                         rng.uniform(*RANGES["rake"]), rng.uniform(*RANGES["f0"]))
            ijk[s] = (i, j, k)
        return params, ijk

    def source_tables(self, p, ijk):
        """Moment-source tables of one source (Ricker f0, peak at 1.2 / f0)."""
        L = wk._setup_sigs()
        m6 = np.zeros(6, np.float32)
        L.cgfdh_moment_from_angles(p[3], p[4], p[5], 3.0e10, 1.0, 1.0e6, wk._p(m6))
        f0 = float(p[6])
        tpeak = 1.2 / f0
        return wk.moment_source(self.ni, self.nj, self.nk, int(ijk[0]), int(ijk[1]), int(ijk[2]),
                                self.dt, 0.0, m6, f0, tpeak, 2.0 * tpeak)


def moment_tensors(params):
    """[n, 6] moment tensors of the sources' mechanisms, exactly as the FD
    source tables take them (cgfdh_moment_from_angles), divided by the
    scalar moment."""
    L = wk._setup_sigs()
    out = np.zeros((len(params), 6), np.float32)
    m6 = np.zeros(6, np.float32)
    m0 = 3.0e10 * 1.0 * 1.0e6
    for n, p in enumerate(params):
        L.cgfdh_moment_from_angles(p[3], p[4], p[5], 3.0e10, 1.0, 1.0e6, wk._p(m6))
        out[n] = m6 / m0
    return out


def features(params, grid, mechanism=True):
    """Network input rows (float32, 32 wide).  mechanism=True: the 7
    parameters scaled to [-1, 1], the unit moment tensor of the mechanism (6
    values) and the sines and cosines of strike, dip and rake; zero padded.
    mechanism=False (a moment-tensor network, which takes the mechanism
    through its output contraction): x, y, depth and f0 scaled to [-1, 1]
    and their sines and cosines at 1, 2 and 4 half-periods over the range
    (28 values), which lets the MLP represent arrival times that move with
    the source."""
    p = np.asarray(params, np.float64)
    n = p.shape[0]
    x = np.zeros((n, 32), np.float32)

    def sc(v, lo, hi):
        return 2.0 * (v - lo) / (hi - lo) - 1.0
    x[:, 0] = sc(p[:, 0], *grid.x_range)
    x[:, 1] = sc(p[:, 1], *grid.y_range)
    x[:, 2] = sc(p[:, 2], *grid.depth_range)
    x[:, 3] = sc(p[:, 3], *RANGES["strike"])
    x[:, 4] = sc(p[:, 4], *RANGES["dip"])
    x[:, 5] = sc(p[:, 5], *RANGES["rake"])
    x[:, 6] = sc(p[:, 6], *RANGES["f0"])
    s, d, r = np.radians(p[:, 3]), np.radians(p[:, 4]), np.radians(p[:, 5])
    # double-couple moment tensor (Aki & Richards, x north / y east / z down)
    mxx = -(np.sin(d) * np.cos(r) * np.sin(2 * s) + np.sin(2 * d) * np.sin(r) * np.sin(s) ** 2)
    mxy = np.sin(d) * np.cos(r) * np.cos(2 * s) + 0.5 * np.sin(2 * d) * np.sin(r) * np.sin(2 * s)
    mxz = -(np.cos(d) * np.cos(r) * np.cos(s) + np.cos(2 * d) * np.sin(r) * np.sin(s))
    myy = np.sin(d) * np.cos(r) * np.sin(2 * s) - np.sin(2 * d) * np.sin(r) * np.cos(s) ** 2
    myz = -(np.cos(d) * np.cos(r) * np.sin(s) - np.cos(2 * d) * np.sin(r) * np.cos(s))
    mzz = np.sin(2 * d) * np.sin(r)
    x[:, 7:13] = np.stack([mxx, myy, mzz, myz, mxz, mxy], axis=1)
    x[:, 13:19] = np.stack([np.sin(s), np.cos(s), np.sin(d), np.cos(d), np.sin(r), np.cos(r)],
                           axis=1)
    if not mechanism:
        base = np.stack([x[:, 0], x[:, 1], x[:, 2], x[:, 6]], axis=1).astype(np.float64)
        enc = [base]
        for f in (1.0, 2.0, 4.0):
            enc += [np.sin(np.pi * f * base), np.cos(np.pi * f * base)]
        x = np.zeros((n, 32), np.float32)
        x[:, :28] = np.concatenate(enc, axis=1)
    return x


def receiver_xy(grid):
    """[nrecv, 2] horizontal coordinates of the surface receivers."""
    nx, ny = grid.ni + 6, grid.nj + 6
    xyz = grid.wl.xyz.reshape(3, grid.nk + 6, ny, nx)
    out = np.zeros((grid.nrecv, 2))
    for r, idx in enumerate(grid.wl.recv_idx):
        p = int(idx[0])
        i, j, k = p % nx, (p // nx) % ny, p // (nx * ny)
        out[r] = xyz[0, k, j, i], xyz[1, k, j, i]
    return out


def receiver_features(params, grid):
    """Input rows of the per-receiver network, one per (source, receiver):
    source x, y, depth, f0 and receiver x, y scaled to [-1, 1]; the offset
    (receiver minus source) and its length scaled by the grid's extent; and
    sines and cosines of the offsets and depth at 1, 2 and 4 half-periods
    over their range.  [n * nrecv, 32] float32, receivers fastest."""
    p = np.asarray(params, np.float64)
    n = len(p)
    rxy = receiver_xy(grid)
    nr = len(rxy)
    Lx = (grid.ni - 1) * grid.dh
    Ly = (grid.nj - 1) * grid.dh

    def sc(v, lo, hi):
        return 2.0 * (v - lo) / (hi - lo) - 1.0
    xs = np.repeat(p[:, 0], nr)
    ys = np.repeat(p[:, 1], nr)
    dep = np.repeat(sc(p[:, 2], *grid.depth_range), nr)
    f0 = np.repeat(sc(p[:, 6], *RANGES["f0"]), nr)
    xr = np.tile(rxy[:, 0], n)
    yr = np.tile(rxy[:, 1], n)
    dx, dy = (xr - xs) / Lx, (yr - ys) / Ly
    dist = np.sqrt(dx * dx + dy * dy)
    base = [sc(xs, *grid.x_range), sc(ys, *grid.y_range), dep, f0, sc(xr, 0, Lx),
            sc(yr, 0, Ly), dx, dy, dist]
    enc = []
    for f in (1.0, 2.0, 4.0):
        for v in (dx, dy, dep):
            enc += [np.sin(np.pi * f * v), np.cos(np.pi * f * v)]
    x = np.zeros((n * nr, 32), np.float32)
    x[:, :27] = np.stack(base + enc, axis=1)
    return x


def arrival_shifts(params, grid, vmax=8500.0, relief=600.0):
    """Per (source, receiver) row, receivers fastest: a lower bound of the
    first arrival in samples, floor(distance / vmax / dt), the distance
    bounded below by the horizontal offset and the source depth less the
    topography's relief (so the bound never passes the true P arrival of a
    medium slower than vmax).  [n * nrecv] int."""
    p = np.asarray(params, np.float64)
    rxy = receiver_xy(grid)
    dx = rxy[None, :, 0] - p[:, None, 0]
    dy = rxy[None, :, 1] - p[:, None, 1]
    dz = np.maximum(p[:, None, 2] - relief, 0.0)
    dist = np.sqrt(dx * dx + dy * dy + dz * dz)
    s = np.floor(dist / vmax / grid.dt).astype(np.int64)
    return np.clip(s, 0, grid.nt - 1).reshape(-1)


def shift_rows(rows, shifts, back=False):
    """rows [m, nt, c] moved earlier by shifts[m] samples (zeros after the
    end), or with back later by them (zeros before the start): the
    arrival-aligned form and its inverse."""
    out = np.zeros_like(rows)
    nt = rows.shape[1]
    for m, s in enumerate(shifts):
        s = int(s)
        if back:
            out[m, s:] = rows[m, :nt - s]
        else:
            out[m, :nt - s] = rows[m, s:]
    return out


def to_receiver_layout(seis):
    """[n, nrecv, 3, nt] -> [n * nrecv, nt, 3] (one row block per trace set)."""
    n, nr, nc, nt = seis.shape
    return np.ascontiguousarray(seis.transpose(0, 1, 3, 2).reshape(n * nr, nt, nc))


def to_target_layout(seis):
    """[n, nrecv, 3, nt] solver seismograms -> [n, nt, 3 * nrecv] (the
    network's channels-last output: channel = component * nrecv + receiver)."""
    n, nr, nc, nt = seis.shape
    return np.ascontiguousarray(seis.transpose(0, 3, 2, 1).reshape(n, nt, nc * nr))


def from_target_layout(y, nrecv):
    """[n, nt, 3 * nrecv] -> [n, nrecv, 3, nt]."""
    n, nt, c = y.shape
    return np.ascontiguousarray(y.reshape(n, nt, c // nrecv, nrecv).transpose(0, 3, 2, 1))


class FdGenerator:
    """Runs batches of sources on the fixed grid with the batched solver.
    Solvers are built once per batch slot and reused: before each batch the
    wavefield and PML memory are reset and the sources replaced."""

    def __init__(self, grid, batch, variant=None, autotune=True, progress=False):
        self.grid, self.batch, self.variant, self.autotune = grid, batch, variant, autotune
        self.progress = progress
        self.solvers = []
        self._zero_w = None
        self._zero_aux = None
        self.fd_seconds = 0.0
        self.runs = 0

    def _slots(self, n):
        while len(self.solvers) < n:
            self.solvers.append(self.grid.wl.build_solver(variant=self.variant))
        return self.solvers[:n]

    def run(self, params, ijk):
        """Seismograms [n, nrecv, 3, nt] of the given sources."""
        g = self.grid
        out = np.zeros((len(params), g.nrecv, 3, g.nt), np.float32)
        for c0 in range(0, len(params), self.batch):
            idx = range(c0, min(len(params), c0 + self.batch))
            sv = self._slots(len(idx))
            for s, q in zip(sv, idx):
                if self._zero_w is None:
                    self._zero_w = np.zeros(s.ncomp * s.nx * s.ny * s.nz, np.float32)
                    self._zero_aux = np.zeros(s.aux_len(), np.float32)
                s.set_wavefield(self._zero_w)
                if self._zero_aux.size:
                    s.set_pml_aux(self._zero_aux)
                s.set_moment_sources(**g.source_tables(params[q], ijk[q]))
            b = Batch(sv)
            if self.autotune and c0 == 0:
                b.autotune(0)
                for s in sv:
                    s.set_wavefield(self._zero_w)
                    if self._zero_aux.size:
                        s.set_pml_aux(self._zero_aux)
            t = time.perf_counter()
            b.run(0, g.nt)
            self.fd_seconds += b.last_ms() / 1e3
            for s, q in zip(sv, idx):
                out[q] = s.seismograms()[:, :3, : g.nt]
            b.close()
            self.runs += len(idx)
            if self.progress:
                print(f"FD runs {idx.stop}/{len(params)}: {time.perf_counter() - t:.1f} s for "
                      f"{len(idx)} sources", file=sys.stderr, flush=True)
        return out

    def close(self):
        for s in self.solvers:
            s.close()
        self.solvers = []


def compress_runs(seis, dt, fmax=8.0):
    """Band-limited storage of FD seismograms [n, nrecv, 3, nt]: the real
    FFT of every trace up to fmax Hz as float16 (real and imaginary parts).
    The sources' Ricker wavelets (f0 <= 3 Hz) carry a fraction of 1e-4 of
    their energy above 8 Hz."""
    n, nr, nc, nt = seis.shape
    nb = int(fmax * nt * dt) + 1
    F = np.fft.rfft(seis.astype(np.float64), axis=3)[..., :nb]
    return np.stack([F.real, F.imag], axis=-1).astype(np.float16), nt


def decompress_runs(z, nt):
    """Inverse of compress_runs: [n, nrecv, 3, nt] float32."""
    F = z[..., 0].astype(np.float64) + 1j * z[..., 1].astype(np.float64)
    full = np.zeros(F.shape[:-1] + (nt // 2 + 1,), np.complex128)
    full[..., :F.shape[-1]] = F
    return np.fft.irfft(full, n=nt, axis=3).astype(np.float32)


def relative_misfit(pred, ref):
    """Per-source relative L2 misfit |pred - ref| / |ref| over all traces."""
    n = ref.shape[0]
    d = (pred.reshape(n, -1).astype(np.float64) - ref.reshape(n, -1))
    return np.sqrt((d ** 2).sum(axis=1) / (ref.reshape(n, -1).astype(np.float64) ** 2).sum(axis=1))
