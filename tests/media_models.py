"""Layered models for the media tests: .md3lay files of every media type
(interfaces with topography, per-layer values, gradients and power laws),
a curvilinear grid that crosses two interfaces with points exactly on one,
and the oracle's grid-level discretisation (oracle/media_oracle.c)."""
import ctypes as C

import numpy as np

from oracle import pyoracle as po
from tests import cases as cs

# targets and methods of cgfdh_layer2model_grid / orc_layer2model_grid
ONECMP, EL_ISO, EL_VTI, EL_ANISO = 0, 1, 2, 3
METHODS = {"loc": 0, "har": 1, "ari": 2}
NOUT = {ONECMP: 1, EL_ISO: 3, EL_VTI: 6, EL_ANISO: 22}
REF_FN = {ONECMP: "media_layer2model_onecmp", EL_ISO: "media_layer2model_el_iso",
          EL_VTI: "media_layer2model_el_vti", EL_ANISO: "media_layer2model_el_aniso"}

# (per-layer base values, gradient per metre) of each property
RHO = ([2200.0, 2500.0, 2800.0], 0.2)
VP = ([3000.0, 4500.0, 5600.0], 0.5)
VS = ([1700.0, 2500.0, 3200.0], 0.28)
AZIMUTH = ([30.0, 45.0, 60.0], 1e-3)
DIP = ([20.0, 10.0, 35.0], 1e-3)
C11 = ([2.6e10, 5.5e10, 9.0e10], 1.2e6)
C33 = ([2.3e10, 5.0e10, 8.5e10], 1.0e6)
C55 = ([6.4e9, 1.5e10, 2.8e10], 3e5)
C66 = ([7.1e9, 1.7e10, 3.0e10], 3e5)
C13 = ([8.0e9, 2.0e10, 3.0e10], 4e5)


def _small(a, b, c):
    return ([a, b, c], 1e-5)


def _full_stiffness():
    """21 coefficients per layer of a positive definite stiffness with every
    entry non-zero."""
    rng = np.random.default_rng(99)
    mats = []
    for L in range(3):
        A = rng.standard_normal((6, 6)) * 2e9
        mats.append(A @ A.T + np.eye(6) * (2e10 + 2e10 * L))
    return [([float(m[i, j]) for m in mats], 1e-5 * abs(float(mats[0][i, j])))
            for i in range(6) for j in range(i, 6)]


THOMSEN = [RHO, VP, VS, _small(0.1, 0.15, 0.05), _small(0.05, 0.08, 0.02),
           _small(0.08, 0.1, 0.04)]
KINDS = {
    "one_component": (0, [RHO]),
    "elastic_isotropic": (21, [RHO, VP, VS]),
    "elastic_vti_prem": (22, [RHO, ([3100.0, 4650.0, 5800.0], 0.5), VP,
                              ([1800.0, 2600.0, 3300.0], 0.28), VS, _small(0.95, 0.9, 0.97)]),
    "elastic_vti_thomsen": (23, THOMSEN),
    "elastic_vti_cij": (24, [RHO, C11, C33, C55, C66, C13]),
    "elastic_tti_thomsen": (25, THOMSEN + [AZIMUTH, DIP]),
    "elastic_tti_bond": (26, [RHO, C11, C33, C55, C66, C13, AZIMUTH, DIP]),
    "elastic_aniso_cij": (27, [RHO] + _full_stiffness()),
}


def model(kind, seed=0, mx=31, my=27, dx=50.0, dy=60.0, x0=-200.0, y0=-300.0):
    """Three interfaces (the top with topography) over an mx x my mesh:
    (type code, fields [1+3 nprop][3][my][mx], geo = minx, miny, dx, dy)."""
    code, props = KINDS[kind]
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(x0 + dx * np.arange(mx), y0 + dy * np.arange(my))
    top = 150.0 * np.exp(-((X - X.mean()) ** 2 + (Y - Y.mean()) ** 2) / (2 * 400.0 ** 2))
    elev = np.stack([top + 5.0 * rng.standard_normal(X.shape), -600.0 + 0.05 * X,
                     -1400.0 + 0.02 * Y])
    f = np.zeros((1 + 3 * len(props), 3, my, mx), np.float32)
    f[0] = elev
    for q, (base, grad) in enumerate(props):
        for L in range(3):
            f[1 + 3 * q, L] = base[L] * (1 + 0.02 * rng.standard_normal(X.shape))
            f[2 + 3 * q, L] = grad * (1 + 0.05 * rng.standard_normal(X.shape))
            f[3 + 3 * q, L] = 1.0 + 0.04 * L + 0.01 * rng.standard_normal(X.shape)
    return code, f, (np.float32(x0), np.float32(y0), np.float32(dx), np.float32(dy))


def write_md3lay(path, kind, f, geo):
    nf, ni, my, mx = f.shape
    with open(path, "w") as fp:
        fp.write("# layered model written by the tests\n")
        fp.write(f"{kind}\n{ni}\n{mx} {my} {geo[0]:.9g} {geo[1]:.9g} {geo[2]:.9g} {geo[3]:.9g}\n")
        fp.write("# samples\n")
        flat = f.reshape(nf, -1)
        for p in range(flat.shape[1]):
            fp.write(" ".join(f"{v:.9g}" for v in flat[:, p]) + "\n")


def grid_points(f, geo, ni=18, nj=15, nk=16, seed=1):
    """A curvilinear grid inside the mesh from 1700 m deep to above the top,
    with one row of points exactly on the middle interface's samples:
    (3, nk, nj, ni)."""
    co = cs.curvilinear_coords(ni, nj, nk, dh=40.0, seed=seed, hill=0.2, wiggle=0.1)
    co[0] += float(geo[0]) + 150.0
    co[1] += float(geo[1]) + 150.0
    z = co[2]
    co[2] = -1700.0 + (z - z.min()) / (z.max() - z.min()) * 2000.0
    for n in range(6):
        i, j = 4 + n, 3 + n
        co[0, 5, 4, 5 + n] = geo[0] + geo[2] * i
        co[1, 5, 4, 5 + n] = geo[1] + geo[3] * j
        co[2, 5, 4, 5 + n] = f[0, 1, j, i]
    return co.astype(np.float32)


def oracle_grid(code, target, method, f, geo, co):
    """orc_layer2model_grid: (rc, the target's arrays of the grid's shape)."""
    lib = po.load_oracle()
    nz, ny, nx = co.shape[1:]
    xyz = [np.ascontiguousarray(co[d], np.float32) for d in range(3)]
    outs = [np.zeros(xyz[0].shape, np.float32) for _ in range(NOUT[target])]
    arr = (C.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
    dims = np.array([f.shape[1], f.shape[3], f.shape[2]], np.int32)
    g = np.array(geo, np.float32)
    fc = np.ascontiguousarray(f, np.float32)
    rc = lib.orc_layer2model_grid(code, target, METHODS[method], po.ptr(dims), po.ptr(g),
                                  po.ptr(fc), nx, ny, nz, *[po.ptr(a) for a in xyz],
                                  C.cast(arr, C.c_void_p))
    return rc, outs


# ---- grid models (.md3grd): surfaces grouped into layers -----------------

def grid_model(kind, seed=0, ngz=(3, 4, 3), mx=31, my=27, dx=50.0, dy=60.0, x0=-200.0,
               y0=-300.0):
    """A grid model of len(ngz) layers: each layer's top follows the
    layered model's interfaces and its last surface is the next layer's top
    (the reference requires the two to coincide), the surfaces between
    evenly spaced, the last layer's stepping down 80 m; values per surface
    from the layer's base plus a depth gradient.  Returns (type code, ngz,
    fields [1+nprop][sum ngz][my][mx], geo of the full mesh)."""
    code, props = KINDS[kind]
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(x0 + dx * np.arange(mx), y0 + dy * np.arange(my))
    top = 150.0 * np.exp(-((X - X.mean()) ** 2 + (Y - Y.mean()) ** 2) / (2 * 400.0 ** 2))
    tops = [top + 5.0 * rng.standard_normal(X.shape), -600.0 + 0.05 * X, -1400.0 + 0.02 * Y]
    ns = sum(ngz)
    f = np.zeros((1 + len(props), ns, my, mx), np.float32)
    s = 0
    for L, g in enumerate(ngz):
        for q in range(g):
            if L + 1 < len(ngz):
                elev = tops[L] + (tops[L + 1] - tops[L]) * (q / (g - 1) if g > 1 else 0.0)
                if q == g - 1:
                    elev = tops[L + 1]
            else:
                elev = tops[L] - 80.0 * q
            f[0, s] = elev
            depth = tops[L] - elev
            for p, (base, grad) in enumerate(props):
                f[1 + p, s] = (base[L % 3] * (1 + 0.02 * rng.standard_normal(X.shape)) +
                               grad * depth)
            s += 1
    return code, list(ngz), f, (np.float32(x0), np.float32(y0), np.float32(dx), np.float32(dy))


def write_md3grd(path, kind, ngz, f, geo):
    """The file read_grid_file reads: type, layers, surfaces per layer, the
    mesh, then per surface per sample the elevation and the values (no
    comment lines: the reference's reader takes none)."""
    nf, ns, my, mx = f.shape
    with open(path, "w") as fp:
        fp.write(f"{kind}\n{len(ngz)}\n{' '.join(str(g) for g in ngz)}\n")
        fp.write(f"{mx} {my} {geo[0]:.9g} {geo[1]:.9g} {geo[2]:.9g} {geo[3]:.9g}\n")
        flat = f.reshape(nf, -1)
        for p in range(flat.shape[1]):
            fp.write(" ".join(f"{v:.9g}" for v in flat[:, p]) + "\n")


def grid_points_on(surface, geo, ni=18, nj=15, nk=16, seed=1):
    """grid_points with the row of points placed on the given surface's
    samples (a surface that begins a layer puts them on a layer boundary)."""
    co = cs.curvilinear_coords(ni, nj, nk, dh=40.0, seed=seed, hill=0.2, wiggle=0.1)
    co[0] += float(geo[0]) + 150.0
    co[1] += float(geo[1]) + 150.0
    z = co[2]
    co[2] = -1700.0 + (z - z.min()) / (z.max() - z.min()) * 2000.0
    for n in range(6):
        i, j = 4 + n, 3 + n
        co[0, 5, 4, 5 + n] = geo[0] + geo[2] * i
        co[1, 5, 4, 5 + n] = geo[1] + geo[3] * j
        co[2, 5, 4, 5 + n] = surface[j, i]
    return co.astype(np.float32)


def extent(co):
    """The grid's x and y extent as the main program passes it."""
    return (float(co[0].min()), float(co[0].max()), float(co[1].min()), float(co[1].max()))


def oracle_grid2model(code, target, method, ngz, f, geo, co):
    """orc_grid2model_grid on a cropped mesh (fields, geo as
    cgfdh_io_read_md3grd gives them): (rc, the target's arrays)."""
    lib = po.load_oracle()
    nz, ny, nx = co.shape[1:]
    xyz = [np.ascontiguousarray(co[d], np.float32) for d in range(3)]
    outs = [np.zeros(xyz[0].shape, np.float32) for _ in range(NOUT[target])]
    arr = (C.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
    dims = np.array([f.shape[1], f.shape[3], f.shape[2]], np.int32)
    g = np.array(geo, np.float32)
    gz = np.array(ngz, np.int32)
    fc = np.ascontiguousarray(f, np.float32)
    rc = lib.orc_grid2model_grid(code, target, METHODS[method], po.ptr(dims), po.ptr(g), len(ngz),
                                 po.ptr(gz), po.ptr(fc), nx, ny, nz, *[po.ptr(a) for a in xyz],
                                 C.cast(arr, C.c_void_p))
    return rc, outs


# ---- binary volumes (media_bin2model) --------------------------------------

def bin_model(order=(2, 0, 1), seed=0, n=(40, 33, 29), spacing=(50.0, 60.0, -70.0),
              origin=(-200.0, -300.0, 300.0)):
    """Smooth random volumes rho, vp, vs (and Qp, Qs) on a regular mesh of
    n = (nx, ny, nz) samples, spacing and origin per direction (z may
    decrease), stored with axis a holding direction order[a], the first
    axis fastest.  Returns (vols (5, nz, ny, nx) in x-fastest layout, the
    stored arrays (5, ...) to write, size, spacing and origin per stored
    axis)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    base = np.array([2500.0, 4500.0, 2500.0, 80.0, 40.0])
    amp = np.array([200.0, 600.0, 300.0, 20.0, 10.0])
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    vols = np.zeros((5, nz, ny, nx), np.float32)
    for q in range(5):
        ph = rng.uniform(0, 2 * np.pi, 3)
        vols[q] = base[q] + amp[q] * (np.sin(0.21 * xx + ph[0]) * np.cos(0.17 * yy + ph[1]) +
                                      0.5 * np.sin(0.13 * zz + ph[2]))
    vols[1] = np.maximum(vols[1], 1.6 * vols[2])        # vp > sqrt 2 vs
    # stored layout: axis a is direction order[a]; numpy's last axis fastest
    perm = [2 - order[a] for a in (2, 1, 0)]             # (nz, ny, nx) axes -> stored
    stored = np.stack([np.ascontiguousarray(np.transpose(v, perm)) for v in vols])
    size = [n[order[a]] for a in range(3)]
    sp = [spacing[order[a]] for a in range(3)]
    org = [origin[order[a]] for a in range(3)]
    return vols, stored, size, sp, org


def write_bins(tmp_path, stored, names=("rho", "vp", "vs", "qp", "qs")):
    paths = []
    for v, nm in zip(stored, names):
        p = str(tmp_path / f"{nm}.bin")
        np.ascontiguousarray(v, np.float32).tofile(p)
        paths.append(p)
    return paths


def oracle_bin(order, size, spacing, origin, extent, vols_window, co, nvol):
    """orc_bin_window + orc_bin2model_grid: (rc, outputs)."""
    lib = po.load_oracle()
    x, y, z = (np.ascontiguousarray(co[d], np.float32).reshape(-1) for d in range(3))
    start = np.zeros(3, np.int32); end = np.zeros(3, np.int32); n3 = np.zeros(3, np.int32)
    xv = np.zeros(4096, np.float32); yv = np.zeros(4096, np.float32); zv = np.zeros(4096, np.float32)
    o, s, sp, org = (np.ascontiguousarray(a, t) for a, t in ((order, np.int32), (size, np.int32),
                                                             (spacing, np.float32),
                                                             (origin, np.float32)))
    rc = lib.orc_bin_window(po.ptr(o), po.ptr(s), po.ptr(sp), po.ptr(org), *map(float, extent),
                            float(z.min()), float(z.max()), po.ptr(start), po.ptr(end),
                            po.ptr(xv), po.ptr(yv), po.ptr(zv), po.ptr(n3))
    if rc:
        return rc, None, None
    outs = [np.zeros(x.shape, np.float32) for _ in range(nvol)]
    arr = (C.c_void_p * nvol)(*[a.ctypes.data for a in outs])
    vw = np.ascontiguousarray(vols_window(start, end), np.float32)
    rc = lib.orc_bin2model_grid(nvol, po.ptr(n3), po.ptr(xv), po.ptr(yv), po.ptr(zv), po.ptr(vw),
                                x.size, po.ptr(x), po.ptr(y), po.ptr(z), C.cast(arr, C.c_void_p))
    return rc, [a.reshape(co[0].shape) for a in outs], (start, end, xv[:n3[0]], yv[:n3[1]],
                                                        zv[:n3[2]])
