"""Python view of one engine instance (include/cgfdh.h): upload the static
fields, advance steps, read back seismograms and state.

The arrays use the reference layouts: grids with 3 ghost points per side,
x fastest; wavefield components Vx,Vy,Vz,Txx,Tyy,Tzz,Tyz,Txz,Txy.
"""
import ctypes as C

import numpy as np

from ._lib import CgfdhDesc, check, lib

MEDIUM_ELASTIC_ISO = 2
MEDIUM_ELASTIC_VTI = 3
MEDIUM_ELASTIC_ANISO = 4
MEDIUM_VISCO_ISO = 5
SRC_POINT = 1
SRC_GAUSSIAN = 2


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def estimate_dt(xyz, medium_type, vols, cfl, nmaxwell=0):
    """Stable time step of a curvilinear grid on the GPU (blk_dt_esti_curv,
    blk_t.c:2840-2930).  xyz: (3, nz, ny, nx) coordinates with ghosts; vols:
    the medium volumes cgfdh_estimate_dt takes for medium_type.  Returns
    (dtmax, vp, L, (i, j, k))."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    vols = np.ascontiguousarray(vols, dtype=np.float32)
    nz, ny, nx = xyz.shape[-3:]
    out = np.zeros(3, np.float32)
    ijk = np.zeros(3, np.int32)
    check(lib().cgfdh_estimate_dt(nx - 6, ny - 6, nz - 6, _p(xyz), int(medium_type), _p(vols),
                                  int(nmaxwell), float(cfl), _p(out[0:]), _p(out[1:]),
                                  _p(out[2:]), _p(ijk)), "estimate_dt")
    return float(out[0]), float(out[1]), float(out[2]), tuple(int(v) for v in ijk)


def locate_points(xyz, pts):
    """Point location on the GPU (cgfdh_locate_points): xyz (3, nz, ny, nx)
    coordinates with ghosts, pts (n, 3).  Returns (is_here[n], ijk[n, 3],
    inc[n, 3])."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
    nz, ny, nx = xyz.shape[-3:]
    n = len(pts)
    here = np.zeros(n, np.int32)
    ijk = np.zeros((n, 3), np.int32)
    inc = np.zeros((n, 3), np.float32)
    check(lib().cgfdh_locate_points(nx - 6, ny - 6, nz - 6, _p(xyz), n, _p(pts), _p(here),
                                    _p(ijk), _p(inc)), "locate_points")
    return here, ijk, inc


def depth_to_axis(xyz, xyd):
    """Depth below the top surface to z on the GPU (cgfdh_depth_to_axis):
    xyd (n, 3) = x, y, depth.  Returns (found[n], z[n])."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    xyd = np.ascontiguousarray(xyd, dtype=np.float32).reshape(-1, 3)
    nz, ny, nx = xyz.shape[-3:]
    n = len(xyd)
    found = np.zeros(n, np.int32)
    z = np.zeros(n, np.float32)
    check(lib().cgfdh_depth_to_axis(nx - 6, ny - 6, nz - 6, _p(xyz), n, _p(xyd), _p(found),
                                    _p(z)), "depth_to_axis")
    return found, z


def gen_grid_layer(ncell, equal, layer_xyz, resample, start, ni, nj, nk):
    """A one-block grid from interface samples on the GPU
    (cgfdh_gen_grid_layer): layer_xyz (3, n_if, ny_l, nx_l).  Returns the
    coordinates (3, nk+6, nj+6, ni+6)."""
    L = np.ascontiguousarray(layer_xyz, np.float32)
    n_if, ny_l, nx_l = L.shape[1:]
    nc = np.ascontiguousarray(ncell, np.int32)
    eq = np.ascontiguousarray(equal, np.int32)
    rs = np.ascontiguousarray(resample, np.int32)
    st = np.ascontiguousarray(start, np.int32)
    out = np.zeros((3, nk + 6, nj + 6, ni + 6), np.float32)
    check(lib().cgfdh_gen_grid_layer(n_if, _p(nc), _p(eq), nx_l, ny_l, _p(L), _p(rs), _p(st),
                                     ni, nj, nk, ni, nj, nk, 0, 0, 0, _p(out)), "gen_grid_layer")
    return out


def layer2model(nprop, fields, geo, xyz):
    """A layered model onto grid points on the GPU with local values
    (cgfdh_layer2model): fields [1+3 nprop][n_if][ny][nx] and geo from
    output.read_md3lay, xyz (3, ...) the points.  Returns [var] or
    [lambda, mu, rho], each of the points' shape."""
    f = np.ascontiguousarray(fields, np.float32)
    g = np.ascontiguousarray(geo, np.float32)
    x, y, z = (np.ascontiguousarray(xyz[d], np.float32) for d in range(3))
    n = x.size
    outs = [np.zeros(x.shape, np.float32) for _ in range(nprop)]
    pp = lambda k: _p(outs[k]) if k < nprop else None  # noqa: E731
    check(lib().cgfdh_layer2model(nprop, f.shape[1], f.shape[3], f.shape[2], _p(g), _p(f), n,
                                  _p(x), _p(y), _p(z), pp(0), pp(1), pp(2)), "layer2model")
    return outs


# targets and methods of the layered-model discretisation
L2M_ONECMP, L2M_EL_ISO, L2M_EL_VTI, L2M_EL_ANISO = 0, 1, 2, 3
_L2M_NOUT = {L2M_ONECMP: 1, L2M_EL_ISO: 3, L2M_EL_VTI: 6, L2M_EL_ANISO: 22}
_L2M_METHODS = {"loc": 0, "har": 1, "ari": 2}
_G2M_FILE_FN = {L2M_ONECMP: "cgfdh_media_grid2model_onecmp",
                L2M_EL_ISO: "cgfdh_media_grid2model_el_iso",
                L2M_EL_VTI: "cgfdh_media_grid2model_el_vti",
                L2M_EL_ANISO: "cgfdh_media_grid2model_el_aniso"}
_L2M_FILE_FN = {L2M_ONECMP: "cgfdh_media_layer2model_onecmp",
                L2M_EL_ISO: "cgfdh_media_layer2model_el_iso",
                L2M_EL_VTI: "cgfdh_media_layer2model_el_vti",
                L2M_EL_ANISO: "cgfdh_media_layer2model_el_aniso"}


def layer2model_grid(media_type, target, method, fields, geo, xyz):
    """A layered model onto a curvilinear grid on the GPU
    (cgfdh_layer2model_grid): media_type, fields and geo as
    output.read_md3lay(path, with_type=True) gives them, target one of
    L2M_ONECMP / _EL_ISO / _EL_VTI / _EL_ANISO, method "loc", "har" or "ari",
    xyz (3, nz, ny, nx).  Returns the target's arrays (var | lambda, mu, rho |
    rho, c11, c33, c55, c66, c13 | rho, c11, c12, .., c66), each (nz, ny, nx)."""
    f = np.ascontiguousarray(fields, np.float32)
    g = np.ascontiguousarray(geo, np.float32)
    x, y, z = (np.ascontiguousarray(xyz[d], np.float32) for d in range(3))
    nz, ny, nx = x.shape
    outs = [np.zeros(x.shape, np.float32) for _ in range(_L2M_NOUT[target])]
    arr = (C.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
    check(lib().cgfdh_layer2model_grid(media_type, target, _L2M_METHODS.get(method, -1),
                                       f.shape[1], f.shape[3], f.shape[2], _p(g), _p(f), nx, ny,
                                       nz, _p(x), _p(y), _p(z), C.cast(arr, C.c_void_p)),
          "layer2model_grid")
    return outs


def media_layer2model(target, path, xyz, method="loc", grid_type=3):
    """The reference-shaped file entry points cgfdh_media_layer2model_onecmp /
    _el_iso / _el_vti / _el_aniso: the .md3lay file at path onto the grid
    xyz (3, nz, ny, nx).  Returns the target's arrays as layer2model_grid."""
    x, y, z = (np.ascontiguousarray(xyz[d], np.float32) for d in range(3))
    nz, ny, nx = x.shape
    outs = [np.zeros(x.shape, np.float32) for _ in range(_L2M_NOUT[target])]
    fn = getattr(lib(), _L2M_FILE_FN[target])
    check(fn(*[_p(o) for o in outs], _p(x), _p(y), _p(z), nx, ny, nz, grid_type, path.encode(),
             method.encode(), 0), "media_layer2model")
    return outs


def bin2model_grid(vols, xv, yv, zv, xyz):
    """Binary volumes sampled at grid points on the GPU
    (cgfdh_bin2model_grid): vols (3 or 2, nz, ny, nx) -- rho, vp, vs giving
    rho, lambda, mu, or Qp, Qs as they are -- on the mesh vectors xv, yv, zv;
    xyz (3, ...) the points.  Returns the outputs, each of the points' shape."""
    v = np.ascontiguousarray(vols, np.float32)
    nvol = v.shape[0]
    n3 = np.array([v.shape[3], v.shape[2], v.shape[1]], np.int32)
    xv, yv, zv = (np.ascontiguousarray(a, np.float32) for a in (xv, yv, zv))
    x, y, z = (np.ascontiguousarray(xyz[d], np.float32) for d in range(3))
    outs = [np.zeros(x.shape, np.float32) for _ in range(nvol)]
    arr = (C.c_void_p * nvol)(*[o.ctypes.data for o in outs])
    check(lib().cgfdh_bin2model_grid(nvol, _p(n3), _p(xv), _p(yv), _p(zv), _p(v), x.size, _p(x),
                                     _p(y), _p(z), C.cast(arr, C.c_void_p)), "bin2model_grid")
    return outs


def media_bin2model(files, xyz, extent, order, size, spacing, origin, grid_type=3):
    """The reference-shaped entry points cgfdh_media_bin2model_el_iso (three
    files rho, vp, vs: returns rho, lambda, mu) and _vis_iso (two files Qp,
    Qs): the binary model onto the grid xyz (3, nz, ny, nx) of x/y extent
    (xmin, xmax, ymin, ymax)."""
    x, y, z = (np.ascontiguousarray(xyz[d], np.float32) for d in range(3))
    nz, ny, nx = x.shape
    outs = [np.zeros(x.shape, np.float32) for _ in files]
    order, size = (np.ascontiguousarray(a, np.int32) for a in (order, size))
    spacing, origin = (np.ascontiguousarray(a, np.float32) for a in (spacing, origin))
    fn = lib().cgfdh_media_bin2model_el_iso if len(files) == 3 else \
        lib().cgfdh_media_bin2model_vis_iso
    check(fn(*[_p(o) for o in outs], _p(x), _p(y), _p(z), nx, ny, nz,
             *[float(e) for e in extent], grid_type, _p(order), _p(size), _p(spacing),
             _p(origin), *[f.encode() for f in files]), "media_bin2model")
    return outs


def grid2model_grid(media_type, target, method, ngz, fields, geo, xyz):
    """A grid model (.md3grd) onto a curvilinear grid on the GPU
    (cgfdh_grid2model_grid): media_type, ngz, fields and geo as
    output.read_md3grd gives them, target and method as layer2model_grid
    (the el_iso target returns rho, lambda, mu here, the reference's order)."""
    f = np.ascontiguousarray(fields, np.float32)
    g = np.ascontiguousarray(geo, np.float32)
    gz = np.ascontiguousarray(ngz, np.int32)
    x, y, z = (np.ascontiguousarray(xyz[d], np.float32) for d in range(3))
    nz, ny, nx = x.shape
    outs = [np.zeros(x.shape, np.float32) for _ in range(_L2M_NOUT[target])]
    arr = (C.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
    check(lib().cgfdh_grid2model_grid(media_type, target, _L2M_METHODS.get(method, -1), len(gz),
                                      _p(gz), f.shape[3], f.shape[2], _p(g), _p(f), nx, ny, nz,
                                      _p(x), _p(y), _p(z), C.cast(arr, C.c_void_p)),
          "grid2model_grid")
    return outs


def media_grid2model(target, path, xyz, extent, method="loc", grid_type=3):
    """The reference-shaped file entry points cgfdh_media_grid2model_*: the
    .md3grd file at path onto the grid xyz (3, nz, ny, nx) whose x/y extent
    is (xmin, xmax, ymin, ymax)."""
    x, y, z = (np.ascontiguousarray(xyz[d], np.float32) for d in range(3))
    nz, ny, nx = x.shape
    outs = [np.zeros(x.shape, np.float32) for _ in range(_L2M_NOUT[target])]
    fn = getattr(lib(), _G2M_FILE_FN[target])
    check(fn(*[_p(o) for o in outs], _p(x), _p(y), _p(z), nx, ny, nz, *[float(e) for e in extent],
             grid_type, path.encode(), method.encode(), 0), "media_grid2model")
    return outs


class PinnedBuffer:
    """Page-locked host memory (cgfdh_host_alloc) viewed as a float32 array,
    the destination that lets cgfdh_subvolume_async copies overlap steps."""

    def __init__(self, shape):
        n = int(np.prod(shape))
        p = C.c_void_p()
        check(lib().cgfdh_host_alloc(max(n, 1) * 4, C.byref(p)), "host_alloc")
        self.ptr = p
        buf = (C.c_float * max(n, 1)).from_address(p.value)
        self.array = np.frombuffer(buf, np.float32, count=n).reshape(shape)

    def close(self):
        if self.ptr:
            self.array = None
            lib().cgfdh_host_free(self.ptr)
            self.ptr = None


class Solver:
    """Engine for one grid block: ni x nj x nk physical points."""

    def __init__(self, ni, nj, nk, dt, nt_total, free_top=True, pml_on=(0,) * 6,
                 pml_layers=(0,) * 6, medium_type=MEDIUM_ELASTIC_ISO, nmaxwell=0):
        d = CgfdhDesc()
        d.ni, d.nj, d.nk = ni, nj, nk
        d.medium_type = medium_type
        d.nmaxwell = nmaxwell
        # wavefield components: 9, plus 6 memory variables per Maxwell body
        self.ncomp = 9 + 6 * nmaxwell if medium_type == MEDIUM_VISCO_ISO else 9
        d.free_top = int(free_top)
        for f in range(6):
            d.pml_on[f] = int(pml_on[f])
            d.pml_layers[f] = int(pml_layers[f]) if pml_on[f] else 0
        d.dt = dt
        d.nt_total = nt_total
        self.desc = d
        self.ni, self.nj, self.nk = ni, nj, nk
        self.nx, self.ny, self.nz = ni + 6, nj + 6, nk + 6
        self.vol = self.nx * self.ny * self.nz
        self.nt_total = nt_total
        self.nrecv = 0
        h = C.c_void_p()
        check(lib().cgfdh_create(C.byref(d), C.byref(h)), "cgfdh_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().cgfdh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- static inputs ----
    def set_metric(self, metric10):
        m = _f32(metric10)
        assert m.size == 10 * self.vol
        check(lib().cgfdh_set_metric(self.h, _p(m)), "set_metric")

    def set_medium_iso(self, lam, mu, slw):
        a, b, c = _f32(lam), _f32(mu), _f32(slw)
        assert a.size == b.size == c.size == self.vol
        check(lib().cgfdh_set_medium_iso(self.h, _p(a), _p(b), _p(c)), "set_medium_iso")

    def set_medium_vti(self, c11, c13, c33, c55, c66, slw):
        v = [_f32(x) for x in (c11, c13, c33, c55, c66, slw)]
        assert all(x.size == self.vol for x in v)
        check(lib().cgfdh_set_medium_vti(self.h, *[_p(x) for x in v]), "set_medium_vti")

    def set_medium_aniso(self, cij21, slw):
        c, s = _f32(cij21), _f32(slw)
        assert c.size == 21 * self.vol and s.size == self.vol
        check(lib().cgfdh_set_medium_aniso(self.h, _p(c), _p(s)), "set_medium_aniso")

    def set_medium_visco(self, lam, mu, slw, ylam_ymu, wl):
        a, b, c = _f32(lam), _f32(mu), _f32(slw)
        y, w = _f32(ylam_ymu), _f32(wl)
        check(lib().cgfdh_set_medium_visco(self.h, _p(a), _p(b), _p(c), _p(y), _p(w)),
              "set_medium_visco")

    def set_medium(self, medium_type, volumes, wl=None):
        """volumes: the medium's arrays in include/cgfdh.h order (for the
        visco-elastic medium lambda, mu, slw, Ylam_0, Ymu_0, ..., with wl)."""
        v = _f32(volumes).reshape(-1, self.vol)
        if medium_type == MEDIUM_VISCO_ISO:
            self.set_medium_visco(v[0], v[1], v[2], v[3:], wl)
        elif medium_type == MEDIUM_ELASTIC_VTI:
            self.set_medium_vti(*v)
        elif medium_type == MEDIUM_ELASTIC_ANISO:
            self.set_medium_aniso(v[:21], v[21])
        else:
            self.set_medium_iso(*v)

    def set_pml_face(self, face, A, B, D):
        a, b, d = _f32(A), _f32(B), _f32(D)
        check(lib().cgfdh_set_pml_face(self.h, face, _p(a), _p(b), _p(d)), "set_pml_face")

    def set_free_surface(self, mats):
        """matVx2Vz2, matVy2Vz2, matF2Vz2 (and matD as a 4th array)."""
        m = _f32(mats).reshape(-1, 9 * self.nx * self.ny)
        check(lib().cgfdh_set_free_surface(self.h, _p(m[0].copy()), _p(m[1].copy()),
                                           _p(m[2].copy())), "set_free_surface")
        if len(m) > 3:
            check(lib().cgfdh_set_free_surface_matD(self.h, _p(m[3].copy())),
                  "set_free_surface_matD")

    def set_moment_sources(self, ijk, it_begin, it_end, max_nt, mij, spatial=SRC_POINT,
                           ext_half=0, ext_coefs=None):
        ijk = np.ascontiguousarray(ijk, np.int32).reshape(-1)
        nsrc = ijk.size // 3
        itb = np.ascontiguousarray(it_begin, np.int32)
        ite = np.ascontiguousarray(it_end, np.int32)
        m = _f32(mij)
        ec = None if ext_coefs is None else _f32(ext_coefs)
        check(lib().cgfdh_set_moment_sources(self.h, nsrc, _p(ijk), _p(itb), _p(ite), max_nt,
                                             _p(m), spatial, ext_half,
                                             None if ec is None else _p(ec)),
              "set_moment_sources")

    def set_sources(self, ijk, it_begin, it_end, max_nt, mij=None, fxyz=None,
                    spatial=SRC_POINT, ext_half=0, ext_coefs=None, strict=False,
                    surf_idx=None, frate=None):
        """Force and/or moment sources (cgfdh_set_sources): mij the 6
        moment tables, fxyz the 3 force tables ([nsrc][max_nt][4] each);
        strict with surf_idx / frate the strict surface forces."""
        ijk = np.ascontiguousarray(ijk, np.int32).reshape(-1)
        nsrc = ijk.size // 3
        itb = np.ascontiguousarray(it_begin, np.int32)
        ite = np.ascontiguousarray(it_end, np.int32)
        m = None if mij is None else _f32(mij)
        f = None if fxyz is None else _f32(fxyz)
        ec = None if ext_coefs is None else _f32(ext_coefs)
        si = None if surf_idx is None else np.ascontiguousarray(surf_idx, np.int32)
        fr = None if frate is None else _f32(frate)
        nsurf = 0 if si is None else si.size
        pp = lambda a: None if a is None else _p(a)  # noqa: E731
        check(lib().cgfdh_set_sources(self.h, nsrc, _p(ijk), _p(itb), _p(ite), max_nt, pp(m),
                                      pp(f), spatial, ext_half, pp(ec), int(bool(strict)),
                                      nsurf, pp(si), pp(fr)), "set_sources")

    def set_dd_sources(self, indx, max_nt, nt_cap=None, vi=True, mij=True):
        """dd sources at the 1-D array indices indx (cgfdh_set_dd_sources);
        nt_cap steps of values on the device (default max_nt)."""
        idx = np.ascontiguousarray(indx, np.uint64).reshape(-1)
        self.ndd = idx.size
        cap = max_nt if nt_cap is None else nt_cap
        check(lib().cgfdh_set_dd_sources(self.h, idx.size, _p(idx), max_nt, cap, int(bool(vi)),
                                         int(bool(mij))), "set_dd_sources")

    def put_dd_values(self, it0, vi=None, mij=None):
        """Values of steps it0.. (cgfdh_put_dd_values): vi [nt][4][ndd][3],
        mij [nt][4][ndd][6]."""
        v = None if vi is None else _f32(vi)
        m = None if mij is None else _f32(mij)
        per = 4 * self.ndd
        nt = v.size // (3 * per) if v is not None else m.size // (6 * per)
        pp = lambda a: None if a is None else _p(a)  # noqa: E731
        check(lib().cgfdh_put_dd_values(self.h, it0, nt, pp(v), pp(m)), "put_dd_values")

    def set_receivers(self, indx1d, dijk):
        idx = np.ascontiguousarray(indx1d, np.int32).reshape(-1)
        w = _f32(dijk).reshape(-1)
        self.nrecv = idx.size // 8
        check(lib().cgfdh_set_receivers(self.h, self.nrecv, _p(idx), _p(w)), "set_receivers")

    # ---- state ----
    def set_wavefield(self, w9):
        w = _f32(w9)
        assert w.size == self.ncomp * self.vol
        check(lib().cgfdh_set_wavefield(self.h, _p(w)), "set_wavefield")

    def wavefield(self):
        w = np.empty(self.ncomp * self.vol, np.float32)
        check(lib().cgfdh_get_wavefield(self.h, _p(w)), "get_wavefield")
        return w

    def check_nan(self):
        """(icmp, iptr) of the first NaN of the current state, or None
        (cgfdh_check_nan)."""
        c, p = C.c_int(), C.c_long()
        check(lib().cgfdh_check_nan(self.h, C.byref(c), C.byref(p)), "check_nan")
        return None if c.value < 0 else (c.value, p.value)

    def aux_len(self):
        return int(lib().cgfdh_pml_aux_len(self.h))

    def set_pml_aux(self, aux):
        a = _f32(aux)
        check(lib().cgfdh_set_pml_aux(self.h, _p(a)), "set_pml_aux")

    def pml_aux(self):
        a = np.empty(max(self.aux_len(), 1), np.float32)
        check(lib().cgfdh_get_pml_aux(self.h, _p(a)), "get_pml_aux")
        return a[: self.aux_len()]

    # ---- time stepping ----
    def run(self, it0, nsteps):
        check(lib().cgfdh_run(self.h, it0, nsteps), "run")

    def seismograms(self):
        s = np.empty((max(self.nrecv, 1), self.ncomp, self.nt_total), np.float32)
        check(lib().cgfdh_get_seismograms(self.h, _p(s)), "get_seismograms")
        return s[: self.nrecv]

    def peak_ground_motion(self):
        """(15, ny, nx) peak maps and (3, ny, nx) accumulated displacement."""
        pg = np.empty((15, self.ny, self.nx), np.float32)
        dis = np.empty((3, self.ny, self.nx), np.float32)
        check(lib().cgfdh_get_peak_ground_motion(self.h, _p(pg), _p(dis)), "peak_ground_motion")
        return pg, dis

    def set_ablexp(self, Ex, Ey, Ez, blocks):
        """ablexp sponge: masks Ex[nx], Ey[ny], Ez[nz] and 6 blocks
        (enable, ni1, ni2, nj1, nj2, nk1, nk2), applied after every step."""
        ex, ey, ez = _f32(Ex), _f32(Ey), _f32(Ez)
        assert ex.size == self.nx and ey.size == self.ny and ez.size == self.nz
        b = np.ascontiguousarray(np.asarray(blocks, dtype=np.int32).reshape(6, 7))
        check(lib().cgfdh_set_ablexp(self.h, _p(ex), _p(ey), _p(ez), _p(b)), "set_ablexp")

    def set_graves_qs(self, Qs, freq):
        """Graves-Qs attenuation: quality factor per point, applied after every step."""
        q = _f32(Qs)
        assert q.size == self.vol
        check(lib().cgfdh_set_graves_qs(self.h, _p(q), float(freq)), "set_graves_qs")

    # ---- output sub-volumes (snapshots, slices) ----
    @staticmethod
    def _box(box):
        b = np.ascontiguousarray(np.asarray(box, dtype=np.int32).reshape(9))
        return b, (int(b[1]), int(b[4]), int(b[7]))

    def subvolume(self, box, comps):
        """Strided box (i1, ni, di, j1, nj, dj, k1, nk, dk) of the given
        wavefield components, packed on the device: (ncomp, nk, nj, ni)."""
        b, (ni, nj, nk) = self._box(box)
        c = np.ascontiguousarray(np.asarray(comps, dtype=np.int32).reshape(-1))
        out = np.empty((max(c.size, 1), max(nk, 1), max(nj, 1), max(ni, 1)), np.float32)
        check(lib().cgfdh_get_subvolume(self.h, _p(b), int(c.size), _p(c), _p(out)),
              "get_subvolume")
        return out

    def run_async(self, it0, nsteps):
        """Enqueue steps without waiting (cgfdh_run_async)."""
        check(lib().cgfdh_run_async(self.h, it0, nsteps), "run_async")

    def subvolume_async(self, box, comps, out):
        """Pack a box of the state after every enqueued step into out (an
        array of the subvolume's shape, kept alive by the caller, best in
        pinned memory from PinnedBuffer) on a copy stream; returns the
        ticket for wait (cgfdh_subvolume_async)."""
        b, (ni, nj, nk) = self._box(box)
        c = np.ascontiguousarray(np.asarray(comps, dtype=np.int32).reshape(-1))
        assert out.dtype == np.float32 and out.flags.c_contiguous
        assert out.size == c.size * ni * nj * nk
        t = C.c_int()
        check(lib().cgfdh_subvolume_async(self.h, _p(b), int(c.size), _p(c), _p(out),
                                          C.byref(t)), "subvolume_async")
        return t.value

    def wait(self, ticket):
        check(lib().cgfdh_wait(self.h, int(ticket)), "wait")

    def sync(self):
        check(lib().cgfdh_sync(self.h), "sync")

    def strain_subvolume(self, box):
        """Exx, Eyy, Ezz, Exz, Eyz, Exy of the box: (6, nk, nj, ni)."""
        b, (ni, nj, nk) = self._box(box)
        out = np.empty((6, max(nk, 1), max(nj, 1), max(ni, 1)), np.float32)
        check(lib().cgfdh_get_strain_subvolume(self.h, _p(b), _p(out)), "get_strain_subvolume")
        return out

    def onestage(self, it, ipair, istage, w_cur, aux_cur=None):
        w = _f32(w_cur)
        rhs = np.empty(self.ncomp * self.vol, np.float32)
        n = self.aux_len()
        arhs = np.empty(max(n, 1), np.float32)
        a = None if aux_cur is None else _f32(aux_cur)
        check(lib().cgfdh_onestage(self.h, it, ipair, istage, _p(w),
                                   None if a is None else _p(a), _p(rhs), _p(arhs)),
              "onestage")
        return rhs, arhs[:n]

    def last_timing(self):
        run_ms, st_ms, n = C.c_double(), C.c_double(), C.c_int()
        check(lib().cgfdh_last_timing(self.h, C.byref(run_ms), C.byref(st_ms), C.byref(n)),
              "last_timing")
        return run_ms.value, st_ms.value, n.value

    def set_kernel_variant(self, v):
        check(lib().cgfdh_set_kernel_variant(self.h, v), "set_kernel_variant")

    def set_tile_group(self, gx, gy):
        check(lib().cgfdh_set_tile_group(self.h, gx, gy), "set_tile_group")

    def kernel_choice(self):
        v, gx, gy = C.c_int(), C.c_int(), C.c_int()
        check(lib().cgfdh_get_kernel_choice(self.h, C.byref(v), C.byref(gx), C.byref(gy)),
              "get_kernel_choice")
        return v.value, gx.value, gy.value


class Batch:
    """Independent solvers advanced together (cgfdh_batch_*): 8-step cycles of
    all of them captured once into a hipGraph and replayed."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        h = C.c_void_p()
        check(lib().cgfdh_batch_create(len(self.solvers), arr, C.byref(h)), "batch_create")
        self.h = h

    def autotune(self, it0=0):
        check(lib().cgfdh_batch_autotune(self.h, it0), "batch_autotune")

    def run(self, it0, nsteps):
        check(lib().cgfdh_batch_run(self.h, it0, nsteps), "batch_run")

    def last_ms(self):
        ms = C.c_double()
        check(lib().cgfdh_batch_last_timing(self.h, C.byref(ms)), "batch_last_timing")
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            lib().cgfdh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
