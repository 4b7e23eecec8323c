"""ctypes access to the CPU oracle (liboracle.so / liboracle64.so) and, where
it was built, to the reference harness (_ref/libcgfd3d_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = C.c_void_p


def ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


class OrcCfg(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int),
                ("free_top", C.c_int),
                ("pml_on", C.c_int * 6), ("pml_layers", C.c_int * 6),
                ("pml_A", _P * 6), ("pml_B", _P * 6), ("pml_D", _P * 6),
                ("nsrc", C.c_int), ("src_gauss", C.c_int), ("src_ext_half", C.c_int),
                ("src_ext_coef", C.c_float),
                ("src_ijk", _P), ("src_inc", _P), ("src_it_begin", _P), ("src_it_end", _P),
                ("src_max_nt", C.c_int), ("src_mij", _P), ("medium_type", C.c_int),
                ("nmaxwell", C.c_int), ("visco_wl", _P), ("xyz", _P), ("src_fxyz", _P),
                ("src_force_strict", C.c_int), ("nsurf", C.c_int), ("src_surf_idx", _P),
                ("src_frate", _P), ("ndd", C.c_int), ("dd_max_nt", C.c_int),
                ("dd_indx", _P), ("dd_vi", _P), ("dd_mij", _P)]


MEDIUM_ISO, MEDIUM_VTI, MEDIUM_ANISO, MEDIUM_VISCO = 2, 3, 4, 5


class RefCfg(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int), ("free_top", C.c_int),
                ("pml_on", C.c_int * 6), ("pml_layers", C.c_int * 6),
                ("pml_alpha", C.c_float * 6), ("pml_beta", C.c_float * 6),
                ("pml_vel", C.c_float * 6),
                ("nsrc", C.c_int), ("src_max_nt", C.c_int), ("src_ijk", _P),
                ("src_it_begin", _P), ("src_it_end", _P), ("src_mij", _P),
                ("medium_type", C.c_int), ("nmaxwell", C.c_int), ("visco_wl", _P),
                ("src_fxyz", _P), ("src_force_strict", C.c_int), ("surf_slices", _P),
                ("surf_nt", C.c_int), ("ndd", C.c_int), ("dd_max_nt", C.c_int),
                ("dd_indx", _P), ("dd_vi", _P), ("dd_mij", _P)]


def load_oracle(fp64=False):
    path = os.path.join(HERE, "liboracle64.so" if fp64 else "liboracle.so")
    lib = C.CDLL(path)
    lib.orc_create.restype = _P
    lib.orc_create.argtypes = [_P, _P, _P]
    lib.orc_destroy.argtypes = [_P]
    lib.orc_aux_len.restype = C.c_size_t
    lib.orc_aux_len.argtypes = [_P]
    lib.orc_vol.restype = C.c_size_t
    lib.orc_vol.argtypes = [_P]
    lib.orc_onestage.argtypes = [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]
    lib.orc_run.argtypes = [_P, C.c_float, C.c_int, C.c_int, _P, _P, C.c_int, _P, _P,
                            C.c_int, _P]
    lib.orc_fd_ops.argtypes = [_P] * 6
    lib.orc_pml_profile.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_int, _P, _P, _P]
    lib.orc_pml_avg_dh.restype = C.c_float
    lib.orc_pml_avg_dh.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, C.c_int, _P]
    lib.orc_pml_region.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P, _P]
    lib.orc_dvh2dvz_iso.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P, _P]
    lib.orc_dvh2dvz.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P]
    lib.orc_dvh2dvz_vis.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]
    lib.orc_visco_gmb.argtypes = [C.c_long, C.c_int, C.c_float, C.c_float, C.c_float,
                                  _P, _P, _P, _P, _P, _P, _P]
    lib.orc_ncomp.restype = C.c_int
    lib.orc_ncomp.argtypes = [_P]
    lib.orc_metric.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P]
    lib.orc_graves_qs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, C.c_float,
                                  C.c_float]
    lib.orc_ablexp_apply.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P]
    lib.orc_dist_point2plane.restype = C.c_float
    lib.orc_dist_point2plane.argtypes = [_P, _P, _P, _P]
    lib.orc_dt_estimate.argtypes = [C.c_int, C.c_int, C.c_int, _P, C.c_int, _P, C.c_int,
                                    C.c_float, _P, _P, _P, _P]
    lib.orc_gauss_weights.argtypes = [_P, C.c_float, C.c_float, C.c_float, C.c_float,
                                      C.c_int, C.c_int]
    lib.orc_ricker.restype = C.c_float
    lib.orc_ricker.argtypes = [C.c_float, C.c_float, C.c_float]
    lib.orc_angle2moment.argtypes = [C.c_float, C.c_float, C.c_float, _P]
    lib.orc_ricker_deriv.restype = C.c_float
    lib.orc_ricker_deriv.argtypes = [C.c_float, C.c_float, C.c_float]
    lib.orc_surface_force_layer.argtypes = [_P, C.c_int, C.c_int, _P, _P]
    lib.orc_point_in_hexa.restype = C.c_int
    lib.orc_point_in_hexa.argtypes = [C.c_float] * 3 + [_P] * 3
    lib.orc_sample_shift.argtypes = [C.c_float] * 3 + [_P] * 3 + [C.c_int] * 3 + [_P]
    lib.orc_locate_point.restype = C.c_int
    lib.orc_locate_point.argtypes = [C.c_int] * 3 + [_P] * 3 + [C.c_float] * 3 + [_P, _P, _P]
    lib.orc_spline.restype = C.c_int
    lib.orc_spline.argtypes = [C.c_int] + [_P] * 5
    lib.orc_spline_at.restype = C.c_int
    lib.orc_spline_at.argtypes = [C.c_int, _P, _P, C.c_int, _P, _P]
    lib.orc_grid_z.argtypes = [_P, _P, _P, _P, C.c_int]
    lib.orc_gen_layer.restype = C.c_int
    lib.orc_gen_layer.argtypes = ([C.c_int, _P, _P, C.c_int, C.c_int, _P, _P, _P] +
                                  [C.c_int] * 9 + [_P] * 3)
    lib.orc_layer2model.restype = C.c_int
    lib.orc_layer2model.argtypes = [C.c_int, _P, _P, _P, C.c_size_t] + [_P] * 6
    lib.orc_layer2model_grid.restype = C.c_int
    lib.orc_layer2model_grid.argtypes = [C.c_int] * 3 + [_P] * 3 + [C.c_int] * 3 + [_P] * 4
    lib.orc_bin_window.restype = C.c_int
    lib.orc_bin_window.argtypes = [_P] * 4 + [C.c_float] * 6 + [_P] * 6
    lib.orc_bin2model_grid.restype = C.c_int
    lib.orc_bin2model_grid.argtypes = [C.c_int] + [_P] * 5 + [C.c_size_t] + [_P] * 4
    lib.orc_grid2model_grid.restype = C.c_int
    lib.orc_grid2model_grid.argtypes = ([C.c_int] * 3 + [_P] * 2 + [C.c_int] + [_P] * 2 +
                                        [C.c_int] * 3 + [_P] * 4)
    lib.orc_depth_to_axis.restype = C.c_int
    lib.orc_depth_to_axis.argtypes = [C.c_int] * 3 + [_P] * 3 + [C.c_float] * 2 + [_P]
    return lib


def locate_points(coords, ni, nj, nk, pts, cells=False):
    """orc_locate_point for each point: (found[n], ijk[n,3], inc[n,3]) and,
    with cells, the holding cells [n,3] too."""
    lib = load_oracle()
    co = np.ascontiguousarray(coords, np.float32).reshape(3, -1)
    pts = np.asarray(pts, np.float32).reshape(-1, 3)
    n = len(pts)
    found = np.zeros(n, np.int32)
    ijk = np.full((n, 3), -1, np.int32)
    cell = np.full((n, 3), -1, np.int32)
    inc = np.zeros((n, 3), np.float32)
    xs, ys, zs = (np.ascontiguousarray(co[d]) for d in range(3))
    for q in range(n):
        a = np.zeros(3, np.int32)
        b = np.zeros(3, np.float32)
        c = np.zeros(3, np.int32)
        found[q] = lib.orc_locate_point(ni, nj, nk, ptr(xs), ptr(ys), ptr(zs), float(pts[q, 0]),
                                        float(pts[q, 1]), float(pts[q, 2]), ptr(a), ptr(b),
                                        ptr(c))
        if found[q]:
            ijk[q], inc[q], cell[q] = a, b, c
    return (found, ijk, inc, cell) if cells else (found, ijk, inc)


def gen_layer(ncell, equal, layer_xyz, resample, start, ni, nj, nk):
    """orc_gen_layer for a one-block grid: layer_xyz (3, n_if, ny_l, nx_l).
    Returns (rc, coords (3, nk+6, nj+6, ni+6))."""
    lib = load_oracle()
    L = np.ascontiguousarray(layer_xyz, np.float32)
    n_if, ny_l, nx_l = L.shape[1:]
    nc = np.ascontiguousarray(ncell, np.int32)
    eq = np.ascontiguousarray(equal, np.int32)
    rs = np.ascontiguousarray(resample, np.int32)
    st = np.ascontiguousarray(start, np.int32)
    out = np.zeros((3, nk + 6, nj + 6, ni + 6), np.float32)
    rc = lib.orc_gen_layer(n_if, ptr(nc), ptr(eq), nx_l, ny_l, ptr(L), ptr(rs), ptr(st), ni, nj,
                           nk, ni, nj, nk, 0, 0, 0, ptr(out[0]), ptr(out[1]), ptr(out[2]))
    return rc, out


def depth_to_axis(coords, ni, nj, nk, xyd):
    """orc_depth_to_axis for each (x, y, depth): (found[n], z[n])."""
    lib = load_oracle()
    co = np.ascontiguousarray(coords, np.float32).reshape(3, -1)
    xyd = np.asarray(xyd, np.float32).reshape(-1, 3)
    xs, ys, zs = (np.ascontiguousarray(co[d]) for d in range(3))
    found = np.zeros(len(xyd), np.int32)
    z = np.empty(len(xyd), np.float32)
    for q in range(len(xyd)):
        v = np.array([xyd[q, 2]], np.float32)
        found[q] = lib.orc_depth_to_axis(ni, nj, nk, ptr(xs), ptr(ys), ptr(zs), float(xyd[q, 0]),
                                         float(xyd[q, 1]), ptr(v))
        z[q] = v[0]
    return found, z


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libcgfd3d_ref.so"))


def load_ref():
    lib = C.CDLL(os.path.join(HERE, "_ref", "libcgfd3d_ref.so"), mode=os.RTLD_LAZY)
    lib.refh_fd_ops.restype = C.c_int
    lib.refh_fd_ops.argtypes = [_P] * 6
    lib.refh_pml.argtypes = [_P, _P, C.c_int, _P, _P, _P, _P]
    lib.refh_dvh2dvz.argtypes = [_P, _P, _P, _P]
    lib.refh_dvh2dvz_vis.argtypes = [_P, _P, _P, _P, _P]
    lib.refh_onestage.restype = C.c_long
    lib.refh_onestage.argtypes = [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]
    lib.refh_run.argtypes = [_P, _P, _P, _P, C.c_float, C.c_int, _P, C.c_int, _P, _P,
                             _P, _P, _P, _P]
    lib.refh_graves_qs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, C.c_float,
                                   C.c_float]
    lib.refh_ablexp_apply.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P]
    lib.refh_dt_esti.argtypes = [C.c_int, C.c_int, C.c_int, _P, C.c_int, _P, C.c_int, C.c_float,
                                 _P, _P]
    return lib


def set_dd_fields(cfg, dd, keep):
    """dd sources at points into an OrcCfg or RefCfg: dd holds indx (1-D
    indices), max_nt and the tables vi [max_nt][4][ndd][3] and/or mij
    [max_nt][4][ndd][6]."""
    if dd is None or len(dd["indx"]) == 0:
        return
    idx = np.ascontiguousarray(dd["indx"], np.int64)
    keep.append(idx)
    cfg.ndd = idx.size
    cfg.dd_max_nt = int(dd["max_nt"])
    cfg.dd_indx = idx.ctypes.data
    for key in ("vi", "mij"):
        if dd.get(key) is not None:
            t = np.ascontiguousarray(dd[key], np.float32)
            keep.append(t)
            setattr(cfg, "dd_" + key, t.ctypes.data)


class Oracle:
    """One oracle solver instance over a fixed grid, medium and boundary set.

    case: a dict built by tests (see tests/cases.py) holding ni,nj,nk, metric,
    medium_type and med (its volumes, slowness last), free_top, pml_on, pml_layers, pml A/B/D tables and the source tables.
    """

    def __init__(self, case, fp64=False):
        self.lib = load_oracle(fp64)
        self.dtype = np.float64 if fp64 else np.float32
        self.case = case
        cfg = OrcCfg()
        cfg.ni, cfg.nj, cfg.nk = case["ni"], case["nj"], case["nk"]
        cfg.free_top = int(case.get("free_top", 0))
        cfg.medium_type = int(case.get("medium_type", MEDIUM_ISO))
        self._keep = []
        if cfg.medium_type == MEDIUM_VISCO:
            wl = np.ascontiguousarray(case["visco_wl"], np.float32)
            xyz = np.ascontiguousarray(case["coords"], np.float32)
            self._keep += [wl, xyz]
            cfg.nmaxwell = len(wl)
            cfg.visco_wl = wl.ctypes.data
            cfg.xyz = xyz.ctypes.data
        for f in range(6):
            on = int(case.get("pml_on", [0] * 6)[f])
            cfg.pml_on[f] = on
            cfg.pml_layers[f] = int(case.get("pml_layers", [0] * 6)[f]) if on else 0
            if on:
                for name, fld in (("pml_A", cfg.pml_A), ("pml_B", cfg.pml_B),
                                  ("pml_D", cfg.pml_D)):
                    a = np.ascontiguousarray(case[name][f], np.float32)
                    self._keep.append(a)
                    fld[f] = a.ctypes.data
        src = case.get("src")
        if src is not None and src["nsrc"] > 0:
            cfg.nsrc = src["nsrc"]
            cfg.src_gauss = int(src.get("gauss", 0))
            cfg.src_ext_half = int(src.get("ext_half", 3))
            cfg.src_ext_coef = float(src.get("ext_coef", 1.5))
            arrs = dict(ijk=np.ascontiguousarray(src["ijk"], np.int32),
                        inc=np.ascontiguousarray(src.get("inc", np.zeros(3 * src["nsrc"])),
                                                 np.float32),
                        itb=np.ascontiguousarray(src["it_begin"], np.int32),
                        ite=np.ascontiguousarray(src["it_end"], np.int32))
            if src.get("mij") is not None:
                arrs["mij"] = np.ascontiguousarray(src["mij"], np.float32)
            if src.get("fxyz") is not None:     # force sources: Fx, Fy, Fz tables
                arrs["fxyz"] = np.ascontiguousarray(src["fxyz"], np.float32)
            self._keep += list(arrs.values())
            cfg.src_ijk = arrs["ijk"].ctypes.data
            cfg.src_inc = arrs["inc"].ctypes.data
            cfg.src_it_begin = arrs["itb"].ctypes.data
            cfg.src_it_end = arrs["ite"].ctypes.data
            cfg.src_max_nt = int(src["max_nt"])
            cfg.src_mij = arrs["mij"].ctypes.data if "mij" in arrs else None
            cfg.src_fxyz = arrs["fxyz"].ctypes.data if "fxyz" in arrs else None
            # strict surface force: the sources listed in surf_idx carry rate
            # tables frate (Fx_rate, Fy_rate, Fz_rate, [nsurf][max_nt][4])
            cfg.src_force_strict = int(src.get("strict", 0))
            if src.get("surf_idx") is not None and len(src["surf_idx"]) > 0:
                si = np.ascontiguousarray(src["surf_idx"], np.int32)
                fr = np.ascontiguousarray(src["frate"], np.float32)
                self._keep += [si, fr]
                cfg.nsurf = len(si)
                cfg.src_surf_idx = si.ctypes.data
                cfg.src_frate = fr.ctypes.data
        set_dd_fields(cfg, case.get("dd"), self._keep)
        self.cfg = cfg
        self.metric = np.ascontiguousarray(case["metric"], np.float32)
        self.med = np.ascontiguousarray(case["med"], np.float32)
        self.h = self.lib.orc_create(C.byref(cfg), ptr(self.metric), ptr(self.med))
        self.vol = self.lib.orc_vol(self.h)
        self.ncomp = self.lib.orc_ncomp(self.h)
        self.aux_len = self.lib.orc_aux_len(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def onestage(self, it, ipair, istage, w_cur, aux_cur=None):
        w = np.ascontiguousarray(w_cur, self.dtype)
        rhs = np.zeros(self.ncomp * self.vol, self.dtype)
        aux = np.ascontiguousarray(aux_cur if aux_cur is not None
                                   else np.zeros(self.aux_len), self.dtype)
        arhs = np.zeros(max(self.aux_len, 1), self.dtype)
        self.lib.orc_onestage(self.h, it, ipair, istage, ptr(w), ptr(aux), ptr(rhs), ptr(arhs))
        return rhs, arhs[: self.aux_len]

    def surface_force_layer(self, it, istage):
        """(TxSrc,TySrc,TzSrc) and (VxSrc,VySrc,VzSrc) of step it, stage
        istage, each (3, ny, nx)."""
        nx, ny = self.case["ni"] + 6, self.case["nj"] + 6
        ts = np.zeros((3, ny, nx), self.dtype)
        vs = np.zeros((3, ny, nx), self.dtype)
        self.lib.orc_surface_force_layer(self.h, it, istage, ptr(ts), ptr(vs))
        return ts, vs

    def surface_slices(self, nt):
        """The surface-force slices of steps 0..nt-1, every stage, in the
        reference harness's layout [nt][4][6][ny][nx] (float32)."""
        nx, ny = self.case["ni"] + 6, self.case["nj"] + 6
        out = np.zeros((nt, 4, 6, ny, nx), np.float32)
        for it in range(nt):
            for s in range(4):
                ts, vs = self.surface_force_layer(it, s)
                out[it, s, :3] = ts
                out[it, s, 3:] = vs
        return out

    def run(self, dt, nt, w=None, aux=None, it0=0, recv_idx=None, recv_w=None, nt_total=None):
        w = (np.zeros(self.ncomp * self.vol, self.dtype) if w is None
             else np.array(w, self.dtype, copy=True))
        aux = (np.zeros(max(self.aux_len, 1), self.dtype) if aux is None
               else np.array(aux, self.dtype, copy=True))
        nrecv = 0 if recv_idx is None else len(recv_idx)
        nt_total = nt_total or (it0 + nt)
        seismo = np.zeros((max(nrecv, 1), self.ncomp, nt_total), self.dtype)
        ri = None if recv_idx is None else np.ascontiguousarray(recv_idx, np.int32)
        rw = None if recv_w is None else np.ascontiguousarray(recv_w, np.float32)
        self.lib.orc_run(self.h, dt, it0, nt, ptr(w), ptr(aux), nrecv, ptr(ri), ptr(rw),
                         nt_total, ptr(seismo))
        return w, aux[: self.aux_len], seismo[:nrecv]


def oracle_pg(orc):
    """Peak ground motion maps (15, ny, nx) and displacement (3, ny, nx) of an
    Oracle after run() (free surface only)."""
    lib = orc.lib
    lib.orc_get_pg.argtypes = [_P, _P, _P]
    nx, ny = orc.case["ni"] + 6, orc.case["nj"] + 6
    pg = np.zeros((15, ny, nx), orc.dtype)
    dis = np.zeros((3, ny, nx), orc.dtype)
    lib.orc_get_pg(orc.h, ptr(pg), ptr(dis))
    return pg, dis
