"""Python view of the output-file writers (include/cgfdh_io.h): the
reference's slice, snapshot, peak-ground-motion and debug-dump netCDF files
and its SAC receiver traces, written by the library's own netCDF-3 classic
and SAC code (no netCDF library needed)."""
import ctypes as C

import numpy as np

from ._lib import lib

_P = C.c_void_p


class OutputError(RuntimeError):
    pass


def _io():
    L = lib()
    if not getattr(L, "_io_sigs", False):
        L.cgfdh_io_last_error.restype = C.c_char_p
        L.cgfdh_io_slice_create.argtypes = [C.c_char_p, C.c_char, C.c_int, C.c_int, C.c_int, _P,
                                            C.c_int, _P, _P]
        L.cgfdh_io_slice_put.argtypes = [_P, C.c_int, C.c_float, C.c_int, _P]
        L.cgfdh_io_snap_create.argtypes = [C.c_char_p] + [C.c_int] * 6 + [_P] * 5
        L.cgfdh_io_snap_put.argtypes = [_P, C.c_int, C.c_float, _P, _P, _P]
        L.cgfdh_nc_close.argtypes = [_P]
        L.cgfdh_io_pg_output.argtypes = [C.c_char_p, C.c_char_p] + [C.c_int] * 6 + [_P, _P]
        L.cgfdh_io_var3d_export.argtypes = [C.c_char_p, C.c_char_p, _P, C.c_int, _P, _P, _P,
                                            C.c_int, _P, C.c_int, C.c_int, C.c_int]
        L.cgfdh_sac_export_1c1r.argtypes = [C.c_char_p, _P] + [C.c_float] * 9 + [C.c_int]
        L.cgfdh_io_recv_output_sac.argtypes = [C.c_char_p, C.c_char_p, C.c_int, _P, _P, _P,
                                               C.c_int, C.c_int, _P, C.c_int, C.c_float]
        L.cgfdh_io_grid_layer_dims.argtypes = [C.c_char_p, _P, _P, _P]
        L.cgfdh_io_read_grid_layer.argtypes = [C.c_char_p, _P, _P, _P]
        L.cgfdh_io_md3lay_dims.argtypes = [C.c_char_p, _P, _P, _P, _P, _P]
        L.cgfdh_io_md3lay_info.argtypes = [C.c_char_p, _P, _P, _P, _P, _P, _P]
        L.cgfdh_io_md3grd_info.argtypes = ([C.c_char_p] + [C.c_float] * 4 + [_P] * 4 +
                                           [C.c_int] + [_P] * 4)
        L.cgfdh_io_read_md3grd.argtypes = [C.c_char_p] + [C.c_float] * 4 + [_P]
        L.cgfdh_io_read_bin_window.argtypes = [C.c_char_p] + [_P] * 5
        L.cgfdh_io_read_md3lay.argtypes = [C.c_char_p, _P]
        L._io_sigs = True
    return L


def _check(rc, what):
    if rc != 0:
        raise OutputError(f"{what}: {_io().cgfdh_io_last_error().decode()}")


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _strings(names):
    bufs = [C.create_string_buffer(n.encode()) for n in names]
    arr = (C.c_char_p * len(bufs))(*[C.cast(b, C.c_char_p) for b in bufs])
    return arr, bufs


class SliceFile:
    """One slice file of io_slice_nc_create / io_slice_nc_put."""

    def __init__(self, path, axis, n1, n2, index, names, topoid=(0, 0)):
        self.names = list(names)
        arr, self._keep = _strings(self.names)
        h = C.c_void_p()
        _check(_io().cgfdh_io_slice_create(path.encode(), axis.encode(), n1, n2, index,
                                           _i32(topoid).ctypes.data_as(_P), len(self.names),
                                           C.cast(arr, _P), C.byref(h)), "slice create")
        self.h, self.n1, self.n2 = h, n1, n2

    def put(self, it, time, planes):
        p = _f32(planes)
        assert p.size == len(self.names) * self.n1 * self.n2
        _check(_io().cgfdh_io_slice_put(self.h, it, time, len(self.names), p.ctypes.data_as(_P)),
               "slice put")

    def close(self):
        if self.h:
            _check(_io().cgfdh_nc_close(self.h), "slice close")
            self.h = None


class SnapshotFile:
    """One snapshot file of io_snap_nc_create / io_snap_nc_put."""

    def __init__(self, path, nk, nj, ni, vel=True, stress=True, strain=False,
                 g_start=(0, 0, 0), l_start=(3, 3, 3), stride=(1, 1, 1), topoid=(0, 0)):
        h = C.c_void_p()
        _check(_io().cgfdh_io_snap_create(path.encode(), nk, nj, ni, int(vel), int(stress),
                                          int(strain), _i32(g_start).ctypes.data_as(_P),
                                          _i32(l_start).ctypes.data_as(_P),
                                          _i32(stride).ctypes.data_as(_P),
                                          _i32(topoid).ctypes.data_as(_P), C.byref(h)),
               "snapshot create")
        self.h = h

    def put(self, record, time, vel=None, stress=None, strain=None):
        arrs = [None if a is None else _f32(a) for a in (vel, stress, strain)]
        ptrs = [None if a is None else a.ctypes.data_as(_P) for a in arrs]
        _check(_io().cgfdh_io_snap_put(self.h, record, time, *ptrs), "snapshot put")

    def close(self):
        if self.h:
            _check(_io().cgfdh_nc_close(self.h), "snapshot close")
            self.h = None


def write_pg(output_dir, frame_coords, pg, ni, nj, gni1=0, gnj1=0, topoid=(0, 0)):
    """PG_slice_output: pg (15, ny, nx)."""
    p = _f32(pg)
    _, ny, nx = p.shape
    _check(_io().cgfdh_io_pg_output(output_dir.encode(), frame_coords.encode(), nx, ny, ni, nj,
                                    gni1, gnj1, _i32(topoid).ctypes.data_as(_P),
                                    p.ctypes.data_as(_P)), "PG output")


def write_var3d(output_dir, prefix, it, w4d, names, dim_names=("i", "j", "k"), topoid=(0, 0)):
    """The whole-wavefield debug dump: w4d (ncomp, nz, ny, nx)."""
    w = _f32(w4d)
    nc, nz, ny, nx = w.shape
    pos = np.ascontiguousarray(np.arange(nc, dtype=np.uint64) * (nz * ny * nx))
    na, k1 = _strings(names)
    nd, k2 = _strings(dim_names)
    _check(_io().cgfdh_io_var3d_export(output_dir.encode(), prefix.encode(),
                                       _i32(topoid).ctypes.data_as(_P), it, w.ctypes.data_as(_P),
                                       pos.ctypes.data_as(_P), C.cast(na, _P), nc,
                                       C.cast(nd, _P), nx, ny, nz), "var3d export")


def write_sac(path, trace, delta, stla=0.0, stlo=0.0, stel=0.0, stdp=0.0, evla=0.0, evlo=0.0,
              evel=0.0, evdp=0.0):
    t = _f32(trace)
    _check(_io().cgfdh_sac_export_1c1r(path.encode(), t.ctypes.data_as(_P), evla, evlo, evel,
                                       evdp, stla, stlo, stel, stdp, delta, t.size), "SAC")


def write_receivers_sac(output_dir, evtnm, names, xyz, seismo, cmp_names, dt, ncomp=None):
    """io_recv_output_sac: seismo (nrecv, ncomp_all, nt)."""
    s = _f32(seismo)
    nr, nca, nt = s.shape
    nc = ncomp or len(cmp_names)
    na, k1 = _strings(names)
    cn, k2 = _strings(cmp_names)
    _check(_io().cgfdh_io_recv_output_sac(output_dir.encode(), evtnm.encode(), nr, C.cast(na, _P),
                                          _f32(xyz).ctypes.data_as(_P), s.ctypes.data_as(_P),
                                          nca, nc, C.cast(cn, _P), nt, dt), "receiver SAC")


def read_grid_layer(path):
    """The grid layer file gd_curv_gen_layer reads: (ncell, equal,
    layer_xyz (3, n_interfaces, ny, nx))."""
    n_if, nx, ny = C.c_int(), C.c_int(), C.c_int()
    _check(_io().cgfdh_io_grid_layer_dims(path.encode(), C.byref(n_if), C.byref(nx),
                                          C.byref(ny)), "grid layer file")
    ncell = np.zeros(n_if.value - 1, np.int32)
    equal = np.zeros(n_if.value - 1, np.int32)
    xyz = np.zeros((3, n_if.value, ny.value, nx.value), np.float32)
    _check(_io().cgfdh_io_read_grid_layer(path.encode(), ncell.ctypes.data_as(_P),
                                          equal.ctypes.data_as(_P), xyz.ctypes.data_as(_P)),
           "grid layer file")
    return ncell, equal, xyz


def read_md3lay(path, with_type=False):
    """The layered-model file media_layer2model reads (every media type):
    (nprop, fields [1+3 nprop][n_interfaces][ny][nx], geo = minx, miny, dx,
    dy), with the type code first when with_type."""
    mt, nprop, n_if, nx, ny = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
    geo = np.zeros(4, np.float32)
    _check(_io().cgfdh_io_md3lay_info(path.encode(), C.byref(mt), C.byref(nprop), C.byref(n_if),
                                      C.byref(nx), C.byref(ny), geo.ctypes.data_as(_P)),
           "md3lay file")
    fields = np.zeros((1 + 3 * nprop.value, n_if.value, ny.value, nx.value), np.float32)
    _check(_io().cgfdh_io_read_md3lay(path.encode(), fields.ctypes.data_as(_P)), "md3lay file")
    if with_type:
        return mt.value, nprop.value, fields, geo
    return nprop.value, fields, geo


def read_bin_window(path, order, size, start, end):
    """The samples start[a]..end[a] of each stored axis of a media_bin2model
    volume (float32, the first axis fastest, axis a holding direction
    order[a]), laid out (nz, ny, nx)."""
    order, size, start, end = (np.ascontiguousarray(v, np.int32) for v in (order, size, start, end))
    n = [0, 0, 0]
    for a in range(3):
        n[order[a]] = int(end[a] - start[a] + 1)
    out_ = np.zeros((n[2], n[1], n[0]), np.float32)
    _check(_io().cgfdh_io_read_bin_window(path.encode(), order.ctypes.data_as(_P),
                                          size.ctypes.data_as(_P), start.ctypes.data_as(_P),
                                          end.ctypes.data_as(_P), out_.ctypes.data_as(_P)),
           "binary model")
    return out_


def read_md3grd(path, xmin, xmax, ymin, ymax):
    """The grid-model file media_grid2model reads, cropped to the window
    covering [xmin, xmax] x [ymin, ymax] as the reference crops it: (type
    code, nprop, surfaces per layer, fields [1+nprop][surfaces][ny][nx], geo
    of the cropped mesh)."""
    mt, nprop, nl, ns, nx, ny = (C.c_int() for _ in range(6))
    ngz = np.zeros(4096, np.int32)
    geo = np.zeros(4, np.float32)
    _check(_io().cgfdh_io_md3grd_info(path.encode(), xmin, xmax, ymin, ymax, C.byref(mt),
                                      C.byref(nprop), C.byref(nl), ngz.ctypes.data_as(_P),
                                      len(ngz), C.byref(ns), C.byref(nx), C.byref(ny),
                                      geo.ctypes.data_as(_P)), "md3grd file")
    fields = np.zeros((1 + nprop.value, ns.value, ny.value, nx.value), np.float32)
    _check(_io().cgfdh_io_read_md3grd(path.encode(), xmin, xmax, ymin, ymax,
                                      fields.ctypes.data_as(_P)), "md3grd file")
    return mt.value, nprop.value, [int(g) for g in ngz[:nl.value]], fields, geo
