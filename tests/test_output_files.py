"""CPU: the library's own output-file writers (include/cgfdh_io.h) against
independent readers and the reference's own code.

netCDF: there is no netCDF library in this image to write reference files
with, so the files are read back with scipy.io.netcdf_file, an independent
implementation of the netCDF-3 classic format, and checked against the
layout io_funcs.c defines: dimension names, sizes and order, the unlimited
time dimension, variable names and order, global attributes and their
values (io_slice_nc_create io_funcs.c:780-883, io_snap_nc_create :885-995,
PG_slice_output :1944-2003, io_var3d_export_nc :121-163), the data written
at each record, the fill value of records not yet written, and the record
count.  SAC: byte-identical to the reference's own sacExport1C1R
(lib/sacLib.c, compiled into oracle/_ref), plus the header fields the
reference sets, read back by the published SAC header layout."""
import ctypes as C
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from cgfd3d_amd import output as out
from oracle import pyoracle as po

FILL = np.float32(9.9692099683868690e+36)


def _read(path):
    f = netcdf_file(path, "r", mmap=False)
    return f


def test_slice_files_layout_and_records(tmp_path):
    names = ["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"]
    rng = np.random.default_rng(0)
    nk, nj, ni = 7, 5, 6
    for axis, (n1, n2), dims, att in (("x", (nk, nj), ("k", "j"), "i_index"),
                                      ("y", (nk, ni), ("k", "i"), "j_index"),
                                      ("z", (nj, ni), ("j", "i"), "k_index")):
        p = str(tmp_path / f"slice_{axis}.nc")
        s = out.SliceFile(p, axis, n1, n2, 4, names, topoid=(1, 2))
        planes = [rng.standard_normal((9, n1, n2)).astype(np.float32) for _ in range(3)]
        for it in (0, 1, 3):                  # record 2 never written
            s.put(it, 0.5 * it, planes[min(it, 2)])
        s.close()
        f = _read(p)
        assert list(f.dimensions) == ["time", dims[0], dims[1]]
        assert f.dimensions["time"] is None
        assert (f.dimensions[dims[0]], f.dimensions[dims[1]]) == (n1, n2)
        assert list(f.variables) == ["time"] + names
        assert f.variables["Vx"].dimensions == ("time",) + dims
        assert int(getattr(f, f"{att}_with_ghosts_in_this_thread")) == 4
        assert list(getattr(f, "coords_of_mpi_topo")) == [1, 2]
        assert f.variables["time"].shape[0] == 4
        assert np.array_equal(f.variables["time"][:], np.array([0, 0.5, FILL, 1.5], np.float32))
        for v, name in enumerate(names):
            d = f.variables[name][:]
            assert np.array_equal(d[0], planes[0][v]) and np.array_equal(d[1], planes[1][v])
            assert np.array_equal(d[3], planes[2][v])
            assert np.all(d[2] == FILL)
        f.close()


def test_snapshot_file_layout(tmp_path):
    nk, nj, ni = 4, 3, 5
    p = str(tmp_path / "snap.nc")
    s = out.SnapshotFile(p, nk, nj, ni, vel=True, stress=False, strain=True,
                         g_start=(10, 20, 30), l_start=(13, 23, 33), stride=(2, 1, 3),
                         topoid=(0, 1))
    rng = np.random.default_rng(1)
    vel = rng.standard_normal((2, 3, nk, nj, ni)).astype(np.float32)
    strain = rng.standard_normal((2, 6, nk, nj, ni)).astype(np.float32)
    for r in range(2):
        s.put(r, 0.1 * (r + 1), vel=vel[r], strain=strain[r])
    s.close()
    f = _read(p)
    assert list(f.dimensions) == ["time", "k", "j", "i"]
    assert list(f.variables) == ["time", "Vx", "Vy", "Vz", "Exx", "Eyy", "Ezz", "Exz", "Eyz",
                                 "Exy"]
    assert list(f.first_index_to_snapshot_output) == [10, 20, 30]
    assert list(f.first_index_in_this_thread_with_ghosts) == [13, 23, 33]
    assert list(f.index_stride_in_this_thread) == [2, 1, 3]
    assert list(f.coords_of_mpi_topo) == [0, 1]
    for r in range(2):
        for q, n in enumerate(["Vx", "Vy", "Vz"]):
            assert np.array_equal(f.variables[n][r], vel[r, q])
        for q, n in enumerate(["Exx", "Eyy", "Ezz", "Exz", "Eyz", "Exy"]):
            assert np.array_equal(f.variables[n][r], strain[r, q])
    f.close()


def test_pg_and_debug_dump_files(tmp_path):
    rng = np.random.default_rng(2)
    ny, nx = 9, 11
    pg = rng.standard_normal((15, ny, nx)).astype(np.float32)
    out.write_pg(str(tmp_path), "px0_py0", pg, ni=nx - 6, nj=ny - 6, gni1=0, gnj1=0,
                 topoid=(0, 0))
    f = _read(str(tmp_path / "PG_V_A_D_px0_py0.nc"))
    assert list(f.dimensions) == ["j", "i"] and f.dimensions["j"] == ny
    names = ["PGV", "PGVh", "PGVx", "PGVy", "PGVz", "PGA", "PGAh", "PGAx", "PGAy", "PGAz",
             "PGD", "PGDh", "PGDx", "PGDy", "PGDz"]
    assert list(f.variables) == names
    assert list(f.global_index_of_first_physical_points) == [0, 0]
    assert list(f.count_index_of_physical_points) == [nx - 6, ny - 6]
    for q, n in enumerate(names):
        assert np.array_equal(f.variables[n][:], pg[q])
    f.close()
    w = rng.standard_normal((3, 4, 5, 6)).astype(np.float32)
    out.write_var3d(str(tmp_path), "w3d", 17, w, ["Vx", "Vy", "Vz"], topoid=(2, 3))
    f = _read(str(tmp_path / "w3d_2_3_it17.nc"))
    assert list(f.dimensions) == ["i", "j", "k"]          # defined x, y, z
    assert f.variables["Vy"].dimensions == ("k", "j", "i")
    assert np.array_equal(f.variables["Vz"][:], w[2])
    f.close()


def _sac_fields(path):
    b = open(path, "rb").read()
    hf = np.frombuffer(b[:280], np.float32)
    hi = np.frombuffer(b[280:440], np.int32)
    hc = b[440:632]
    data = np.frombuffer(b[632:], np.float32)
    return hf, hi, hc, data


def test_sac_header_fields(tmp_path):
    tr = np.sin(np.arange(100) * 0.1).astype(np.float32)
    p = str(tmp_path / "t.sac")
    out.write_sac(p, tr, 0.005, stla=100.0, stlo=200.0, stel=-30.0, stdp=0.005)
    hf, hi, hc, data = _sac_fields(p)
    assert np.array_equal(data, tr)
    assert hf[0] == np.float32(0.005) and hf[5] == 0.0                 # delta, b
    assert hf[6] == np.float32(np.float32(0.005) * np.float32(99))     # e
    assert (hf[31], hf[32], hf[33], hf[34]) == (100.0, 200.0, -30.0, np.float32(0.005))
    assert hf[1] == -12345.0 and hf[7] == -12345.0                     # depmin, o undefined
    assert (hi[6], hi[9], hi[15], hi[35]) == (6, 100, 1, 1)            # nvhdr npts iftype leven
    assert hi[0] == -12345
    assert hc[:8] == b"-12345\0\0" and hc[160:168] == b"-12345\0\0"    # kstnm, kcmpnm
    assert hc[8:24] == b"-12345  " + b"\0" * 8                         # kevnm
    assert hc[24:32] == b"-12345  "


@pytest.mark.skipif(not po.ref_available(), reason="reference build absent")
def test_sac_bytes_equal_reference(tmp_path):
    """The reference's own sacExport1C1R (lib/sacLib.c) writes the same bytes."""
    ref = po.load_ref()
    f = ref.sacExport1C1R
    f.argtypes = [C.c_char_p, C.c_void_p] + [C.c_float] * 9 + [C.c_int, C.c_char_p]
    f.restype = C.c_int
    rng = np.random.default_rng(3)
    for npts, dt in ((1, 0.01), (257, 0.004), (1001, 0.005)):
        tr = rng.standard_normal(npts).astype(np.float32)
        hdr = (1.5, -2.0, 3.25, 4.0, 1000.0, -2000.0, 300.0, dt)
        pa, pb = str(tmp_path / "ours.sac"), str(tmp_path / "ref.sac")
        out.write_sac(pa, tr, dt, evla=hdr[0], evlo=hdr[1], evel=hdr[2], evdp=hdr[3],
                      stla=hdr[4], stlo=hdr[5], stel=hdr[6], stdp=hdr[7])
        msg = C.create_string_buffer(512)
        assert f(pb.encode(), tr.ctypes.data, *hdr, dt, npts, msg) == 0
        assert open(pa, "rb").read() == open(pb, "rb").read(), npts


def test_receiver_sac_files(tmp_path):
    rng = np.random.default_rng(4)
    seis = rng.standard_normal((2, 27, 50)).astype(np.float32)   # a GMB run: 27 components
    names = ["r1", "st02"]
    cmp_names = ["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"]
    xyz = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]], np.float32)
    out.write_receivers_sac(str(tmp_path), "evt", names, xyz, seis, cmp_names, 0.01, ncomp=9)
    files = sorted(os.listdir(tmp_path))
    assert len(files) == 18 and "evt.st02.Txy.sac" in files
    hf, hi, hc, data = _sac_fields(str(tmp_path / "evt.st02.Txz.sac"))
    assert np.array_equal(data, seis[1, 7])
    assert (hf[31], hf[32], hf[33], hf[34]) == (4.0, 5.0, 6.0, np.float32(0.01))   # stdp = dt
