"""The C ABI library on the CPU: it loads, exports every entry point
include/cgfdh.h declares, rejects bad input before touching a device, and its
host table builders reproduce the reference's tables (checked against the
oracle and, where built, the reference itself) bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tests.conftest import gpu_available
from oracle import pyoracle as po
from tests import cases as cs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    from cgfd3d_amd import workloads
    return workloads._setup_sigs()


def _declared(header):
    hdr = open(os.path.join(ROOT, "include", header)).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(cgfd[hs]_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    from cgfd3d_amd._lib import LIB_PATH
    headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert headers == ["cgfdh.h", "cgfdh_io.h", "cgfds.h"]
    names = [n for h in headers for n in _declared(h)]
    assert len(names) >= 30
    lib = C.CDLL(LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_surrogate_binding_covers_its_header():
    from cgfd3d_amd._lib import SIGNATURES
    bound = {s[0] for s in SIGNATURES}
    assert set(_declared("cgfds.h")) <= bound


def test_python_binding_covers_the_header():
    from cgfd3d_amd._lib import SIGNATURES
    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "cgfdh.h")).read(),
                 flags=re.S)
    solver_calls = set(re.findall(r"\b(cgfdh_[a-z_]+)\s*\(\s*cgfdh_solver", hdr))
    assert solver_calls <= {s[0] for s in SIGNATURES} | {
        "cgfdh_compute_metric", "cgfdh_compute_free_surface", "cgfdh_get_metric",
        "cgfdh_get_free_surface"}


def test_bad_descriptor_rejected_before_the_device():
    from cgfd3d_amd._lib import CgfdhDesc, lib
    d = CgfdhDesc()
    d.ni, d.nj, d.nk, d.medium_type, d.nt_total, d.dt = 4, 20, 20, 2, 10, 0.01
    h = C.c_void_p()
    assert lib().cgfdh_create(C.byref(d), C.byref(h)) == -1
    assert b"minimum" in lib().cgfdh_last_error()
    d.ni, d.medium_type = 20, 6     # no medium with this code
    assert lib().cgfdh_create(C.byref(d), C.byref(h)) == -1
    assert b"medium" in lib().cgfdh_last_error()
    d.medium_type = 5               # visco-elastic needs 1..4 Maxwell bodies
    for nm in (0, 5):
        d.nmaxwell = nm
        assert lib().cgfdh_create(C.byref(d), C.byref(h)) == -1
    d.medium_type, d.nmaxwell = 2, 0
    d.pml_on[0], d.pml_layers[0] = 1, 25
    assert lib().cgfdh_create(C.byref(d), C.byref(h)) == -1
    d.pml_layers[0] = 5
    d.free_top, d.pml_on[5], d.pml_layers[5] = 1, 1, 5
    assert lib().cgfdh_create(C.byref(d), C.byref(h)) == -1


def test_dt_estimate_rejects_bad_arguments_before_the_device():
    """cgfdh_estimate_dt checks its arguments on the host before any copy."""
    from cgfd3d_amd._lib import lib
    xyz = np.zeros((3, 14, 14, 14), np.float32)
    vols = np.ones((3, 14, 14, 14), np.float32)
    out = np.zeros(3, np.float32)
    ijk = np.zeros(3, np.int32)
    call = lambda ni, mt, nmw, cfl: lib().cgfdh_estimate_dt(
        ni, 8, 8, po.ptr(xyz), mt, po.ptr(vols), nmw, cfl, po.ptr(out[0:]), po.ptr(out[1:]),
        po.ptr(out[2:]), po.ptr(ijk))
    assert call(8, 2, 0, 0.0) == -1 and b"cfl" in lib().cgfdh_last_error()
    assert call(8, 2, 0, float("nan")) == -1
    assert call(0, 2, 0, 1.0) == -1 and b"grid" in lib().cgfdh_last_error()
    assert call(8, 7, 0, 1.0) == -1 and b"medium_type" in lib().cgfdh_last_error()
    assert call(8, 5, 0, 1.0) == -1 and b"nmaxwell" in lib().cgfdh_last_error()
    assert call(8, 5, 5, 1.0) == -1
    assert lib().cgfdh_estimate_dt(8, 8, 8, None, 2, po.ptr(vols), 0, 1.0, po.ptr(out[0:]),
                                   po.ptr(out[1:]), po.ptr(out[2:]), po.ptr(ijk)) == -1


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_create_without_device_fails_loudly():
    from cgfd3d_amd import CgfdhError, Solver
    with pytest.raises(CgfdhError, match="no HIP device"):
        Solver(16, 16, 16, 0.01, 10)


@pytest.mark.parametrize("layers", [4, 9])
def test_pml_face_coefs_match_oracle_and_reference(layers):
    case = cs.make_case(18, 16, 14, seed=13, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=layers)
    L = _lib()
    xyz = np.ascontiguousarray(case["coords"])
    for f in range(5):
        A = np.zeros(layers + 1, np.float32); B = np.zeros_like(A); D = np.zeros_like(A)
        rc = L.cgfdh_pml_face_coefs(18, 16, 14, po.ptr(xyz), f, layers, case["pml_vel"],
                                    case["pml_alpha"], case["pml_beta"], po.ptr(A), po.ptr(B),
                                    po.ptr(D))
        assert rc == 0
        assert np.array_equal(A, case["pml_A"][f]) and np.array_equal(B, case["pml_B"][f])
        assert np.array_equal(D, case["pml_D"][f])
    if po.ref_available():
        ref = po.load_ref()
        cfg, keep = cs.ref_cfg(case)
        Ar = np.zeros((6, layers + 1), np.float32); Br = np.zeros_like(Ar); Dr = np.zeros_like(Ar)
        reg = np.zeros((6, 6), np.int32)
        ref.refh_pml(C.byref(cfg), po.ptr(xyz), layers + 1, po.ptr(Ar), po.ptr(Br), po.ptr(Dr),
                     po.ptr(reg))
        for f in range(5):
            A = np.zeros(layers + 1, np.float32); B = np.zeros_like(A); D = np.zeros_like(A)
            L.cgfdh_pml_face_coefs(18, 16, 14, po.ptr(xyz), f, layers, case["pml_vel"],
                                   case["pml_alpha"], case["pml_beta"], po.ptr(A), po.ptr(B),
                                   po.ptr(D))
            assert np.array_equal(A, Ar[f]) and np.array_equal(B, Br[f]) and np.array_equal(D, Dr[f])


def test_source_tables_match_oracle():
    L, orc = _lib(), po.load_oracle()
    for t in np.linspace(-0.3, 1.7, 41, dtype=np.float32):
        assert L.cgfdh_ricker(float(t), 2.0, 0.5) == orc.orc_ricker(float(t), 2.0, 0.5)
    for sdr in [(30.0, 60.0, 90.0), (211.0, 33.0, -47.0)]:
        m = np.zeros(6, np.float32); u = np.zeros(6, np.float32)
        L.cgfdh_moment_from_angles(*sdr, 3.0e10, 1.5, 2.0e6, po.ptr(m))
        orc.orc_angle2moment(*sdr, po.ptr(u))
        M0 = np.float32(np.float32(3.0e10) * np.float32(1.5)) * np.float32(2.0e6)
        assert np.array_equal(m, (M0 * u).astype(np.float32))
    for inc, kb in [((0.0, 0.0, 0.0), 3), ((0.31, -0.12, 0.4), 1)]:
        a = np.zeros(343, np.float32); b = np.zeros(343, np.float32)
        assert L.cgfdh_gauss_weights(po.ptr(a), *inc, 1.5, 3, kb) == 0
        orc.orc_gauss_weights(po.ptr(b), *inc, 1.5, 3, kb)
        assert np.array_equal(a, b)


def test_ricker_moment_table_matches_reference_layout():
    """it_begin/it_end/max_nt and the [it][stage] table of src_put_into_struct
    (start at 0, t0 = 0), against the layout the parity tests feed the oracle."""
    L = _lib()
    dt, nt = 0.004, 60
    m6 = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0], np.float32)  # Mxx Myy Mzz Myz Mxz Mxy
    itb, ite, mnt = C.c_int(), C.c_int(), C.c_int()
    table = np.zeros((6, nt * 4), np.float32)
    assert L.cgfdh_ricker_moment_table(0.0, dt, nt * dt, 0.0, 10.0, 0.12, po.ptr(m6),
                                       C.byref(itb), C.byref(ite), C.byref(mnt),
                                       po.ptr(table), nt * 4) == 0
    assert (itb.value, ite.value, mnt.value) == (0, nt - 1, nt)
    case = cs.make_case(10, 9, 8)
    ref = cs.ricker_point_source(case, 6, 6, 6, nt, dt, fc=10.0, t0=0.12, m0=1.0,
                                 mech=(1, 1, 1, 1, 1, 1))["mij"].reshape(6, nt * 4)
    # engine order Mxx,Myy,Mzz,Mxz,Myz,Mxy from the .src order Mxx,Myy,Mzz,Myz,Mxz,Mxy
    for m, src_slot in enumerate([0, 1, 2, 4, 3, 5]):
        assert np.array_equal(table[m], (ref[m] * m6[src_slot]).astype(np.float32))


@pytest.mark.skipif(not os.path.isdir("/root/reference/forward"), reason="reference headers absent")
def test_reference_shim_compiles():
    """The drop-in drv_rk_curv_col_allstep compiles against the reference's
    own headers (the structs it unpacks)."""
    import subprocess
    import tempfile
    shim = os.path.join(ROOT, "cgfd3d_amd", "integration", "drv_rk_curv_col_cgfdh.c")
    inc = ["-I/opt/conda/include", "-I/root/reference/lib", "-I/root/reference/forward",
           "-I" + os.path.join(ROOT, "include")]
    r = subprocess.run(["gcc", "-fsyntax-only", "-std=gnu99", "-Wall", "-Werror"] + inc + [shim],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # and to an object, optimised, with every warning an error
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run(["gcc", "-c", "-O2", "-std=gnu99", "-Wall", "-Wextra",
                            "-Wno-unused-parameter", "-Werror"] + inc +
                           [shim, "-o", os.path.join(d, "shim.o")],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
