"""GPU: cgfdh_get_subvolume_noalloc (the sub-volume read through the
library's fixed device buffer, no allocation) against cgfdh_get_subvolume,
called through the C ABI: bit-identical packed values on two calls in a row,
the second with a smaller box, and a request beyond the buffer's size
rejected with an error."""
import ctypes as C

import numpy as np
import pytest

from tests import cases as cs

pytestmark = pytest.mark.gpu


def _call(fn, s, box, comps):
    from cgfd3d_amd._lib import lib
    b = np.ascontiguousarray(box, np.int32)
    c = np.ascontiguousarray(comps, np.int32)
    n = int(b[1]) * int(b[4]) * int(b[7]) * c.size
    out = np.full(n, np.nan, np.float32)
    rc = getattr(lib(), fn)(s.h, b.ctypes.data_as(C.c_void_p), int(c.size),
                            c.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return rc, out


def test_noalloc_subvolume_equals_subvolume():
    case = cs.make_case(40, 36, 30, seed=21, free_top=True, pml=(1, 1, 1, 1, 1, 0), layers=3)
    case["src"] = cs.ricker_point_source(case, 23, 21, 18, 12, 0.004, fc=10.0, t0=0.05)
    s = cs.make_solver(case, 0.004, 12, variant=12)
    try:
        s.set_wavefield(cs.random_wavefield(case, seed=22))
        s.run(0, 3)
        nx, ny, nz = 46, 42, 36
        big = [0, nx, 1, 0, ny, 1, 0, nz, 1]           # every point, every component
        small = [5, 7, 3, 4, 6, 2, 10, 5, 4]           # a strided box, 3 components
        for box, comps in ((big, list(range(9))), (small, [8, 0, 4])):
            rc0, ref = _call("cgfdh_get_subvolume", s, box, comps)
            rc1, got = _call("cgfdh_get_subvolume_noalloc", s, box, comps)
            print("noalloc box", box, "comps", comps, "values", got.size, flush=True)
            assert rc0 == 0 and rc1 == 0
            assert np.abs(ref).max() > 0
            assert np.array_equal(got, ref)
    finally:
        s.close()


def test_noalloc_subvolume_rejects_oversize():
    from cgfd3d_amd._lib import lib
    case = cs.make_case(130, 130, 60, seed=23, free_top=False, pml=(0,) * 6, layers=2)
    s = cs.make_solver(case, 0.004, 2, variant=12)
    try:
        box = [0, 136, 1, 0, 136, 1, 0, 66, 1]        # 9 x 1.22 M values > 8 Mi
        rc, _ = _call("cgfdh_get_subvolume_noalloc", s, box, list(range(9)))
        assert rc != 0
        msg = lib().cgfdh_last_error().decode()
        print("rejected:", msg, flush=True)
        assert "exceed" in msg
        rc, _ = _call("cgfdh_get_subvolume_noalloc", s, box, [0])   # 1.22 M values fit
        assert rc == 0
    finally:
        s.close()
