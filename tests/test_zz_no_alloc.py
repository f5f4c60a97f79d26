"""ph_heat2d_noalloc (the allocation-free variant of ph_heat2d) called through
the C-ABI, against ph_heat2d on the same grids: the fields and the FP64
energies must agree bit for bit, on two calls in a row, the second on a
smaller grid (the fixed partial-sum array is shared by every call)."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib(ph):
    lib = ctypes.CDLL(os.path.join(ph.LIB_DIR, "libphhip.so"))
    lib.ph_heat2d_noalloc.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_int64, ctypes.c_double, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.c_void_p]
    lib.ph_heat2d_noalloc.restype = ctypes.c_int
    lib.ph_last_error.restype = ctypes.c_char_p
    return lib


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_heat2d_noalloc_matches_heat2d(ph, dt):
    lib = _lib(ph)
    code = 2 if dt == np.float64 else 1
    D = ph.DType.F64 if dt == np.float64 else ph.DType.F32
    rng = np.random.default_rng(5)
    for ny, nx, steps, every in [(1000, 2048, 40, 8), (130, 250, 37, 12)]:
        print(f"noalloc heat {dt.__name__} {ny}x{nx} steps={steps}", flush=True)
        u0 = (rng.standard_normal((ny, nx)) * 5).astype(dt)
        ref = ph.NArray.from_numpy(u0)
        e_ref = np.asarray(ph.heat2d(ref, 0.2, steps, every=every))
        u = ph.NArray.from_numpy(u0)
        tmp = ph.NArray.empty([ny, nx], D)
        ne = steps // every + 1
        e = ph.NArray.empty([ne], ph.DType.F64)
        rc = lib.ph_heat2d_noalloc(code, u.device_ptr, tmp.device_ptr, ny, nx, 0.2, steps, every, e.device_ptr)
        assert rc == 0, lib.ph_last_error()
        assert np.array_equal(u.to_numpy(), ref.to_numpy()), (ny, nx)
        assert np.array_equal(e.to_numpy(), e_ref), (ny, nx)


def test_heat2d_noalloc_errors(ph):
    lib = _lib(ph)
    u = ph.NArray.empty([8, 8], ph.DType.F64)
    rc = lib.ph_heat2d_noalloc(2, u.device_ptr, u.device_ptr, 8, 8, 0.2, 3, 0, None)
    assert rc != 0 and b"distinct" in lib.ph_last_error()
    rc = lib.ph_heat2d_noalloc(0, u.device_ptr, ph.NArray.empty([8, 8], ph.DType.F64).device_ptr, 8, 8, 0.2, 3, 0,
                               None)
    assert rc != 0 and b"dtype" in lib.ph_last_error()
