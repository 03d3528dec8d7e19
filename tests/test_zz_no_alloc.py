"""tis_bdhi_factor_noalloc (the factorisation of tis_bdhi_tensors with its
scratch in the library's static device array) against tis_bdhi_tensors'
own factor, bit for bit, on two calls in a row, the second smaller; and the
size limit refused with TIS_ERR_ARG before any device work.  The caller's
device buffers come from the HIP runtime that libtis_hip.so itself links."""
import ctypes as C

import numpy as np
import pytest

from tis_amd.bdhi import bead_spring_chain
from tis_amd.engine import BDEngine, load_library

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    load_library()
    h = C.CDLL("libamdhip64.so.7")     # already loaded by libtis_hip.so: the same runtime
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    for fn in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
        getattr(h, fn).restype = C.c_int
    return h


def _reference_factor(n, seed):
    s, x = bead_spring_chain(n, seed=seed)
    e = BDEngine(s, key=1, seed=2)
    e.set_coords(x)
    e.tensors()
    D, L = e.tensor(0), e.tensor(1)
    e.close()
    return D, L


def test_noalloc_factor_equals_tensors_factor():
    hip = _hip()
    lib = load_library()
    f = lib.tis_bdhi_factor_noalloc
    f.restype = C.c_int32
    f.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_int32)]
    for n in (300, 100):
        print(f"[noalloc] n={n} (3n={3 * n})", flush=True)
        D, Lref = _reference_factor(n, seed=11 + n)
        n3 = 3 * n
        nbytes = n3 * n3 * 8
        Dd, Ld = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(Dd), nbytes) == 0 and hip.hipMalloc(C.byref(Ld), nbytes) == 0
        try:
            Dc = np.asfortranarray(D)                       # column-major (n3, n3)
            assert hip.hipMemcpy(Dd, Dc.ctypes.data, nbytes, H2D) == 0
            junk = np.full((n3, n3), np.nan, order="F")     # L is fully overwritten
            assert hip.hipMemcpy(Ld, junk.ctypes.data, nbytes, H2D) == 0
            info = C.c_int32(-1)
            rc = f(0, Dd, Ld, n3, None, C.byref(info))
            assert rc == 0 and info.value == 0
            assert hip.hipDeviceSynchronize() == 0
            L = np.zeros((n3, n3), order="F")
            assert hip.hipMemcpy(L.ctypes.data, Ld, nbytes, D2H) == 0
        finally:
            hip.hipFree(Dd)
            hip.hipFree(Ld)
        assert np.array_equal(L, Lref)
        assert np.abs(np.triu(L, 1)).max() == 0.0
    # beyond the scratch array: refused, nothing launched
    big = 9216 + 3
    assert f(0, C.c_void_p(1), C.c_void_p(1), big, None, None) == -1
