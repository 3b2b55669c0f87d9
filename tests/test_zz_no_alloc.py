"""pz_get_ddr_noalloc, the allocation-free variant of pz_get_dr/pz_get_ddr on
device buffers: its results equal the host-buffer entry points' bit for bit
(the same transform kernel), on two calls in a row, the second with a
smaller input, with and without the second derivative, and sizes beyond its
staging limit are rejected before anything is launched.

The device buffers come from hipMalloc of the HIP runtime the library is
linked against (looked up by its soname once the library is loaded), so the
test holds one HIP runtime whatever else the process has imported."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    from pizza_amd.runtime import lib
    lib()
    hip = ctypes.CDLL("libamdhip64.so.7")
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
    hip.hipFree.argtypes = [vp]
    hip.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
    hip.hipMemset.argtypes = [vp, ctypes.c_int, sz]
    for f in (hip.hipMalloc, hip.hipFree, hip.hipMemcpy, hip.hipMemset, hip.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return hip


class _Device:
    """One hipMalloc'd buffer of complex(8) values."""

    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), 16 * n) == 0
        assert hip.hipMemset(self.ptr, 0, 16 * n) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.size == self.n
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape):
        a = np.empty(shape, dtype=np.complex128)
        assert a.size == self.n
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.ptr)


def test_get_ddr_noalloc_equals_get_ddr():
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Run, lib
    run = Run(W.c2_production())
    hip = _hip()
    rng = np.random.default_rng(3)
    L = lib()
    for n_m, l_dct_in, second in ((run.n_m, 1, True), (7, 0, True), (5, 1, False)):
        print(f"case: n_m {n_m}, l_dct_in {l_dct_in}, second derivative {second}", flush=True)
        f = rng.standard_normal((n_m, run.n_r)) + 1j * rng.standard_normal((n_m, run.n_r))
        if second:
            d1, d2 = run.get_ddr(f, bool(l_dct_in))
        else:
            d1, d2 = run.get_dr(f, bool(l_dct_in)), None
        # device arrays in the reference layout f(n_m, n_r), m fastest
        df, o1, o2 = (_Device(hip, n_m * run.n_r) for _ in range(3))
        df.put(f.T)
        assert hip.hipDeviceSynchronize() == 0
        rc = L.pz_get_ddr_noalloc(run.h, df.ptr, o1.ptr, o2.ptr if second else None, n_m, l_dct_in)
        assert rc == 0, L.pz_last_error(run.h)
        assert hip.hipDeviceSynchronize() == 0
        assert np.array_equal(o1.get((run.n_r, n_m)).T, d1)
        if second:
            assert np.array_equal(o2.get((run.n_r, n_m)).T, d2)
        for b in (df, o1, o2):
            b.free()


def test_get_ddr_noalloc_rejects_oversize():
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Run, lib
    run = Run(W.c2_production())
    n_m = 262144 // run.n_r + 1
    dummy = ctypes.c_void_p(1)
    rc = lib().pz_get_ddr_noalloc(run.h, dummy, dummy, dummy, n_m, 1)
    assert rc == -2
    assert b"staging limit" in lib().pz_last_error(run.h)
