"""The allocation-free binning entry point (mpm_bin_positions_noalloc)
through the C ABI: device buffers in and out, the library's fixed device
array for every intermediate. Its results equal those of
mpm_bin_positions (the allocating entry point) bit for bit, on two calls in
a row, the second with fewer positions (the array keeps the first call's
contents)."""
import ctypes as C

import numpy as np
import pytest

from mpm_mi355x import capi

pytestmark = pytest.mark.gpu


def positions(n, seed, lo, hi, nelem):
    rng = np.random.default_rng(seed)
    lo, hi = np.array(lo), np.array(hi)
    x = lo + rng.random((n, 3)) * (hi - lo)
    h = (hi - lo) / np.array(nelem)
    k = n // 20
    x[:k] = lo + np.floor((x[:k] - lo) / h) * h       # on cell faces: the higher cell
    x[k:k + 10] = hi - 1e-15                            # last cell
    x[k + 10:k + 20] = [hi[0], lo[1], lo[2]]            # outside (upper face)
    x[k + 20:k + 30] = lo - 1e-12                       # outside (below)
    x[k + 30:3 * k] = x[k + 30]                         # crowded cell (stability)
    return x


class Hip:
    """The HIP runtime through ctypes: device buffers for the test (the entry
    point under test allocates nothing itself)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]

    def ok(self, rc):
        assert rc == 0, "HIP error %d" % rc

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.lib.hipMalloc(C.byref(p), max(int(nbytes), 1)))
        return p

    def h2d(self, dst, a):
        self.ok(self.lib.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, 1))

    def d2h(self, a, src):
        self.ok(self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, 2))


def noalloc(grid, x):
    hip = Hip()
    n = len(x)
    x = np.ascontiguousarray(x, np.float64)
    dx, dc, do = hip.alloc(x.nbytes), hip.alloc(4 * n), hip.alloc(4 * n)
    try:
        hip.h2d(dx, x)
        hip.ok(hip.lib.hipMemset(dc, 0xFF, 4 * n))
        hip.ok(hip.lib.hipMemset(do, 0xFF, 4 * n))
        hip.ok(hip.lib.hipDeviceSynchronize())
        capi.check(capi.load().mpm_bin_positions_noalloc(C.byref(grid), n, dx, dc, do, None))
        hip.ok(hip.lib.hipDeviceSynchronize())
        cell = np.zeros(n, np.int32)
        order = np.zeros(n, np.int32)
        hip.d2h(cell, dc)
        hip.d2h(order, do)
    finally:
        for p in (dx, dc, do):
            hip.lib.hipFree(p)
    return cell, order


def test_noalloc_binning_equals_the_allocating_entry_point():
    nelem = (37, 23, 29)
    lo, hi = (0.0, -0.5, 1.0), (1.0, 0.7, 1.3)
    g = capi.grid_desc(lo, hi, nelem)
    for n, seed in ((600_000, 1), (70_001, 2)):   # the second call is smaller
        print("noalloc binning n=%d seed=%d" % (n, seed), flush=True)
        x = positions(n, seed, lo, hi, nelem)
        want_cell, want_order = capi.bin_positions(g, x)
        cell, order = noalloc(g, x)
        assert np.array_equal(cell, want_cell)
        assert np.array_equal(order, want_order)
        assert np.sum(cell < 0) >= 20


def test_noalloc_limits_are_checked_on_the_host():
    g = capi.grid_desc((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (4, 4, 4))
    lib = capi.load()
    rc = lib.mpm_bin_positions_noalloc(C.byref(g), capi.NOALLOC_MAX_POSITIONS + 1, C.c_void_p(8), C.c_void_p(8),
                                       C.c_void_p(8), None)
    assert rc == capi.MPM_ERR_INVALID_ARGUMENT
    assert lib.mpm_bin_positions_noalloc(C.byref(g), 0, None, None, None, None) == 0
