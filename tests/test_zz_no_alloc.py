"""The allocation-free column solve (exo_implicit_columns_noalloc) through the C
ABI against the allocating shim (exo_implicit_column), bit for bit, on two
calls in a row, the second with fewer and shorter columns; and the fixed
scratch's size limit is an error, not a fault."""
import ctypes as C

import numpy as np
import pytest

from helpers import atm_config

pytestmark = pytest.mark.gpu


def columns(rng, ncol, nx1):
    nc1 = nx1 + 6
    dz = 30000.0 / nx1
    z = (np.arange(nc1) - 3 + 0.5) * dz
    w = np.zeros((ncol, 5, nc1))
    w[:, 0] = 1.2 * np.exp(-z / 7300.0) * (1 + 0.01 * rng.standard_normal((ncol, nc1)))
    w[:, 4] = 1e5 * np.exp(-z / 7300.0) * (1 + 0.01 * rng.standard_normal((ncol, nc1)))
    w[:, 1:4] = rng.normal(scale=5.0, size=(ncol, 3, nc1))
    zf = np.arange(nx1 + 1) * dz
    area = np.tile(1.0e10 * (1 + zf / 6.371e6) ** 2, (ncol, 1))
    vol = 1.0e10 * dz * (1 + rng.uniform(0, 1e-3, (ncol, nx1)))
    du = rng.normal(scale=1e-2, size=(ncol, 5, nc1)) * np.abs(w) + 1e-6
    return w, du, area, vol


class DeviceArrays:
    """Device copies of host arrays through the HIP runtime the library uses
    (the test allocates; the entry point under test does not)."""

    def __init__(self, arrays):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.ptrs = []
        for a in arrays:
            a = np.ascontiguousarray(a)
            p = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)) == 0
            assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes),
                                      1) == 0   # hipMemcpyHostToDevice
            self.ptrs.append(p)

    def get(self, q, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptrs[q],
                                  C.c_size_t(out.nbytes), 2) == 0   # hipMemcpyDeviceToHost
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


def test_noalloc_columns_match_the_allocating_shim(ex):
    rng = np.random.default_rng(21)
    g = ex.Solver(atm_config(ex, nx=6, nx1=8))
    for ncol, nx1, dt in ((70, 32, 40.0), (5, 16, 7.5)):
        w, du, area, vol = columns(rng, ncol, nx1)
        dev = DeviceArrays((w, du, area, vol))
        rc = g.L.exo_implicit_columns_noalloc(g.h, ncol, nx1, *dev.ptrs, dt)
        assert rc == 0, g.L.exo_last_error(g.h)
        assert g.L.exo_sync(g.h) == 0
        got = dev.get(1, du)
        dev.free()
        for c in range(ncol):
            ref = du[c].copy()
            assert g.L.exo_implicit_column(g.h, nx1, np.ascontiguousarray(w[c]), ref,
                                           np.ascontiguousarray(area[c]),
                                           np.ascontiguousarray(vol[c]), dt) == 0
            assert np.array_equal(got[c], ref), (ncol, nx1, c)
        assert not np.array_equal(got, du)
    # more coefficient scratch than the fixed array holds: refused before launch
    w, du, area, vol = columns(rng, 65, 200)
    dev = DeviceArrays((w, du, area, vol))
    rc = g.L.exo_implicit_columns_noalloc(g.h, 65, 200, *dev.ptrs, 1.0)
    assert rc == -1
    assert b"scratch" in g.L.exo_last_error(g.h)
    assert g.L.exo_sync(g.h) == 0
    assert np.array_equal(dev.get(1, du), du)
    dev.free()
    g.close()
