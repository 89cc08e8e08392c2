"""The allocation-free rendering entry point (sph_splash_map_2d_noalloc)
against sph_splash_map_2d on the same inputs: bitwise the same grid, on two
calls in a row, the second with fewer particles and pixels (the fixed
scratch keeps the first call's contents).  The test's own device arrays
come from the HIP runtime the library links (plumbing only); the entry
point itself allocates nothing."""
import ctypes

import numpy as np
import pytest

from pyticles_amd import _lib, configs

pytestmark = pytest.mark.gpu

_D = ctypes.POINTER(ctypes.c_double)


def _fn():
    return _lib.lib.sph_splash_map_2d_noalloc  # bound with device pointers as c_void_p (_lib._SIGS)


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libpyticles_hip.so
    uses (the same loaded instance)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, shape):
        out = np.empty(shape)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.rt.hipDeviceSynchronize()
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _case(ns, L, npx, npy, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = ns * ns
    r = configs.lattice(n, L, rng, jitter=0.3)
    m = rng.uniform(0.8, 1.2, size=n)
    rho = rng.uniform(0.9, 1.1, size=n)
    a = rng.normal(size=n)
    h = np.full(n, 2.0)
    x = (np.arange(npx) + 0.5) * (L / npx)
    y = (np.arange(npy) + 0.5) * (L / npy)
    bounds = np.array([0.0, L, 0.0, L])
    return r, m, rho, a, h, x, y, bounds


def _reference(r, m, rho, a, h, x, y, bounds, cutoff):
    Z = np.empty((x.size, y.size))
    _lib.check(_lib.lib.sph_splash_map_2d(_lib.context(), r.shape[0], _lib.dptr(r), _lib.dptr(m), _lib.dptr(rho),
                                          _lib.dptr(a), _lib.dptr(h), x.size, _lib.dptr(x), y.size, _lib.dptr(y),
                                          _lib.dptr(bounds), cutoff, _lib.dptr(Z)))
    return Z


def _noalloc(r, m, rho, a, h, x, y, bounds, cutoff):
    hip = _Hip()
    try:
        d = [hip.put(v) for v in (r, m, rho, a, h, x, y)]
        Z = hip.put(np.full((x.size, y.size), np.nan))
        assert hip.rt.hipDeviceSynchronize() == 0  # inputs on the device before the library's stream reads them
        rc = _fn()(_lib.context(), r.shape[0], d[0], d[1], d[2], d[3], d[4], x.size, d[5], y.size, d[6],
                   bounds.ctypes.data_as(_D), cutoff, Z)
        _lib.check(rc)
        return hip.get(Z, (x.size, y.size))  # waits for the device: the call returns without waiting
    finally:
        hip.free()


def test_noalloc_render_equals_allocating_render_twice():
    big = _case(200, 200.0, 160, 150, seed=1)    # 40,000 particles
    small = _case(60, 60.0, 50, 40, seed=2)      # 3,600 particles, fewer pixels
    for case in (big, small):
        ref = _reference(*case, cutoff=2.0)
        got = _noalloc(*case, cutoff=2.0)
        assert np.array_equal(got, ref)
        assert np.all(np.isfinite(got))


def test_noalloc_render_rejects_sizes_beyond_its_scratch():
    # 66,049 particles, one record over the scratch's 65,536: refused before
    # any launch (real device arrays, so even a launch would read in bounds)
    r, m, rho, a, h, x, y, bounds = _case(257, 257.0, 8, 8, seed=3)
    hip = _Hip()
    try:
        d = [hip.put(v) for v in (r, m, rho, a, h, x, y)]
        Z = hip.put(np.zeros((x.size, y.size)))
        rc = _fn()(_lib.context(), r.shape[0], d[0], d[1], d[2], d[3], d[4], x.size, d[5], y.size, d[6],
                   bounds.ctypes.data_as(_D), 2.0, Z)
        assert rc == _lib.SPH_ERR_ARG and "at most 65536 particles" in _lib.last_error()
    finally:
        hip.free()
