"""Device-resident box: the NumPy-facing handle over sph_box (DeviceBox).

A Box owns one periodic 2-D system in HBM.  Every method is a C-ABI call
into libpyticles_hip.so; host arrays are copied in and out only when asked.
"""
import ctypes

import numpy as np

from . import _lib
from .configs import to_struct


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError("expected shape %s, got %s" % (shape, a.shape))
    return a


class DeviceBox:
    FIELDS = ("r", "v", "u", "rho", "rho_lr", "T", "p", "pco", "acc", "udot", "gradv", "jq", "P")

    def __init__(self, n, prm, device=0):
        self.n = int(n)
        self.prm = dict(prm)
        self.ctx = _lib.context(device)
        self._h = _lib._P()
        _lib.check(_lib.lib.sph_box_create(self.ctx, self.n, ctypes.byref(to_struct(self.prm)),
                                           ctypes.byref(self._h)))

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if self._h:
            _lib.lib.sph_box_destroy(self._h)
            self._h = _lib._P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state ------------------------------------------------------------
    def set_params(self, **kw):
        self.prm.update(kw)
        _lib.check(_lib.lib.sph_box_set_params(self._h, ctypes.byref(to_struct(self.prm))))

    def upload(self, r=None, v=None, m=None, u=None):
        n = self.n
        r = None if r is None else _f64(r, (n, 2))
        v = None if v is None else _f64(v, (n, 2))
        m = None if m is None else _f64(m, (n,))
        u = None if u is None else _f64(u, (n,))
        _lib.check(_lib.lib.sph_box_upload(self._h, _lib.dptr(r), _lib.dptr(v), _lib.dptr(m), _lib.dptr(u)))

    def download(self, *names):
        """Arrays in the caller's particle order, e.g. box.download('r', 'rho')."""
        shapes = {"r": (2,), "v": (2,), "acc": (2,), "jq": (2,), "gradv": (2, 2), "P": (2, 2)}
        out = {}
        ptrs = []
        for f in self.FIELDS:
            if f in names:
                a = np.empty((self.n,) + shapes.get(f, ()))
                out[f] = a
                ptrs.append(_lib.dptr(a))
            else:
                ptrs.append(None)
        for nm in names:
            if nm not in self.FIELDS:
                raise KeyError(nm)
        _lib.check(_lib.lib.sph_box_download(self._h, *ptrs))
        return out[names[0]] if len(names) == 1 else out

    # -- passes -----------------------------------------------------------
    def density(self):
        _lib.check(_lib.lib.sph_density(self._h))

    def gradients(self):
        _lib.check(_lib.lib.sph_gradients(self._h))

    def rates(self, terms=_lib.ALL_TERMS):
        _lib.check(_lib.lib.sph_rates(self._h, int(terms)))

    def init_thermal(self, T):
        if np.isscalar(T):
            _lib.check(_lib.lib.sph_init_thermal(self._h, None, float(T)))
        else:
            T = _f64(T, (self.n,))
            _lib.check(_lib.lib.sph_init_thermal(self._h, _lib.dptr(T), 0.0))

    def step(self, nsteps):
        _lib.check(_lib.lib.sph_step(self._h, int(nsteps)))

    def advance(self, nsteps=1):
        """nsteps steps of an open run (one rate evaluation each); v and u
        stay at the half step until sync() or any call that reads them."""
        _lib.check(_lib.lib.sph_box_advance(self._h, int(nsteps)))

    def sync(self):
        """Close an open run (closing half kick)."""
        _lib.check(_lib.lib.sph_box_sync(self._h))

    def run_stats(self):
        """Counters of the current or last run: steps, builds, extends, open."""
        k, b, e = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        o = ctypes.c_int()
        _lib.check(_lib.lib.sph_box_run_stats(self._h, ctypes.byref(k), ctypes.byref(b), ctypes.byref(e),
                                              ctypes.byref(o)))
        return dict(steps=k.value, builds=b.value, extends=e.value, open=bool(o.value))

    def energy(self):
        ke, ue, et = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.lib.sph_energy(self._h, ctypes.byref(ke), ctypes.byref(ue), ctypes.byref(et)))
        return ke.value, ue.value, et.value

    def set_graph(self, on):
        _lib.check(_lib.lib.sph_box_set_graph(self._h, 1 if on else 0))

    def snapshot(self, bufs):
        """Start an asynchronous copy of r, v, u, m into bufs (a PinnedState);
        steps enqueued afterwards overlap the transfer."""
        _lib.check(_lib.lib.sph_box_snapshot(self._h, _lib.dptr(bufs.r), _lib.dptr(bufs.v), _lib.dptr(bufs.u),
                                             _lib.dptr(bufs.m)))

    def snapshot_wait(self):
        _lib.check(_lib.lib.sph_box_snapshot_wait(self._h))

    def set_schedule(self, unroll, rates_unroll=0, block=0, list_bits=0, fused=1):
        """List entries in flight per lane in the pair passes (1, 2, 4; 0 =
        default), in the rate pass alone (0 = the same), the pair passes'
        block size (128, 256, 512; 0 = 256), the list entry width (16,
        32; 0 = leave it) and whether the leapfrog step runs inside the rate
        pass (1) or as a launch of its own (0).  unroll may be a schedule
        dict from autotune.schedule_for."""
        if isinstance(unroll, dict):
            unroll, rates_unroll, block, list_bits, fused = (
                unroll["unroll"], unroll.get("rates_unroll", 0), unroll.get("block", 0),
                unroll.get("list_bits", 0), unroll.get("fused", 1))
        _lib.check(_lib.lib.sph_box_set_schedule(self._h, int(unroll)))
        _lib.check(_lib.lib.sph_box_set_rates_unroll(self._h, int(rates_unroll)))
        _lib.check(_lib.lib.sph_box_set_pass_block(self._h, int(block)))
        if list_bits:
            self.set_list_bits(list_bits)
        _lib.check(_lib.lib.sph_box_set_fused_step(self._h, 1 if fused else 0))

    def set_list_bits(self, bits):
        """Entry width of the lists the next builds write: 16 (default;
        falls back to 32 by itself where offsets do not fit) or 32."""
        _lib.check(_lib.lib.sph_box_set_list_bits(self._h, int(bits)))

    def list_bits(self):
        """Entry width of the box's lists in memory (after its first build)."""
        b = ctypes.c_int()
        _lib.check(_lib.lib.sph_box_list_bits(self._h, ctypes.byref(b)))
        return b.value

    def timing(self, reset=False):
        out = np.zeros(7)
        _lib.check(_lib.lib.sph_box_timing(self._h, _lib.dptr(out), 7, 1 if reset else 0))
        return dict(zip(("build_ms", "density_ms", "gradients_ms", "rates_ms", "pairs_ms", "steps",
                         "integrate_ms"), out))

    PASSES = ("density", "gradients", "rates")
    BUILD = "build"  # time_pass of the list build stage (reorders slots; not one of PASSES)

    def time_pass(self, name, reps=50):
        """Average microseconds of one launch of a pair pass ('density',
        'gradients' or 'rates') at the current state: events around reps
        back-to-back launches.  Writes only derived fields (the next step
        re-derives them); the box must have been stepped.  'build' times the
        list build stage (it may reorder the slots; the next run builds
        first)."""
        us = ctypes.c_double()
        idx = 3 if name == self.BUILD else self.PASSES.index(name)
        _lib.check(_lib.lib.sph_box_time_pass(self._h, idx, int(reps), ctypes.byref(us)))
        return us.value

    # -- (a) neighbour list ----------------------------------------------
    def pairs(self, cap=None):
        """Every unordered pair (i<j) with rsq < rc^2 + tol^2, lexicographic."""
        if cap is None:
            cap = max(64, 16 * self.n)
        while True:
            iap = np.empty((cap, 2), dtype=np.int64)
            nip = ctypes.c_int64()
            rc = _lib.lib.sph_nlist_build(self._h, cap, _lib.i64ptr(iap), ctypes.byref(nip))
            if rc == _lib.SPH_ERR_CAP:
                cap = int(nip.value)
                continue
            _lib.check(rc)
            return iap[: nip.value].copy()

    def separations(self, iap):
        iap = np.ascontiguousarray(iap, dtype=np.int64)
        nip = iap.shape[0]
        drij = np.empty((nip, 2))
        dv = np.empty((nip, 2))
        rij = np.empty(nip)
        rsq = np.empty(nip)
        _lib.check(_lib.lib.sph_separations(self._h, nip, _lib.i64ptr(iap), _lib.dptr(drij), _lib.dptr(rij),
                                            _lib.dptr(rsq), _lib.dptr(dv)))
        return drij, rij, rsq, dv


class PinnedState:
    """r[n][2], v[n][2], u[n], m[n] in pinned host memory (hipHostMalloc),
    the target of DeviceBox.snapshot."""

    def __init__(self, n):
        self.n = int(n)
        nbytes = 8 * 6 * self.n
        self._p = _lib._P()
        _lib.check(_lib.lib.sph_host_alloc(nbytes, ctypes.byref(self._p)))
        if nbytes:
            buf = (ctypes.c_double * (6 * self.n)).from_address(self._p.value)
            flat = np.ctypeslib.as_array(buf)
        else:
            flat = np.empty(0)
        self.r = flat[: 2 * n].reshape(n, 2)
        self.v = flat[2 * n: 4 * n].reshape(n, 2)
        self.u = flat[4 * n: 5 * n]
        self.m = flat[5 * n: 6 * n]

    def close(self):
        if self._p:
            self.r = self.v = self.u = self.m = None
            _lib.lib.sph_host_free(self._p)
            self._p = _lib._P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def from_system(sysm, device=0):
    """Box holding a configs.System with u = kB T - a rho set on device."""
    b = DeviceBox(sysm.n, sysm.prm, device)
    b.upload(r=sysm.r, v=sysm.v, m=sysm.m, u=np.zeros(sysm.n))
    b.init_thermal(sysm.T)
    return b
