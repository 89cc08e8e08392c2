"""Many boxes in one device-resident arena (batched and served forms).

An Arena holds the state of many independent periodic boxes in one slot
pool in HBM.  step() advances every running box by one leapfrog step per
launch sequence (replayed from a captured HIP graph); boxes of any size
share the launches.  Each box keeps its own neighbour-list schedule,
decided on the device from its own particles, so a box's trajectory is
bitwise the same alone or batched with others.  Boxes can be added while
others run (they initialise, u = kB T - a rho, during the next step) and
removed when done; their slots are reused.

No counterpart in the reference, which runs one system per process.
"""
import ctypes

import numpy as np

from . import _lib
from .configs import to_struct


def footprint(capacity, precision=_lib.FP32C):
    """Device bytes of an arena with `capacity` particle slots."""
    b = ctypes.c_int64()
    _lib.check(_lib.lib.sph_arena_footprint(int(capacity), int(precision), ctypes.byref(b)))
    return int(b.value)


def capacity_for(budget_bytes, precision=_lib.FP32C):
    """Largest slot capacity whose arena fits in budget_bytes of HBM."""
    lo, hi = 0, 1
    while footprint(hi, precision) <= budget_bytes:
        lo, hi = hi, hi * 2
        if hi > 2 ** 31 - 1:
            hi = 2 ** 31 - 1
            break
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if footprint(mid, precision) <= budget_bytes:
            lo = mid
        else:
            hi = mid
    return lo


class Arena:
    def __init__(self, capacity, precision=_lib.FP32C, device=0, graph=True):
        self.capacity = int(capacity)
        self.precision = int(precision)
        self.ctx = _lib.context(device)
        self._h = _lib._P()
        _lib.check(_lib.lib.sph_arena_create(self.ctx, self.capacity, self.precision, ctypes.byref(self._h)))
        self.sizes = {}
        if not graph:
            self.set_graph(False)

    def close(self):
        if self._h:
            _lib.lib.sph_arena_destroy(self._h)
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

    def set_graph(self, on):
        _lib.check(_lib.lib.sph_arena_set_graph(self._h, 1 if on else 0))

    def timing(self, reset=False):
        """Accumulated per-pass GPU ms while context timing is on (keys as
        device.DeviceBox.timing)."""
        out = np.zeros(7)
        _lib.check(_lib.lib.sph_arena_timing(self._h, _lib.dptr(out), 7, 1 if reset else 0))
        keys = ("build_ms", "density_ms", "gradients_ms", "rates_ms", "pairs_ms", "steps", "integrate_ms")
        return dict(zip(keys, out))

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
        _lib.check(_lib.lib.sph_arena_set_schedule(self._h, int(unroll)))
        _lib.check(_lib.lib.sph_arena_set_rates_unroll(self._h, int(rates_unroll)))
        _lib.check(_lib.lib.sph_arena_set_pass_block(self._h, int(block)))
        if list_bits:
            self.set_list_bits(list_bits)
        _lib.check(_lib.lib.sph_arena_set_fused_step(self._h, 1 if fused else 0))

    def set_list_bits(self, bits):
        """Entry width of the lists the boxes' next builds write (16 or 32)."""
        _lib.check(_lib.lib.sph_arena_set_list_bits(self._h, int(bits)))

    def list_bits(self, box_id):
        """Entry width of a box's lists in memory."""
        b = ctypes.c_int()
        _lib.check(_lib.lib.sph_arena_list_bits(self._h, int(box_id), ctypes.byref(b)))
        return b.value

    def add(self, sysm, nsteps=None, T0=None):
        """Admit a configs.System; returns its box id.  Raises SphError with
        code SPH_ERR_NOMEM when the arena has no room."""
        prm = dict(sysm.prm, precision=self.precision)
        n = sysm.n
        bid = ctypes.c_int()
        _lib.check(_lib.lib.sph_arena_add(
            self._h, n, ctypes.byref(to_struct(prm)), _lib.dptr(np.ascontiguousarray(sysm.r, dtype=np.float64)),
            _lib.dptr(np.ascontiguousarray(sysm.v, dtype=np.float64)),
            _lib.dptr(np.ascontiguousarray(sysm.m, dtype=np.float64)),
            float(sysm.T if T0 is None else T0), int(sysm.steps if nsteps is None else nsteps), ctypes.byref(bid)))
        self.sizes[bid.value] = n
        return bid.value

    def step(self, nsteps=1):
        _lib.check(_lib.lib.sph_arena_step(self._h, int(nsteps)))

    def poll_all(self, nmax=None):
        nmax = (max(self.sizes) + 1) if (nmax is None and self.sizes) else (nmax or 0)
        st = np.zeros(nmax, dtype=np.int32)
        k = np.zeros(nmax, dtype=np.int64)
        hl = np.zeros(nmax, dtype=np.int32)
        if nmax:
            _lib.check(_lib.lib.sph_arena_poll_all(self._h, nmax, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                   k.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                   hl.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return st, k, hl

    def done(self, box_id):
        steps, d = ctypes.c_int64(), ctypes.c_int()
        _lib.check(_lib.lib.sph_arena_poll(self._h, int(box_id), ctypes.byref(steps), ctypes.byref(d)))
        return bool(d.value), int(steps.value)

    def result(self, box_id):
        n = self.sizes[box_id]
        r, v, u, rho = np.empty((n, 2)), np.empty((n, 2)), np.empty(n), np.empty(n)
        e0, e1 = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.lib.sph_arena_result(self._h, int(box_id), _lib.dptr(r), _lib.dptr(v), _lib.dptr(u),
                                             _lib.dptr(rho), ctypes.byref(e0), ctypes.byref(e1)))
        return dict(r=r, v=v, u=u, rho=rho, E0=e0.value, E1=e1.value)

    def remove(self, box_id):
        _lib.check(_lib.lib.sph_arena_remove(self._h, int(box_id)))
        self.sizes.pop(box_id, None)

    def stats(self):
        used, cap = ctypes.c_int64(), ctypes.c_int64()
        running, boxes = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib.sph_arena_stats(self._h, ctypes.byref(used), ctypes.byref(cap), ctypes.byref(running),
                                            ctypes.byref(boxes)))
        return dict(used=used.value, capacity=cap.value, running=running.value, boxes=boxes.value)
