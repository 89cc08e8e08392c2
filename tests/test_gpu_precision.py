"""The FP32-compensated form against the FP64 oracle where it is hardest: on
relaxed (phase-separating) liquid states, where the pressure, viscous and
heat-flux pair terms cancel to a small net rate and the vdW EOS amplifies a
density error about tenfold (T = (u + a rho)/kB with a rho >> T).  The
bound is 1e-6 normwise (max |x - x_ref| / max |x_ref|) for rho, rho_lr,
acc and udot.  Also: the compensated FP32 accumulator beats a plain FP32 sum
on a cancelling input."""
import ctypes
import math

import numpy as np
import pytest

from oracle import oracle as O
from pyticles_amd import _lib, configs
from pyticles_amd.device import from_system

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "rho_lr", "acc", "udot")


def rel_norm(a, b):
    s = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / (s if s > 0 else 1.0))


def _relaxed(nside, steps):
    """A C3-style quench (T 1.2 -> 0.4 at step 0, eta = 1, k = 5) stepped
    `steps` steps by the FP64 device path; returns (system, r, v, u)."""
    s = configs.c3(nside=nside, precision=_lib.FP64)
    with from_system(s) as b:
        b.step(steps)
        fin = b.download("r", "v", "u", "rho")
    print("[precision] %dx%d quench after %d steps: rho %.3f .. %.3f" % (
        nside, nside, steps, fin["rho"].min(), fin["rho"].max()), flush=True)
    return s, fin["r"], fin["v"], fin["u"]


def _fp32c_errors(s, r, v, u, **pkw):
    prm = dict(s.prm, precision=_lib.FP32C, **pkw)
    ref = O.derivatives(r, v, s.m, u, s.prm, terms=prm["terms"])
    s32 = configs.System(r, v, s.m, s.T, prm, 0, s.seed)
    with from_system(s32) as b:
        b.upload(u=u)
        b.rates(prm["terms"])
        got = b.download(*FIELDS)
    return {k: rel_norm(got[k], ref[k]) for k in FIELDS}


def test_relaxed_256_fp32c_within_1e6():
    s, r, v, u = _relaxed(256, 2000)
    err = _fp32c_errors(s, r, v, u)
    print("[precision] 256x256 relaxed FP32C errors %s" % {k: "%.2e" % e for k, e in err.items()}, flush=True)
    for k in FIELDS:
        assert err[k] < 1e-6, (k, err[k])


def test_full_size_c3_relaxed_fp32c_within_1e6():
    # C3 itself: 1,048,576 particles after the 2,000 quench steps
    s, r, v, u = _relaxed(1024, 2000)
    err = _fp32c_errors(s, r, v, u)
    print("[precision] C3 relaxed FP32C errors %s" % {k: "%.2e" % e for k, e in err.items()}, flush=True)
    for k in FIELDS:
        assert err[k] < 1e-6, (k, err[k])


def _accumulate(x, unroll=4):
    x = np.ascontiguousarray(x, dtype=np.float64)
    k, p = ctypes.c_double(), ctypes.c_double()
    _lib.check(_lib.lib.sph_fp32_accumulate(_lib.context(), x.size, _lib.dptr(x), unroll, ctypes.byref(k),
                                            ctypes.byref(p)))
    return k.value, p.value


@pytest.mark.parametrize("case", ["cancelling_offset", "many_positive"])
def test_compensated_sum_beats_plain_fp32(case):
    rng = np.random.Generator(np.random.PCG64(17))
    if case == "cancelling_offset":
        # a large term cancelled at the end: while the running total is
        # large, plain FP32 rounds every small term added to it
        x = np.concatenate([[1.0e4], np.full(20000, 0.3), [-1.0e4]])
    else:
        # many terms of one sign, as the density sums over a list
        x = rng.uniform(0.5, 1.5, size=200000)
    xf = x.astype(np.float32).astype(np.float64)
    exact = math.fsum(xf)
    cond = np.sum(np.abs(xf)) / abs(exact)
    k, p = _accumulate(x)
    ek, ep = abs(k - exact) / abs(exact), abs(p - exact) / abs(exact)
    print("[precision] %s (condition %.1f): compensated %.2e, plain FP32 %.2e" % (case, cond, ek, ep), flush=True)
    assert ek <= 2 * 2.0 ** -24 * cond  # the compensated sum's bound, 2 eps times the condition number
    assert ep > 100 * max(ek, 1e-12)
