"""Synthetic inputs for the five workloads, generated on the host with
NumPy PCG64 so every consumer (the HIP path, the tests, the CPU baseline)
sees identical bytes.

    C1  FP64, L=32,   N=1,024,     inviscid, non-conducting, 100 steps, seed 1
    C2  FP64, L=512,  N=262,144,   eta=1, k=5,               1,000 steps, seed 2
    C3  FP32C, L=1024, N=1,048,576, quench T 1.2 -> 0.4,     2,000 steps, seed 3
    C4  512 boxes, N_b log-uniform [256, 16384], T_b in [0.3, 1], 500 steps, seed 1000+b
    C5  Poisson job stream (200 jobs/s, seed 5), job j: N log-uniform
        [1024, 65536], steps in [100, 2000], T in [0.3, 1.2], seed j

Physics shared by all (the reference's driver values, toybox.py:37-56 and
properties.py:18-20): m = 1, h = 2, h_l = 4, VerletList cutoff 4 with
tolerance 1 (list radius^2 = 17), a = 2, b = 0.5, kB = 1, leapfrog dt = 0.005.
"""
import math

import numpy as np

from . import _lib

H, HL, RC, TOL = 2.0, 4.0, 4.0, 1.0
A_VDW, B_VDW, KB = 2.0, 0.5, 1.0
DT = 0.005
SKIN = 0.5  # neighbour-list skin (lists cover r_cut + skin)


def params(L, eta=1.0, zeta=0.0, kcond=5.0, terms=_lib.ALL_TERMS, precision=_lib.FP64,
           h=H, hl=HL, rc=RC, tol=TOL, a=A_VDW, b=B_VDW, kB=KB, dt=DT, Ly=None, skin=SKIN):
    """sph_params for a periodic square (or Lx x Ly) box."""
    return dict(Lx=float(L), Ly=float(L if Ly is None else Ly), h=h, hl=hl, rc=rc, tol=tol,
                a=a, b=b, kB=kB, eta=eta, zeta=zeta, kcond=kcond, dt=dt, skin=skin,
                terms=int(terms), precision=int(precision))


def to_struct(p):
    s = _lib.Params()
    for k, v in p.items():
        setattr(s, k, v)
    return s


def lattice(n, L, rng, jitter=0.1):
    """First n sites of a ceil(sqrt(n))-column square lattice of spacing
    L / cols at cell centres, each jittered by U(-jitter, jitter) * spacing
    per component and wrapped into [0, L)."""
    cols = int(math.ceil(math.sqrt(n)))
    sp = L / cols
    k = np.arange(n)
    r = np.empty((n, 2))
    r[:, 0] = (k % cols + 0.5) * sp
    r[:, 1] = (k // cols + 0.5) * sp
    if jitter:
        r += rng.uniform(-jitter, jitter, size=(n, 2)) * sp
    r = np.mod(r, L)
    r[r >= L] -= L  # np.mod can round up to L
    return np.ascontiguousarray(r)


def velocities(n, T, rng, m=1.0, kB=KB):
    """Gaussian velocities at temperature T with the net momentum removed."""
    v = rng.normal(0.0, math.sqrt(kB * T / m), size=(n, 2))
    v -= v.mean(axis=0)
    return np.ascontiguousarray(v)


class System:
    """Initial state of one box: r, v, m, the temperature used to set
    u = kB T - a rho on device, and its parameters."""

    def __init__(self, r, v, m, T, prm, steps, seed):
        self.r, self.v, self.m, self.T, self.prm = r, v, m, T, prm
        self.steps, self.seed = steps, seed
        self.n = r.shape[0]


def square_system(nside, T, seed, steps, **pkw):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = nside * nside
    L = float(nside)
    r = lattice(n, L, rng)
    v = velocities(n, T, rng)
    return System(r, v, np.ones(n), T, params(L, **pkw), steps, seed)


def c1():
    """Plain FP64, L=32, N=1024, inviscid and non-conducting, 100 steps."""
    return square_system(32, 0.9, 1, 100, eta=0.0, zeta=0.0, kcond=0.0,
                         terms=_lib.PRESSURE | _lib.COHESION)


def c2(nside=512):
    """Plain FP64 on one GPU, L=512, N=262,144, eta=1, k=5, 1000 steps."""
    return square_system(nside, 0.9, 2, 1000)


def c3(nside=1024, precision=_lib.FP32C):
    """Quench: initial state at T=1.2, velocities rescaled by sqrt(0.4/1.2)
    at step 0 and u reset to kB*0.4 - a*rho (the system starts at T=0.4 with
    the T=1.2 configuration)."""
    s = square_system(nside, 1.2, 3, 2000, precision=precision)
    s.v = np.ascontiguousarray(s.v * math.sqrt(0.4 / 1.2))
    s.T = 0.4
    return s


def lognormal_box(seed, nlo, nhi, tlo, thi, steps=None, slo=None, shi=None, precision=_lib.FP32C):
    """One box of the batched / served workloads (density 1, L = sqrt(N))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(round(math.exp(rng.uniform(math.log(nlo), math.log(nhi)))))
    if steps is None:
        steps = int(rng.integers(slo, shi + 1))
    T = float(rng.uniform(tlo, thi))
    L = math.sqrt(n)
    r = lattice(n, L, rng)
    v = velocities(n, T, rng)
    return System(r, v, np.ones(n), T, params(L, precision=precision), steps, seed)


def c4_box(b, precision=_lib.FP32C):
    return lognormal_box(1000 + b, 256, 16384, 0.3, 1.0, steps=500, precision=precision)


def c5_job(j, precision=_lib.FP32C):
    return lognormal_box(j, 1024, 65536, 0.3, 1.2, slo=100, shi=2000, precision=precision)


def c5_arrivals(njobs, rate=200.0, seed=5):
    """Arrival times (s) of a Poisson stream."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.cumsum(rng.exponential(1.0 / rate, size=njobs))
