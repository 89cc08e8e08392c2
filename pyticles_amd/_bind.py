"""Glue between the reference's duck-typed particle / neighbour-list objects
and a device box.

The reference keeps state in NumPy arrays on a particle object
(particles.py:106-131, 325-344) and pair data on a neighbour list
(neighbour_list.py:29-44).  Every surface call here uploads the particle
state that the call reads, runs the pass on the GPU through the C ABI and
writes the results back into the caller's arrays in place, as the
reference's Cython and Fortran wrappers do (c_forces.pyx:113-115,
spam_complete_force.py:171-183).
"""
import math

import numpy as np

from . import _lib
from .device import DeviceBox

# EOS constants used when a particle object does not carry its own
# (properties.py:18-20)
ADASH, BDASH, KBDASH = 2.0, 0.5, 1.0
INF = float("inf")


def box_sides(p):
    """(Lx, Ly) of the particle system's box (box.py:18-22)."""
    bx = getattr(p, "box", None)
    if bx is None:
        raise ValueError("particle system has no box; the GPU path needs a periodic box")
    return float(bx.xmax), float(bx.ymax)


def check_dim(p):
    if getattr(p, "dim", 2) != 2:
        raise NotImplementedError("the GPU path is 2-D (particle arrays [maxn, 2]); got dim=%s" % p.dim)


def uniform(a, n, name):
    a = np.asarray(a[:n], dtype=np.float64)
    if n and np.any(a != a[0]):
        raise NotImplementedError("per-particle %s must be uniform on the GPU path" % name)
    return float(a[0]) if n else 1.0


def list_radius(nl):
    """(rc, tol) of a neighbour list: VerletList cutoff/tolerance
    (neighbour_list.py:153-157) or the unbounded brute list (:46-56)."""
    if nl is not None and hasattr(nl, "cutoff_radius"):
        return float(nl.cutoff_radius), math.sqrt(float(nl.tolerance_sq))
    return INF, 0.0


def params_for(p, nl=None, terms=_lib.ALL_TERMS, eta=0.0, zeta=0.0, kcond=0.0, h=None, hl=None,
               a=None, b=None, kB=None, dt=None):
    check_dim(p)
    n = p.n
    Lx, Ly = box_sides(p)
    rc, tol = list_radius(nl)
    hh = uniform(p.h, n, "h") if h is None else float(np.asarray(h).ravel()[0])
    hhl = uniform(p.hlr, n, "hlr") if hl is None else float(np.asarray(hl).ravel()[0])
    if math.isinf(rc):
        # an unbounded list still only interacts within the kernel supports
        rc, tol = max(hh, hhl) * 2.0, 0.0
        pair_rc = INF
    else:
        pair_rc = rc
    return dict(Lx=Lx, Ly=Ly, h=hh, hl=hhl, rc=rc, tol=tol,
                a=ADASH if a is None else a, b=BDASH if b is None else b, kB=KBDASH if kB is None else kB,
                eta=eta, zeta=zeta, kcond=kcond, dt=float(getattr(p, "dt", 0.0) if dt is None else dt),
                skin=0.5, terms=int(terms), precision=_lib.FP64), pair_rc


def wrapped(r, Lx, Ly):
    """Positions folded into [0, L) (a PeriodicBox may leave a coordinate
    exactly at L, box.py:36-47)."""
    out = np.array(r, dtype=np.float64, copy=True)
    out[:, 0] = np.mod(out[:, 0], Lx)
    out[:, 1] = np.mod(out[:, 1], Ly)
    out[out[:, 0] >= Lx, 0] -= Lx
    out[out[:, 1] >= Ly, 1] -= Ly
    return np.ascontiguousarray(out)


def device_box(p, prm):
    """The DeviceBox cached on particle object p: recreated when the particle
    count changes, re-parameterised (no reallocation) when only the
    parameters do."""
    cur = getattr(p, "_sph_dev", None)
    if cur is not None and cur[0] == p.n:
        dev = cur[1]
        if dev.prm != prm:
            dev.set_params(**prm)
        return dev
    if cur is not None:
        cur[1].close()
    dev = DeviceBox(p.n, prm)
    p._sph_dev = (p.n, dev)
    return dev


def upload(p, dev, u=None, T=None):
    """State into the device box: r, v, m, and either u or T (u = kB T - a rho
    set on device, properties.py:119)."""
    n = p.n
    Lx, Ly = box_sides(p)
    dev.upload(r=wrapped(p.r[:n], Lx, Ly), v=np.ascontiguousarray(p.v[:n], dtype=np.float64),
               m=np.ascontiguousarray(p.m[:n], dtype=np.float64),
               u=np.ascontiguousarray((p.u if u is None else u)[:n], dtype=np.float64))
    if T is not None:
        dev.init_thermal(np.ascontiguousarray(np.asarray(T, dtype=np.float64)[:n]))


def ensure_pair_capacity(nl, nip):
    """Grow the neighbour list's pair arrays to hold nip pairs (the reference
    preallocates maxn^2/2 - 1 rows, neighbour_list.py:32, which is impossible
    at the sizes this path serves)."""
    cap = nl.iap.shape[0]
    if nip <= cap:
        return
    new = max(nip, int(cap * 1.3) + 16)
    dim = nl.drij.shape[1]

    def grow(a, shape):
        b = np.zeros(shape, dtype=a.dtype)
        b[: a.shape[0]] = a
        return b

    nl.iap = grow(nl.iap, (new, 2))
    nl.rij = grow(nl.rij, (new,))
    nl.rsq = grow(nl.rsq, (new,))
    nl.drij = grow(nl.drij, (new, dim))
    nl.dv = grow(nl.dv, (new, dim))
    nl.wij = grow(nl.wij, (new,))
    nl.wij_lr = grow(nl.wij_lr, (new,))
    nl.dwij = grow(nl.dwij, (new, dim))
    nl.dwij_lr = grow(nl.dwij_lr, (new, dim))
    nl.max_interactions = new


def pair_kernels(nl, h, hl):
    """wij, dwij (h) and wij_lr, dwij_lr (h_l) for the listed pairs from
    nl.rij and nl.drij (properties.py:88, f_properties.py:100-103)."""
    ni = nl.nip
    if ni == 0:
        return
    ctx = _lib.context()
    r = np.ascontiguousarray(nl.rij[:ni], dtype=np.float64)
    dx = np.ascontiguousarray(nl.drij[:ni, :2], dtype=np.float64)
    for hh, w, g in ((h, nl.wij, nl.dwij), (hl, nl.wij_lr, nl.dwij_lr)):
        wo = np.empty(ni)
        go = np.empty((ni, 2))
        _lib.check(_lib.lib.sph_kernel_lucy(ctx, ni, _lib.dptr(r), _lib.dptr(dx), float(hh), 2,
                                            _lib.dptr(wo), _lib.dptr(go)))
        w[:ni] = wo
        g[:ni, :2] = go
