"""TEST INFRASTRUCTURE ONLY -- ctypes binding of the FP64 CPU oracle
(oracle/oracle.c).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg may import it; the product path never does.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
# Same source built with OpenMP: pair loops summed per thread, so results
# differ from the serial checker's by rounding.  Used only for the all-cores
# CPU baseline; every check uses the serial library.
LIB_OMP_PATH = os.path.join(HERE, "_build", "liboracle_omp.so")

PRESSURE, COHESION, VISCOSITY, CONDUCTION = 1, 2, 4, 8
ALL_TERMS = 15


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("Lx", "Ly", "h", "hl", "rc", "tol", "a", "b", "kB", "eta", "zeta", "kcond", "dt")]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_libs = {}
_override = []


@contextlib.contextmanager
def library(path):
    """Route the serial checker's calls to another build of oracle.c at
    path while the block runs (the mutation tests build copies of the
    source with one planted mistake each and check that the closed-form
    tests reject them)."""
    _override.append(path)
    try:
        yield
    finally:
        _override.pop()


def lib(threaded=False):
    path = LIB_OMP_PATH if threaded else (_override[-1] if _override else LIB_PATH)
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        D = ctypes.POINTER(ctypes.c_double)
        I = ctypes.POINTER(ctypes.c_int64)
        P = ctypes.POINTER(Params)
        i64 = ctypes.c_int64
        sig = {
            "orc_min_image": (None, [D, D, ctypes.c_double, ctypes.c_double]),
            "orc_pairs_brute": (i64, [i64, D, ctypes.c_double, ctypes.c_double, ctypes.c_double, i64, I]),
            "orc_pairs_cell": (i64, [i64, D, ctypes.c_double, ctypes.c_double, ctypes.c_double, i64, I]),
            "orc_separations": (None, [D, D, ctypes.c_double, ctypes.c_double, i64, I, D, D, D, D]),
            "orc_lucy": (None, [i64, D, D, ctypes.c_double, ctypes.c_int, D, D]),
            "orc_density": (None, [i64, D, i64, I, D, ctypes.c_double, ctypes.c_double, D, D]),
            "orc_eos": (None, [i64, D, D, D, ctypes.c_double, ctypes.c_double, ctypes.c_double, D, D, D]),
            "orc_gradients": (None, [i64, D, D, D, D, i64, I, D, D, ctypes.c_double, D, D]),
            "orc_derivatives": (i64, [i64, D, D, D, D, P, ctypes.c_uint, D, D, D, D, D, D, D, D]),
            "orc_run": (ctypes.c_int, [i64, D, D, D, D, P, ctypes.c_uint, i64, D]),
            "orc_init_u": (None, [i64, D, D, D, P, D]),
            "orc_num_threads": (ctypes.c_int, []),
            "orc_splash_map_2d": (None, [i64, D, D, D, D, D, ctypes.c_int, D, ctypes.c_int, D, D,
                                         ctypes.c_double, D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _libs[path] = L
    return _libs[path]


def _d(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def params(prm):
    p = Params()
    for k in ("Lx", "Ly", "h", "hl", "rc", "tol", "a", "b", "kB", "eta", "zeta", "kcond", "dt"):
        setattr(p, k, float(prm[k]))
    return p


def pairs(r, Lx, Ly, R2, method="cell", threaded=False):
    r = _c(r)
    n = r.shape[0]
    f = lib(threaded).orc_pairs_cell if method == "cell" else lib(threaded).orc_pairs_brute
    cap = max(16, 16 * n)
    while True:
        iap = np.empty((cap, 2), dtype=np.int64)
        nip = f(n, _d(r), Lx, Ly, R2, cap, _i(iap))
        if nip <= cap:
            return iap[:nip].copy()
        cap = nip


def separations(r, v, Lx, Ly, iap):
    r, v, iap = _c(r), _c(v), _c(iap, np.int64)
    nip = iap.shape[0]
    drij, dv, rij, rsq = np.empty((nip, 2)), np.empty((nip, 2)), np.empty(nip), np.empty(nip)
    lib().orc_separations(_d(r), _d(v), Lx, Ly, nip, _i(iap), _d(drij), _d(rij), _d(rsq), _d(dv))
    return drij, rij, rsq, dv


def min_image(dr, Lx, Ly):
    out = _c(dr).copy()
    x = ctypes.c_double()
    y = ctypes.c_double()
    for k in range(out.shape[0]):
        x.value, y.value = out[k, 0], out[k, 1]
        lib().orc_min_image(ctypes.byref(x), ctypes.byref(y), Lx, Ly)
        out[k] = (x.value, y.value)
    return out


def lucy(r, dx, h, dim):
    r = _c(r)
    dx = _c(dx).reshape(r.shape[0], dim)
    w, dwdx = np.empty(r.shape[0]), np.empty((r.shape[0], dim))
    lib().orc_lucy(r.shape[0], _d(r), _d(dx), h, dim, _d(w), _d(dwdx))
    return w, dwdx


def density(m, iap, rij, h, hl):
    m, iap, rij = _c(m), _c(iap, np.int64), _c(rij)
    n = m.shape[0]
    rho, rl = np.empty(n), np.empty(n)
    lib().orc_density(n, _d(m), iap.shape[0], _i(iap), _d(rij), h, hl, _d(rho), _d(rl))
    return rho, rl


def eos(rho, rho_lr, u, a, b, kB):
    rho, rho_lr, u = _c(rho), _c(rho_lr), _c(u)
    n = rho.shape[0]
    T, p, pco = np.empty(n), np.empty(n), np.empty(n)
    lib().orc_eos(n, _d(rho), _d(rho_lr), _d(u), a, b, kB, _d(T), _d(p), _d(pco))
    return T, p, pco


def gradients(m, rho, v, T, iap, rij, drij, h):
    m, rho, v, T = _c(m), _c(rho), _c(v), _c(T)
    iap, rij, drij = _c(iap, np.int64), _c(rij), _c(drij)
    n = m.shape[0]
    gv, gT = np.empty((n, 2, 2)), np.empty((n, 2))
    lib().orc_gradients(n, _d(m), _d(rho), _d(v), _d(T), iap.shape[0], _i(iap), _d(rij), _d(drij), h,
                        _d(gv), _d(gT))
    return gv, gT


def derivatives(r, v, m, u, prm, terms=ALL_TERMS):
    """All derived fields of one evaluation at the given state."""
    r, v, m, u = _c(r), _c(v), _c(m), _c(u)
    n = r.shape[0]
    out = dict(rho=np.empty(n), rho_lr=np.empty(n), T=np.empty(n), acc=np.empty((n, 2)),
               udot=np.empty(n), gradv=np.empty((n, 2, 2)), jq=np.empty((n, 2)), P=np.empty((n, 2, 2)))
    p = params(prm)
    nip = lib().orc_derivatives(n, _d(r), _d(v), _d(m), _d(u), ctypes.byref(p), terms, _d(out["rho"]),
                                _d(out["rho_lr"]), _d(out["T"]), _d(out["acc"]), _d(out["udot"]),
                                _d(out["gradv"]), _d(out["jq"]), _d(out["P"]))
    if nip < 0:
        raise MemoryError("oracle allocation failed")
    out["nip"] = nip
    rho, rl, T = out["rho"], out["rho_lr"], out["T"]
    out["p"] = (rho * prm["kB"] * T) / (1.0 - rho * prm["b"])
    out["pco"] = -prm["a"] * rl * rl
    return out


def init_u(r, m, T, prm):
    r, m = _c(r), _c(m)
    n = r.shape[0]
    T = _c(np.broadcast_to(np.asarray(T, dtype=np.float64), (n,)))
    u = np.empty(n)
    p = params(prm)
    lib().orc_init_u(n, _d(r), _d(m), _d(T), ctypes.byref(p), _d(u))
    return u


def run(r, v, m, u, prm, nsteps, terms=ALL_TERMS, energy=False, threaded=False):
    """nsteps leapfrog steps from a synchronised state; returns (r, v, u[, E]).
    threaded=True runs the OpenMP build (CPU baseline only)."""
    r, v, m, u = _c(r).copy(), _c(v).copy(), _c(m), _c(u).copy()
    n = r.shape[0]
    E = np.empty(nsteps + 1) if energy else None
    p = params(prm)
    rc = lib(threaded).orc_run(n, _d(r), _d(v), _d(m), _d(u), ctypes.byref(p), terms, nsteps, _d(E))
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return (r, v, u, E) if energy else (r, v, u)


def num_threads(threaded=False):
    """Threads the oracle's pair loops use (1 for the serial checker)."""
    return int(lib(threaded).orc_num_threads())


def splash_map_2d(x, y, r, m, rho, a, h, bounds, cutoff):
    """interpolate.splash_map_2d restated (Z[nx, ny])."""
    r, m, rho, a = _c(r), _c(m), _c(rho), _c(a)
    n = r.shape[0]
    h = _c(np.broadcast_to(np.asarray(h, dtype=np.float64), (n,)))
    x, y = _c(x), _c(y)
    Z = np.empty((x.size, y.size))
    lib().orc_splash_map_2d(n, _d(r), _d(m), _d(rho), _d(a), _d(h), x.size, _d(x), y.size, _d(y),
                            _d(_c(bounds)), float(cutoff), _d(Z))
    return Z
