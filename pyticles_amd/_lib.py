"""ctypes binding of libpyticles_hip.so (the C ABI in include/pyticles_hip.h).

The library is the only compute path: if it is missing or fails to load,
importing this module raises.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PYTICLES_HIP_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("PYTICLES_HIP_LIB") or os.path.join(HERE, "libpyticles_hip.so")

SPH_OK = 0
SPH_ERR_ARG = 1
SPH_ERR_HIP = 2
SPH_ERR_NOMEM = 3
SPH_ERR_CAP = 4
SPH_ERR_STATE = 5
SPH_ERR_NUMERIC = 6

PRESSURE = 1
COHESION = 2
VISCOSITY = 4
CONDUCTION = 8
ALL_TERMS = 15

FP64 = 0
FP32C = 1

BOX_FREE, BOX_INIT, BOX_RUN, BOX_DONE = 0, 1, 2, 3


class SphError(RuntimeError):
    """A non-zero status from the C ABI, with the library's message."""

    def __init__(self, code, msg):
        super().__init__("pyticles_hip error %d: %s" % (code, msg))
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [(name, ctypes.c_double) for name in
                ("Lx", "Ly", "h", "hl", "rc", "tol", "a", "b", "kB", "eta", "zeta", "kcond", "dt",
                 "skin")] + \
               [("terms", ctypes.c_uint), ("precision", ctypes.c_int)]


_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.c_int64

# name -> argtypes (all return int unless listed in _RESTYPES)
_SIGS = {
    "sph_ctx_create": [ctypes.c_int, ctypes.POINTER(_P)],
    "sph_ctx_destroy": [_P],
    "sph_last_error": [],
    "sph_version": [],
    "sph_ctx_set_timing": [_P, ctypes.c_int],
    "sph_box_create": [_P, _I64, ctypes.POINTER(Params), ctypes.POINTER(_P)],
    "sph_box_destroy": [_P],
    "sph_box_set_params": [_P, ctypes.POINTER(Params)],
    "sph_box_upload": [_P, _D, _D, _D, _D],
    "sph_box_download": [_P] + [_D] * 13,
    "sph_nlist_build": [_P, _I64, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)],
    "sph_separations": [_P, _I64, ctypes.POINTER(ctypes.c_int64), _D, _D, _D, _D],
    "sph_minimum_image": [_P, _I64, _D, ctypes.c_double, ctypes.c_double],
    "sph_fp32_accumulate": [_P, _I64, _D, ctypes.c_int, _D, _D],
    "sph_kernel_lucy": [_P, _I64, _D, _D, ctypes.c_double, ctypes.c_int, _D, _D],
    "sph_density": [_P],
    "sph_gradients": [_P],
    "sph_rates": [_P, ctypes.c_uint],
    "sph_init_thermal": [_P, _D, ctypes.c_double],
    "sph_step": [_P, _I64],
    "sph_box_advance": [_P, _I64],
    "sph_box_sync": [_P],
    "sph_box_run_stats": [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                          ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)],
    "sph_energy": [_P, _D, _D, _D],
    "sph_box_timing": [_P, _D, ctypes.c_int, ctypes.c_int],
    "sph_box_time_pass": [_P, ctypes.c_int, ctypes.c_int, _D],
    "sph_box_set_graph": [_P, ctypes.c_int],
    "sph_arena_create": [_P, _I64, ctypes.c_int, ctypes.POINTER(_P)],
    "sph_arena_destroy": [_P],
    "sph_arena_add": [_P, _I64, ctypes.POINTER(Params), _D, _D, _D, ctypes.c_double, _I64,
                      ctypes.POINTER(ctypes.c_int)],
    "sph_arena_step": [_P, _I64],
    "sph_arena_poll": [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)],
    "sph_arena_poll_all": [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64),
                           ctypes.POINTER(ctypes.c_int)],
    "sph_arena_result": [_P, ctypes.c_int, _D, _D, _D, _D, _D, _D],
    "sph_arena_remove": [_P, ctypes.c_int],
    "sph_arena_stats": [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "sph_arena_set_graph": [_P, ctypes.c_int],
    "sph_arena_timing": [_P, _D, ctypes.c_int, ctypes.c_int],
    "sph_arena_set_schedule": [_P, ctypes.c_int],
    "sph_box_set_schedule": [_P, ctypes.c_int],
    "sph_arena_set_rates_unroll": [_P, ctypes.c_int],
    "sph_box_set_rates_unroll": [_P, ctypes.c_int],
    "sph_arena_set_pass_block": [_P, ctypes.c_int],
    "sph_box_set_pass_block": [_P, ctypes.c_int],
    "sph_arena_set_fused_step": [_P, ctypes.c_int],
    "sph_box_set_fused_step": [_P, ctypes.c_int],
    "sph_arena_set_list_bits": [_P, ctypes.c_int],
    "sph_box_set_list_bits": [_P, ctypes.c_int],
    "sph_arena_list_bits": [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    "sph_box_list_bits": [_P, ctypes.POINTER(ctypes.c_int)],
    "sph_arena_footprint": [_I64, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)],
    "sph_host_alloc": [_I64, ctypes.POINTER(_P)],
    "sph_host_free": [_P],
    "sph_box_snapshot": [_P, _D, _D, _D, _D],
    "sph_box_snapshot_wait": [_P],
    "sph_splash_map_2d": [_P, _I64, _D, _D, _D, _D, _D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_double,
                          _D],
    # device pointers as integers (c_void_p); bounds on the host
    "sph_splash_map_2d_noalloc": [_P, _I64, _P, _P, _P, _P, _P, ctypes.c_int, _P, ctypes.c_int, _P, _D,
                                  ctypes.c_double, _P],
}
_RESTYPES = {"sph_last_error": ctypes.c_char_p, "sph_version": ctypes.c_char_p}

# every symbol include/pyticles_hip.h declares
EXPORTED = tuple(_SIGS)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "pyticles_amd: %s is not built; run `python -m pyticles_amd.build` "
            "(or __graft_entry__.build()) first.  There is no CPU fallback." % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, ctypes.c_int)
    return lib


lib = _load()


def last_error():
    return lib.sph_last_error().decode()


def check(rc):
    if rc != SPH_OK:
        raise SphError(rc, last_error())
    return rc


def dptr(a):
    """Pointer to a C-contiguous float64 array (None -> NULL)."""
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_D)


def i64ptr(a):
    if a is None:
        return None
    if a.dtype != np.int64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous int64 array")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


_ctx = {}


def context(device=0):
    """Process-wide context (device + stream) per device, created on first use."""
    c = _ctx.get(device)
    if c is None:
        h = _P()
        check(lib.sph_ctx_create(device, ctypes.byref(h)))
        c = h
        _ctx[device] = c
    return c
