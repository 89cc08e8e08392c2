"""The reference's Cython property module surface (c_properties.pyx),
computed on the GPU in FP64.

    spam_properties(p, nl, h)          c_properties.pyx:83-211
    lucy_kernel_3d(r, dx, h)           c_properties.pyx:50-80 (c_properties.pxd:1)
    vdw(rho, t)                        c_properties.pyx:40-43

The Cython module computes in float32, accumulates its drk separation
across pairs without resetting it (c_properties.pyx:150) and uses rho_j for
particle i in the velocity gradient (:166-188); those are not reproduced.
"""
import numpy as np

from . import properties as _props
from .spkernel import lucy_kernel_array

H = 3.0
ADASH = _props.ADASH
BDASH = _props.BDASH
KBDASH = _props.KBDASH
RHONAUGHT = _props.RHONAUGHT
ADKE = False
vdw = _props.vdw
calc_pressure = vdw
ideal_isothermal = _props.ideal_isothermal
art_water = _props.art_water


def lucy_kernel_3d(r, dx, h):
    """3-D Lucy kernel: (w, dwdx[3]) for one separation."""
    dx = np.asarray(dx, dtype=np.float64).reshape(1, 3)
    w, g = lucy_kernel_array(np.array([abs(float(r))]), dx, h)
    return float(w[0]), g[0].copy()


def spam_properties(p, nl, h):
    """Densities, EOS and velocity gradient with smoothing length h."""
    _props.spam_properties(p, nl, hs=np.full(max(p.n, 1), float(h)), hl=None)
