"""Smoothing kernels with the reference's surface (spkernel.py).

    kernel(r, dx, h, type)     spkernel.py:22-37
    lucy_kernel(r, dx, h)      spkernel.py:86-118   (on the GPU)
    lucy_w(r, h)               spkernel.py:54-63    (on the GPU, 2-D)
    lucy_w3d(r, h)             spkernel.py:65-83    (on the GPU, 3-D)
    gauss_kernel, debrun_kernel  spkernel.py:39-52, 120-153

The Lucy kernel is the only kernel on the hot path; lucy_kernel evaluates it
with the device kernel (sph_kernel_lucy) and keeps the reference's return
convention: (w, dwdx) with dwdx a list of dim components, or the scalar 0.0
at r == 0 (spkernel.py:108-109).  lucy_kernel_array is the vectorised form
(np pairs in one launch).  The Gaussian and Debrun kernels are not on the
path and are evaluated with NumPy as in the reference.
"""
import math

import numpy as np

from . import _lib

ktable = {1: "gaussian", 2: "lucy", 3: "debrun"}


def lucy_kernel_array(r, dx, h):
    """W and dW/dx for arrays r[np], dx[np, dim] (dim 1..3) on the GPU."""
    r = np.ascontiguousarray(np.atleast_1d(r), dtype=np.float64)
    dx = np.asarray(dx, dtype=np.float64)
    if dx.ndim == 1:
        dx = dx.reshape(r.shape[0], -1)
    dx = np.ascontiguousarray(dx)
    npair, dim = dx.shape
    w = np.empty(npair)
    g = np.empty((npair, dim))
    _lib.check(_lib.lib.sph_kernel_lucy(_lib.context(), npair, _lib.dptr(r), _lib.dptr(dx), float(h), dim,
                                        _lib.dptr(w), _lib.dptr(g)))
    return w, g


def lucy_kernel(r, dx, h):
    """The Lucy kernel for one pair.  Returns W, dwdx; works for one, two and
    three dimensions (dim = len(dx))."""
    dx = np.atleast_1d(np.asarray(dx, dtype=np.float64))
    dim = dx.size
    if dim not in (1, 2, 3):
        raise ValueError("lucy_kernel: dim must be 1, 2 or 3 (got %d)" % dim)
    rr = abs(float(r))
    w, g = lucy_kernel_array(np.array([rr]), dx.reshape(1, dim), h)
    if rr < h and rr == 0:
        return float(w[0]), 0.0
    if rr >= h:
        return 0, [0] * dim
    return float(w[0]), [float(x) for x in g[0]]


def lucy_w(r, h):
    """The Lucy kernel W only, 2-D."""
    w, _ = lucy_kernel_array(np.array([abs(float(r))]), np.zeros((1, 2)), h)
    return float(w[0]) if abs(r) < h else 0


def lucy_w3d(r, h):
    """The Lucy kernel W only, 3-D, as a one-element array."""
    w, _ = lucy_kernel_array(np.array([float(r)]), np.array([[float(r), 0.0, 0.0]]), h)
    return w


def gauss_kernel(r, dx, h):
    """Gaussian kernel, any dimension.  Returns w, dwdx (spkernel.py:39-52,
    including its sign convention)."""
    dx = np.atleast_1d(dx)
    dim = dx.size
    q = r / h
    factor = 1.e0 / ((h ** dim) * (np.pi ** (dim / 2.)))
    w = factor * math.exp(-q * q)
    dwdx = [w * ((-2.e0) * (-1.e0) * dx[i] / (h * h)) for i in range(dim)]
    return w, dwdx


def debrun_kernel(r, dx, h):
    """Debrun spiky kernel (spkernel.py:120-153)."""
    dx = np.atleast_1d(dx)
    dim = dx.size
    if r < 0:
        r = abs(r)
    if r < h:
        if dim == 1:
            q = (((-7. * (h ** 4) / 4) + 2 * (h ** 4)) ** (-1)) / 2
        elif dim == 2:
            q = 10.0 / (math.pi * (h ** 6))
        else:
            q = 15.0 / (math.pi * (h ** 6))
        w = q * ((h - r) ** 3)
        if r == 0:
            dwdx = 0.0
        else:
            dwdx = [q * (-3.0 * (h ** 2) + 6.0 * h * r - 3.0 * (r ** 2)) * (dx[i] / r) for i in range(dim)]
    else:
        dwdx = [0] * dim
        w = 0
    return w, dwdx


def kernel(r, dx, h, type):
    """Kernel and kernel gradient for one pair; type 'gaussian', 'lucy' or 'debrun'."""
    if type == "gaussian":
        return gauss_kernel(r, dx, h)
    if type == "lucy":
        return lucy_kernel(r, dx, h)
    if type == "debrun":
        return debrun_kernel(r, dx, h)
    raise ValueError("unknown kernel type %r" % (type,))
