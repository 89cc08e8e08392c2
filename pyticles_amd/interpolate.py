"""SPH interpolation onto grids with the reference's surface (interpolate.py).

    smooth_point(r, dr, x, vol, h, type)                 interpolate.py:18-37
    splash_map_2d(x, y, r, m, rho, a, h, bounds, cutoff)  interpolate.py:39-88 (on the GPU)

splash_map_2d renders a particle field onto the grid points (x[i], y[j]):
Z[i, j] = sum_p a_p m_p / rho_p W(|dr|, h_p) over particles within the
cutoff, dr = (x_p - x[i], y[j] - y_p) with the minimum image of periods
(xmax, ymax) (the reference's ``neighbours.minimum_image``, absent from its
tree, is taken to be neighbour_list.py:105-123).  The device gathers per
pixel over a cell list (sph_splash_map_2d); the reference scatters per
particle over half-open pixel ranges, which drop the last pixel row and
column of each particle's range.  The two agree when cutoff >= max(h) +
pixel width (the dropped pixels then have zero weight); below that the
device also adds those edge pixels.
"""
import numpy as np

from . import _lib
from . import spkernel


def smooth_point(r, dr, x, vol, h, type):
    """Smoothed value <A> = sum_j vol_j A_j W(r_j, h_j) at one point, from
    its neighbours' distances r, separations dr, values x and volumes vol."""
    total = 0.0
    for j, rj in enumerate(np.atleast_1d(r)):
        total += x[j] * vol[j] * spkernel.kernel(rj, dr[j], h[j], type)[0]
    return total


def splash_map_2d(x, y, r, m, rho, a, h, bounds, cutoff):
    """Z[nx, ny] of field a on the grid (x, y); bounds = (xmin, xmax, ymin,
    ymax); h scalar or per particle."""
    x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
    y = np.ascontiguousarray(np.atleast_1d(y), dtype=np.float64)
    nx, ny = x.size, y.size
    if cutoff > nx:
        print("cutoff is too large this won't work")  # interpolate.py:59
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64)[:, :2])
    n = r.shape[0]
    m = np.ascontiguousarray(np.broadcast_to(np.asarray(m, dtype=np.float64), (n,)))
    rho = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, dtype=np.float64), (n,)))
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)))
    h = np.ascontiguousarray(np.broadcast_to(np.asarray(h, dtype=np.float64), (n,)))
    b = np.ascontiguousarray(np.asarray(bounds, dtype=np.float64)[:4])
    Z = np.empty((nx, ny))
    _lib.check(_lib.lib.sph_splash_map_2d(_lib.context(), n, _lib.dptr(r), _lib.dptr(m), _lib.dptr(rho), _lib.dptr(a),
                                          _lib.dptr(h), nx, _lib.dptr(x), ny, _lib.dptr(y), _lib.dptr(b),
                                          float(cutoff), _lib.dptr(Z)))
    return Z
