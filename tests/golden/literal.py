"""Pure-Python restatement of the reference's FP64 Python path, written line
by line after the reference's own loops (2-D).  Test infrastructure only: it
produces the golden vectors in this folder and pins the C oracle on small
systems.  Independent of oracle/oracle.c (different enumeration, the
reference's polynomial kernel gradient, Python floats throughout).

    VerletList.build                 neighbour_list.py:160-189
    NeighbourList.minimum_image      neighbour_list.py:105-123
    NeighbourList.separations        neighbour_list.py:63-83
    spkernel.lucy_kernel             spkernel.py:86-118
    properties.spam_properties (+ long h as spam_properties_ls)
                                     properties.py:63-120, 138-161
    vdw, vdw_energy, vdw_temp        properties.py:38-49
    SpamForce / CohesiveSpamForce    forces.py:327-405 (vdot with m_j)
"""
from math import pi, sqrt


def minimum_image(dr, xmax, ymax):
    drx, dry = dr
    if drx > xmax / 2.:
        drx = drx - xmax
    if dry > ymax / 2.:
        dry = dry - ymax
    if drx < -xmax / 2.:
        drx = drx + xmax
    if dry < -ymax / 2.:
        dry = dry + ymax
    return [drx, dry]


def verlet_build(r, xmax, ymax, cutoff, tolerance):
    cutoff_radius_sq = cutoff ** 2
    tolerance_sq = tolerance * tolerance
    iap = []
    n = len(r)
    for i in range(n):
        for j in range(i + 1, n):
            d = minimum_image([r[j][0] - r[i][0], r[j][1] - r[i][1]], xmax, ymax)
            rsquared = d[0] * d[0] + d[1] * d[1]
            if rsquared < (cutoff_radius_sq + tolerance_sq):
                iap.append((i, j))
    return iap


def separations(r, v, iap, xmax, ymax):
    drij, rij, dv = [], [], []
    for (i, j) in iap:
        d = minimum_image([r[j][0] - r[i][0], r[j][1] - r[i][1]], xmax, ymax)
        rsquared = d[0] * d[0] + d[1] * d[1]
        drij.append(d)
        rij.append(sqrt(rsquared))
        dv.append([v[j][0] - v[i][0], v[j][1] - v[i][1]])
    return drij, rij, dv


def lucy_kernel(r, dx, h):
    dim = len(dx)
    if dim == 1:
        q = 5. / (4. * h)
    elif dim == 2:
        q = 5. / (pi * h ** 2)
    else:
        q = 105. / (pi * 16. * (h ** 3))
    if r < 0:
        r = abs(r)
    if r < h:
        w = q * (1 + 3. * r / h) * ((1. - r / h)) ** 3
        if r == 0:
            dwdx = [0.0] * dim
        else:
            dwdx = [q * ((-12. / (h ** 4)) * (r ** 3) + (24. / (h ** 3)) * (r ** 2)
                         - (12. * r / (h ** 2))) * dx[c] / r for c in range(dim)]
    else:
        w = 0
        dwdx = [0] * dim
    return w, dwdx


def vdw(rho, t, a, b, kb):
    return (rho * kb * t) / (1 - rho * b), -a * rho * rho


def vdw_energy(rho, t, a, kb):
    return t * kb - a * rho


def vdw_temp(rho, u, a, kb):
    return (u + a * rho) / kb


def properties(m, iap, rij, drij, u, h, hl, a, b, kb):
    n = len(m)
    zerokern = lucy_kernel(0.0, (0.0, 0.0), h)[0]
    zerokern_lr = lucy_kernel(0.0, (0.0, 0.0), hl)[0]
    rho = [zerokern] * n
    rho_lr = [zerokern_lr] * n
    wij, dwij, wij_lr, dwij_lr = [], [], [], []
    for k, (i, j) in enumerate(iap):
        w, dw = lucy_kernel(rij[k], drij[k], h)
        wl, dwl = lucy_kernel(rij[k], drij[k], hl)
        wij.append(w); dwij.append(dw); wij_lr.append(wl); dwij_lr.append(dwl)
        rho[i] += w * m[j]
        rho[j] += w * m[i]
        rho_lr[i] += wl * m[j]
        rho_lr[j] += wl * m[i]
    T, p, pco = [], [], []
    for i in range(n):
        t = vdw_temp(rho[i], u[i], a, kb)
        t = t if t > 0.0 else 0.0
        T.append(t)
        pp, _ = vdw(rho[i], t, a, b, kb)
        p.append(pp)
        pco.append(-a * rho_lr[i] * rho_lr[i])
    return rho, rho_lr, T, p, pco, dwij, dwij_lr


def spam_forces(m, iap, dv, rho, rho_lr, p, pco, dwij, dwij_lr):
    """SpamForce + CohesiveSpamForce accumulation (forces.py:338-405)."""
    n = len(m)
    vdot = [[0.0, 0.0] for _ in range(n)]
    udot = [0.0] * n
    for k, (i, j) in enumerate(iap):
        for (pi_, pj_, ri, rj, dw) in ((p[i], p[j], rho[i], rho[j], dwij[k]),
                                       (pco[i], pco[j], rho_lr[i], rho_lr[j], dwij_lr[k])):
            ps = (pi_ / ri ** 2 + pj_ / rj ** 2)
            ax = ps * dw[0]
            ay = ps * dw[1]
            vdot[i][0] += ax * m[j]
            vdot[i][1] += ay * m[j]
            vdot[j][0] -= ax * m[i]
            vdot[j][1] -= ay * m[i]
            du = 0.5 * (ax * dv[k][0] + ay * dv[k][1])
            udot[i] += du * m[j]
            udot[j] += du * m[i]
    return vdot, udot
