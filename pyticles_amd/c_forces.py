"""The reference's Cython force module surface (c_forces.pyx), on the GPU.

    SpamForce(p, nl).apply()          c_forces.pyx:54-122
    CohesiveSpamForce(p, nl).apply()  c_forces.pyx:125-182
    SpamConduction(p, nl).apply()     c_forces.pyx:185-239

SpamForce.apply first calls the 4-argument spam_properties with the
particle smoothing lengths (c_forces.pyx:66-67), then adds the pressure
force to p.vdot and its work to p.udot and sets p.rdot = p.v
(c_forces.pyx:113-115).  SpamConduction adds the heat-conduction term
u_i += sum_j m_j (q_i/rho_i^2 + q_j/rho_j^2) . dW_ij/dx_j-x_i with
q = -thermalk grad T: the sign that makes heat flow from hot to cold (the
Cython loop subtracts with the opposite gradient convention,
c_forces.pyx:232-237, while the Fortran heat flux it reads is fed -dwdx,
f_properties.py:130).
"""
from . import _lib
from . import forces as _f
from . import properties


class Force(_f.Force):
    def __init__(self, particles1, particles2, nl):
        _f.Force.__init__(self, particles1, nl)
        self.p1 = particles1
        self.p2 = particles2


class SpamForce(_f.Force):
    terms = _lib.PRESSURE

    def __init__(self, particles, nl):
        _f.Force.__init__(self, particles, nl)

    def apply(self):
        p = self.p
        n = p.n
        properties.spam_properties(p, self.nl, p.h[0:n], p.hlr[0:n])
        _f.Force.apply(self)
        p.rdot[0:n, :] = p.v[0:n, :]

    def apply_force(self, k):
        pass


class CohesiveSpamForce(_f.Force):
    terms = _lib.COHESION

    def __init__(self, particles, nl):
        _f.Force.__init__(self, particles, nl)

    def apply(self):
        p = self.p
        n = p.n
        _f.Force.apply(self)
        p.rdot[0:n, :] = p.v[0:n, :]


class SpamConduction(_f.Force):
    terms = _lib.CONDUCTION

    def __init__(self, particles, nl):
        _f.Force.__init__(self, particles, nl)

    def apply(self):
        out = self._rates(_lib.CONDUCTION, kcond=properties.thermalk)
        if out is None:
            return
        n = self.p.n
        self.p.udot[:n] += out["udot"]
