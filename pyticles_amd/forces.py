"""Pair forces with the reference's FP64 Python surface (forces.py), on the GPU.

    Force(particles, nl, cutoff=100.0)                    forces.py:22-52
    SpamForce(particles, nl, cutoff=5.0)                  forces.py:321-368
    CohesiveSpamForce(particles, nl, cutoff=10.0)         forces.py:371-405
    SpamForce2d / CohesiveSpamForce2d                     forces.py:246-318

apply() accumulates into p.vdot and p.udot like the reference's per-pair
apply_force (the caller zeroes them, particles.py:548-550).  The pressures
p.p / p.pco and densities come from the particle object's state on the
device: the rates pass recomputes density and EOS from r and u.
Acceleration carries the neighbour mass m_j, as the momentum equation
requires; the reference omits it (forces.py:357-362), which is identical at
m = 1, the mass of every configuration.
HookesForce, CollisionForce and Gravity are not smooth-particle forces and
are not provided.
"""
from . import _bind, _lib
from . import properties as _props


class Force:
    """A generic pairwise force on the particles of a neighbour list."""

    terms = 0

    def __init__(self, particles, nl, cutoff=100.0):
        self.p = particles
        self.nl = nl
        self.cutoff = cutoff
        self.cutoffsq = cutoff * cutoff

    def _rates(self, terms, kcond=0.0):
        p = self.p
        n = p.n
        if n == 0:
            return None
        prm, _ = _bind.params_for(p, self.nl, terms=terms, kcond=kcond, a=_props.ADASH, b=_props.BDASH,
                                  kB=_props.KBDASH)
        dev = _bind.device_box(p, prm)
        _bind.upload(p, dev)
        dev.rates(terms)
        return dev.download("acc", "udot", "rho", "rho_lr", "T", "jq")

    def apply(self):
        """Apply the force to all particles in the neighbour list."""
        out = self._rates(self.terms)
        if out is None:
            return
        n = self.p.n
        self.p.vdot[:n, :2] += out["acc"]
        self.p.udot[:n] += out["udot"]
        return out

    def apply_sorted(self):
        return self.apply()


class SpamForce(Force):
    """Repulsive vdW pressure force and its work."""

    terms = _lib.PRESSURE

    def __init__(self, particles, neighbour_list, cutoff=5.0):
        Force.__init__(self, particles, neighbour_list, cutoff=cutoff)


class CohesiveSpamForce(Force):
    """Cohesive (long smoothing length) pressure force and its work."""

    terms = _lib.COHESION

    def __init__(self, particles, neighbour_list, cutoff=10.0):
        Force.__init__(self, particles, neighbour_list, cutoff=cutoff)


SpamForce2d = SpamForce
CohesiveSpamForce2d = CohesiveSpamForce
