"""SpamComplete with the reference's surface (spam_complete_force.py:19-183),
on the GPU.

SpamComplete(particles, neighbour_list, adash=2.0, bdash=0.5, kbdash=1.0,
sigma=0.0, rcoef=0.0, cgrad=1.0, eta=1.0, zeta=0.1, kernel_type=2,
cutoff=5.0).apply() computes, in one device evaluation: short and long
summation densities, T from u (clamped at 0), the repulsive pressure
tensor p I (h), the cohesive pco I (h_l), the viscous stress
Pi = -eta (G + G^T - tr G I) - zeta tr G I from the velocity gradient G,
the heat flux q = -thermalk grad T, and the rates

    a_i    = sum_j m_j (P_i/rho_i^2 + P_j/rho_j^2) . dW_h + (pco terms) . dW_hl
    udot_i = sum_j m_j [ dv . f / 2 + (q_i/rho_i^2 + q_j/rho_j^2) . dW_h ]

It overwrites p.vdot and p.udot and writes rho, rho_lr, P (= p_rev +
p_rev_lr + pi_irr), p, pco, t, jq and the pair kernels, as
spam_complete_force.py:171-183 does.  The Fortran sphforce3d it replaces
is not in the reference tree, so these formulas are this library's; the
capillary (cgrad) and repulsive-core (sigma, rcoef) terms are not
implemented, and a non-zero value only triggers a warning.
"""
import warnings

from . import _bind, _lib
from . import properties
from .forces import Force

thermalk = properties.thermalk


class SpamComplete(Force):
    """All smooth-particle properties and rates in one device pass."""

    def __init__(self, particles, neighbour_list, adash=2.0, bdash=0.5, kbdash=1.0, sigma=0.0, rcoef=0.0,
                 cgrad=1.0, eta=1.0, zeta=0.1, kernel_type=2, cutoff=5.0, thermalk=None):
        Force.__init__(self, particles, neighbour_list, cutoff=cutoff)
        if kernel_type != 2:
            raise NotImplementedError("only the Lucy kernel (kernel_type=2) is on the GPU path")
        if sigma or rcoef or cgrad:
            warnings.warn("SpamComplete: the capillary (cgrad) and repulsive-core (sigma, rcoef) terms are "
                          "not implemented on the GPU path and are ignored", stacklevel=2)
        self.adash, self.bdash, self.kbdash = adash, bdash, kbdash
        self.sigma, self.rcoef, self.cgrad = sigma, rcoef, cgrad
        self.eta = eta
        self.zeta = zeta
        self.kernel_type = 2
        self.thermalk = properties.thermalk if thermalk is None else thermalk

    def device_params(self, dt=None):
        prm, _ = _bind.params_for(self.p, self.nl, terms=_lib.ALL_TERMS, eta=self.eta, zeta=self.zeta,
                                  kcond=self.thermalk, a=self.adash, b=self.bdash, kB=self.kbdash, dt=dt)
        return prm

    def apply(self):
        p, nl = self.p, self.nl
        n = p.n
        if n == 0:
            return
        prm = self.device_params()
        dev = _bind.device_box(p, prm)
        _bind.upload(p, dev)
        dev.rates(_lib.ALL_TERMS)
        out = dev.download("rho", "rho_lr", "udot", "P", "p", "pco", "acc", "T", "jq")
        p.rho[0:n] = out["rho"]
        p.rho_lr[0:n] = out["rho_lr"]
        p.udot[0:n] = out["udot"]
        p.P[0:n, :2, :2] = out["P"]
        p.p[0:n] = out["p"]
        p.pco[0:n] = out["pco"]
        p.vdot[0:n, :2] = out["acc"]
        p.t[0:n] = out["T"]
        p.jq[0:n, :2] = out["jq"]
        if nl is not None and getattr(nl, "nip", 0):
            _bind.pair_kernels(nl, prm["h"], prm["hl"])
