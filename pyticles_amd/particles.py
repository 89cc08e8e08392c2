"""Particle systems with the reference's attribute contract (particles.py),
stepped on the GPU.

    ParticleSystem(n, d=2, maxn, xmax, ymax, ...)        particles.py:69-253
    SmoothParticleSystem(n, d=2, maxn, ..., hshort, hlong, temperature)
                                                         particles.py:256-575

The arrays the hot path reads and writes keep the reference's names and
shapes ([maxn, dim] for r, v, rdot, vdot, jq; [maxn] for m, rho, rho_lr,
t, u, udot, h, hlr, p, pco; [maxn, dim, dim] for gradv and P), so the
reference's neighbour lists and forces drive them unchanged.
derivatives() is the reference's composition (particles.py:544-570):
zero the rates, refresh the separations of every list, apply every force.

update(dt) advances one step with the device leapfrog (kick-drift-kick,
periodic wrap) of the physics the registered forces select; run(nsteps, dt)
advances many steps without leaving the device.  The reference's
euler / ieuler / rk4 host integrators (integrator.py) and mirror walls are
outside the hot path and are not provided for GPU stepping.  The systems
are 2-D and periodic (box.PeriodicBox).
"""
import time

import numpy as np

from . import _bind, _lib
from . import box as _box
from . import properties

XMAX = 5
YMAX = 5
ZMAX = 1
VMAX = 0.1
SPROPS = False


class ParticleSystem:
    def __init__(self, n, d=2, maxn=125, controllers=(), xmax=XMAX, ymax=YMAX, zmax=ZMAX, vmax=VMAX,
                 simbox=None, mass=1.0, rinit=None, source=None, side=(5, 5, 5), integrator="leapfrog",
                 spacing=0.1, seed=None):
        if d != 2:
            raise NotImplementedError("the GPU path is 2-D (d=2)")
        self.n = n
        self.dim = d
        self.maxn = maxn
        self.dt = 0.0
        self.steps = 0
        self.box = simbox if simbox is not None else _box.PeriodicBox(p=self, xmax=xmax, ymax=ymax, zmax=zmax)
        if integrator not in ("leapfrog",):
            raise NotImplementedError("GPU stepping uses the device leapfrog; integrator=%r is not provided"
                                      % (integrator,))
        self.integrator = integrator
        rng = np.random.default_rng(seed)
        self.r = self.box.xmax * rng.random([self.maxn, self.dim])
        self.r[:, 1] = self.box.ymax * rng.random(self.maxn)
        self.m = np.zeros(self.maxn, dtype=float)
        self.v = vmax * (rng.random([self.maxn, self.dim]) - 0.5)
        self.rdot = np.zeros(self.r.shape)
        self.vdot = np.zeros(self.v.shape)
        self.mdot = np.zeros(self.m.shape)
        if rinit == "grid":
            k = np.arange(n)
            cols = int(np.ceil(np.sqrt(n)))
            sp = spacing if spacing else min(self.box.xmax, self.box.ymax) / cols
            self.r[:n, 0] = (k % cols + 0.5) * sp
            self.r[:n, 1] = (k // cols + 0.5) * sp
        self.m[:] = mass
        self.colour = 1.0, 0.0, 0.0
        self.nlists = []
        self.forces = []
        self.controllers = list(controllers)
        self.timing = {"force time": -1, "deriv time": -1, "pairsep time": -1, "update time": -1,
                       "integrate time": -1}

    def rebuild_lists(self):
        for nl in self.nlists:
            if nl.rebuild_list:
                nl.build()

    def derivatives(self):
        """Rates of change of every particle (particles.py:235-253)."""
        self.rdot = self.v
        self.vdot[:, :] = 0.0
        for nl in self.nlists:
            nl.separations()
        for force in self.forces:
            force.apply()
        for controller in self.controllers:
            controller.apply()


class SmoothParticleSystem(ParticleSystem):
    """A long / short smoothing-length smooth particle system."""

    def __init__(self, n, d=2, maxn=100, controllers=(), xmax=XMAX, ymax=YMAX, zmax=ZMAX, vmax=VMAX,
                 rinit=None, side=None, mass=1.0, spacing=None, temperature=1.0, thermostat_temp=1.0,
                 thermostat=False, hshort=1.0, hlong=3.0, source=None, integrator="leapfrog", simbox=None,
                 seed=None):
        ParticleSystem.__init__(self, n=n, d=d, xmax=xmax, ymax=ymax, zmax=zmax, rinit=rinit, side=side,
                                mass=mass, source=source, spacing=spacing, integrator=integrator, vmax=vmax,
                                maxn=maxn, simbox=simbox, seed=seed)
        self.rho = np.zeros(self.maxn)
        self.rho_lr = np.zeros(self.maxn)
        self.rhodot = np.zeros(self.rho.shape)
        self.gradv = np.zeros([self.maxn, self.dim, self.dim])
        self.jq = np.zeros([self.maxn, self.dim])
        self.t = np.ones(self.maxn)
        self.t[:] = temperature
        self.thermostat = thermostat
        self.u = np.ones(self.maxn, dtype=float)
        self.udot = np.zeros(self.maxn, dtype=float)
        self.h = np.zeros(self.maxn)
        self.hlr = np.zeros(self.maxn)
        self.h[:] = hshort
        self.hlr[:] = hlong
        self.p = np.zeros([self.maxn])
        self.pco = np.zeros([self.maxn])
        self.P = np.zeros([self.maxn, self.dim, self.dim])
        self.thermostat_temp = thermostat_temp
        self.timing["SPAM time"] = -1
        self.nl_default = None

    def apply_thermostat(self, target_temp):
        """Scale every temperature to the target mean and reset
        u = kB T - a rho (particles.py:450-457)."""
        n = self.n
        tav = self.t[:n].mean()
        self.t[:n] = self.t[:n] * (target_temp / tav)
        self.u[:n] = properties.vdw_energy(self.rho[:n], self.t[:n])

    def derivatives(self):
        """Zero the rates, refresh separations, [properties], apply the forces
        (particles.py:544-570), with the same timing keys."""
        self.rdot = self.v
        self.vdot[:, :] = 0.0
        self.udot[:] = 0.0
        t = time.time()
        for nl in self.nlists:
            nl.separations()
        self.timing["pairsep time"] = time.time() - t
        t = time.time()
        if SPROPS and self.nl_default is not None:
            properties.spam_properties(self, self.nl_default, self.h[0:self.n], self.hlr[0:self.n])
        self.timing["SPAM time"] = time.time() - t
        t = time.time()
        for force in self.forces:
            force.apply()
        self.timing["force time"] = time.time() - t

    # -- device stepping ----------------------------------------------------
    def _step_params(self, dt):
        terms, eta, zeta, kcond = 0, 0.0, 0.0, 0.0
        a, b, kB = properties.ADASH, properties.BDASH, properties.KBDASH
        nl = self.nl_default if self.nl_default is not None else (self.nlists[0] if self.nlists else None)
        for f in self.forces:
            name = type(f).__name__
            if name == "SpamComplete":
                terms |= _lib.ALL_TERMS
                eta, zeta, kcond = f.eta, f.zeta, f.thermalk
                a, b, kB = f.adash, f.bdash, f.kbdash
            elif name == "SpamConduction":
                terms |= _lib.CONDUCTION
                kcond = properties.thermalk
            else:
                terms |= int(getattr(f, "terms", 0))
        if terms == 0:
            raise ValueError("no smooth-particle force registered in p.forces")
        prm, _ = _bind.params_for(self, nl, terms=terms, eta=eta, zeta=zeta, kcond=kcond, a=a, b=b, kB=kB,
                                  dt=dt)
        return prm

    def run(self, nsteps, dt):
        """nsteps leapfrog steps on the device; the state arrays are updated
        at the end (r wrapped into the box)."""
        n = self.n
        t1 = time.time()
        self.dt = dt
        if n == 0:
            self.steps += int(nsteps)
            return
        prm = self._step_params(dt)
        if not getattr(self.box, "periodic", False):
            raise NotImplementedError("GPU stepping needs a periodic box")
        dev = _bind.device_box(self, prm)
        _bind.upload(self, dev)
        t = time.time()
        dev.step(int(nsteps))
        self.timing["integrate time"] = time.time() - t
        out = dev.download("r", "v", "u", "rho", "rho_lr", "T", "p", "pco", "acc", "udot")
        self.r[:n] = out["r"]
        self.v[:n] = out["v"]
        self.u[:n] = out["u"]
        self.rho[:n] = out["rho"]
        self.rho_lr[:n] = out["rho_lr"]
        self.t[:n] = out["T"]
        self.p[:n] = out["p"]
        self.pco[:n] = out["pco"]
        self.vdot[:n] = out["acc"]
        self.udot[:n] = out["udot"]
        self.rdot = self.v
        self.steps += int(nsteps)
        for nl in self.nlists:
            if hasattr(nl, "ponder_rebuild"):
                nl.ponder_rebuild()
        if self.thermostat:
            self.apply_thermostat(self.thermostat_temp)
        self.timing["update time"] = time.time() - t1

    def update(self, dt):
        """One device leapfrog step (particles.py:459-494 for the periodic
        case)."""
        self.rebuild_lists()
        self.run(1, dt)
