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
advances many steps in one call.  The state stays resident on the device
between calls: update() costs one rate evaluation and no transfer, and k
calls of update() give the bits of run(k).  The state arrays are fetched
when read: p.r alone (the reference's driver checks it every step,
toybox.py:101-106) keeps the device run open; reading any other state or
derived array closes it (the closing half kick synchronises v and u).  A
state array changed on the host (in place or by assignment) is uploaded at
the next step.  The device keeps its own neighbour lists, so a device step
does not touch the host lists; they are pondered (ponder_rebuild) when the
whole state is fetched.  The reference's euler / ieuler / rk4 host
integrators (integrator.py) and mirror walls are outside the hot path and
are not provided for GPU stepping.  The systems are 2-D and periodic
(box.PeriodicBox).
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


class _Resident:
    """A state array of a SmoothParticleSystem that its resident device run
    may hold newer values of.  Reading fetches them first (positions alone
    keep the run open); the read also marks the array as exposed, so the
    next step compares it with what the device holds and uploads it if the
    caller changed it."""

    def __init__(self, name, positions=False, device=True):
        self.name = name
        self.key = "_a_" + name
        self.positions = positions
        self.device = device  # False: the device never writes it (only exposure is tracked)

    def __get__(self, obj, cls=None):
        if obj is None:
            return self
        st = obj.__dict__.get("_res")
        if st is not None:
            if self.device:
                obj._pull(self.positions)
            st["exposed"].add(self.name)
        return obj.__dict__[self.key]

    def __set__(self, obj, value):
        st = obj.__dict__.get("_res")
        if st is not None:
            obj._pull(False)
            st["exposed"].add(self.name)
        obj.__dict__[self.key] = value


# state the device run writes back: (attribute, DeviceBox field)
_PULLED = (("v", "v"), ("u", "u"), ("rho", "rho"), ("rho_lr", "rho_lr"), ("t", "T"), ("p", "p"),
           ("pco", "pco"), ("vdot", "acc"), ("udot", "udot"))
# state uploaded to the device (compared against what it holds before a step)
_PUSHED = ("r", "v", "u", "m")


class SmoothParticleSystem(ParticleSystem):
    """A long / short smoothing-length smooth particle system."""

    r = _Resident("r", positions=True)
    v = _Resident("v")
    u = _Resident("u")
    m = _Resident("m")
    rho = _Resident("rho")
    rho_lr = _Resident("rho_lr")
    t = _Resident("t")
    p = _Resident("p")
    pco = _Resident("pco")
    vdot = _Resident("vdot")
    udot = _Resident("udot")
    h = _Resident("h", device=False)
    hlr = _Resident("hlr", device=False)

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

    def _step_key(self, dt):
        """What the stepping parameters depend on besides h and hlr."""
        fk = tuple((type(f).__name__, id(f), getattr(f, "eta", None), getattr(f, "zeta", None),
                    getattr(f, "thermalk", None), getattr(f, "adash", None), getattr(f, "bdash", None),
                    getattr(f, "kbdash", None), getattr(f, "terms", None)) for f in self.forces)
        nl = self.nl_default if self.nl_default is not None else (self.nlists[0] if self.nlists else None)
        nk = None if nl is None else (id(nl), getattr(nl, "cutoff_radius", None), getattr(nl, "tolerance_sq", None))
        return (float(dt), fk, nk, self.n, float(self.box.xmax), float(self.box.ymax))

    def _cached_step_params(self, dt):
        """_step_params, recomputed only when the forces, the list, the box,
        dt or the smoothing-length arrays (read or assigned since) change:
        the per-particle uniformity check of h and hlr is O(n)."""
        st = self.__dict__.get("_res")
        key = self._step_key(dt)
        if st is not None and st.get("prm_key") == key and not ({"h", "hlr"} & st["exposed"]):
            return st["prm"]
        prm = self._step_params(dt)
        st = self.__dict__.get("_res")
        if st is not None:
            st["exposed"] -= {"h", "hlr"}
            st["prm_key"], st["prm"] = key, prm
        return prm

    def _pull(self, positions):
        """Fetch what the device run holds newer: positions only (the run
        stays open), or the whole state (the run is closed first)."""
        st = self.__dict__.get("_res")
        if st is None or not st["ahead"]:
            return
        n = st["n"]
        d = self.__dict__
        dev = st["dev"]
        if positions:
            if not st["pos_fresh"]:
                r = dev.download("r")
                d["_a_r"][:n] = r
                st["snap"]["r"] = r
                st["pos_fresh"] = True
            return
        st["ahead"] = False  # before the downloads: the setters below may read
        out = dev.download("r", *[f for _, f in _PULLED])
        d["_a_r"][:n] = out["r"]
        for attr, f in _PULLED:
            d["_a_" + attr][:n] = out[f]
        self.rdot = d["_a_v"]
        st["snap"].update(r=out["r"].copy(), v=out["v"].copy(), u=out["u"].copy())
        st["pos_fresh"] = True
        for nl in self.nlists:
            if hasattr(nl, "ponder_rebuild"):
                nl.ponder_rebuild()

    def _resident(self, prm):
        """The device box holding this system's state for stepping; uploads
        the host state when the box is new or the caller changed a state
        array since the device last held it."""
        n = self.n
        d = self.__dict__
        st = d.get("_res")
        if st is not None and st["n"] != n:
            self._pull(False)  # the old box goes away with the old particle count
            st = d["_res"] = None
        dev = _bind.device_box(self, prm)
        fresh = st is None or st["dev"] is not dev or st["n"] != n
        changed = []
        if not fresh:
            for name in st["exposed"]:
                if name in _PUSHED and not np.array_equal(d["_a_" + name][:n], st["snap"][name]):
                    changed.append(name)
            st["exposed"] &= {"h", "hlr"}  # the step-parameter cache tracks these
        if fresh or changed:
            if st is not None and st["ahead"]:
                # keep what the device holds newer for the arrays not changed
                dev.sync()
                keep = [f for a, f in _PULLED if a not in changed]
                out = dev.download(*keep) if keep else {}
                if len(keep) == 1:
                    out = {keep[0]: out}
                for attr, f in _PULLED:
                    if f in out:
                        d["_a_" + attr][:n] = out[f]
                if "r" not in changed:
                    d["_a_r"][:n] = dev.download("r")
            Lx, Ly = _bind.box_sides(self)
            r = _bind.wrapped(d["_a_r"][:n], Lx, Ly)
            snap = dict(r=r, v=np.array(d["_a_v"][:n], dtype=np.float64),
                        u=np.array(d["_a_u"][:n], dtype=np.float64), m=np.array(d["_a_m"][:n], dtype=np.float64))
            dev.upload(r=snap["r"], v=snap["v"], m=snap["m"], u=snap["u"])
            d["_a_r"][:n] = r
            d["_res"] = dict(dev=dev, n=n, ahead=False, pos_fresh=True, snap=snap, exposed=set(),
                             prm_key=None if st is None else st.get("prm_key"),
                             prm=None if st is None else st.get("prm"))
        return dev

    def run(self, nsteps, dt):
        """nsteps leapfrog steps on the device (continuing a run update()
        left open); the state arrays are updated at the end (r wrapped into
        the box)."""
        n = self.n
        t1 = time.time()
        self.dt = dt
        if n == 0:
            self.steps += int(nsteps)
            return
        prm = self._step_params(dt)
        if not getattr(self.box, "periodic", False):
            raise NotImplementedError("GPU stepping needs a periodic box")
        dev = self._resident(prm)
        t = time.time()
        dev.step(int(nsteps))
        self.timing["integrate time"] = time.time() - t
        st = self.__dict__["_res"]
        st["ahead"], st["pos_fresh"] = True, False
        self._pull(False)
        self.steps += int(nsteps)
        if self.thermostat:
            self.apply_thermostat(self.thermostat_temp)
        self.timing["update time"] = time.time() - t1

    def update(self, dt):
        """One device leapfrog step (particles.py:459-494 for the periodic
        case): one rate evaluation, the state left resident on the device."""
        t1 = time.time()
        self.rebuild_lists()
        n = self.n
        self.dt = dt
        if n == 0:
            self.steps += 1
            return
        prm = self._cached_step_params(dt)
        if not getattr(self.box, "periodic", False):
            raise NotImplementedError("GPU stepping needs a periodic box")
        dev = self._resident(prm)
        t = time.time()
        dev.advance(1)
        self.timing["integrate time"] = time.time() - t
        st = self.__dict__["_res"]
        st["ahead"], st["pos_fresh"] = True, False
        self.steps += 1
        if self.thermostat:
            self.apply_thermostat(self.thermostat_temp)
        self.timing["update time"] = time.time() - t1
