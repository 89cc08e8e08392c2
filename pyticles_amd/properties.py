"""SPH properties with the reference's surface (properties.py, and the
4-argument form of f_properties.spam_properties).

    spam_properties(p, nl[, hs, hl])   properties.py:63-120, f_properties.py:16-147
    vdw, vdw_energy, vdw_temp, ideal_isothermal, art_water, calc_pressure
                                       properties.py:24-52
    hamiltonian(p)                     properties.py:55-60
    ADASH, BDASH, KBDASH, RHONAUGHT, ADKE  properties.py:18-22

spam_properties runs on the GPU: summation density with h and h_l (self
term W(0, h) not times m_i), u = kB T - a rho from the particle
temperature (properties.py:119), T, p, pco from the vdW EOS, the velocity
gradient and the heat flux q = -thermalk grad T, and the per-pair kernel
values of the neighbour list.  The EOS helpers are the reference's
pointwise formulas (the device evaluates the same ones inside the density
pass).
"""
from . import _bind, _lib

ADASH = 2.0
BDASH = 0.5
KBDASH = 1.0
RHONAUGHT = 1.0
ADKE = False
thermalk = 1.0  # heat conductivity of the heat-flux output (f_properties.py:14)


def ideal_isothermal(rho, t):
    """Kinetic (isothermal) equation of state."""
    return rho * KBDASH


def art_water(rho, t):
    """Isothermal EOS with a reference density."""
    return (rho - RHONAUGHT) * KBDASH


def vdw(rho, t):
    """Van der Waals repulsive and cohesive pressure."""
    return (rho * KBDASH * t) / (1 - rho * BDASH), - ADASH * rho * rho


def vdw_energy(rho, t):
    """Van der Waals internal energy."""
    return t * KBDASH - ADASH * rho


def vdw_temp(rho, u):
    return (u + ADASH * rho) / KBDASH


calc_pressure = vdw


def hamiltonian(p):
    """Prints sum m v^2 / 2 + sum u (properties.py:55-60)."""
    n = p.n
    H = 0.5 * float((p.m[:n, None] * p.v[:n] * p.v[:n]).sum()) + float(p.u[:n].sum())
    print(H)
    return H


def spam_properties(p, nl, hs=None, hl=None):
    """Densities, EOS, velocity gradient, heat flux and pair kernels of
    particle system p over neighbour list nl, on the GPU."""
    n = p.n
    if n == 0:
        return
    prm, _ = _bind.params_for(p, nl, terms=_lib.ALL_TERMS, eta=0.0, zeta=0.0, kcond=thermalk, h=hs, hl=hl,
                              a=ADASH, b=BDASH, kB=KBDASH)
    dev = _bind.device_box(p, prm)
    _bind.upload(p, dev, T=p.t)
    dev.gradients()
    out = dev.download("rho", "rho_lr", "T", "p", "pco", "u", "gradv", "jq")
    p.rho[:n] = out["rho"]
    if hasattr(p, "rho_lr"):
        p.rho_lr[:n] = out["rho_lr"]
    p.p[:n] = out["p"]
    p.pco[:n] = out["pco"]
    p.u[:n] = out["u"]
    p.t[:n] = out["T"]
    p.gradv[:n, :2, :2] = out["gradv"]
    if hasattr(p, "jq"):
        p.jq[:n, :2] = out["jq"]
    if nl is not None and getattr(nl, "nip", 0):
        _bind.pair_kernels(nl, prm["h"], prm["hl"])
