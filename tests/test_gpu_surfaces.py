"""The reference's Python surfaces (neighbour_list, spkernel, properties,
c_properties, forces, c_forces, spam_complete_force, particles) driven the
way the reference's own tests and drivers drive them, on the GPU, checked
against the oracle.  The reference's tests are 3-D; these use the same
geometry in 2-D (test/neighbour_list_test.py, test/force_test.py)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from pyticles_amd import c_forces, c_properties, configs, forces, neighbour_list, particles, properties
from pyticles_amd import spam_complete_force, spkernel

pytestmark = pytest.mark.gpu


def rel_norm(a, b):
    s = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / (s if s > 0 else 1.0))


def three_particles(maxn=5):
    p = particles.SmoothParticleSystem(3, d=2, maxn=maxn, xmax=5, ymax=5)
    p.r[0, :] = (0.0, 0.0)
    p.r[1, :] = (1.0, 0.0)
    p.r[2, :] = (0.0, 1.0)
    return p


def test_nlist_reference_unit_test():
    # test/neighbour_list_test.py:8-24
    p = three_particles()
    nl = neighbour_list.NeighbourList(p)
    nl.build()
    nl.separations()
    k = nl.find_pair(0, 1)
    assert nl.rij[k] == 1.0
    assert nl.nip == 3 and nl.find_pair(1, 0) == -1


def test_verlet_list_reference_unit_test():
    # test/neighbour_list_test.py:26-53
    p = three_particles()
    nl = neighbour_list.VerletList(p, cutoff=10, tolerance=2)
    nl.build()
    nl.compress()
    nl.separations()
    assert nl.rij[0] == 1.0
    nl.ponder_rebuild()
    assert nl.rebuild_list is False
    p.r[0, :] = (100.0, 100.0)
    nl.ponder_rebuild()
    assert nl.rebuild_list is True


def _system(ns=24, seed=3, T=0.9, L=None):
    rng = np.random.Generator(np.random.PCG64(seed))
    L = float(ns) if L is None else L
    n = ns * ns
    p = particles.SmoothParticleSystem(n, d=2, maxn=n + 7, xmax=L, ymax=L, hshort=2.0, hlong=4.0,
                                       temperature=T)
    p.r[:n] = configs.lattice(n, L, rng)
    p.v[:n] = configs.velocities(n, T, rng)
    p.m[:] = 1.0
    return p, rng


def test_verlet_build_matches_oracle_and_sorted_list():
    p, _ = _system()
    n = p.n
    nl = neighbour_list.VerletList(p, cutoff=4.0, tolerance=1.0)
    nl.build()
    ref = O.pairs(p.r[:n], 24.0, 24.0, 17.0)
    assert np.array_equal(nl.iap[: nl.nip], ref)
    d0, r0, q0, v0 = O.separations(p.r[:n], p.v[:n], 24.0, 24.0, ref)
    assert np.array_equal(nl.drij[: nl.nip], d0) and np.array_equal(nl.rsq[: nl.nip], q0)
    nl.separations()
    assert np.array_equal(nl.rij[: nl.nip], r0) and np.array_equal(nl.dv[: nl.nip], v0)
    snl = neighbour_list.SortedVerletList(p, cutoff=4.0, tolerance=1.0)
    snl.build()
    snl.separations()
    assert np.all(np.diff(snl.rsq[: snl.nip]) <= 0)
    got = {tuple(x) for x in snl.iap[: snl.nip].tolist()}
    assert got == {tuple(x) for x in ref.tolist()}


def test_spkernel_surface():
    for h in (2.0, 4.0):
        w, g = spkernel.lucy_kernel(0.0, [0.0, 0.0], h)
        assert w == pytest.approx(5.0 / (math.pi * h * h), rel=1e-15) and g == 0.0
    w, g = spkernel.lucy_kernel(1.0, [1.0, 0.0], 2.0)
    w0, g0 = O.lucy(np.array([1.0]), np.array([[1.0, 0.0]]), 2.0, 2)
    assert w == pytest.approx(w0[0], rel=1e-14) and g[0] == pytest.approx(g0[0, 0], rel=1e-13)
    assert spkernel.kernel(2.5, [2.5, 0.0], 2.0, "lucy") == (0, [0, 0])
    w3, g3 = c_properties.lucy_kernel_3d(0.0, [0.0, 0.0, 0.0], 2.0)
    assert w3 == pytest.approx(105.0 / (16.0 * math.pi * 8.0), rel=1e-15)
    assert spkernel.lucy_w(0.5, 2.0) == pytest.approx(O.lucy(np.array([0.5]), np.zeros((1, 2)), 2.0, 2)[0][0])


def test_spam_properties_matches_oracle():
    p, _ = _system(seed=5)
    n = p.n
    nl = neighbour_list.VerletList(p, cutoff=4.0, tolerance=1.0)
    nl.build()
    nl.separations()
    T0 = 0.9 + 0.2 * np.sin(p.r[:n, 0] * 2 * np.pi / 24.0)  # a temperature gradient: non-zero heat flux
    p.t[:n] = T0
    properties.spam_properties(p, nl)
    prm = configs.params(24.0, eta=0.0, kcond=properties.thermalk)
    u0 = 1.0 * T0 - 2.0 * p.rho[:n]
    ref = O.derivatives(p.r[:n], p.v[:n], p.m[:n], u0, prm, terms=15)
    assert rel_norm(p.rho[:n], ref["rho"]) < 1e-13
    assert rel_norm(p.rho_lr[:n], ref["rho_lr"]) < 1e-13
    np.testing.assert_allclose(p.t[:n], T0, rtol=1e-13)
    assert rel_norm(p.p[:n], ref["p"]) < 1e-12 and rel_norm(p.pco[:n], ref["pco"]) < 1e-13
    assert rel_norm(p.gradv[:n], ref["gradv"]) < 1e-12
    assert rel_norm(p.jq[:n], ref["jq"]) < 1e-12
    w0, g0 = O.lucy(nl.rij[: nl.nip], nl.drij[: nl.nip], 2.0, 2)
    assert rel_norm(nl.wij[: nl.nip], w0) < 1e-14 and rel_norm(nl.dwij[: nl.nip], g0) < 1e-13


def test_two_particle_spam_force():
    # test/force_test.py:10-25 in 2-D, with the result checked by hand
    p = particles.SmoothParticleSystem(2, d=2, maxn=5, xmax=20, ymax=20, hshort=2.0, hlong=4.0, temperature=1.0)
    p.r[0, :] = (5.0, 5.0)
    p.r[1, :] = (6.0, 5.0)
    p.v[:2] = 0.0
    nl = neighbour_list.VerletList(p, cutoff=10, tolerance=2)
    nl.build()
    nl.compress()
    nl.separations()
    properties.spam_properties(p, nl)
    p.vdot[:] = 0.0
    p.udot[:] = 0.0
    forces.SpamForce(p, nl).apply()
    q = 5.0 / (math.pi * 4.0)
    W1 = q * (1 + 3 * 0.5) * 0.5 ** 3  # Lucy W at r = 1, h = 2
    rho = q + W1
    pr = rho * 1.0 / (1 - 0.5 * rho)
    dW = -12 * q * 1.0 * 0.25 / 4.0  # dW/dr at r = 1, h = 2
    f = 2 * pr / rho ** 2 * dW
    assert p.rho[0] == pytest.approx(rho, rel=1e-15)
    assert p.vdot[0, 0] == pytest.approx(f, rel=1e-13) and p.vdot[1, 0] == pytest.approx(-f, rel=1e-13)
    assert abs(p.vdot[0, 1]) < 1e-15


def test_c_forces_and_spam_complete_match_oracle():
    p, _ = _system(seed=8)
    n = p.n
    nl = neighbour_list.VerletList(p, cutoff=4.0, tolerance=1.0)
    nl.build()
    p.t[:n] = 0.8 + 0.2 * np.cos(p.r[:n, 1] * 2 * np.pi / 24.0)
    p.vdot[:] = 0.0
    p.udot[:] = 0.0
    c_forces.SpamForce(p, nl).apply()          # sets u from T, adds pressure
    c_forces.CohesiveSpamForce(p, nl).apply()  # adds cohesion
    c_forces.SpamConduction(p, nl).apply()     # adds conduction
    u = p.u[:n].copy()
    prm = configs.params(24.0, eta=0.0, kcond=properties.thermalk)
    ref = O.derivatives(p.r[:n], p.v[:n], p.m[:n], u, prm, terms=O.PRESSURE | O.COHESION | O.CONDUCTION)
    assert rel_norm(p.vdot[:n], ref["acc"]) < 1e-12
    assert rel_norm(p.udot[:n], ref["udot"]) < 1e-12
    assert np.array_equal(p.rdot[:n], p.v[:n])
    sc = spam_complete_force.SpamComplete(p, nl, eta=1.0, zeta=0.1, cgrad=0.0)
    sc.apply()
    prm = configs.params(24.0, eta=1.0, zeta=0.1, kcond=properties.thermalk)
    ref = O.derivatives(p.r[:n], p.v[:n], p.m[:n], u, prm, terms=15)
    assert rel_norm(p.vdot[:n], ref["acc"]) < 1e-12
    assert rel_norm(p.udot[:n], ref["udot"]) < 1e-12
    assert rel_norm(p.P[:n], ref["P"]) < 1e-12
    assert rel_norm(p.jq[:n], ref["jq"]) < 1e-12
    assert rel_norm(p.t[:n], ref["T"]) < 1e-13


def test_particles_run_matches_oracle():
    p, _ = _system(ns=20, seed=11)
    n = p.n
    nl = neighbour_list.VerletList(p, cutoff=4.0, tolerance=1.0)
    p.nlists.append(nl)
    p.nl_default = nl
    p.forces.append(spam_complete_force.SpamComplete(p, nl, eta=1.0, zeta=0.0, cgrad=0.0, thermalk=5.0))
    nl.build()
    p.t[:n] = 0.9
    properties.spam_properties(p, nl)
    r0, v0, u0 = p.r[:n].copy(), p.v[:n].copy(), p.u[:n].copy()
    p.run(12, 0.005)
    prm = configs.params(20.0, eta=1.0, zeta=0.0, kcond=5.0)
    r1, v1, u1 = O.run(r0, v0, p.m[:n], u0, prm, 12, terms=15)
    dr = p.r[:n] - r1
    dr = np.where(dr > 10, dr - 20, np.where(dr < -10, dr + 20, dr))
    assert np.max(np.abs(dr)) < 1e-10
    assert rel_norm(p.v[:n], v1) < 1e-11 and rel_norm(p.u[:n], u1) < 1e-11
    assert p.steps == 12
    # the reference's derivatives() composition on the same objects
    p.derivatives()
    assert np.all(np.isfinite(p.vdot[:n]))


def test_c_level_lucy_kernel_3d():
    # the reference's cimportable C kernel (c_properties.pxd:1) in FP32,
    # against the oracle's FP64 3-D Lucy kernel and W(0, h) = 105/(16 pi h^3)
    import ctypes
    from pyticles_amd import _lib
    f = ctypes.CDLL(_lib.LIB_PATH).lucy_kernel_3d
    F3 = ctypes.c_float * 3
    f.argtypes = [ctypes.c_float, F3, ctypes.c_float, ctypes.POINTER(ctypes.c_float), F3]
    f.restype = None
    rng = np.random.Generator(np.random.PCG64(41))
    for h in (2.0, 4.0):
        w, g = ctypes.c_float(), F3()
        f(0.0, F3(0.0, 0.0, 0.0), h, ctypes.byref(w), g)
        assert w.value == pytest.approx(105.0 / (16.0 * math.pi * h ** 3), rel=1e-6) and list(g) == [0.0] * 3
        for _ in range(50):
            dx = rng.uniform(-h, h, size=3) * 0.7
            r = float(np.linalg.norm(dx))
            f(r, F3(*dx), h, ctypes.byref(w), g)
            w0, g0 = O.lucy(np.array([r]), dx.reshape(1, 3), h, 3)
            assert w.value == pytest.approx(w0[0], rel=1e-5, abs=1e-7)
            assert np.allclose(np.array(g[:]), g0[0], rtol=1e-5, atol=1e-7)
        f(1.01 * h, F3(1.01 * h, 0.0, 0.0), h, ctypes.byref(w), g)
        assert w.value == 0.0 and list(g) == [0.0] * 3
