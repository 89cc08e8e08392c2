"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle on
the same seeded inputs.  Pair sets and separations must be bit-exact; FP64
fields within 1e-12 (normwise where sums cancel); FP32-compensated fields
within 1e-6."""
import ctypes
import math

import numpy as np
import pytest

from oracle import oracle as O
from pyticles_amd import _lib, configs
from pyticles_amd.device import DeviceBox as Box, from_system

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def rel_norm(a, b):
    """max |a - b| / max |b| (normwise; per-element relative error is
    ill-conditioned where pair sums cancel)."""
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (s if s > 0 else 1.0))


def _system(ns, seed, T=0.9, L=None, jit=0.1, **pkw):
    rng = np.random.Generator(np.random.PCG64(seed))
    L = float(ns) if L is None else L
    n = ns * ns
    r = configs.lattice(n, L, rng, jitter=jit)
    v = configs.velocities(n, T, rng)
    return r, v, np.ones(n), configs.params(L, **pkw)


def _box(r, v, m, prm, u=None, T=0.9):
    b = Box(r.shape[0], prm)
    b.upload(r=r, v=v, m=m, u=np.zeros(r.shape[0]) if u is None else u)
    if u is None:
        b.init_thermal(T)
    return b


# ------------------------------------------------------------------ (a) pairs
@pytest.mark.parametrize("ns,L,rc,tol", [(32, 32.0, 4.0, 1.0), (64, 64.0, 4.0, 1.0), (13, 13.0, 4.0, 1.0),
                                         (8, 8.0, 4.0, 1.0), (10, 5.0, 2.0, 1.0), (20, 20.0, 3.0, 0.5)])
def test_pairs_bit_exact(ns, L, rc, tol):
    r, v, m, prm = _system(ns, seed=ns, L=L, rc=rc, tol=tol)
    R2 = rc * rc + tol * tol
    ref = O.pairs(r, L, L, R2)
    with _box(r, v, m, prm, u=np.zeros(ns * ns)) as b:
        got = b.pairs()
        assert got.shape == ref.shape
        assert np.array_equal(got, ref)
        drij, rij, rsq, dv = b.separations(ref)
    d0, r0, q0, v0 = O.separations(r, v, L, L, ref)
    assert np.array_equal(drij, d0) and np.array_equal(rsq, q0)
    assert np.array_equal(rij, r0) and np.array_equal(dv, v0)


def test_pairs_lattice_ties_and_reference_answer():
    ns = 24
    r = configs.lattice(ns * ns, float(ns), None, jitter=0)
    prm = configs.params(ns)
    with _box(r, np.zeros_like(r), np.ones(ns * ns), prm, u=np.zeros(ns * ns)) as b:
        iap = b.pairs()
        assert iap.shape[0] == 24 * ns * ns  # the 8 neighbours at r^2 == 17 are excluded
        assert np.array_equal(iap, O.pairs(r, ns, ns, 17.0))
        b.density()
        np.testing.assert_allclose(b.download("rho"), 1.02006720065032, rtol=1e-14)
        np.testing.assert_allclose(b.download("rho_lr"), 1.0006325056255552, rtol=1e-14)
    # test/neighbour_list_test.py:15-24 in 2-D
    r = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    prm = configs.params(5.0, rc=10.0, tol=2.0)
    with _box(r, np.zeros_like(r), np.ones(3), prm, u=np.zeros(3)) as b:
        iap = b.pairs()
        assert iap.tolist() == [[0, 1], [0, 2], [1, 2]]
        _, rij, _, _ = b.separations(iap)
        assert rij[0] == 1.0


def test_pairs_after_moves_and_capacity():
    r, v, m, prm = _system(40, seed=9)
    with _box(r, v, m, prm) as b:
        b.step(5)
        r5 = b.download("r")
        got = b.pairs()
        assert np.array_equal(got, O.pairs(r5, 40.0, 40.0, 17.0))
        # caller capacity below the count: SPH_ERR_CAP with the true count
        import ctypes
        iap = np.empty((10, 2), dtype=np.int64)
        nip = ctypes.c_int64()
        rc = _lib.lib.sph_nlist_build(b._h, 10, _lib.i64ptr(iap), ctypes.byref(nip))
        assert rc == _lib.SPH_ERR_CAP and nip.value == got.shape[0]
        assert np.array_equal(iap, got[:10])


def test_minimum_image_and_lucy_kernel():
    rng = np.random.Generator(np.random.PCG64(1))
    dr = rng.uniform(-12, 12, size=(1000, 2))
    dr[:4] = [[5.0, -5.0], [5.0000001, -5.0000001], [10.0, -10.0], [0.0, 0.0]]
    import ctypes
    got = dr.copy()
    _lib.check(_lib.lib.sph_minimum_image(_lib.context(), 1000, _lib.dptr(got), 10.0, 10.0))
    assert np.array_equal(got, O.min_image(dr, 10.0, 10.0))
    for dim in (1, 2, 3):
        r = rng.uniform(0, 2.5, size=500)
        r[:2] = [0.0, 2.0]
        dx = rng.normal(size=(500, dim))
        w, g = np.empty(500), np.empty((500, dim))
        _lib.check(_lib.lib.sph_kernel_lucy(_lib.context(), 500, _lib.dptr(r), _lib.dptr(dx), 2.0, dim,
                                            _lib.dptr(w), _lib.dptr(g)))
        w0, g0 = O.lucy(r, dx, 2.0, dim)
        assert rel_norm(w, w0) < 1e-14 and rel_norm(g, g0) < 1e-13
        assert w[1] == 0.0 and np.all(g[0] == 0.0)


# ------------------------------------------------------------------ (b) density, EOS
@pytest.mark.parametrize("ns,seed", [(32, 1), (48, 2), (9, 3)])
def test_density_eos_parity(ns, seed):
    r, v, m, prm = _system(ns, seed=seed)
    rng = np.random.Generator(np.random.PCG64(seed + 100))
    m = rng.uniform(0.8, 1.2, size=ns * ns)
    u = rng.uniform(-1.6, -0.6, size=ns * ns)
    ref = O.derivatives(r, v, m, u, prm, terms=15)
    with _box(r, v, m, prm, u=u) as b:
        b.density()
        got = b.download("rho", "rho_lr", "T", "p", "pco")
    for k in ("rho", "rho_lr", "T", "p", "pco"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-13, err_msg=k)


# ------------------------------------------------------------------ (c) rates
@pytest.mark.parametrize("terms", [3, 1, 2, 7, 11, 15])
def test_rates_parity(terms):
    r, v, m, prm = _system(40, seed=terms, terms=terms)
    rng = np.random.Generator(np.random.PCG64(terms + 7))
    m = rng.uniform(0.8, 1.2, size=r.shape[0])
    u = rng.uniform(-1.6, -0.6, size=r.shape[0])
    ref = O.derivatives(r, v, m, u, prm, terms=terms)
    with _box(r, v, m, prm, u=u) as b:
        b.rates(terms)
        got = b.download("acc", "udot", "rho", "P", "jq", "gradv" if terms & 12 else "rho")
    assert rel_norm(got["acc"], ref["acc"]) < TOL64
    assert rel_norm(got["udot"], ref["udot"]) < TOL64
    assert rel_norm(got["P"], ref["P"]) < TOL64
    if terms & 8:
        assert rel_norm(got["jq"], ref["jq"]) < TOL64
    if terms & 12:
        assert rel_norm(got["gradv"], ref["gradv"]) < TOL64
    # conservation on the device result itself
    mom = (m[:, None] * got["acc"]).sum(axis=0)
    assert np.all(np.abs(mom) < 1e-12 * np.sum(np.abs(m[:, None] * got["acc"])))


def test_full_size_c2_pairs_and_rates():
    # the 512 x 512 box of 262,144 particles at its real size: the pair set
    # (6.8M pairs) bit-exact, FP64 density and rates within the parity
    # tolerance, and the device accelerations conserving momentum
    s = configs.c2()
    prm = s.prm
    ref_pairs = O.pairs(s.r, prm["Lx"], prm["Ly"], prm["rc"] ** 2 + prm["tol"] ** 2)
    u0 = O.init_u(s.r, s.m, s.T, prm)
    ref = O.derivatives(s.r, s.v, s.m, u0, prm, terms=prm["terms"])
    with from_system(s) as b:
        got_pairs = b.pairs()
        assert got_pairs.shape == ref_pairs.shape
        assert np.array_equal(got_pairs, ref_pairs)
        b.rates(prm["terms"])
        got = b.download("rho", "rho_lr", "acc", "udot")
    for k in ("rho", "rho_lr"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-12, err_msg=k)
    assert rel_norm(got["acc"], ref["acc"]) < TOL64
    assert rel_norm(got["udot"], ref["udot"]) < TOL64
    mom = (s.m[:, None] * got["acc"]).sum(axis=0)
    assert np.all(np.abs(mom) < 1e-12 * np.sum(np.abs(s.m[:, None] * got["acc"])))


def test_full_size_c3_fp32c_within_1e6_of_fp64_oracle():
    # the 1024 x 1024 quench box of 1,048,576 particles in the FP32
    # compensated form (cell-relative mirror, the box being over 512 wide):
    # one rate evaluation within 1e-6 of the FP64 oracle, normwise
    s = configs.c3()
    prm = s.prm
    u0 = O.init_u(s.r, s.m, s.T, prm)
    ref = O.derivatives(s.r, s.v, s.m, u0, prm, terms=prm["terms"])
    with from_system(s) as b:
        b.upload(u=u0)
        b.rates(prm["terms"])
        got = b.download("rho", "rho_lr", "acc", "udot")
    for k in ("rho", "rho_lr", "acc", "udot"):
        assert rel_norm(got[k], ref[k]) < 1e-6, k


def test_gradients_api():
    r, v, m, prm = _system(30, seed=4)
    u = np.full(r.shape[0], -1.1)
    ref = O.derivatives(r, v, m, u, prm, terms=15)
    with _box(r, v, m, prm, u=u) as b:
        b.gradients()
        got = b.download("gradv", "jq")
    assert rel_norm(got["gradv"], ref["gradv"]) < TOL64
    assert rel_norm(got["jq"], ref["jq"]) < TOL64


# ------------------------------------------------------------------ stepping
@pytest.mark.parametrize("cfg,nsteps", [("c1", 20), ("full", 10), ("full_mass", 10)])
def test_step_parity(cfg, nsteps):
    if cfg == "c1":
        s = configs.c1()
    else:
        s = configs.square_system(48, 0.9, 2, nsteps)
    if cfg == "full_mass":  # non-uniform masses: the per-neighbour mass loads and m-weighted energy
        m = np.random.Generator(np.random.PCG64(9)).uniform(0.8, 1.2, size=s.n)
        s = configs.System(s.r, s.v, m, s.T, s.prm, s.steps, s.seed)
    terms = s.prm["terms"]
    u0 = O.init_u(s.r, s.m, s.T, s.prm)
    r1, v1, u1, E = O.run(s.r, s.v, s.m, u0, s.prm, nsteps, terms=terms, energy=True)
    with from_system(s) as b:
        ug = b.download("u")
        np.testing.assert_allclose(ug, u0, rtol=1e-13)
        b.step(nsteps)
        got = b.download("r", "v", "u")
        e = b.energy()
    L = s.prm["Lx"]
    dr = got["r"] - r1
    dr = np.where(dr > L / 2, dr - L, dr)
    dr = np.where(dr < -L / 2, dr + L, dr)
    assert np.max(np.abs(dr)) < 1e-10
    assert rel_norm(got["v"], v1) < 1e-11
    assert rel_norm(got["u"], u1) < 1e-11
    assert e[2] == pytest.approx(E[-1], rel=1e-11)


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.FP32C])
def test_fused_and_separate_leapfrog_step_give_the_same_bits(precision):
    # the leapfrog step inside the rate pass or as a launch of its own: the
    # same arithmetic on the same rates, with the state buffers swapping
    # every step (an odd and an even number of steps, through both the step
    # graphs and direct launches)
    s = configs.square_system(40, 0.9, 7, 0, precision=precision)
    outs = []
    for fused in (1, 0):
        with from_system(s) as b:
            b.set_schedule(0, fused=fused)
            b.step(37)
            b.step(24)
            b.set_graph(False)
            b.step(5)
            outs.append(b.download("r", "v", "u", "rho"))
    for k in ("r", "v", "u", "rho"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.parametrize("nsteps", [8, 150])
def test_step_graph_equals_direct_launches(nsteps):
    # 150 steps: runs of plain steps between builds long enough to replay the
    # 8-, 4- and 2-step graphs
    s = configs.square_system(24, 0.9, 5, 8)
    outs = []
    for graph in (True, False):
        with from_system(s) as b:
            b.set_graph(graph)
            b.step(nsteps)
            outs.append(b.download("r", "v", "u"))
    for k in ("r", "v", "u"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_time_pass_leaves_state_unchanged():
    # the roofline timing relaunches the pair passes after a step: the
    # particle state and the densities must be bit-identical afterwards (the
    # rates are re-derived at the synchronised velocities, which the next
    # step re-derives anyway), and stepping on must give what the same
    # calls give untimed
    s = configs.square_system(24, 0.9, 6, 5)
    with from_system(s) as b:
        with pytest.raises(_lib.SphError):
            b.time_pass("density", reps=2)  # not stepped yet
        b.step(5)
        before = b.download("r", "v", "u", "rho", "rho_lr")
        for p in b.PASSES:
            assert b.time_pass(p, reps=3) > 0.0
        after = b.download("r", "v", "u", "rho", "rho_lr")
        for k in before:
            assert np.array_equal(before[k], after[k]), k
        b.step(2)  # and stepping on gives what the same calls give untimed
        got = b.download("r", "v", "u")
    with from_system(s) as c:
        c.step(5)
        c.step(2)
        ref = c.download("r", "v", "u")
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k


def test_eos_pole_is_reported():
    # lattice spacing 0.6: rho*b > 1, the vdW pole (properties.py:41)
    ns = 16
    L = ns * 0.6
    r = configs.lattice(ns * ns, L, None, jitter=0)
    prm = configs.params(L)
    with _box(r, np.zeros_like(r), np.ones(ns * ns), prm, u=np.zeros(ns * ns)) as b:
        with pytest.raises(_lib.SphError) as ei:
            b.density()
        assert ei.value.code == _lib.SPH_ERR_NUMERIC


def test_positions_outside_box_rejected():
    r = np.array([[0.5, 0.5], [5.5, 1.0]])
    with Box(2, configs.params(5.0)) as b:
        with pytest.raises(_lib.SphError):
            b.upload(r=r)


# ------------------------------------------------------------------ FP32-compensated form
def test_fp32_compensated_error_vs_fp64():
    r, v, m, prm = _system(64, seed=12)
    prm32 = dict(prm, precision=_lib.FP32C)
    u = np.full(r.shape[0], -1.1)
    ref = O.derivatives(r, v, m, u, prm, terms=15)
    with _box(r, v, m, prm32, u=u) as b:
        b.rates(15)
        got = b.download("rho", "rho_lr", "acc", "udot")
    assert rel_norm(got["rho"], ref["rho"]) < 1e-6
    assert rel_norm(got["rho_lr"], ref["rho_lr"]) < 1e-6
    assert rel_norm(got["acc"], ref["acc"]) < 1e-6
    assert rel_norm(got["udot"], ref["udot"]) < 1e-6
    print("fp32c errors rho %.2e rho_lr %.2e acc %.2e udot %.2e" % (
        rel_norm(got["rho"], ref["rho"]), rel_norm(got["rho_lr"], ref["rho_lr"]),
        rel_norm(got["acc"], ref["acc"]), rel_norm(got["udot"], ref["udot"])))


# ------------------------------------------------------------------ box shapes and edge sizes
def _rect_system(nx, ny, Lx, Ly, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    r = np.stack([(i.ravel() + 0.5) * Lx / nx, (j.ravel() + 0.5) * Ly / ny], axis=1)
    r += rng.uniform(-0.1, 0.1, size=r.shape)
    r[:, 0] = np.mod(r[:, 0], Lx)
    r[:, 1] = np.mod(r[:, 1], Ly)
    n = nx * ny
    v = configs.velocities(n, 0.9, rng)
    m = rng.uniform(0.9, 1.1, size=n)
    u = rng.uniform(-1.5, -0.8, size=n)
    return np.ascontiguousarray(r), v, m, u


@pytest.mark.parametrize("nx,ny,Lx,Ly", [(40, 13, 40.0, 13.0), (64, 8, 64.0, 8.0), (6, 50, 6.0, 50.0),
                                         (30, 30, 30.0, 30.5)])
def test_rectangular_boxes(nx, ny, Lx, Ly):
    # elongated boxes, and boxes with a single cell along one axis (minimum
    # image instead of cell offsets, in the FP64 build and the FP32 mirrors)
    r, v, m, u = _rect_system(nx, ny, Lx, Ly, seed=nx * ny)
    prm = configs.params(Lx, Ly=Ly)
    ref_pairs = O.pairs(r, Lx, Ly, 17.0)
    ref = O.derivatives(r, v, m, u, prm, terms=15)
    with _box(r, v, m, prm, u=u) as b:
        assert np.array_equal(b.pairs(), ref_pairs)
        b.rates(15)
        got = b.download("rho", "acc", "udot")
    assert rel_norm(got["rho"], ref["rho"]) < 1e-13
    assert rel_norm(got["acc"], ref["acc"]) < TOL64 and rel_norm(got["udot"], ref["udot"]) < TOL64
    with _box(r, v, m, dict(prm, precision=_lib.FP32C), u=u) as b:
        b.rates(15)
        got = b.download("rho", "acc")
    assert rel_norm(got["rho"], ref["rho"]) < 1e-6 and rel_norm(got["acc"], ref["acc"]) < 2e-6


def test_empty_and_single_particle():
    with pytest.raises(_lib.SphError) as ei:
        Box(0, configs.params(10.0))
    assert ei.value.code == _lib.SPH_ERR_ARG
    from pyticles_amd import neighbour_list, particles, properties
    p = particles.SmoothParticleSystem(0, d=2, maxn=4, xmax=10, ymax=10)
    nl = neighbour_list.VerletList(p, cutoff=4.0, tolerance=1.0)
    nl.build()
    nl.separations()
    properties.spam_properties(p, nl)
    assert nl.nip == 0
    # one particle: only the self term, no force
    r = np.array([[3.0, 4.0]])
    prm = configs.params(10.0)
    with _box(r, np.array([[0.1, -0.2]]), np.array([2.0]), prm, u=np.array([-1.0])) as b:
        assert b.pairs().shape == (0, 2)
        b.rates(15)
        got = b.download("rho", "rho_lr", "acc", "udot")
    assert got["rho"][0] == pytest.approx(5.0 / (math.pi * 4.0), rel=1e-15)
    assert got["rho_lr"][0] == pytest.approx(5.0 / (math.pi * 16.0), rel=1e-15)
    assert np.all(got["acc"] == 0.0) and got["udot"][0] == 0.0


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.FP32C])
def test_reused_lists_stay_complete_over_long_runs(precision):
    # A small skin makes the device review and postpone many list builds; at
    # the end the densities from the last step's (reused) lists must equal a
    # fresh build's at the same state to rounding: a pair missing from a list
    # would move a density by ~1e-3.
    s = configs.square_system(40, 1.0, 21, 300, skin=0.2, precision=precision)
    with from_system(s) as b:
        b.step(300)
        got = b.download("rho", "rho_lr")
        b.rates(15)  # rebuilds every list from the current positions
        fresh = b.download("rho", "rho_lr")
    tol = 1e-13 if precision == _lib.FP64 else 1e-6
    assert rel_norm(got["rho"], fresh["rho"]) < tol
    assert rel_norm(got["rho_lr"], fresh["rho_lr"]) < tol


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.FP32C])
def test_energy_totals_at_a_fixed_state(precision):
    # sph_energy at an uploaded state (no step): sum m v^2 / 2, sum m u and
    # their sum against exactly rounded host sums, within 1e-12 (the totals
    # are FP64 in both forms)
    r, v, m, prm = _system(40, seed=23)
    m = np.random.Generator(np.random.PCG64(24)).uniform(0.8, 1.2, size=r.shape[0])
    u = np.random.Generator(np.random.PCG64(25)).uniform(-1.5, -0.8, size=r.shape[0])
    with _box(r, v, m, dict(prm, precision=precision), u=u) as b:
        ke, ue, et = b.energy()
    ke0 = math.fsum(0.5 * m * (v[:, 0] ** 2 + v[:, 1] ** 2))
    ue0 = math.fsum(m * u)
    assert abs(ke - ke0) <= 1e-12 * abs(ke0)
    assert abs(ue - ue0) <= 1e-12 * abs(ue0)
    assert abs(et - (ke0 + ue0)) <= 1e-12 * (abs(ke0) + abs(ue0))
