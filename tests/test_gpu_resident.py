"""Resident stepping: an open device run (sph_box_advance) continued over
many calls gives the bits of one sph_step of the same steps, and the
reference-style per-tick loop (particles.update, particles.py:459-494, as
toybox.py:101-106 drives it) keeps the state on the device: one rate
evaluation per update() and no transfer."""
import time

import numpy as np
import pytest

from pyticles_amd import _lib, configs, neighbour_list, particles, spam_complete_force
from pyticles_amd.device import from_system

pytestmark = pytest.mark.gpu


def _state(b):
    return b.download("r", "v", "u")


def _same(a, b):
    for k in ("r", "v", "u"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.FP32C])
def test_open_run_in_pieces_equals_one_step_call(precision):
    s = configs.square_system(48, 0.9, 21, 0, precision=precision)
    nsteps = 90
    with from_system(s) as a:
        a.step(nsteps)
        ref = _state(a)
        st = a.run_stats()
        assert st["builds"] >= 3 and not st["open"]  # several list builds inside the run
    with from_system(s) as b:  # one step per call
        for _ in range(nsteps):
            b.advance(1)
        assert b.run_stats()["open"]
        b.sync()
        _same(_state(b), ref)
    with from_system(s) as c:  # uneven pieces, then sph_step continues the open run and closes it
        for k in (7, 13, 1, 1, 19):
            c.advance(k)
        c.step(nsteps - 41)
        assert not c.run_stats()["open"]
        _same(_state(c), ref)


def test_position_reads_keep_the_run_open():
    s = configs.square_system(40, 0.9, 22, 0)
    with from_system(s) as a:
        a.step(50)
        ref = _state(a)
    with from_system(s) as b:
        for _ in range(50):
            b.advance(1)
            r = b.download("r")  # positions only: the run stays open
            assert np.all(np.isfinite(r))
        assert b.run_stats()["open"]
        _same(_state(b), ref)


def test_following_runs_continue_the_list_schedule():
    # runs that follow each other keep the box's lists and their schedule;
    # an upload (even of the same values) drops them, so the next run
    # builds at its first step
    s = configs.square_system(48, 0.9, 23, 0)
    with from_system(s) as a, from_system(s) as c:
        a.step(5)
        c.step(5)
        builds_a = builds_c = 0
        for _ in range(10):
            a.step(3)
            builds_a += a.run_stats()["builds"]
            st = c.download("r", "v", "u")
            c.upload(r=st["r"], v=st["v"], u=st["u"])
            c.step(3)
            builds_c += c.run_stats()["builds"]
        assert builds_c >= 10 and builds_a < builds_c / 2, (builds_a, builds_c)
        # the same trajectory up to the summation order of other lists
        ga, gc = _state(a), _state(c)
        for k in ("r", "v", "u"):
            assert np.max(np.abs(ga[k] - gc[k])) <= 1e-9 * np.max(np.abs(gc[k])), k


def _surface_system(ns, seed=31):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = ns * ns
    L = float(ns)
    p = particles.SmoothParticleSystem(n, d=2, maxn=n, xmax=L, ymax=L, hshort=2.0, hlong=4.0, temperature=0.9)
    p.r[:n] = configs.lattice(n, L, rng)
    p.v[:n] = configs.velocities(n, 0.9, rng)
    p.m[:] = 1.0
    p.u[:n] = 0.9 - 2.0 * 1.02
    nl = neighbour_list.VerletList(p, cutoff=4.0, tolerance=1.0)
    p.nlists.append(nl)
    p.nl_default = nl
    p.forces.append(spam_complete_force.SpamComplete(p, nl, eta=1.0, zeta=0.0, cgrad=0.0, thermalk=5.0))
    return p


def test_update_calls_equal_run_bitwise():
    n_steps = 100
    a = _surface_system(64)
    a.run(n_steps, 0.005)
    b = _surface_system(64)
    for _ in range(n_steps):
        b.update(0.005)
    n = a.n
    assert b.steps == a.steps == n_steps
    for k in ("r", "v", "u", "rho", "udot"):
        assert np.array_equal(getattr(b, k)[:n], getattr(a, k)[:n]), k
    # the reference driver's loop: p.r checked every step (toybox.py:101-106);
    # position reads leave the device run open, so the bits do not change
    c = _surface_system(64)
    for _ in range(n_steps):
        c.update(0.005)
        assert not np.isnan(c.r).any()
    for k in ("r", "v", "u"):
        assert np.array_equal(getattr(c, k)[:n], getattr(a, k)[:n]), k


def test_host_changes_are_uploaded_between_updates():
    a = _surface_system(32, seed=33)
    b = _surface_system(32, seed=33)
    n = a.n
    for p in (a, b):
        for _ in range(5):
            p.update(0.005)
    b.v[:n] *= 0.5  # read (closes b's run), changed in place: the next step starts from it
    for _ in range(5):
        b.update(0.005)
    a.v[:n] = a.v[:n] * 0.5  # the same change through assignment
    a.run(5, 0.005)
    # both restarted from the same synchronised, rescaled state
    for k in ("r", "v", "u"):
        assert np.array_equal(getattr(b, k)[:n], getattr(a, k)[:n]), k


def test_update_costs_what_run_costs():
    # one box of C2's size: 100 update() calls against one run(100), three
    # times each, alternating; the best of each (a single pair of ~20 ms
    # windows is within a few ms of the bound on host noise alone)
    n_steps = 100
    a = _surface_system(512, seed=35)
    a.run(5, 0.005)  # device box, graphs and schedule set up
    b = _surface_system(512, seed=35)
    b.run(5, 0.005)
    t_run, t_upd = [], []
    for _ in range(3):
        t = time.perf_counter()
        a.run(n_steps, 0.005)
        t_run.append(time.perf_counter() - t)
        # the fetch after run() or a read pondered the host list (never
        # built, so flagged): the reference's update() would rebuild it
        # first (particles.py:471-472); do that before the timed loop, which
        # then holds device steps only
        b.rebuild_lists()
        assert not b.nlists[0].rebuild_list
        t = time.perf_counter()
        for _ in range(n_steps):
            b.update(0.005)
        _ = b.v  # closes the run and fetches the state, as run() does
        t_upd.append(time.perf_counter() - t)
    print("[resident] run(100) %s ms, 100 x update() %s ms"
          % (["%.2f" % (1e3 * x) for x in t_run], ["%.2f" % (1e3 * x) for x in t_upd]), flush=True)
    assert min(t_upd) < 1.2 * min(t_run), (t_upd, t_run)
