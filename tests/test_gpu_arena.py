"""Batched and served forms: boxes of uneven size share one arena; each
box's result must be bitwise the one it gets alone (same per-box list
schedule, same summation order), including boxes admitted while others run
and boxes placed in reused slots."""
import numpy as np
import pytest

from oracle import oracle as O
from pyticles_amd import _lib, configs
from pyticles_amd.arena import Arena
from pyticles_amd.device import from_system

pytestmark = pytest.mark.gpu


def alone(sysm, precision):
    s = configs.System(sysm.r, sysm.v, sysm.m, sysm.T, dict(sysm.prm, precision=precision), sysm.steps, sysm.seed)
    with from_system(s) as b:
        e_start = b.energy()[2]
        b.step(sysm.steps)
        out = b.download("r", "v", "u", "rho")
        out["E0"], out["E1"] = e_start, b.energy()[2]
    return out


def small_boxes():
    return [configs.lognormal_box(seed, 256, 2048, 0.4, 1.0, steps=st) for seed, st in
            ((101, 30), (102, 17), (103, 41))]


def run_to_completion(ar, ids, limit=500):
    for _ in range(limit):
        st, k, hl = ar.poll_all()
        assert not np.any(hl[ids]), "health flags %s" % hl
        if all(st[i] == _lib.BOX_DONE for i in ids):
            return
        ar.step(1)
    raise AssertionError("boxes did not finish")


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.FP32C])
def test_batched_equals_alone(precision):
    boxes = small_boxes()
    ref = [alone(s, precision) for s in boxes]
    with Arena(sum(s.n for s in boxes) + 1000, precision=precision) as ar:
        ids = [ar.add(s) for s in boxes]
        run_to_completion(ar, ids)
        for i, s, rf in zip(ids, boxes, ref):
            got = ar.result(i)
            for k in ("r", "v", "u", "rho"):
                assert np.array_equal(got[k], rf[k]), (precision, i, k)
            assert got["E0"] == pytest.approx(rf["E0"], rel=1e-12)
            assert got["E1"] == pytest.approx(rf["E1"], rel=1e-12)


def test_full_c4_arena_equals_alone():
    # All 512 boxes of the batched workload (1.97M particles, N from 256 to
    # 16,384, T from 0.3 to 1.0) in one FP32C arena, 40 steps each; six of
    # them, the smallest and the largest included, must be bitwise what they
    # give alone.
    boxes = []
    for b in range(512):
        s = configs.c4_box(b)
        s.steps = 40
        boxes.append(s)
    ns = np.array([s.n for s in boxes])
    pick = sorted({int(np.argmin(ns)), int(np.argmax(ns)), 0, 137, 300, 511})
    ref = {i: alone(boxes[i], _lib.FP32C) for i in pick}
    with Arena(int(ns.sum()), precision=_lib.FP32C) as ar:
        ids = [ar.add(s) for s in boxes]
        run_to_completion(ar, ids)
        for i in pick:
            got = ar.result(ids[i])
            for k in ("r", "v", "u", "rho"):
                assert np.array_equal(got[k], ref[i][k]), (i, k)


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.FP32C])
def test_boxes_smaller_than_a_tile(precision):
    # Boxes of 16-60 slots with a freed gap among them: one 64-slot tile
    # holds parts of three or more boxes and free slots, and each box
    # rebuilds its lists on its own schedule, so the list build's tile plan
    # meets tiles shared by building and resting boxes.
    boxes = [configs.lognormal_box(seed, 16, 60, 0.3, 1.0, steps=st) for seed, st in
             ((201, 23), (202, 31), (203, 17), (204, 40), (205, 29), (206, 35))]
    boxes.append(configs.lognormal_box(207, 300, 301, 0.8, 0.9, steps=26))
    ref = [alone(s, precision) for s in boxes]
    gap = configs.lognormal_box(208, 20, 21, 0.5, 0.6, steps=5)
    with Arena(sum(s.n for s in boxes) + gap.n + 200, precision=precision) as ar:
        ids = [ar.add(s) for s in boxes[:3]]
        g = ar.add(gap)
        ids += [ar.add(s) for s in boxes[3:]]
        ar.remove(g)
        run_to_completion(ar, ids)
        for i, rf in zip(ids, ref):
            got = ar.result(i)
            for k in ("r", "v", "u", "rho"):
                assert np.array_equal(got[k], rf[k]), (precision, i, k)


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.FP32C])
def test_box_sort_and_radix_sort_agree(precision):
    # Alone, a small box is ordered by the per-box counting sort; beside a
    # box beyond its limits (more than 16,384 slots and 1,024 cells) the
    # arena uses the global radix sort for all boxes.  Both are stable by (cell, slot), so results agree bitwise.
    boxes = small_boxes()
    ref = [alone(s, precision) for s in boxes]
    big = configs.lognormal_box(110, 44000, 44001, 0.6, 0.7, steps=45)
    assert big.n > 16384 and (int(big.prm["Lx"] / 4.5) ** 2) > 1024
    with Arena(sum(s.n for s in boxes) + big.n + 64, precision=precision) as ar:
        ids = [ar.add(s) for s in boxes]
        ib = ar.add(big)
        run_to_completion(ar, ids + [ib])
        for i, rf in zip(ids, ref):
            got = ar.result(i)
            for k in ("r", "v", "u", "rho"):
                assert np.array_equal(got[k], rf[k]), (precision, i, k)


def test_served_admission_mid_run_and_slot_reuse():
    prec = _lib.FP32C
    a, b, c = small_boxes()
    ra, rb, rc = (alone(s, prec) for s in (a, b, c))
    with Arena(a.n + b.n + 64, precision=prec) as ar:
        ia = ar.add(a)
        ar.step(7)
        ib = ar.add(b)  # joins while a runs
        run_to_completion(ar, [ia, ib])
        for i, rf in ((ia, ra), (ib, rb)):
            got = ar.result(i)
            assert np.array_equal(got["r"], rf["r"]) and np.array_equal(got["u"], rf["u"])
        ar.remove(ia)
        big = configs.lognormal_box(105, a.n + 200, a.n + 201, 0.5, 0.6, steps=5)
        with pytest.raises(_lib.SphError) as ei:
            ar.add(big)  # no free run of slots is long enough
        assert ei.value.code == _lib.SPH_ERR_NOMEM
        small = configs.lognormal_box(104, 256, 300, 0.5, 0.6, steps=12)
        rs = alone(small, prec)
        isml = ar.add(small)  # reuses a's slots
        run_to_completion(ar, [isml])
        got = ar.result(isml)
        assert np.array_equal(got["r"], rs["r"]) and np.array_equal(got["v"], rs["v"])


def test_arena_fp64_matches_oracle():
    s = configs.lognormal_box(7, 400, 500, 0.7, 0.8, steps=10, precision=_lib.FP64)
    with Arena(s.n + 10, precision=_lib.FP64) as ar:
        i = ar.add(s)
        run_to_completion(ar, [i])
        got = ar.result(i)
    u0 = O.init_u(s.r, s.m, s.T, s.prm)
    r1, v1, u1 = O.run(s.r, s.v, s.m, u0, s.prm, 10)
    L = s.prm["Lx"]
    dr = got["r"] - r1
    dr = np.where(dr > L / 2, dr - L, np.where(dr < -L / 2, dr + L, dr))
    assert np.max(np.abs(dr)) < 1e-10
    assert np.max(np.abs(got["v"] - v1)) < 1e-10 * np.max(np.abs(v1))


def test_fp32c_trajectory_error_vs_fp64():
    s = configs.lognormal_box(9, 1500, 1600, 0.6, 0.7, steps=50, precision=_lib.FP64)
    r64 = alone(s, _lib.FP64)
    r32 = alone(s, _lib.FP32C)
    L = s.prm["Lx"]
    dr = r32["r"] - r64["r"]
    dr = np.where(dr > L / 2, dr - L, np.where(dr < -L / 2, dr + L, dr))
    assert np.max(np.abs(dr)) < 1e-5
    assert np.max(np.abs(r32["rho"] - r64["rho"])) / np.max(r64["rho"]) < 1e-5
    # conserved total energy: drift of the compensated form stays at the FP64 level
    d64 = abs(r64["E1"] - r64["E0"]) / abs(r64["E0"])
    d32 = abs(r32["E1"] - r32["E0"]) / abs(r32["E0"])
    assert d32 < 10 * d64 + 1e-6
