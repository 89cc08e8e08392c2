"""Serving loop, launch schedules and the autotuner on the device.

A served stream admits jobs while others run (they share every launch);
each job's result must be bitwise the one it gets alone.  Launch schedules
change how neighbours are walked, never which pairs interact."""
import numpy as np
import pytest

from pyticles_amd import _lib, autotune, configs
from pyticles_amd.device import DeviceBox, from_system
from pyticles_amd.serve import Server, latency_summary

pytestmark = pytest.mark.gpu


def alone(sysm, precision, unroll=0, skin=None, rates_unroll=0, block=0):
    prm = dict(sysm.prm, precision=precision)
    if skin is not None:
        prm["skin"] = skin
    # the schedule is set before the initial density pass (u = kB T - a rho):
    # the FP32 form's partial sums follow the unroll factor
    with DeviceBox(sysm.n, prm) as b:
        b.set_schedule(unroll, rates_unroll, block)
        b.upload(r=sysm.r, v=sysm.v, m=sysm.m, u=np.zeros(sysm.n))
        b.init_thermal(sysm.T)
        b.step(sysm.steps)
        return b.download("r", "v", "u", "rho")


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_served_stream_overlaps_and_equals_alone():
    prec = _lib.FP32C
    jobs = [configs.lognormal_box(300 + j, 256, 1500, 0.3, 1.2, slo=8, shi=40) for j in range(10)]
    # arrivals 0.3 ms apart: each job runs 8-40 steps, so later jobs arrive
    # while earlier ones are on the device
    arr = np.linspace(0.0, 0.003, len(jobs))
    sched = dict(unroll=2, skin=0.75)
    srv = Server(2e9, precision=prec, max_resident=4000, max_job=1500, chunk=3, keep_results=True,
                 schedule=sched)
    try:
        recs = srv.run(jobs, arr)
    finally:
        srv.close()
    assert all(r.finished is not None and r.health == 0 for r in recs), [(r.job_id, r.health) for r in recs]
    assert srv.max_running >= 2, "jobs never overlapped on the device"
    for rec in recs:
        job = jobs[rec.job_id]
        ref = alone(job, prec, unroll=sched["unroll"], skin=sched["skin"])
        for k in ("r", "v", "u", "rho"):
            assert np.array_equal(rec.result[k], ref[k]), (rec.job_id, k)
    s = latency_summary(recs)
    assert s["jobs"] == len(jobs) and s["p99_s"] >= s["p50_s"] > 0


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.FP32C])
def test_unroll_schedules_agree(precision):
    # Same pairs in the same order; the compiler may fuse multiply-adds
    # differently per unroll, so the FP32 form agrees to FP32 rounding of the
    # rates propagated over 25 steps (measured ~1e-9 relative), the FP64
    # form to FP64 rounding.
    tol = 1e-13 if precision == _lib.FP64 else 1e-7
    s = configs.square_system(24, 0.9, 7, 25)
    base = alone(s, precision, unroll=0)
    for u, ru in ((1, 0), (2, 0), (4, 0), (4, 2), (2, 4), (1, 4)):
        got = alone(s, precision, unroll=u, rates_unroll=ru)
        for k in ("r", "v", "u", "rho"):
            assert rel(got[k], base[k]) <= tol, (u, ru, k)
    # the pair passes' block size changes which work-items share a block,
    # never a slot's own walk: the same bits
    for bs in (128, 512):
        got = alone(s, precision, block=bs)
        for k in ("r", "v", "u", "rho"):
            assert np.array_equal(got[k], base[k]), (bs, k)


def test_skin_changes_lists_not_physics():
    s = configs.square_system(24, 0.9, 8, 30)
    a = alone(s, _lib.FP64, skin=0.3)
    b = alone(s, _lib.FP64, skin=0.75)
    for k in ("r", "v", "u", "rho"):
        assert rel(a[k], b[k]) <= 1e-11, k


def test_bad_schedule_rejected():
    s = configs.square_system(8, 0.9, 9, 1)
    with from_system(s) as b:
        with pytest.raises(_lib.SphError) as ei:
            b.set_schedule(3)
        assert ei.value.code == _lib.SPH_ERR_ARG
        with pytest.raises(_lib.SphError) as ei:
            b.set_schedule(2, 3)
        assert ei.value.code == _lib.SPH_ERR_ARG


def test_autotune_small_workload(tmp_path):
    p = str(tmp_path / "sched.json")
    s = configs.square_system(32, 0.9, 10, 10)
    best, table = autotune.tune(s, _lib.FP64, "box", steps=3, warm=1, path=p)
    assert autotune.valid(best)
    assert len(table) == (len(autotune.candidates()) + len(autotune.UNROLLS) + len(autotune.BLOCKS)
                          + len(autotune.LIST_BITS) + len(autotune.FUSED))
    assert best["rates_unroll"] in (0,) + autotune.UNROLLS and best["block"] in (0,) + autotune.BLOCKS
    assert best["list_bits"] in autotune.LIST_BITS
    assert all(e["rate"] > 0 for e in table)
    assert autotune.schedule_for(_lib.FP64, "box", path=p)["source"] == "tuned"
    boxes = [configs.lognormal_box(400 + j, 256, 600, 0.4, 0.9, steps=20) for j in range(4)]
    best, table = autotune.tune(boxes, _lib.FP32C, "batched", steps=3, warm=1, path=p)
    assert autotune.valid(best)
    assert set(autotune.load_cache(p)) == {"fp64:box", "fp32c:batched"}


def test_c5_jobs_served_equal_alone():
    # jobs of the served workload itself (N log-uniform in [1024, 65536],
    # T in [0.3, 1.2], their own seeds) with their step counts cut to 20-60
    # so the test stays short; they arrive while others run
    prec = _lib.FP32C
    jobs = []
    for j in range(24):
        s = configs.c5_job(j)
        s.steps = 20 + (7 * j) % 41
        jobs.append(s)
    assert max(s.n for s in jobs) > 40000
    arr = np.linspace(0.0, 0.02, len(jobs))
    sched = autotune.schedule_for(prec, "batched")
    srv = Server(200e9, precision=prec, max_resident=4_000_000, chunk=4, keep_results=True, schedule=sched)
    try:
        recs = srv.run(jobs, arr)
    finally:
        srv.close()
    assert all(r.finished is not None and r.health == 0 for r in recs), [(r.job_id, r.health) for r in recs]
    assert srv.max_running >= 2
    print("[serve] c5 jobs: N %d..%d, %d resident at most" % (
        min(s.n for s in jobs), max(s.n for s in jobs), srv.peak_resident), flush=True)
    for rec in recs:
        ref = alone(jobs[rec.job_id], prec, unroll=sched["unroll"], skin=sched["skin"],
                    rates_unroll=sched.get("rates_unroll", 0), block=sched.get("block", 0))
        for k in ("r", "v", "u", "rho"):
            assert np.array_equal(rec.result[k], ref[k]), (rec.job_id, k)


def test_budget_admission_binds_on_device():
    # an HBM budget that holds ~7,000 particle slots: jobs of 2,000-4,000
    # particles queue until earlier ones leave, the resident count never
    # passes what the budget holds, and every result is what it is alone
    from pyticles_amd.arena import capacity_for, footprint
    prec = _lib.FP32C
    jobs = [configs.lognormal_box(900 + j, 2000, 4000, 0.4, 1.0, steps=15 + 3 * j) for j in range(10)]
    budget = footprint(7000, prec)
    srv = Server(budget, precision=prec, chunk=2, keep_results=True)
    try:
        recs = srv.run(jobs, np.zeros(len(jobs)))
    finally:
        srv.close()
    assert srv.capacity == capacity_for(budget, prec)
    assert srv.peak_resident <= srv.capacity and footprint(srv.peak_resident, prec) <= budget
    waited = [r for r in recs if r.admitted > r.arrival + 1e-4]
    assert len(waited) >= 5, "the budget never held a job back"
    assert all(r.health == 0 and r.error is None for r in recs)
    for rec in recs:
        ref = alone(jobs[rec.job_id], prec)
        for k in ("r", "v", "u", "rho"):
            assert np.array_equal(rec.result[k], ref[k]), (rec.job_id, k)
