"""Host logic of the serving loop and the autotuner cache, on the CPU: a
stand-in arena with the Arena interface records what the server does."""
import json

import numpy as np
import pytest

from pyticles_amd import _lib, autotune, serve


class FakeArena:
    """First-fit slot ranges; each step advances running boxes by one."""

    def __init__(self, capacity):
        self.capacity = capacity
        self.boxes = {}
        self.next_id = 0
        self.steps_run = 0
        self.max_used = 0

    def _used(self):
        return sum(b["n"] for b in self.boxes.values())

    def add(self, sysm, nsteps=None):
        if self._used() + sysm.n > self.capacity:
            raise _lib.SphError(_lib.SPH_ERR_NOMEM, "full")
        bid = self.next_id
        self.next_id += 1
        self.boxes[bid] = dict(n=sysm.n, steps=nsteps, k=-1, skin=sysm.prm.get("skin"))
        self.max_used = max(self.max_used, self._used())
        return bid

    def step(self, k):
        self.steps_run += k
        for b in self.boxes.values():
            b["k"] = min(b["steps"], b["k"] + k)

    def poll_all(self, nmax):
        st = np.zeros(nmax, dtype=np.int32)
        k = np.zeros(nmax, dtype=np.int64)
        hl = np.zeros(nmax, dtype=np.int32)
        for bid, b in self.boxes.items():
            if bid < nmax:
                st[bid] = 3 if b["k"] >= b["steps"] else 2
                k[bid] = b["k"]
        return st, k, hl

    def result(self, bid):
        return dict(n=self.boxes[bid]["n"])

    def remove(self, bid):
        del self.boxes[bid]

    def close(self):
        pass


class Job:
    def __init__(self, n, steps):
        self.n, self.steps = n, steps
        self.r = self.v = self.m = None
        self.T = 1.0
        self.prm = dict(skin=0.5)


class Clock:
    """Simulated time: advances by dt per call."""

    def __init__(self, dt=1e-3):
        self.t, self.dt = 0.0, dt

    def __call__(self):
        self.t += self.dt
        return self.t


def test_server_admits_fifo_within_limits_and_finishes_all():
    rng = np.random.default_rng(0)
    jobs = [Job(int(rng.integers(100, 1000)), int(rng.integers(5, 40))) for _ in range(60)]
    arr = np.cumsum(rng.exponential(0.002, size=60))
    fa = FakeArena(3000)
    srv = serve.Server(1e12, max_resident=2500, max_job=1000, chunk=2, keep_results=True, arena=fa,
                       schedule=dict(unroll=2, skin=0.75))
    recs = srv.run(jobs, arr, clock=Clock())
    assert len(recs) == 60 and all(r.finished is not None for r in recs)
    assert fa.max_used <= 2500 and srv.resident == 0 and not fa.boxes
    adm = [r.admitted for r in sorted(recs, key=lambda r: r.job_id)]
    assert all(a <= b for a, b in zip(adm, adm[1:]))  # FIFO admission
    assert all(r.admitted >= r.arrival for r in recs)
    assert all(r.result["n"] == r.n for r in recs)
    s = serve.latency_summary(recs, warmup=10)
    assert s["jobs"] == 50 and s["p50_s"] <= s["p99_s"] <= s["max_s"]
    assert srv.particle_steps == sum(j.n * j.steps for j in jobs)


def test_server_budget_limits_residency():
    jobs = [Job(500, 10) for _ in range(8)]
    fa = FakeArena(10 ** 6)
    budget = serve.footprint(1200)  # room for two jobs at a time
    srv = serve.Server(budget, max_resident=None, arena=fa, chunk=1)
    srv.run(jobs, np.zeros(8), clock=Clock())
    assert fa.max_used == 1000


def test_server_rejects_job_that_never_fits():
    fa = FakeArena(100)
    srv = serve.Server(1e12, max_resident=50, max_job=50, arena=fa)
    with pytest.raises(RuntimeError, match="never fit"):
        srv.run([Job(80, 5)], [0.0], clock=Clock())


def test_footprint_monotone_and_capacity_inverse():
    f1 = serve.footprint(1 << 16, _lib.FP32C)
    f2 = serve.footprint(1 << 17, _lib.FP32C)
    assert 0 < f1 < f2 and serve.footprint(1 << 16, _lib.FP64) < f1
    from pyticles_amd.arena import capacity_for
    c = capacity_for(f2, _lib.FP32C)
    assert c >= 1 << 17 and serve.footprint(c + 1, _lib.FP32C) > f2


def test_autotune_cache_roundtrip(tmp_path):
    p = str(tmp_path / "sched.json")
    d = autotune.schedule_for(_lib.FP64, "box", path=p)
    assert d["source"] == "default" and d["unroll"] == 2
    autotune.save_cache({autotune.key(_lib.FP32C, "batched"): {"best": {"unroll": 1, "skin": 0.3}}}, p)
    got = autotune.schedule_for(_lib.FP32C, "batched", path=p)
    # a cache without rates_unroll, block or list_bits: their defaults
    assert got == dict(unroll=1, skin=0.3, rates_unroll=0, block=0, list_bits=0, fused=1, source="tuned")
    assert autotune.pick_fused([dict(fused=1, rate=1.0), dict(fused=0, rate=1.005)]) == 1
    assert autotune.pick_fused([dict(fused=1, rate=1.0), dict(fused=0, rate=1.05)]) == 0
    # the list entry width picks 16 unless 32 is clearly faster
    assert autotune.pick_list_bits([dict(list_bits=16, rate=1.0), dict(list_bits=32, rate=1.005)]) == 16
    assert autotune.pick_list_bits([dict(list_bits=16, rate=1.0), dict(list_bits=32, rate=1.05)]) == 32
    # an invalid entry falls back to the default
    autotune.save_cache({autotune.key(_lib.FP32C, "box"): {"best": {"unroll": 3, "skin": 0.3}}}, p)
    assert autotune.schedule_for(_lib.FP32C, "box", path=p)["source"] == "default"
    with pytest.raises(ValueError):
        autotune.key(_lib.FP64, "huge")
    assert len(autotune.candidates()) == len(autotune.UNROLLS) * len(autotune.SKINS)
    json.load(open(p))


def test_committed_schedule_cache_is_valid():
    data = autotune.load_cache()
    for k, e in data.items():
        assert k.split(":")[0] in ("fp64", "fp32c") and k.split(":")[1] in autotune.CLASSES
        assert autotune.valid(e["best"])


def test_autotune_pick_prefers_smaller_skin_on_near_ties():
    t = [dict(unroll=4, skin=0.5, rate=1.00e9), dict(unroll=4, skin=0.75, rate=1.01e9),
         dict(unroll=4, skin=1.25, rate=1.025e9), dict(unroll=2, skin=0.5, rate=0.90e9)]
    assert autotune.pick(t) == dict(unroll=4, skin=0.5)  # 1.25 and 0.75 are within 3 % of 0.5
    t.append(dict(unroll=4, skin=1.0, rate=1.2e9))  # a clear winner is kept
    assert autotune.pick(t) == dict(unroll=4, skin=1.0)
    # equal skins: the faster unroll
    assert autotune.pick([dict(unroll=2, skin=0.5, rate=1.0e9), dict(unroll=4, skin=0.5, rate=1.005e9)]) == \
        dict(unroll=4, skin=0.5)


def test_autotune_pick_rates():
    # the rate pass keeps the overall unroll unless another depth is clearly faster
    t = [dict(rates_unroll=1, rate=0.90e9), dict(rates_unroll=2, rate=1.10e9), dict(rates_unroll=4, rate=1.00e9)]
    assert autotune.pick_rates(t, 4) == 2
    t[1]["rate"] = 1.005e9  # within 1 % of the overall unroll's own rate
    assert autotune.pick_rates(t, 4) == 0
    assert autotune.pick_rates(t, 2) == 0
