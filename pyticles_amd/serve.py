"""Serving loop: a stream of simulation jobs on one device-resident arena.

Jobs (box of N particles, temperature, step count) arrive over time.  A job
is admitted as soon as its slots fit: the resident jobs' device footprint
(Arena memory for their particles) must stay within the HBM budget, and the
resident particle count within max_resident (the point past which more
resident work adds latency but no throughput).  Admitted jobs join the
running ones in the same launches at the next step: one arena step advances
every resident box by one leapfrog step, so jobs that arrive while others
run overlap with them on the device (continuous batching, no waiting for a
batch to drain).  A finished job's state is copied out, its slots are freed
and the next queued jobs take them.

Each box keeps its own neighbour-list schedule and its own slot range, so a
job's result does not depend on which other jobs shared its launches.

No counterpart in the reference, which runs one system per process.
"""
import collections
import time

import numpy as np

from . import _lib
from .arena import Arena, capacity_for, footprint  # noqa: F401  (footprint re-exported)

DONE = 3


class JobRecord:
    __slots__ = ("job_id", "n", "steps", "arrival", "admitted", "finished", "box", "result", "health", "error")

    def __init__(self, job_id, n, steps, arrival):
        self.job_id, self.n, self.steps, self.arrival = job_id, n, steps, arrival
        self.admitted = self.finished = None
        self.box = None
        self.result = None
        self.health = 0      # health bits of a job evicted on a numerical failure
        self.error = None    # message of a job refused at admission

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class Server:
    """Continuous-batching job server over one Arena.

    budget_bytes   HBM the arena may hold (its capacity is the largest slot
                   count whose footprint fits, capped by what max_resident
                   and the largest job can ever use)
    max_resident   resident particles beyond which jobs wait in the queue
    slack_jobs     slot headroom beyond max_resident, in largest jobs: a job
                   needs one contiguous slot range, and with a single job of
                   headroom, freed ranges too short for the queue's head
                   stalled admission (first-fit, FIFO) while the resident
                   count sat well under max_resident.  Launches cover slots
                   up to the last live box only, so headroom costs memory,
                   not work.
    chunk          arena steps launched between polls (admission granularity)
    keep_results   copy each finished job's state out (r, v, u, rho, E0, E1)
    schedule       dict(unroll=..., skin=...) for every job (autotune output)
    """

    def __init__(self, budget_bytes, precision=_lib.FP32C, max_resident=None, max_job=65536, chunk=4,
                 keep_results=False, schedule=None, arena=None, slack_jobs=16):
        self.precision = int(precision)
        self.chunk = int(chunk)
        self.keep_results = keep_results
        self.schedule = dict(schedule or {})
        cap = capacity_for(budget_bytes, self.precision) if arena is None else arena.capacity
        if max_resident is not None:
            cap = min(cap, int(max_resident) + max(1, int(slack_jobs)) * int(max_job))
        self.max_resident = int(max_resident) if max_resident is not None else cap
        self.budget_bytes = budget_bytes
        if arena is None:
            arena = Arena(cap, precision=self.precision)
            if "unroll" in self.schedule:
                arena.set_schedule(self.schedule)
        self.arena = arena
        self.capacity = cap
        self.resident = 0
        self.queue = collections.deque()
        self.running = {}
        self.records = []
        self.arena_steps = 0
        self.particle_steps = 0
        self.max_running = 0
        self.peak_resident = 0

    def close(self):
        self.arena.close()

    # -- admission ----------------------------------------------------------
    def _fits(self, n):
        if self.resident + n > self.max_resident:
            return False
        return footprint(self.resident + n, self.precision) <= self.budget_bytes

    def _admit(self, now):
        while self.queue:
            rec, sysm = self.queue[0]
            if not self._fits(rec.n):
                break
            if self.schedule.get("skin") is not None:
                sysm = _with_skin(sysm, self.schedule["skin"])
            try:
                bid = self.arena.add(sysm, nsteps=rec.steps)
            except _lib.SphError as e:
                if e.code == _lib.SPH_ERR_NOMEM:  # no contiguous room yet
                    break
                if e.code != _lib.SPH_ERR_ARG:
                    raise
                # a job the arena refuses (invalid or too dense for its
                # lists) fails alone; the stream goes on
                self.queue.popleft()
                rec.error = str(e)
                rec.finished = now
                continue
            self.queue.popleft()
            rec.box = bid
            rec.admitted = now
            self.running[bid] = rec
            self.resident += rec.n
            self.max_running = max(self.max_running, len(self.running))
            self.peak_resident = max(self.peak_resident, self.resident)

    def _retire(self, now):
        if not self.running:
            return
        st, k, hl = self.arena.poll_all(max(self.running) + 1)
        for bid in list(self.running):
            if st[bid] == DONE or hl[bid]:
                rec = self.running.pop(bid)
                rec.health = int(hl[bid])
                if self.keep_results:
                    rec.result = self.arena.result(bid)
                rec.finished = now()
                self.arena.remove(bid)
                self.resident -= rec.n

    # -- the loop -------------------------------------------------------------
    def run(self, jobs, arrivals, clock=time.perf_counter, log=None):
        """Serve `jobs` (configs.System, each with .steps) arriving at times
        `arrivals` (s, non-decreasing, relative to the call).  Returns the
        job records in submission order."""
        pending = collections.deque(zip(range(len(jobs)), jobs, arrivals))
        recs = []
        t0 = clock()

        def now():
            return clock() - t0

        last_log = 0.0
        while pending or self.queue or self.running:
            t = now()
            while pending and pending[0][2] <= t:
                j, sysm, a = pending.popleft()
                rec = JobRecord(j, sysm.n, sysm.steps, a)
                recs.append(rec)
                self.queue.append((rec, sysm))
            self._admit(t)
            if not self.running:
                if pending:  # idle until the next arrival
                    wait = pending[0][2] - now()
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                    continue
                if self.queue:
                    rec = self.queue[0][0]
                    raise RuntimeError("job %d (N=%d) can never fit: capacity %d particles, budget %.3g B"
                                       % (rec.job_id, rec.n, self.capacity, self.budget_bytes))
                break
            self.arena.step(self.chunk)
            self.arena_steps += self.chunk
            self._retire(now)
            if log and now() - last_log > 5.0:
                last_log = now()
                log("serve t=%.1fs done=%d running=%d queued=%d resident=%d" % (
                    last_log, sum(r.finished is not None for r in recs), len(self.running), len(self.queue),
                    self.resident))
        self.records = recs
        self.particle_steps = int(sum(r.n * r.steps for r in recs
                                      if r.finished is not None and not r.health and r.error is None))
        return recs


def _with_skin(sysm, skin):
    class _S:
        pass
    s = _S()
    s.r, s.v, s.m, s.T, s.steps, s.n = sysm.r, sysm.v, sysm.m, sysm.T, sysm.steps, sysm.n
    s.prm = dict(sysm.prm, skin=float(skin))
    return s


def latency_summary(recs, warmup=0):
    """p50 / p99 / mean job latency (s) over the jobs after the first
    `warmup` arrivals; throughput over every finished job, from the first
    arrival to the last completion."""
    done = [r for r in recs if r.finished is not None]
    meas = sorted(recs, key=lambda r: r.arrival)[warmup:]
    lat = np.array([r.latency for r in meas if r.latency is not None])
    if lat.size == 0 or not done:
        return dict(jobs=0)
    first = min(r.arrival for r in recs)
    last = max(r.finished for r in done)
    ps = sum(r.n * r.steps for r in done if not r.health and r.error is None)  # completed work only
    return dict(jobs=int(lat.size), p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), max_s=float(lat.max()), window_s=float(last - first),
                particle_steps=int(ps), particle_steps_per_s=float(ps / max(last - first, 1e-9)),
                failed=int(sum(1 for r in recs if r.health)),
                refused=int(sum(1 for r in recs if r.error is not None)))
