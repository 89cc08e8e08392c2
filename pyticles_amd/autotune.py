"""Launch-schedule autotuner for the pair passes.

These knobs set how the device walks a box's neighbours, and none changes
which pairs interact (only the summation order, so results agree to
rounding; the list entry width not even that):

    unroll  list entries whose loads are in flight per lane in the density,
            gradient and rate passes (1, 2 or 4): more hides more gather
            latency, fewer keeps more waves resident
    rates_unroll
            the same for the rate pass alone, whose walk keeps five records
            per entry in flight (two was best for one large FP64 box, four
            for batched boxes); tuned after the other two
    block   work-items per block of the three pair passes (128, 256 or
            512): how the slots, in cell order, are dealt to the XCDs and
            CUs
    list_bits
            neighbour-list entry width (16: slot offsets, half the list
            bytes, one 16-byte load per eight entries; 32: slots); same
            results either way
    fused   the leapfrog step inside the rate pass (1) or as a launch of
            its own (0); same results; tuned last
    skin    neighbour-list skin: lists hold pairs within r_cut + skin and are
            reused while no particle can have crossed half of it; a larger
            skin rebuilds less often and walks longer lists

The best choice depends on precision (FP64 registers are twice as wide), on
box size (many small boxes share launches with a sort over all of them; a
box of a million particles fills the chip many times over where one of a
quarter million fills it about once) and on density.  tune() times every candidate on a sample of the caller's workload
through the same entry points the workload uses and keeps the fastest;
results are cached per (precision, workload class) in schedules.json beside
this file so later runs start from the measured choice.

No counterpart in the reference (its loops have no launch schedule).
"""
import itertools
import json
import math
import os
import time

from . import _lib

UNROLLS = (1, 2, 4)
BLOCKS = (128, 256, 512)
LIST_BITS = (16, 32)
FUSED = (1, 0)
SKINS = (0.3, 0.5, 0.75, 1.0, 1.25)
CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "schedules.json")

DEFAULT = {_lib.FP64: dict(unroll=2, skin=0.5), _lib.FP32C: dict(unroll=4, skin=0.5)}


def prec_name(precision):
    return "fp32c" if int(precision) == _lib.FP32C else "fp64"


CLASSES = ("box", "large", "batched")
LARGE_N = 500_000  # single boxes above this many particles form their own class


def key(precision, cls):
    """Cache key of a workload class: 'box' (one box per launch), 'large'
    (one box of more than LARGE_N particles) or 'batched' (many boxes
    sharing an arena's launches)."""
    if cls not in CLASSES:
        raise ValueError("workload class must be one of %s, got %r" % (CLASSES, cls))
    return "%s:%s" % (prec_name(precision), cls)


def single_box_class(n):
    """Workload class of one box of n particles."""
    return "large" if int(n) > LARGE_N else "box"


def load_cache(path=CACHE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        data = json.load(f)
    if not isinstance(data, dict):
        raise ValueError("%s: expected a JSON object" % path)
    return data


def save_cache(data, path=CACHE):
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    os.replace(tmp, path)


def valid(s):
    return (isinstance(s, dict) and s.get("unroll") in UNROLLS and isinstance(s.get("skin"), (int, float))
            and 0.0 < float(s["skin"]) <= 2.0 and s.get("rates_unroll", 0) in (0,) + UNROLLS
            and s.get("block", 0) in (0,) + BLOCKS and s.get("list_bits", 0) in (0,) + LIST_BITS
            and s.get("fused", 1) in FUSED)


def schedule_for(precision, cls, path=CACHE):
    """The cached schedule of a workload class, else the precision's default.
    Returns dict(unroll, skin, source)."""
    entry = load_cache(path).get(key(precision, cls))
    if entry and valid(entry.get("best")):
        b = entry["best"]
        return dict(unroll=int(b["unroll"]), skin=float(b["skin"]), rates_unroll=int(b.get("rates_unroll", 0)),
                    block=int(b.get("block", 0)), list_bits=int(b.get("list_bits", 0)), fused=int(b.get("fused", 1)),
                    source="tuned")
    d = DEFAULT[int(precision)]
    return dict(d, rates_unroll=0, block=0, list_bits=0, fused=1, source="default")


def candidates():
    return [dict(unroll=u, skin=s) for u, s in itertools.product(UNROLLS, SKINS)]


def with_schedule(sysm, sched, precision):
    """A copy of a configs.System's parameters carrying the schedule's skin."""
    return dict(sysm.prm, skin=float(sched["skin"]), precision=int(precision))


def _time_box(sysm, sched, precision, steps, warm):
    from .device import from_system
    b = from_system(_Resched(sysm, with_schedule(sysm, sched, precision)))
    try:
        b.set_schedule(sched)
        b.step(warm)
        t = time.perf_counter()
        b.step(steps)
        dt = time.perf_counter() - t
    finally:
        b.close()
    return sysm.n * steps / dt


def _time_batched(systems, sched, precision, steps, warm):
    from .arena import Arena
    total = sum(s.n for s in systems)
    ar = Arena(total, precision=precision)
    try:
        ar.set_schedule(sched)
        for s in systems:
            ar.add(_Resched(s, with_schedule(s, sched, precision)), nsteps=steps + warm + 10)
        ar.step(warm)
        ar.poll_all()  # synchronises
        t = time.perf_counter()
        ar.step(steps)
        ar.poll_all()
        dt = time.perf_counter() - t
    finally:
        ar.close()
    return total * steps / dt


class _Resched:
    """A configs.System view with replaced parameters."""

    def __init__(self, s, prm):
        self.r, self.v, self.m, self.T, self.steps, self.n = s.r, s.v, s.m, s.T, s.steps, s.n
        self.prm = prm


def tune(systems, precision, cls=None, steps=None, warm=100, save=True, path=CACHE, log=None):
    """Time every candidate schedule on `systems` (one large box, or a list
    of boxes run batched) and return (best, table).  With save=True the
    result is stored under key(precision, cls) in the cache.

    The first `warm` steps are not timed: systems that start from a lattice
    rebuild their lists often until the start-up transient has passed, which
    favours large skins that long runs do not (a 300-step sample from the
    start picked skin 0.5 for one FP64 box; over 2000 steps 0.3 was 8%
    faster)."""
    if not isinstance(systems, (list, tuple)):
        systems = [systems]
    if cls is None:
        cls = single_box_class(systems[0].n) if len(systems) == 1 else "batched"
    single = cls != "batched"
    if steps is None:  # many rebuild cycles, so a skin's list cost and rebuild savings both show
        steps = 600 if single else 300
    table = []
    for c in candidates():
        if single:
            rate = _time_box(systems[0], c, precision, steps, warm)
        else:
            rate = _time_batched(systems, c, precision, steps, warm)
        table.append(dict(c, rate=rate))
        if log:
            log("tune %s unroll=%d skin=%.2f  %.3e particle-steps/s" % (key(precision, cls), c["unroll"],
                                                                        c["skin"], rate))
    best = pick(table)
    # then the rate pass's own depth at that schedule
    rtable = []
    for ru in UNROLLS:
        c = dict(best, rates_unroll=ru)
        if single:
            rate = _time_box(systems[0], c, precision, steps, warm)
        else:
            rate = _time_batched(systems, c, precision, steps, warm)
        rtable.append(dict(c, rate=rate))
        if log:
            log("tune %s unroll=%d skin=%.2f rates_unroll=%d  %.3e particle-steps/s" % (
                key(precision, cls), c["unroll"], c["skin"], ru, rate))
    best["rates_unroll"] = pick_rates(rtable, best["unroll"])
    table += rtable
    # then the pair passes' block size at that schedule
    btable = []
    for bs in BLOCKS:
        c = dict(best, block=bs)
        if single:
            rate = _time_box(systems[0], c, precision, steps, warm)
        else:
            rate = _time_batched(systems, c, precision, steps, warm)
        btable.append(dict(c, rate=rate))
        if log:
            log("tune %s unroll=%d skin=%.2f rates_unroll=%d block=%d  %.3e particle-steps/s" % (
                key(precision, cls), c["unroll"], c["skin"], c["rates_unroll"], bs, rate))
    best["block"] = pick_block(btable)
    table += btable
    # then the list entry width
    ltable = []
    for lb in LIST_BITS:
        c = dict(best, list_bits=lb)
        if single:
            rate = _time_box(systems[0], c, precision, steps, warm)
        else:
            rate = _time_batched(systems, c, precision, steps, warm)
        ltable.append(dict(c, rate=rate))
        if log:
            log("tune %s unroll=%d skin=%.2f rates_unroll=%d block=%d list_bits=%d  %.3e particle-steps/s" % (
                key(precision, cls), c["unroll"], c["skin"], c["rates_unroll"], c["block"], lb, rate))
    best["list_bits"] = pick_list_bits(ltable)
    table += ltable
    # then where the leapfrog step runs
    ftable = []
    for fu in FUSED:
        c = dict(best, fused=fu)
        if single:
            rate = _time_box(systems[0], c, precision, steps, warm)
        else:
            rate = _time_batched(systems, c, precision, steps, warm)
        ftable.append(dict(c, rate=rate))
        if log:
            log("tune %s unroll=%d skin=%.2f rates_unroll=%d block=%d list_bits=%d fused=%d  %.3e particle-steps/s" % (
                key(precision, cls), c["unroll"], c["skin"], c["rates_unroll"], c["block"], c["list_bits"], fu, rate))
    best["fused"] = pick_fused(ftable)
    table += ftable
    if save:
        data = load_cache(path)
        data[key(precision, cls)] = dict(
            best=best, table=table, particles=int(sum(s.n for s in systems)), boxes=len(systems),
            steps=int(steps))
        save_cache(data, path)
    return best, table


def pick(table, within=0.03):
    """The candidate with the smallest skin among those within `within` of
    the fastest rate (then the fastest of those): skins whose sampled rates
    differ by timing noise are not worth longer lists."""
    top = max(e["rate"] for e in table)
    near = [e for e in table if e["rate"] >= (1.0 - within) * top]
    best = min(near, key=lambda e: (e["skin"], -e["rate"]))
    return dict(unroll=best["unroll"], skin=best["skin"])


def pick_rates(rtable, unroll, within=0.01):
    """The rate pass's depth: the fastest, unless the overall unroll is
    within `within` of it (then 0, meaning the same as the other passes)."""
    top = max(rtable, key=lambda e: e["rate"])
    same = [e for e in rtable if e["rates_unroll"] == unroll]
    if same and same[0]["rate"] >= (1.0 - within) * top["rate"]:
        return 0
    return int(top["rates_unroll"])


def pick_block(btable, within=0.01):
    """The pair passes' block size: the fastest, unless the default (256)
    is within `within` of it (then 0, the default)."""
    top = max(btable, key=lambda e: e["rate"])
    same = [e for e in btable if e["block"] == 256]
    if same and same[0]["rate"] >= (1.0 - within) * top["rate"]:
        return 0
    return int(top["block"])


def rate_spread(table):
    """max / min rate over a tuning table (how much the choice matters)."""
    r = [e["rate"] for e in table]
    return max(r) / min(r) if r and min(r) > 0 else math.nan


def pick_list_bits(ltable, within=0.01):
    """16-bit entries (half the list traffic) unless 32-bit ones are more
    than `within` faster."""
    r = {e["list_bits"]: e["rate"] for e in ltable}
    if 16 in r and 32 in r and r[32] > (1.0 + within) * r[16]:
        return 32
    return 16


def pick_fused(ftable, within=0.01):
    """The leapfrog step inside the rate pass unless a launch of its own is
    more than `within` faster."""
    r = {e["fused"]: e["rate"] for e in ftable}
    if 1 in r and 0 in r and r[0] > (1.0 + within) * r[1]:
        return 0
    return 1
