"""Benchmark of the smooth-particle hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1|c3]
                    [--precision fp64|fp32c] [--legs c2x,c3,c4,c5] [--no-cpu]
                    [--no-extra] [--tune PATH] [--calib-copy] [--c4-sample STEPS]

Headline: one periodic box 512 x 512, N = 262,144 jittered lattice, T = 0.9,
full physics (pressure, cohesion, viscosity eta = 1, conduction k = 5),
FP64, leapfrog dt = 0.005.  A step is one leapfrog step of the whole hot
path (cell build + sort when due, density + EOS, gradients, rates,
kick/drift); inputs are resident in HBM when the timed region starts.

Further legs (each adds an object to the same JSON line):
    c2x  the same box in the FP32-compensated form, errors of both forms
         against the FP64 oracle
    c3   quench, N = 1,048,576, FP32-compensated vs FP64 on the device:
         per-field error against the FP64 oracle and total-energy drift
    c4   512 boxes of N log-uniform in [256, 16384], 500 steps each,
         batched FP32-compensated arena vs the plain FP64 form one box per
         launch (the >= 5x comparison)
    c5   served Poisson job stream: throughput, p50 / p99 job latency
--tune PATH times every launch schedule for each workload class and writes
the table to PATH (the autotuner's cache format).  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

# Algorithmic HBM bytes per particle of each kernel: the fixed per-pass
# figures (every per-particle array a logical pass must read or write counted
# once at its storage width; neighbour re-reads and list words not counted;
# sort/permute counted): build 124 (FP32C 116), density r, m -> rho, rho_lr
# 40 (28), gradients 104 (72), rates 112 (76; inviscid 64 / 52), leapfrog
# writes 40.  The density kernel also computes the cohesive part of the rates
# and the rate kernel adds it; the figures stay those of the logical passes.
B_PS = {
    ("fp64", True): dict(build=124, density=40, gradients=104, rates=112, integrate=40),
    ("fp64", False): dict(build=124, density=40, gradients=0, rates=64, integrate=40),
    ("fp32c", True): dict(build=116, density=28, gradients=72, rates=76, integrate=40),
    ("fp32c", False): dict(build=116, density=28, gradients=0, rates=52, integrate=40),
}
KERNEL = {"fp64": dict(density="k_density64", gradients="k_grad64", rates="k_rates64"),
          "fp32c": dict(density="k_density32", gradients="k_grad32", rates="k_rates32")}
HBM_PEAK_GBS = 8000.0


def workload(name, precision):
    from pyticles_amd import _lib, configs
    if name == "c1":
        s = configs.c1()
        desc = "C1: FP64 vdW SPH, periodic 32x32, N=1024 jittered lattice, T=0.9, h=2, h_l=4, inviscid, non-conducting"
    elif name == "c3":
        s = configs.c3()
        desc = "C3: periodic 1024x1024, N=1,048,576, quench T 1.2 -> 0.4, full physics"
    else:
        s = configs.c2()
        desc = ("C2: one periodic box 512x512, N=262,144 jittered lattice, T=0.9, h=2, h_l=4, vdW a=2 b=0.5, "
                "eta=1, k=5, leapfrog dt=0.005")
    s.prm["precision"] = _lib.FP32C if precision == "fp32c" else _lib.FP64
    return s, desc


def schedule(s, cls):
    """Apply the autotuned (or default) schedule of a workload class to a
    configs.System: sets its skin and returns (launch schedule, schedule);
    the first goes to set_schedule (pass unrolls)."""
    from pyticles_amd import autotune
    if cls == "box":
        cls = autotune.single_box_class(s.n)
    sc = autotune.schedule_for(s.prm["precision"], cls)
    s.prm["skin"] = sc["skin"]
    return sc, sc


def device_box(s, unroll):
    """A box holding s with the launch schedule set before its first pass."""
    from pyticles_amd.device import DeviceBox
    b = DeviceBox(s.n, s.prm)
    b.set_schedule(unroll)
    b.upload(r=s.r, v=s.v, m=s.m, u=np.zeros(s.n))
    b.init_thermal(s.T)
    return b


def pcie_inclusive(s, steps, unroll):
    """Particle-steps/s of a whole job with host buffers: upload of r, v, m,
    the initial density pass, the steps and the download of r, v, u (device
    allocation excluded)."""
    from pyticles_amd.device import DeviceBox
    b = DeviceBox(s.n, s.prm)
    b.set_schedule(unroll)
    t0 = time.perf_counter()
    b.upload(r=s.r, v=s.v, m=s.m, u=np.zeros(s.n))
    b.init_thermal(s.T)
    b.step(steps)
    b.download("r", "v", "u")
    el = time.perf_counter() - t0
    b.close()
    return {"value": s.n * steps / el, "unit": "particle-steps/s", "seconds": el,
            "includes": "host->device r, v, m; initial density pass; %d steps; device->host r, v, u" % steps}


def run_gpu(s, steps, warmup, timing=True, unroll=0, kernel_steps=200, launch_reps=50):
    """Warm up, then time `steps` steps as the library runs them (captured
    step graphs replayed, no events).  With timing, a second region of
    min(steps, kernel_steps) steps follows with HIP events recorded on the
    context stream around every pass launch (direct launches: events cannot
    sit between the kernels of a replayed graph); its per-pass times are
    returned in tm with tm["steps"] = the steps of that region.  Then each
    pair pass is launched launch_reps times back to back between two events
    (tm["launch_us"]: microseconds per launch)."""
    from pyticles_amd import _lib
    b = device_box(s, unroll)
    if warmup:
        b.step(warmup)
    b.energy()  # synchronises the stream
    E0 = b.energy()[2]
    t0 = time.perf_counter()
    b.step(steps)  # returns after the stream is drained (health check)
    t1 = time.perf_counter()
    run_stats = b.run_stats()  # list builds and postponed builds inside the timed steps
    E1 = b.energy()[2]
    tm = None
    if timing:
        ks = min(steps, kernel_steps)
        _lib.check(_lib.lib.sph_ctx_set_timing(b.ctx, 1))
        b.timing(reset=True)
        b.step(ks)
        tm = b.timing()
        tm["steps"] = ks
        _lib.check(_lib.lib.sph_ctx_set_timing(b.ctx, 0))
        # each pair pass alone: events around back-to-back launches at the
        # state the steps left (no per-launch event overhead)
        tm["launch_us"] = {p: b.time_pass(p, reps=launch_reps) for p in b.PASSES}
        tm["launch_reps"] = launch_reps
        # the list build stage alone, back to back at the same state (it may
        # reorder the slots, so after the pair passes)
        tm["build_us"] = b.time_pass(b.BUILD, reps=20)
    if tm is not None:
        tm["timed_run"] = run_stats
    return b, t1 - t0, tm, (E0, E1)


def host_cpu():
    """The host CPU's model name and the cores this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count()
    return model, ncpu


def _oracle_rate(O, s, u0, nsteps, threaded, runs=5):
    """Median particle-steps/s of `runs` oracle runs of nsteps steps each,
    after a one-step warm-up."""
    O.run(s.r, s.v, s.m, u0, s.prm, 1, terms=s.prm["terms"], threaded=threaded)
    rates = []
    for _ in range(runs):
        t = time.perf_counter()
        O.run(s.r, s.v, s.m, u0, s.prm, nsteps, terms=s.prm["terms"], threaded=threaded)
        rates.append(s.n * nsteps / (time.perf_counter() - t))
    return float(np.median(rates)), rates


def cpu_baseline(s, full=False, max_run_s=3.0):
    """The FP64 oracle (C restatement of the reference's pair-list loops) on
    the same workload: single-threaded (as the reference runs) and with
    OpenMP on every host core the process may use.  Each figure is the
    median of 5 runs after a one-step warm-up; a run is the whole job when
    full=True (C1), else 10 steps, fewer when a step takes over a tenth of
    max_run_s (the 1M-particle box)."""
    from oracle import oracle as O
    u0 = O.init_u(s.r, s.m, s.T, s.prm)
    t = time.perf_counter()
    O.run(s.r, s.v, s.m, u0, s.prm, 1, terms=s.prm["terms"])
    one = time.perf_counter() - t  # 1 step = 2 rate evaluations
    nsteps = int(s.steps) if full else max(1, min(10, int(max_run_s / max(one / 2, 1e-6))))
    v1, runs1 = _oracle_rate(O, s, u0, nsteps, threaded=False)
    vt, runst = _oracle_rate(O, s, u0, nsteps, threaded=True)
    model, ncpu = host_cpu()
    nth = O.num_threads(True)
    return dict(value=v1, unit="particle-steps/s", cores=1, kind="port",
                sample="oracle/oracle.c FP64 on the same initial state, %d-step runs of N=%d (%d rate "
                       "evaluations each), median of 5 after a 1-step warm-up; pair list rebuilt every "
                       "evaluation" % (nsteps, s.n, nsteps + 1),
                host=dict(cpu=model, cores_available=ncpu),
                runs=runs1,
                all_cores=dict(value=vt, threads=nth, runs=runst,
                               note="OpenMP build of the same source (per-thread pair sums)"))


def error_vs_oracle(s, r=None, v=None, u=None, unroll=0):
    """Normwise errors of one rate evaluation of the device path against the
    FP64 oracle, at the initial state of s or at the state (r, v, u):
    max|x - x_ref| / max|x_ref| per field, and the RMS ratio as <field>_rms."""
    from oracle import oracle as O
    r = s.r if r is None else r
    v = s.v if v is None else v
    u = O.init_u(s.r, s.m, s.T, s.prm) if u is None else u
    ref = O.derivatives(r, v, s.m, u, s.prm, terms=s.prm["terms"])
    with device_box(s, unroll) as b:
        b.upload(r=r, v=v, u=u)
        b.rates(s.prm["terms"])
        got = b.download("rho", "rho_lr", "acc", "udot")
    out = {}
    for k in ("rho", "rho_lr", "acc", "udot"):
        d = got[k] - ref[k]
        out[k] = float(np.max(np.abs(d)) / np.max(np.abs(ref[k])))
        out[k + "_rms"] = float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(ref[k] * ref[k])))
    return out


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


# ---- legs ------------------------------------------------------------------

def leg_c2x(value64, steps, warmup):
    from pyticles_amd import _lib
    s32, _ = workload("c2", "fp32c")
    un, sc = schedule(s32, "box")
    b32, el32, _, (e0, e1) = run_gpu(s32, steps, warmup, timing=False, unroll=un)
    b32.close()
    v32 = s32.n * steps / el32
    out = {"value": v32, "unit": "particle-steps/s", "ms_per_step": 1e3 * el32 / steps,
           "speedup_vs_fp64": v32 / value64, "energy_drift_rel": abs(e1 - e0) / abs(e0),
           "schedule": sc}
    out["rel_err_vs_fp64_oracle"] = error_vs_oracle(s32, unroll=un)
    s64, _ = workload("c2", "fp64")
    un64, _ = schedule(s64, "box")
    out["fp64_rel_err_vs_oracle"] = error_vs_oracle(s64, unroll=un64)
    assert s32.prm["precision"] == _lib.FP32C
    return out


def leg_c3(steps):
    """Quench at N = 1,048,576: both device forms run the same steps; errors
    of one rate evaluation against the FP64 oracle at the initial state and
    at the FP64 run's final (phase-separating) state."""
    from pyticles_amd import _lib, configs
    out = {"workload": "C3: periodic 1024x1024, N=1,048,576 jittered lattice at T=1.2, velocities rescaled to "
                       "T=0.4 and u reset at step 0, eta=1, k=5, dt=0.005, %d steps" % steps}
    res = {}
    for name, prec in (("fp32c", _lib.FP32C), ("fp64", _lib.FP64)):
        s = configs.c3(precision=prec)
        un, sc = schedule(s, "box")
        log("c3 %s %d steps" % (name, steps))
        b, el, _, (e0, e1) = run_gpu(s, steps, 0, timing=False, unroll=un)
        fin = b.download("r", "v", "u", "rho")
        b.close()
        res[name] = (s, un, fin)
        out[name] = {"value": s.n * steps / el, "unit": "particle-steps/s", "ms_per_step": 1e3 * el / steps,
                     "energy_drift_rel": abs(e1 - e0) / abs(e0), "E0": e0, "E1": e1, "schedule": sc,
                     "rho_range_final": [float(fin["rho"].min()), float(fin["rho"].max())]}
    out["fp32c"]["speedup_vs_fp64"] = out["fp32c"]["value"] / out["fp64"]["value"]
    s32, un32, _ = res["fp32c"]
    s64, un64, fin64 = res["fp64"]
    log("c3 errors vs oracle")
    out["fp32c"]["rel_err_vs_fp64_oracle_initial"] = error_vs_oracle(s32, unroll=un32)
    out["fp32c"]["rel_err_vs_fp64_oracle_final_state"] = error_vs_oracle(
        s32, r=fin64["r"], v=fin64["v"], u=fin64["u"], unroll=un32)
    out["fp64"]["rel_err_vs_fp64_oracle_final_state"] = error_vs_oracle(
        s64, r=fin64["r"], v=fin64["v"], u=fin64["u"], unroll=un64)
    return out


def sample_errors(systems, states, unroll=0):
    """Per-field FP32C error against the FP64 oracle at the given states
    (r, v, u of each system): the maximum over the sample of each field's
    normwise error."""
    worst = {}
    for s, st in zip(systems, states):
        e = error_vs_oracle(s, r=st["r"], v=st["v"], u=st["u"], unroll=unroll)
        for k, x in e.items():
            worst[k] = max(worst.get(k, 0.0), x)
    return worst


def run_arena(boxes, precision, chunk=25):
    """All boxes in one arena until every box is done; (seconds, results)."""
    from pyticles_amd import autotune
    from pyticles_amd.arena import Arena
    sc = autotune.schedule_for(precision, "batched")
    total = sum(b.n for b in boxes)
    ar = Arena(total, precision=precision)
    ar.set_schedule(sc)
    ids = [ar.add(autotune._Resched(b, autotune.with_schedule(b, sc, precision))) for b in boxes]
    ar.poll_all()
    t = time.perf_counter()
    while True:
        ar.step(chunk)
        st, _, hl = ar.poll_all()
        if np.all((st[ids] == 3) | (hl[ids] != 0)):
            break
    el = time.perf_counter() - t
    res = [ar.result(i) for i in ids]
    _, _, hl = ar.poll_all()
    ar.close()
    return el, res, int(np.count_nonzero(hl[ids])), sc


def leg_c4(nbox=512, plain_every=4):
    """Batched FP32C (and FP64) arena over all boxes vs the plain FP64 form,
    one box per launch sequence, on every plain_every-th box."""
    from pyticles_amd import _lib, configs
    boxes = [configs.c4_box(b) for b in range(nbox)]
    steps = boxes[0].steps
    ps = sum(b.n for b in boxes) * steps
    out = {"workload": "C4: %d periodic boxes, N_b log-uniform in [256, 16384] (total %d), density 1, T_b in "
                       "[0.3, 1.0], eta=1, k=5, %d steps each" % (nbox, sum(b.n for b in boxes), steps)}
    for name, prec in (("batched_fp32c", _lib.FP32C), ("batched_fp64", _lib.FP64)):
        bx = [configs.c4_box(b, precision=prec) for b in range(nbox)] if prec != _lib.FP32C else boxes
        log("c4 %s" % name)
        el, res, failed, sc = run_arena(bx, prec)
        drift = np.array([abs(r["E1"] - r["E0"]) / abs(r["E0"]) for r in res])
        out[name] = {"value": ps / el, "unit": "particle-steps/s", "seconds": el, "failed_boxes": failed,
                     "energy_drift_rel_median": float(np.median(drift)), "energy_drift_rel_max": float(drift.max()),
                     "schedule": sc}
        if prec == _lib.FP32C:
            pick = list(range(0, nbox, max(1, nbox // 8)))
            log("c4 fp32c errors vs oracle on %d boxes" % len(pick))
            out[name]["rel_err_vs_fp64_oracle_final_states"] = sample_errors(
                [bx[i] for i in pick], [res[i] for i in pick], unroll=sc["unroll"])
            out[name]["error_sample"] = "boxes %s at their final (step %d) states, one FP32C rate evaluation each " \
                                        "against the FP64 oracle; maximum over the sample per field" % (pick, steps)
    sample = list(range(0, nbox, plain_every))
    el = 0.0
    pss = 0
    log("c4 plain fp64, %d boxes one at a time" % len(sample))
    for i in sample:
        sb = configs.c4_box(i, precision=_lib.FP64)
        un, _ = schedule(sb, "box")  # the tuned single-box schedule
        b = device_box(sb, un)
        b.step(1)
        t = time.perf_counter()
        b.step(steps)
        el += time.perf_counter() - t
        pss += sb.n * steps
        b.close()
    out["plain_fp64_one_box_per_launch"] = {
        "value": pss / el, "unit": "particle-steps/s", "seconds": el,
        "sample": "boxes %s (every %dth of %d), %d steps each, tuned single-box schedule, timed per box "
                  "excluding set-up" % (
            "0,%d,...,%d" % (plain_every, sample[-1]), plain_every, nbox, steps)}
    out["batched_fp32c_over_plain_fp64"] = out["batched_fp32c"]["value"] / out["plain_fp64_one_box_per_launch"]["value"]
    out["batched_fp64_over_plain_fp64"] = out["batched_fp64"]["value"] / out["plain_fp64_one_box_per_launch"]["value"]
    return out


def leg_c5(njobs, warm, rate, max_resident, budget):
    from pyticles_amd import _lib, autotune, configs
    from pyticles_amd.serve import Server, latency_summary
    sc = autotune.schedule_for(_lib.FP32C, "batched")
    out = {"workload": "C5: Poisson job stream, N log-uniform in [1024, 65536], steps in [100, 2000], T in "
                       "[0.3, 1.2], eta=1, k=5, FP32C", "budget_bytes": budget, "max_resident": max_resident,
           "schedule": sc}

    def serve(nj, wj, lam, tag):
        jobs = [configs.c5_job(j) for j in range(nj)]
        arr = configs.c5_arrivals(nj, rate=lam)
        offered = sum(j.n * j.steps for j in jobs) / arr[-1]
        log("c5 %s: %d jobs at %.1f/s (offered %.3g particle-steps/s)" % (tag, nj, lam, offered))
        srv = Server(budget, precision=_lib.FP32C, max_resident=max_resident, schedule=sc, chunk=4)
        recs = srv.run(jobs, arr, log=log)
        srv.close()
        sm = latency_summary(recs, warmup=wj)
        sm.update(rate_jobs_per_s=lam, jobs_total=nj, warmup_jobs=wj, offered_particle_steps_per_s=offered,
                  arena_capacity=srv.capacity)
        sm["saturated"] = bool(sm["particle_steps_per_s"] < 0.9 * offered)
        return sm

    # per-field FP32C error on jobs of the stream's distribution: four jobs
    # of different sizes stepped by the FP64 device path (their own step
    # counts, at most 400), then one FP32C rate evaluation against the oracle
    from pyticles_amd.device import from_system
    sample, states = [], []
    for j in (3, 11, 29, 47):
        s64 = configs.c5_job(j, precision=_lib.FP64)
        with from_system(s64) as b:
            b.step(min(s64.steps, 400))
            states.append(b.download("r", "v", "u"))
        sample.append(configs.c5_job(j))
    log("c5 fp32c errors vs oracle on jobs %s" % [s.n for s in sample])
    out["rel_err_vs_fp64_oracle"] = sample_errors(sample, states, unroll=sc["unroll"])
    out["error_sample"] = "jobs 3, 11, 29, 47 (N = %s) after min(their steps, 400) FP64 device steps; one FP32C " \
                          "rate evaluation each against the FP64 oracle; maximum per field" % [s.n for s in sample]
    out["stream_200"] = serve(njobs, warm, rate, "stream")
    cap = out["stream_200"]["particle_steps_per_s"]
    mean_work = np.mean([j.n * j.steps for j in (configs.c5_job(k) for k in range(200))])
    lam2 = 0.6 * cap / mean_work
    n2 = max(100, min(600, int(lam2 * 15)))
    out["stream_below_capacity"] = serve(n2, n2 // 10, lam2, "below capacity")
    return out


def tune_all(path):
    from pyticles_amd import _lib, autotune, configs
    res = {}
    for prec in (_lib.FP64, _lib.FP32C):
        s = configs.c2()
        s.prm["precision"] = prec
        best, table = autotune.tune(s, prec, "box", path=path, log=log)
        res[autotune.key(prec, "box")] = best
        big = configs.c3(precision=prec)
        best, table = autotune.tune(big, prec, "large", steps=400, path=path, log=log)
        res[autotune.key(prec, "large")] = best
        boxes = [configs.c4_box(b, precision=prec) for b in range(0, 512, 8)]
        best, table = autotune.tune(boxes, prec, "batched", path=path, log=log)
        res[autotune.key(prec, "batched")] = best
    return res


def load_traffic(key, kernel):
    """HBM bytes per launch of a kernel from the committed counter summary
    (profiles/traffic.json: {workload_precision: {kernel: bytes}})."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        entry = json.load(open(p)).get(key)
    except (OSError, ValueError):
        return None
    return entry.get(kernel) if isinstance(entry, dict) else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32c"])
    ap.add_argument("--legs", default="c2x,c3,c4,c5", help="comma list of extra legs (c2x, c3, c4, c5) or ''")
    ap.add_argument("--c3-steps", type=int, default=2000)
    ap.add_argument("--c5-jobs", type=int, default=2200)
    ap.add_argument("--c5-warmup", type=int, default=200)
    ap.add_argument("--c5-rate", type=float, default=200.0)
    ap.add_argument("--c5-resident", type=int, default=4_000_000)
    ap.add_argument("--c5-budget", type=float, default=200e9)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip every extra leg")
    ap.add_argument("--tune", metavar="PATH", help="run the autotuner for every workload class, write PATH, exit")
    ap.add_argument("--list-bits", type=int, default=0, choices=[0, 16, 32],
                    help="neighbour-list entry width (0: the schedule's, else 16)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a plain device-to-device copy of 256 MiB (HBM counter calibration)")
    ap.add_argument("--c4-sample", type=int, default=0, metavar="STEPS",
                    help="only step the batched FP32C arena of the 512 C4 boxes STEPS times (for profilers)")
    args = ap.parse_args()

    if args.c4_sample:
        from pyticles_amd import _lib, autotune, configs
        from pyticles_amd.arena import Arena
        sc = autotune.schedule_for(_lib.FP32C, "batched")
        boxes = [configs.c4_box(b) for b in range(512)]
        ar = Arena(sum(b.n for b in boxes), precision=_lib.FP32C)
        ar.set_schedule(sc)
        for b in boxes:
            ar.add(autotune._Resched(b, autotune.with_schedule(b, sc, _lib.FP32C)))
        t = time.perf_counter()
        ar.step(args.c4_sample)
        ar.poll_all()
        el = time.perf_counter() - t
        print(json.dumps({"c4_sample_steps": args.c4_sample, "particles": sum(b.n for b in boxes),
                          "ms_per_step": 1e3 * el / args.c4_sample}))
        ar.close()
        return

    if args.calib_copy:
        # reference kernel of known traffic for the FETCH_SIZE / WRITE_SIZE
        # scale factor: reads 256 MiB and writes 256 MiB once
        import torch
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes_read": x.numel() * 4, "calib_copy_bytes_written": y.numel() * 4}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the stepping path runs on one GPU; --gpus %d is not supported (boxes would shard by "
              "whole box across GPUs, which is not implemented)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.tune:
        print(json.dumps({"tuned": tune_all(args.tune), "path": args.tune}), flush=True)
        return

    s, desc = workload(args.workload, args.precision)
    unroll, sched = schedule(s, "box")
    if args.list_bits:
        sched["list_bits"] = args.list_bits
    full = bool(s.prm["terms"] & 12)
    bps = B_PS[(args.precision, full)]
    log("%s, %s, %d steps after %d warmup, schedule %s" % (desc, args.precision, args.steps, args.warmup, sched))
    b, el, tm, (E0, E1) = run_gpu(s, args.steps, args.warmup, unroll=unroll)
    nip = int(b.pairs().shape[0])
    b.close()
    value = s.n * args.steps / el
    evals = args.steps + 1
    kevals = int(tm["steps"]) + 1  # rate evaluations of the kernel-timing region
    # the rate pass takes the leapfrog step in the same launch, unless the
    # schedule runs it as a launch of its own ("leapfrog")
    per = {k: tm[k + "_ms"] / kevals for k in ("build", "density", "gradients", "rates")}
    if tm["integrate_ms"] > 0:
        per["leapfrog"] = tm["integrate_ms"] / kevals
    # the dominant pair-pass kernel by measured time, timed alone
    lus = tm["launch_us"]
    dom = max(("density", "gradients", "rates"), key=lambda k: lus[k])
    dom_ms = lus[dom] * 1e-3
    alg_bytes = bps[dom] * s.n
    achieved = alg_bytes / (dom_ms * 1e-3) / 1e9
    kname = KERNEL[args.precision][dom]
    traffic = load_traffic("%s_%s" % (args.workload, args.precision), kname)
    roof = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 5), traffic=traffic,
                kernel=kname, kernel_avg_ms=round(dom_ms, 5), alg_bytes_per_launch=alg_bytes,
                alg_bytes_per_particle=bps[dom],
                timing="HIP events on the context stream around %d back-to-back launches of the pass at the "
                       "state the timed steps left (the passes write only derived fields)" % int(tm["launch_reps"]),
                passes_us_per_launch={k: round(v, 2) for k, v in lus.items()})
    step_bytes = sum(bps.values())
    line = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+kahan/f64",
        "data": "synthetic (NumPy PCG64 seeded jittered lattice, Gaussian velocities)",
        "config": {"workload": desc, "N": s.n, "box": [s.prm["Lx"], s.prm["Ly"]], "precision": args.precision,
                   "timed": "%d leapfrog steps (%d rate evaluations incl. the closing half kick) in one sph_step "
                            "call; step graphs captured at box set-up, outside the timed region" % (args.steps, evals),
                   "schedule": sched},
        "timed_list_builds": int(tm["timed_run"]["builds"]),
        # the whole list build stage (sort, permutation, lists) at the end
        # state, 20 back-to-back launches between two events
        "build_stage_us": round(tm["build_us"], 2),
        "timed_builds_postponed": int(tm["timed_run"]["extends"]),
        "roofline": roof,
        "pair_interactions_per_s": nip * args.steps / el,
        "nip": nip,
        # events around every pass launch of 200 further steps (includes each
        # launch's dispatch latency; the breakdown of a step)
        "passes_ms_per_eval": {k: round(v, 5) for k, v in per.items()},
        "step_roofline_frac": round(value * step_bytes / (HBM_PEAK_GBS * 1e9), 5),
        "energy_drift_rel": abs(E1 - E0) / abs(E0),
        "pcie_inclusive": pcie_inclusive(s, args.steps, unroll),
    }
    legs = [] if args.no_extra else [x for x in args.legs.split(",") if x]
    is_head = args.workload == "c2" and args.precision == "fp64"
    if "c2x" in legs and is_head:
        log("c2 fp32c")
        line["c2_fp32c"] = leg_c2x(value, args.steps, args.warmup)
    if "c3" in legs:
        line["c3_precision"] = leg_c3(args.c3_steps)
    if "c4" in legs:
        line["c4_batched"] = leg_c4()
    if "c5" in legs:
        line["c5_served"] = leg_c5(args.c5_jobs, args.c5_warmup, args.c5_rate, args.c5_resident, args.c5_budget)
    if not args.no_cpu:
        log("cpu baseline")
        s_cpu, _ = workload(args.workload, "fp64")
        cb = cpu_baseline(s_cpu, full=args.workload == "c1")
        cb["gpu_over_cpu"] = value / cb["value"]
        cb["all_cores"]["gpu_over_cpu"] = value / cb["all_cores"]["value"]
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
