"""Benchmark of the smooth-particle hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1|c3]
                    [--precision fp64|fp32c] [--no-cpu] [--no-extra]

Default workload: one periodic box 512 x 512, N = 262,144 jittered lattice,
T = 0.9, full physics (pressure, cohesion, viscosity eta = 1, conduction
k = 5), FP64, leapfrog dt = 0.005.  A step is one leapfrog step of the whole
hot path (cell build + sort, density + EOS, gradients, rates, kick/drift);
inputs are resident in HBM when the timed region starts.  Prints one JSON
line (rank 0, N = 1 only).
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

# algorithmic HBM bytes per particle-step of each pass (fixed accounting:
# every per-particle array a pass must read or write counted once at its
# storage width; neighbour re-reads not counted; sort/permute counted).  The
# rates pass reads r, v, m, rho, rho_lr, u, grad v, J_q; its leapfrog writes
# (r', v', u', 40 B) are counted on the kick/drift kernel.
B_PS = {
    ("fp64", True): dict(build=124, density=40, gradients=104, rates=112, integrate=40),
    ("fp64", False): dict(build=124, density=40, gradients=0, rates=64, integrate=40),
    ("fp32c", True): dict(build=116, density=28, gradients=72, rates=76, integrate=40),
    ("fp32c", False): dict(build=116, density=28, gradients=0, rates=52, integrate=40),
}
HBM_PEAK_GBS = 8000.0
RATES_KERNEL = {"fp64": "k_rates64", "fp32c": "k_rates32"}


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


def run_gpu(s, steps, warmup, timing=True):
    from pyticles_amd import _lib
    from pyticles_amd.device import from_system
    b = from_system(s)
    if warmup:
        b.step(warmup)
    _lib.check(_lib.lib.sph_ctx_set_timing(b.ctx, 1 if timing else 0))
    b.timing(reset=True)
    b.energy()  # synchronises the stream
    E0 = b.energy()[2]
    t0 = time.perf_counter()
    b.step(steps)  # returns after the stream is drained (health check)
    t1 = time.perf_counter()
    E1 = b.energy()[2]
    tm = b.timing()
    _lib.check(_lib.lib.sph_ctx_set_timing(b.ctx, 0))
    return b, t1 - t0, tm, (E0, E1)


def cpu_baseline(s, target_s=12.0):
    """The FP64 oracle (single-threaded C restatement of the reference's
    pair-list loops) on a bounded sample of the same workload."""
    from oracle import oracle as O
    u0 = O.init_u(s.r, s.m, s.T, s.prm)
    t = time.perf_counter()
    O.run(s.r, s.v, s.m, u0, s.prm, 1, terms=s.prm["terms"])
    one = time.perf_counter() - t  # 1 step = 2 evaluations
    nsteps = max(1, int(target_s / max(one / 2, 1e-6)) - 1)
    t = time.perf_counter()
    O.run(s.r, s.v, s.m, u0, s.prm, nsteps, terms=s.prm["terms"])
    el = time.perf_counter() - t
    return dict(value=s.n * nsteps / el, unit="particle-steps/s", cores=1, kind="port",
                sample="oracle/oracle.c FP64 on the same initial state, %d steps of N=%d (%d rate evaluations, "
                       "%.1f s); single thread, pair list rebuilt every evaluation" % (nsteps, s.n, nsteps + 1, el))


def error_vs_oracle(s, sub=None):
    """Normwise error of one rate evaluation of the GPU path at the initial
    state against the FP64 oracle."""
    from oracle import oracle as O
    from pyticles_amd.device import from_system
    u0 = O.init_u(s.r, s.m, s.T, s.prm)
    ref = O.derivatives(s.r, s.v, s.m, u0, s.prm, terms=s.prm["terms"])
    with from_system(s) as b:
        b.upload(u=u0)
        b.rates(s.prm["terms"])
        got = b.download("rho", "rho_lr", "acc", "udot")
    out = {}
    for k in ("rho", "rho_lr", "acc", "udot"):
        out[k] = float(np.max(np.abs(got[k] - ref[k])) / np.max(np.abs(ref[k])))
    return out


def load_traffic(key):
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get(key)
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32c"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the FP32C comparison and error checks")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a plain device-to-device copy of 256 MiB (HBM counter calibration)")
    args = ap.parse_args()

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
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (boxes shard by whole box "
              "across GPUs, which no run here exercises)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    s, desc = workload(args.workload, args.precision)
    full = bool(s.prm["terms"] & 12)
    bps = B_PS[(args.precision, full)]
    print("[bench] %s, %s, %d steps after %d warmup" % (desc, args.precision, args.steps, args.warmup),
          file=sys.stderr, flush=True)
    b, el, tm, (E0, E1) = run_gpu(s, args.steps, args.warmup)
    nip = int(b.pairs().shape[0])
    b.close()
    value = s.n * args.steps / el
    evals = args.steps + 1
    per = {k: tm[k + "_ms"] / evals for k in ("build", "density", "gradients", "rates", "integrate")}
    rates_s = per["rates"] * 1e-3
    alg_bytes = bps["rates"] * s.n
    achieved = alg_bytes / rates_s / 1e9
    traffic = load_traffic("%s_%s" % (args.workload, args.precision))
    roof = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 5), traffic=traffic,
                kernel=RATES_KERNEL[args.precision], kernel_avg_ms=round(per["rates"], 5),
                alg_bytes_per_launch=alg_bytes)
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
                   "timed": "%d leapfrog steps (%d rate evaluations incl. the closing half kick)" % (args.steps, evals)},
        "roofline": roof,
        "pair_interactions_per_s": nip * args.steps / el,
        "nip": nip,
        "passes_ms_per_eval": {k: round(v, 5) for k, v in per.items()},
        "step_roofline_frac": round(value * step_bytes / (HBM_PEAK_GBS * 1e9), 5),
        "energy_drift_rel": abs(E1 - E0) / abs(E0),
    }
    if not args.no_extra and args.workload == "c2" and args.precision == "fp64":
        s32, _ = workload("c2", "fp32c")
        b32, el32, tm32, (e0, e1) = run_gpu(s32, args.steps, args.warmup, timing=False)
        b32.close()
        v32 = s32.n * args.steps / el32
        line["fp32c"] = {"value": v32, "ms_per_step": 1e3 * el32 / args.steps, "speedup_vs_fp64": v32 / value,
                         "energy_drift_rel": abs(e1 - e0) / abs(e0)}
        try:
            line["fp32c"]["rel_err_vs_fp64_oracle"] = error_vs_oracle(s32)
            line["fp64_rel_err_vs_oracle"] = error_vs_oracle(workload("c2", "fp64")[0])
        except Exception as exc:  # noqa: BLE001
            line["fp32c"]["rel_err_vs_fp64_oracle"] = "failed: %s" % exc
    if not args.no_cpu:
        s_cpu, _ = workload(args.workload, "fp64")
        line["cpu_baseline"] = cpu_baseline(s_cpu)
        line["cpu_baseline"]["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
