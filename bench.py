"""RK3 steps/s of the pseudospectral carrier + sediment substage on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision mixed|fp64|both]
               [--workload turb256] [--no-cpu-baseline]

Workload (default): 256x256x257 oscillatory bottom boundary layer with fine
sediment (Re_delta = 800, settling velocity 5e-4, Sc = 0.5), synthetic
seeded initial state (Stokes layer + windowed Gaussian noise).  A step is one
full RK3 step (three substages of the carrier and the scalar phase) with the
inputs resident in HBM; the loop mirrors FlatChannelSolver::Execute
(TimeStepper::Update with the previous step's CFL feature, three substages,
alternation toggle).  The headline value is the mixed-precision rate; the
plain FP64 rate is measured in the same run.

Roofline: per kernel class, HIP events around every launch on the library's
stream during the timed region give the average launch duration; the
dominant class's algorithmic work per launch (field-transforms x
(w*R + 2w*S) bytes, or contractions x 4*Nz^2*P flops) over that duration is
`achieved`.  cpu_baseline: the FP64 CPU restatement (oracle/) timed on the
host for one step of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from turbid_amd import runtime as RT   # noqa: E402
from turbid_amd import workloads as W  # noqa: E402

HBM_PEAK = 8.0e12          # B/s
MFMA_PEAK = {"fp64": 78.6e12, "fp32": 157.3e12}

CLASSES = {  # name: (library class, kind, precision)
    "fft64": (0, "bytes", 8),
    "zgemm64": (1, "flops", "fp64"),
    "helmholtz": (2, "flops", "fp64"),
    "zgemm32": (4, "flops", "fp32"),
    "fft32": (5, "bytes", 4),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run(cfg, state, steps, warmup, profile=True):
    s = RT.Solver(cfg)
    s.upload(*state)
    st = RT.Stepper(cfg["dt"])
    RT.run_steps(s, st, warmup)
    s.sync()
    if profile:
        s.profile(True)
        s.profile_reset()
    t0 = time.perf_counter()
    feat = RT.run_steps(s, st, steps)
    s.sync()
    t1 = time.perf_counter()
    prof = {}
    if profile:
        for name, (k, _, _) in CLASSES.items():
            prof[name] = s.profile_read(k)
        prof["other"] = s.profile_read(3)
    s.close()
    return (t1 - t0), feat, prof


def roofline(prof, cfg):
    d = W.dims(cfg["nx"], cfg["ny"], cfg["nz"])
    nz = cfg["nz"]
    P = d[1] * d[2]
    S = nz * P
    R = cfg["nx"] * cfg["ny"] * nz
    best = None
    for name, (_, kind, prec) in CLASSES.items():
        n, ms, units = prof[name]
        if n == 0 or ms <= 0:
            continue
        if best is None or ms > prof[best][1]:
            best = name
    n, ms, units = prof[best]
    _, kind, prec = CLASSES[best]
    avg_s = ms / n / 1e3
    if kind == "bytes":
        w = prec
        per_unit = w * R + 2 * w * S
        work = units / n * per_unit
        achieved = work / avg_s
        out = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s"}
    else:
        per_unit = 4.0 * nz * nz * P
        work = units / n * per_unit
        achieved = work / avg_s
        out = {"bound": "mfma", "achieved": achieved / 1e12, "peak": MFMA_PEAK[prec] / 1e12, "unit": "TFLOP/s"}
    out["frac"] = out["achieved"] / out["peak"]
    out["traffic"] = None
    out["kernel"] = best
    out["launches"] = n
    out["avg_launch_ms"] = ms / n
    out["work_per_launch"] = work
    return out


def cpu_baseline(cfg, state, nthreads):
    from oracle import oracle as O
    t0 = time.perf_counter()
    o = O.Oracle(cfg, nthreads=nthreads)
    o.set_state(*state)
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.step()
    dt = time.perf_counter() - t0
    o.close()
    return {"value": 1.0 / dt, "unit": "RK3 steps/s", "cores": nthreads, "kind": "port",
            "sample": "1 RK3 step (3 substages, carrier + scalar) of the same workload after %.1f s setup" % setup}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="both", choices=["mixed", "fp64", "both"])
    ap.add_argument("--workload", default="turb256")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the single-DNS shards and the ensemble members are independent per GPU); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    cfg64 = W.workload(args.workload, precision=0)
    log("bench: workload %s %dx%dx%d, building initial state" % (args.workload, cfg64["nx"], cfg64["ny"], cfg64["nz"]))
    uvw, p, c = W.initial_state(cfg64)
    state = (uvw, p, c)
    results = {}
    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    for prec in precs:
        cfg = dict(cfg64)
        cfg["precision"] = 1 if prec == "mixed" else 0
        log("bench: %s warmup %d, timing %d steps" % (prec, args.warmup, args.steps))
        t, feat, prof = run(cfg, state, args.steps, args.warmup)
        results[prec] = {"seconds": t, "steps_per_s": args.steps / t, "ms_per_step": 1e3 * t / args.steps,
                         "roofline": roofline(prof, cfg), "prof": {k: list(v) for k, v in prof.items()},
                         "cfl": float(feat[19] * cfg["dt"]), "max_u": float(feat[0])}
        log("bench: %s %.2f steps/s (%.3f ms/step)" % (prec, args.steps / t, 1e3 * t / args.steps))
    head = "mixed" if "mixed" in results else "fp64"
    line = {
        "metric": "RK3 steps/s @256x256x257 mixed vs FP64; spin-up GPU-s surrogate vs plain",
        "value": results[head]["steps_per_s"],
        "unit": "RK3 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": results[head]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic (seeded Stokes layer + Gaussian noise)",
        "config": {"workload": "oscillatory BBL + fine sediment, Re_delta=800, %dx%dx%d, dt=T/2000, precision=%s"
                   % (cfg64["nx"], cfg64["ny"], cfg64["nz"], head),
                   "grid": [cfg64["nx"], cfg64["ny"], cfg64["nz"]], "phases": 2},
        "roofline": {k: v for k, v in results[head]["roofline"].items()},
    }
    if "fp64" in results:
        line["fp64_steps_per_s"] = results["fp64"]["steps_per_s"]
        line["fp64_roofline"] = results["fp64"]["roofline"]
    if "mixed" in results and "fp64" in results:
        line["mixed_vs_fp64"] = results["mixed"]["steps_per_s"] / results["fp64"]["steps_per_s"]
    line["kernel_ms_per_step"] = {prec: {k: v[1] / args.steps for k, v in r["prof"].items()}
                                  for prec, r in results.items()}
    if rank == 0 and not args.no_cpu_baseline:
        nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log("bench: cpu baseline (oracle, %d threads)" % nthreads)
        line["cpu_baseline"] = cpu_baseline(cfg64, state, nthreads)
    else:
        line["cpu_baseline"] = None
    if rank == 0:
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
