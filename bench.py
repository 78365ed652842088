"""RK3 steps/s of the pseudospectral carrier + sediment substage on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision mixed|fp64|both]
               [--workload turb256] [--no-cpu-baseline] [--no-autotune]

Workload (default): 256x256x257 oscillatory bottom boundary layer with fine
sediment (Re_delta = 800, settling velocity 5e-4, Sc = 0.5), synthetic
seeded initial state (Stokes layer + windowed Gaussian noise).  A step is one
full RK3 step (three substages of the carrier and the scalar phase) with the
inputs resident in HBM; the loop mirrors FlatChannelSolver::Execute
(TimeStepper::Update with the previous step's CFL feature, three substages,
alternation toggle).  The headline value is the mixed-precision rate; the
plain FP64 rate is measured in the same run.

Kernel variant: before the warmup the library's autotuner (tbd_autotune)
times the step's kernel variants on a scratch copy of the context and keeps
the fastest (reported under `autotune`; --no-autotune keeps the default).

Timing: after the warmup, K steps are timed as the library runs them (one
captured HIP graph per step; `value`, `ms_per_step`; K = 1000 by default,
the workload's step count), then min(K, --profile-steps) further steps run
with HIP events around every launch on the library's stream (eager
substages), which gives each kernel's launch count, average duration and
algorithmic work.  Roofline: the kernel with the largest total time;
`achieved` = its algorithmic work per launch (bytes read plus written by a
transform pass or product kernel, or 4*Nz^2*P flops per contraction of one
component) over its average launch duration.  step_roofline: the whole
step's time at the roofline (minimal algorithmic transform bytes and
contraction flops per step) over the measured time.  spinup_gpu_s: the
metric's second half, GPU seconds to stationary TKE and wall shear from the
plain noisy-laminar start and from the learned proposal on a held-out seed
(turbid_amd.spinup).  ensemble: the 16-member ensemble of uneven grids
(64..128 squared by 65..129, mixed precision) in member-steps/s, advanced
through grouped launches (one launch per kernel instance over all members,
one graph), on the members' own streams, and one member at a time
(turbid_amd.ensemble).  cpu_baseline: the FP64
CPU restatement (oracle/) timed on the host's cores for one warm-up step and
then --cpu-steps steps of the same workload.
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
from turbid_amd.roofline import HBM_PEAK, MFMA_PEAK, step_roofline  # noqa: E402


# profiler class -> (kernel symbol, kind, peak)
KERNELS = {
    "ypass_bwd64": ("k_ypb2<double,...>", "bytes", HBM_PEAK),
    "xpass_bwd64": ("k_xpb2<double,...>", "bytes", HBM_PEAK),
    "xpass_fwd64": ("k_xpf2<double,...>", "bytes", HBM_PEAK),
    "ypass_fwd64": ("k_ypf2<double,...>", "bytes", HBM_PEAK),
    "ypass_bwd32": ("k_ypb2<float,...>", "bytes", HBM_PEAK),
    "xpass_bwd32": ("k_xpb2<float,...>", "bytes", HBM_PEAK),
    "xpass_fwd32": ("k_xpf2<float,...>", "bytes", HBM_PEAK),
    "ypass_fwd32": ("k_ypf2<float,...>", "bytes", HBM_PEAK),
    "xfused64": ("k_xfused2<double,...>", "bytes", HBM_PEAK),
    "xfused32": ("k_xfused2<float,...>", "bytes", HBM_PEAK),
    "zgemm64": ("k_zgemm<double,...>", "flops", MFMA_PEAK["fp64"]),
    "zgemm32": ("k_zgemm<float,...>", "flops", MFMA_PEAK["fp32"]),
    "helmholtz": ("k_helmholtz<1, false>", "flops", MFMA_PEAK["fp64"]),
    "helmholtz_split": ("k_helmholtz_split<1, false>", "flops", MFMA_PEAK["fp64"]),
    "rhs_helmholtz": ("k_rhs_helmholtz<1, false, false>", "flops", MFMA_PEAK["fp64"]),
    "rhs_helmholtz_split": ("k_rhs_helmholtz<1, true, false>", "flops", MFMA_PEAK["fp64"]),
    "pressure_chain": ("k_pressure_chain<1, false>", "flops", MFMA_PEAK["fp64"]),
    "products64": ("k_products_*<double>", "bytes", HBM_PEAK),
    "products32": ("k_products_*<float>", "bytes", HBM_PEAK),
    "other": ("(boundary rows, H update, filter, diagnostics, copies)", None, None),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run(cfg, state, steps, warmup, autotune, variant=None, psteps=None):
    """Pick the kernel variant (the library's autotuner, on a scratch copy of
    the context, outside the timed region), warm up, time `steps` steps as the
    library runs them (one captured HIP graph per step), then time `psteps`
    further steps (default: as many) with HIP events around every launch
    (eager substages) for the per-kernel figures."""
    psteps = steps if psteps is None else psteps
    s = RT.Solver(cfg)
    s.upload(*state)
    tune = None
    if variant is not None:
        s.set_variant(variant)
        tune = {"kept": {k: bool(s.variant() & b) for k, b in RT.Solver.VARIANT_BITS.items()}, "given": True}
    elif autotune:
        try:
            bits, ms = s.autotune(5)
        except RT.TurbidError as e:   # e.g. no device memory for the scratch context
            log("bench: autotune failed (%s); default variant" % e)
        else:
            tune = {"kept": {k: bool(bits & b) for k, b in RT.Solver.VARIANT_BITS.items()},
                    "ms_per_step": {str(i): round(float(v), 3) for i, v in enumerate(ms) if not np.isnan(v)}}
    st = RT.Stepper(cfg["dt"])
    RT.run_steps(s, st, warmup)
    s.sync()
    t0 = time.perf_counter()
    feat = RT.run_steps(s, st, steps)
    s.sync()
    t1 = time.perf_counter()
    s.profile(True)
    s.profile_reset()
    tp0 = time.perf_counter()
    RT.run_steps(s, st, psteps)
    s.sync()
    tp1 = time.perf_counter()
    prof = {name: s.profile_read(name) for name in KERNELS}
    s.profile(False)
    s.close()
    return (t1 - t0), (tp1 - tp0), feat, prof, tune


TRAFFIC = os.path.join(ROOT, "profiles", "turb256_hbm_traffic.json")


def measured_traffic(workload, cls):
    """HBM bytes per dispatch of the roofline kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a plain
    copy), when that summary is for the same workload and kernel."""
    try:
        with open(TRAFFIC) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload or t.get("class") != cls:
        return None
    return float(t["bytes_per_dispatch"])


def roofline(prof):
    best = max((n for n in KERNELS if KERNELS[n][1] and prof[n][0] > 0), key=lambda n: prof[n][1])
    n, ms, work = prof[best]
    sym, kind, peak = KERNELS[best]
    avg_s = ms / n / 1e3
    per_launch = work / n
    achieved = per_launch / avg_s
    if kind == "bytes":
        out = {"bound": "hbm", "achieved": achieved / 1e9, "peak": peak / 1e9, "unit": "GB/s"}
    else:
        out = {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s"}
    out["frac"] = out["achieved"] / out["peak"]
    out["traffic"] = None
    out["kernel"] = sym
    out["class"] = best
    out["launches"] = n
    out["avg_launch_ms"] = ms / n
    out["work_per_launch"] = per_launch
    return out


def class_rates(prof):
    """Achieved rate and fraction of peak of every kernel class with a
    roofline (algorithmic bytes or flops per launch over its average time)."""
    out = {}
    for name, (sym, kind, peak) in KERNELS.items():
        n, ms, work = prof[name]
        if not kind or n == 0 or ms <= 0:
            continue
        rate = work / (ms / 1e3)
        unit, scale = ("GB/s", 1e9) if kind == "bytes" else ("TFLOP/s", 1e12)
        out[name] = {"achieved": round(rate / scale, 2), "unit": unit, "frac": round(rate / peak, 3)}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, state, nthreads, ksteps):
    """The FP64 CPU restatement on the host: one warm-up step, then ksteps
    timed steps of the same workload (wall clock)."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    o = O.Oracle(cfg, nthreads=nthreads)
    o.set_state(*state)
    setup = time.perf_counter() - t0
    o.step()
    t0 = time.perf_counter()
    for _ in range(ksteps):
        o.step()
    dt = time.perf_counter() - t0
    o.close()
    return {"value": ksteps / dt, "unit": "RK3 steps/s", "cores": nthreads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": "%d RK3 steps (3 substages, carrier + scalar) of the same workload after 1 warm-up step "
                      "and %.1f s setup; baseline only" % (ksteps, setup)}


SPINUP_MODEL = os.path.join(ROOT, "turbid_amd", "spinup_model.npz")


def spinup_leg(path, periods, noises):
    """Second half of the metric: GPU seconds to stationary statistics (period
    means of TKE and wall shear in band for two consecutive wave periods)
    from the plain start and from the learned proposal (turbid_amd.spinup),
    on the spin-up workload and a held-out seed, with the model trained
    earlier on this device (its one-time data-generation and training
    seconds come with it).  One entry per noise amplitude of the plain
    start; the first is the headline (plain, surrogate)."""
    from turbid_amd import spinup as SP
    cfg = W.workload("spinup128", precision=1)
    model, rec = SP.load_model(cfg, path)
    starts = {}
    for noise in noises:
        res = SP.evaluate(cfg, model, rec["reference"], periods=periods, noise=noise,
                          log_fn=lambda m: log("bench: " + m))
        s = SP.summary(res, rec)
        s["periods_run"] = {k: v["periods_run"] for k, v in res.items()}
        starts[str(noise)] = s
    head = starts[str(noises[0])]
    return dict(plain=head["plain"], surrogate=head["surrogate"], plain_lower_bound=head["plain_lower_bound"],
                ratio=head.get("ratio"), one_time_s=head["one_time_s"], break_even_runs=head.get("break_even_runs"),
                starts=starts, periods_cap=periods, held_out_seed=7, data_seed=rec["seed_data"],
                snapshots=rec["snapshots"], tolerance=SP.TOL,
                workload="spinup128: Re_delta=1000, box 2x2x2, 128x128x129, dt=T/12000, mixed precision")


def copy_calibration():
    """One plain device-to-device copy of 2^30 bytes after a warm-up copy: the
    FETCH_SIZE / WRITE_SIZE counters of its dispatch read 2^30 bytes each."""
    import torch
    n = 1 << 30
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.copy_(a)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({"copy_bytes": n, "seconds": dt, "GB/s": 2 * n / dt / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed steps per precision (the workload's 1000)")
    ap.add_argument("--profile-steps", type=int, default=100,
                    help="steps with per-launch HIP events after the timed ones (at most --steps)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="both", choices=["mixed", "fp64", "both"])
    ap.add_argument("--workload", default="turb256")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="run the default kernel variant")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=5, help="timed CPU-baseline steps after one warm-up step")
    ap.add_argument("--variant", default=None,
                    help="run this kernel variant (tbd_set_variant bits) instead of autotuning: one value for "
                         "every precision, or FP64,MIXED")
    ap.add_argument("--spinup-model", default=SPINUP_MODEL, help="trained proposal (python -m turbid_amd.spinup build)")
    ap.add_argument("--spinup-periods", type=int, default=8, help="cap on each spin-up run; 0 skips the leg")
    ap.add_argument("--spinup-noise", type=float, nargs="+", default=[0.01],
                    help="noise amplitudes of the plain starts (the first is the headline): 0.01, the weakly "
                         "perturbed laminar layer; 0.1, the start of the 256^3 workload")
    ap.add_argument("--ensemble-steps", type=int, default=20, help="timed steps of the ensemble leg; 0 skips it")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 1 GiB (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path "
              "(the single-DNS shards and the ensemble members are independent per GPU); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    cfg64 = W.workload(args.workload, precision=0)
    log("bench: workload %s %dx%dx%d, building initial state" % (args.workload, cfg64["nx"], cfg64["ny"], cfg64["nz"]))
    uvw, p, c = W.initial_state(cfg64)
    state = (uvw, p, c)
    variants = {}
    if args.variant is not None:
        v = [int(x) for x in str(args.variant).split(",")]
        variants = {"fp64": v[0], "mixed": v[-1]}
    results = {}
    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    for prec in precs:
        cfg = dict(cfg64)
        cfg["precision"] = 1 if prec == "mixed" else 0
        psteps = min(args.steps, args.profile_steps)
        log("bench: %s warmup %d, timing %d steps (graph), then %d steps with per-launch events"
            % (prec, args.warmup, args.steps, psteps))
        t, tp, feat, prof, tune = run(cfg, state, args.steps, args.warmup, not args.no_autotune,
                                      variants.get(prec), psteps)
        if tune:
            log("bench: %s variant kept %s" % (prec, tune["kept"]))
        results[prec] = {"tune": tune, "seconds": t, "steps_per_s": args.steps / t, "ms_per_step": 1e3 * t / args.steps,
                         "profiled_ms_per_step": 1e3 * tp / psteps, "psteps": psteps,
                         "roofline": roofline(prof), "prof": {k: list(v) for k, v in prof.items()},
                         "cfl": float(feat[19] * cfg["dt"]), "max_u": float(feat[0])}
        log("bench: %s %.2f steps/s (%.3f ms/step; %.3f ms/step with events)"
            % (prec, args.steps / t, 1e3 * t / args.steps, 1e3 * tp / psteps))
        if not np.isfinite(feat[0]):
            log("bench: %s state is not finite" % prec)
            sys.exit(3)
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
    line["roofline"]["traffic"] = measured_traffic(args.workload, line["roofline"]["class"])
    if "fp64" in results:
        line["fp64_steps_per_s"] = results["fp64"]["steps_per_s"]
        line["fp64_roofline"] = results["fp64"]["roofline"]
    if "mixed" in results and "fp64" in results:
        line["mixed_vs_fp64"] = results["mixed"]["steps_per_s"] / results["fp64"]["steps_per_s"]
    line["kernel_ms_per_step"] = {prec: {k: round(v[1] / r["psteps"], 4) for k, v in r["prof"].items() if v[0]}
                                  for prec, r in results.items()}
    line["profiled_ms_per_step"] = {prec: r["profiled_ms_per_step"] for prec, r in results.items()}
    line["class_rates"] = {prec: class_rates(r["prof"]) for prec, r in results.items()}
    # kernel variant the autotuner kept (ms per step of every variant it tried,
    # keyed by variant bits: 1 fused x transforms, 2 fused rhs + solve, 4 overlap,
    # 8 wide transform kernels)
    line["autotune"] = {prec: r["tune"] for prec, r in results.items()}
    line["step_roofline"] = {prec: step_roofline(cfg64, prec == "mixed", r["ms_per_step"])
                             for prec, r in results.items()}
    if args.spinup_periods > 0 and os.path.exists(args.spinup_model):
        log("bench: spin-up, plain start vs proposal (noise %s, at most %d periods each)"
            % (args.spinup_noise, args.spinup_periods))
        line["spinup_gpu_s"] = spinup_leg(args.spinup_model, args.spinup_periods, args.spinup_noise)
    else:
        line["spinup_gpu_s"] = None
    if args.ensemble_steps > 0:
        from turbid_amd import ensemble as E
        log("bench: ensemble of 16 members, %d steps per leg" % args.ensemble_steps)
        # the grouped leg tunes its members' variant bits together first, as
        # the single-grid legs run the variant the autotuner keeps
        line["ensemble"] = E.measure(E.member_configs(16, seed=0, precision=1), args.ensemble_steps, 3,
                                     tune=True, log=lambda m: log("bench: " + m))
    else:
        line["ensemble"] = None
    if rank == 0 and not args.no_cpu_baseline:
        # all the host cores this process may use: the OpenMP thread count the
        # host sets for it, else every CPU
        nthreads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1)
        log("bench: cpu baseline (oracle, %d threads, %d steps after a warm-up step)" % (nthreads, args.cpu_steps))
        line["cpu_baseline"] = cpu_baseline(cfg64, state, nthreads, args.cpu_steps)
    else:
        line["cpu_baseline"] = None
    if rank == 0:
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
