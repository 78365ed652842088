"""Phase-averaged statistics of FP64 and mixed-precision runs of the same
start, accumulated on the device (tbd_stats_*), and their difference.

    python -m turbid_amd.phasestats --workload turb256 --periods 10 --bins 12
    python -m turbid_amd.phasestats --workload spinup128 --skip 2 --periods 10

For each precision the run starts from the same seeded state, advances
`periods` wave periods with the reference's step loop, and at the end of
each of the `bins` phase bins adds the state's profiles to that bin: mean
u, v, w, c and the plane means of the squared fluctuations, plus the
mean-mode wall shear (features 10, 11) on the host.  Reported per
quantity: the largest difference between the two precisions over bins and
heights, relative to the FP64 scale of its kind: mean velocities to the
largest |U|, fluctuation variances to the largest u'^2 + v'^2 + w'^2, the
concentration quantities and the wall shear to their own largest magnitude.

In a turbulent flow two trajectories that start a round-off apart separate
as fast as any two, so a finite average carries sampling scatter whatever
the precision.  `--control EPS` adds a third run, FP64 from the same start
with a seeded perturbation of EPS times the start's noise field added to
u, v, w and c, and reports its difference from the FP64 run beside the
mixed one: the control's difference is the scatter an FP64 run itself
shows at that averaging length.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import runtime as RT
from . import workloads as W

NAMES = ["U", "V", "W", "C", "uu", "vv", "ww", "cc"]


def start(cfg, eps=0.0):
    """The seeded start; with eps > 0, plus eps times a second seeded noise
    field of the start's amplitudes (an independent draw, same windows)."""
    uvw, p, c = W.initial_state(cfg)
    if eps > 0.0:
        seed = cfg.get("seed", 0) + 7919
        uvw2, _, c2 = W.initial_state(cfg, seed=seed)
        uvw0, _, c0 = W.initial_state(cfg, noise_u=0.0, noise_c=0.0)
        uvw = uvw + eps * (uvw2 - uvw0)
        c = c + eps * (c2 - c0)
    return uvw, p, c


def run(cfg, periods, bins, log, skip=0, eps=0.0):
    s = RT.Solver(cfg)
    s.upload(*start(cfg, eps))
    s.stats_reset()
    st = RT.Stepper(cfg["dt"])
    spp = int(round(2 * math.pi / cfg["force_amplitude"][1] / cfg["dt"]))
    per_bin = spp // bins
    tau = np.zeros((bins, 2))
    t0 = time.perf_counter()
    for p in range(skip):     # spin-up periods, not averaged
        RT.run_steps(s, st, per_bin * bins)
        log("phasestats: precision %d%s spin-up period %d done (%.1f s)"
            % (cfg["precision"], " (control)" if eps else "", p + 1, time.perf_counter() - t0))
    for p in range(periods):
        for b in range(bins):
            feat = RT.run_steps(s, st, per_bin)
            s.stats_accumulate(b)
            tau[b] += feat[10:12]
        log("phasestats: precision %d%s period %d done (%.1f s)"
            % (cfg["precision"], " (control)" if eps else "", p + 1, time.perf_counter() - t0))
    prof = np.stack([s.stats_read(b)[0] for b in range(bins)])      # (bins, 8, nz)
    s.close()
    return prof, tau / periods, time.perf_counter() - t0


def differences(p64, t64, p, t):
    """Largest difference of (p, t) from the FP64 (p64, t64) per quantity,
    relative to the FP64 scale of its kind."""
    diff = {}
    uscale = np.max(np.abs(p64[:, 0, :]))
    kscale = np.max(np.abs(p64[:, 4, :] + p64[:, 5, :] + p64[:, 6, :]))
    scales = [uscale, uscale, uscale, np.max(np.abs(p64[:, 3, :])), kscale, kscale, kscale,
              np.max(np.abs(p64[:, 7, :]))]
    for i, n in enumerate(NAMES):
        diff[n] = float(np.max(np.abs(p[:, i, :] - p64[:, i, :])) / scales[i]) if scales[i] > 0 else 0.0
    scale = np.max(np.abs(t64))
    diff["wall_shear"] = float(np.max(np.abs(t - t64)) / scale)
    return diff


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="turb256")
    ap.add_argument("--periods", type=int, default=10)
    ap.add_argument("--bins", type=int, default=12)
    ap.add_argument("--skip", type=int, default=0, help="periods run before the averaging starts")
    ap.add_argument("--control", type=float, default=0.0,
                    help="also run FP64 from the start perturbed by this multiple of its noise field")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    p64, t64, s64 = run(W.workload(args.workload, precision=0), args.periods, args.bins, log, args.skip)
    p32, t32, s32 = run(W.workload(args.workload, precision=1), args.periods, args.bins, log, args.skip)
    diff = differences(p64, t64, p32, t32)
    out = dict(workload=args.workload, periods=args.periods, skip=args.skip, bins=args.bins, seconds_fp64=s64,
               seconds_mixed=s32,
               max_relative_difference=diff, max_over_quantities=max(diff.values()))
    arrays = dict(fp64=p64, mixed=p32, tau_fp64=t64, tau_mixed=t32)
    if args.control > 0.0:
        pc, tc, sc = run(W.workload(args.workload, precision=0), args.periods, args.bins, log, args.skip,
                         eps=args.control)
        cdiff = differences(p64, t64, pc, tc)
        out.update(control_eps=args.control, seconds_control=sc, control_relative_difference=cdiff,
                   control_max_over_quantities=max(cdiff.values()))
        arrays.update(control=pc, tau_control=tc)
    line = json.dumps(out)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")
        np.savez(args.out.rsplit(".", 1)[0] + ".npz", **arrays)
    print(line, flush=True)


if __name__ == "__main__":
    main()
