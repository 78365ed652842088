"""Spin-up experiment for the learned proposal (no counterpart in the reference).

    python -m turbid_amd.spinup --workload train128 --spin 10 --sample 20 \\
        --epochs 30 --eval 12 --out results.json

1. Plain start (laminar Stokes layer + seeded noise) run for `spin` wave
   periods, then `sample` more periods during which 12 snapshots per period
   (one per phase bin) are reduced to the network grid.  The statistics of
   those periods (turbulent kinetic energy, mean-mode wall shear, mean
   concentration) are the reference state.
2. The proposal network is trained on the snapshots (device BF16 MFMA).
3. Two fresh runs of `eval` periods each: the plain start, and the proposed
   state (from the same plain start, projected divergence-free).  For each,
   the GPU seconds until the period means of TKE, wall shear and mean
   concentration first stay within a tolerance of the reference for two
   consecutive periods.

Every time step is the DNS's own; the network only sets the initial state.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import runtime as RT
from . import surrogate as SG
from . import workloads as W

PHASE_BINS = 12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Run:
    """One DNS with the reference's step loop and per-bin statistics."""

    def __init__(self, cfg, uvw, p, c, project=False):
        self.cfg = cfg
        self.s = RT.Solver(cfg)
        self.s.upload(uvw, p, c)
        if project:
            self.s.project()
        self.st = RT.Stepper(cfg["dt"])
        self.steps_per_period = int(round(2 * math.pi / cfg["force_amplitude"][1] / cfg["dt"]))
        self.gpu_s = 0.0
        self.hist = []      # per bin: dict(t, tke, tau, cmean)

    def advance_bin(self, parts=1, on_part=None):
        """One phase bin (1/12 period) in `parts` equal pieces; on_part(uvw, c)
        is called with the state after each piece."""
        n = self.steps_per_period // PHASE_BINS
        for q in range(parts):
            t0 = time.perf_counter()
            feat = RT.run_steps(self.s, self.st, n // parts if q + 1 < parts else n - (parts - 1) * (n // parts))
            self.s.sync()
            self.gpu_s += time.perf_counter() - t0
            if on_part is not None or q + 1 == parts:
                uvw, p, c = self.s.state()
                if on_part is not None:
                    on_part(uvw, c)
        st = SG.flow_stats(uvw, c, feat)
        st.update(t=self.st.time, gpu_s=self.gpu_s, finite=bool(np.isfinite(st["tke"])))
        self.hist.append(st)
        return uvw, c, st

    def phase(self):
        return (self.st.time * self.cfg["force_amplitude"][1]) % (2 * math.pi)

    def close(self):
        self.s.close()


def period_means(hist):
    out = []
    for i in range(0, len(hist) - PHASE_BINS + 1, PHASE_BINS):
        seg = hist[i:i + PHASE_BINS]
        out.append(dict(tke=float(np.mean([h["tke"] for h in seg])),
                        tau=float(np.mean([h["tau"] for h in seg])),
                        cmean=float(np.mean([h["cmean"] for h in seg])),
                        gpu_s=seg[-1]["gpu_s"], period=i // PHASE_BINS + 1))
    return out


STATS = ("tke", "tau", "cmean")


def time_to_reference(hist, ref, tol, hold=2):
    """GPU seconds (and period) at the end of the first of `hold` consecutive
    periods whose period means of TKE, wall shear and mean concentration are
    all within tol (relative) of the reference; None if that never happens."""
    pm = period_means(hist)
    ok = [all(abs(p[k] - ref[k]) <= tol * abs(ref[k]) for k in STATS) for p in pm]
    for i in range(len(pm) - hold + 1):
        if all(ok[i:i + hold]):
            return pm[i]["gpu_s"], pm[i]["period"]
    return None, None


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="train128")
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--spin", type=int, default=10)
    ap.add_argument("--sample", type=int, default=20)
    ap.add_argument("--per-bin", type=int, default=1, help="snapshots per phase bin while sampling")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--eval", type=int, default=12)
    ap.add_argument("--channels", type=int, default=32)
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--model-in", default=None)
    ap.add_argument("--model-out", default=None)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)

    cfg = W.workload(args.workload, precision=args.precision)
    uvw0, p0, c0 = W.initial_state(cfg)
    result = dict(workload=args.workload, precision=args.precision, grid=[cfg["nx"], cfg["ny"], cfg["nz"]])

    # 1. developed reference and snapshots
    run = Run(cfg, uvw0, p0, c0)
    snaps, phases = [], []
    t0 = time.perf_counter()
    def keep(uvw, c):
        snaps.append(SG.reduce_state(uvw, c))
        phases.append(run.phase())

    for per in range(args.spin + args.sample):
        for b in range(PHASE_BINS):
            sampling = per >= args.spin
            uvw, c, st = run.advance_bin(args.per_bin if sampling else 1, keep if sampling else None)
            if not st["finite"]:
                log("spinup: state not finite at period %d" % per)
                sys.exit(3)
        pm = period_means(run.hist)[-1]
        log("spinup: reference period %d tke %.4g tau %.4g (%.1f s)" % (per + 1, pm["tke"], pm["tau"],
                                                                        time.perf_counter() - t0))
    ref_hist = run.hist[args.spin * PHASE_BINS:]
    ref = {k: float(np.mean([h[k] for h in ref_hist])) for k in STATS}
    result["reference"] = ref
    result["reference_periods"] = period_means(run.hist)
    result["snapshots"] = len(snaps)
    run.close()

    # 2. proposal network
    model = SG.Proposal(cfg, channels=args.channels, layers=args.layers, seed=args.seed)
    plain_red = SG.reduce_state(uvw0, c0)
    lam0 = SG.laminar_grid(cfg, 0.0)
    model.noise_rms = float(math.sqrt(np.mean((plain_red - lam0)[..., :3] ** 2)))
    if args.model_in:
        model.load(args.model_in)
        result["train_loss"] = None
    else:
        tt = time.perf_counter()
        hist = model.fit(np.asarray(snaps), np.asarray(phases), epochs=args.epochs, log=log)
        result["train_loss"] = hist
        result["train_s"] = time.perf_counter() - tt
    if args.model_out:
        model.save(args.model_out)

    # 3. plain start vs proposed start
    runs = {}
    uvw_p, c_p = model.propose(uvw0, c0, 0.0)
    for name, (u, cc, proj) in {"plain": (uvw0, c0, False), "proposal": (uvw_p, c_p, True)}.items():
        r = Run(cfg, u, p0, cc, project=proj)
        for per in range(args.eval):
            for b in range(PHASE_BINS):
                _, _, st = r.advance_bin()
                if not st["finite"]:
                    break
            pm = period_means(r.hist)[-1]
            log("spinup: %s period %d tke %.4g tau %.4g gpu %.1f s" % (name, per + 1, pm["tke"], pm["tau"],
                                                                          r.gpu_s))
        runs[name] = r.hist
        r.close()
    result["runs"] = {k: period_means(v) for k, v in runs.items()}
    result["time_to_reference"] = {}
    for tol in (0.05, 0.1, 0.2):
        entry = {}
        for k, v in runs.items():
            s, per = time_to_reference(v, ref, tol)
            entry[k] = dict(gpu_s=s, period=per)
        result["time_to_reference"][str(tol)] = entry
    line = json.dumps(result)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")
    print(json.dumps(result["time_to_reference"]), flush=True)


if __name__ == "__main__":
    main()
