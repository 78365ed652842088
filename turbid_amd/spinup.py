"""Spin-up experiment for the learned proposal (no counterpart in the reference).

    python -m turbid_amd.spinup build --out turbid_amd/spinup_model.npz
    python -m turbid_amd.spinup eval --model turbid_amd/spinup_model.npz

Workload `spinup128`: the oscillatory boundary layer at Re_delta = 1000 in a
2 x 2 x 2 box at 128 x 128 x 129 (the box and wave of the reference's own
Re_delta = 800 example, examples/oscillatory_flow_re800/result/
configuration.ini, at the Reynolds number where turbulence is sustained on
this grid), fine sediment with zero-flux walls.

build (the one-time cost of the proposal, all on the GPU):
  1. data generation: a plain start with strong noise (seed A) runs `spin`
     periods to a developed state, then `sample` periods during which
     `per_bin` snapshots per phase bin (12 bins) are reduced to the network
     grid; the period means of turbulent kinetic energy and mean-mode wall
     shear over the sampled periods are the reference statistics;
  2. training: the proposal network on those snapshots (BF16 MFMA kernels).
  The model file keeps the network, the reference statistics and the GPU
  seconds both steps took.

eval (what the network gains, on a held-out seed B):
  the plain start (the laminar Stokes layer plus weak seeded noise, the
  method without the network) and the proposal made from that same start
  (projected divergence-free, then the unchanged DNS), each run until its
  period means of TKE and wall shear stay within tolerance of the reference
  for `hold` consecutive periods, or `periods` periods.  Reported: GPU
  seconds to stationary statistics for both, and the number of spin-ups
  after which the one-time cost of the proposal is paid back.

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
STATS = ("tke", "tau")
# tolerance on the period means: about twice the period-to-period spread of
# the developed flow in this box (TKE within 2.4 %, wall shear within 0.7 %
# of their means over ten periods)
TOL = {"tke": 0.05, "tau": 0.015}


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


def in_band(pm, ref, tol=TOL):
    return all(abs(pm[k] - ref[k]) <= tol[k] * abs(ref[k]) for k in STATS)


def time_to_reference(hist, ref, tol=TOL, hold=2):
    """GPU seconds (and period) at the end of the first of `hold` consecutive
    periods whose period means of TKE and wall shear are within tol of the
    reference; (None, None) if that never happens."""
    pm = period_means(hist)
    ok = [in_band(p, ref, tol) for p in pm]
    for i in range(len(pm) - hold + 1):
        if all(ok[i:i + hold]):
            return pm[i]["gpu_s"], pm[i]["period"]
    return None, None


def build(cfg, spin=2, sample=10, per_bin=14, epochs=12, channels=32, layers=4, seed=1, net_seed=1,
          noise_ref=0.1, noise_eval=0.01, out=None):
    """Data generation and training; returns the Proposal and a record of
    the reference statistics and the GPU seconds spent."""
    rec = dict(grid=[cfg["nx"], cfg["ny"], cfg["nz"]], re=cfg["reynolds"], seed_data=seed,
               spin=spin, sample=sample, per_bin=per_bin, epochs=epochs)
    uvw0, p0, c0 = W.initial_state(cfg, seed=seed, noise_u=noise_ref)
    run = Run(cfg, uvw0, p0, c0)
    snaps, phases = [], []
    t_wall = time.perf_counter()

    def keep(uvw, c):
        snaps.append(SG.reduce_state(uvw, c))
        phases.append(run.phase())

    for per in range(spin + sample):
        sampling = per >= spin
        for b in range(PHASE_BINS):
            _, _, st = run.advance_bin(per_bin if sampling else 1, keep if sampling else None)
            if not st["finite"]:
                raise RuntimeError("spinup: data run not finite at period %d" % (per + 1))
        pm = period_means(run.hist)[-1]
        log("spinup: data period %d tke %.4g tau %.4g (%d snapshots, %.1f GPU-s)"
            % (per + 1, pm["tke"], pm["tau"], len(snaps), run.gpu_s))
    ref_hist = run.hist[spin * PHASE_BINS:]
    rec["reference"] = {k: float(np.mean([h[k] for h in ref_hist])) for k in STATS + ("cmean",)}
    # TKE of the developed flow by wave phase: entry i of the history is the
    # state after i + 1 bins, at phase 2 pi (i + 1)/12
    def by_phase(key):
        return [float(np.mean([h[key] for i, h in enumerate(run.hist)
                               if i >= spin * PHASE_BINS and (i + 1) % PHASE_BINS == k]))
                for k in range(PHASE_BINS)]

    rec["tke_phase"] = by_phase("tke")
    rec["tke_lo_phase"] = by_phase("tke_lo")
    rec["reference_periods"] = period_means(run.hist)
    rec["snapshots"] = len(snaps)
    rec["data_gpu_s"] = run.gpu_s
    rec["data_wall_s"] = time.perf_counter() - t_wall
    run.close()

    model = SG.Proposal(cfg, channels=channels, layers=layers, seed=net_seed)
    model.tke_phase = rec["tke_phase"]
    model.tke_lo_phase = rec["tke_lo_phase"]
    # the input the network sees at inference: the plain start's noise on its
    # grid (training inputs are snapshot deviations scaled to this size)
    u_pl, _, c_pl = W.initial_state(cfg, seed=seed + 1, noise_u=noise_eval)
    model.noise_rms = float(math.sqrt(np.mean((SG.reduce_state(u_pl, c_pl) - SG.laminar_grid(cfg, 0.0))[..., :3] ** 2)))
    rec["noise_eval"] = noise_eval
    t0 = time.perf_counter()
    rec["train_loss"] = model.fit(np.asarray(snaps), np.asarray(phases), epochs=epochs, log=log)
    rec["train_s"] = time.perf_counter() - t0
    if out:
        model.save(out, record=rec)
    return model, rec


def evaluate(cfg, model, ref, periods=8, seed=7, noise=0.01, hold=2, log_fn=log):
    """Plain start vs proposal on a held-out seed.  Each run stops once its
    statistics are in band for `hold` consecutive periods, or after
    `periods` periods."""
    uvw0, p0, c0 = W.initial_state(cfg, seed=seed, noise_u=noise)
    t0 = time.perf_counter()
    uvw_p, c_p = model.propose(uvw0, c0, 0.0)
    propose_s = time.perf_counter() - t0     # network inference and the grid transfers
    out = {}
    for name, (u, cc, proj) in (("plain", (uvw0, c0, False)), ("proposal", (uvw_p, c_p, True))):
        r = Run(cfg, u, p0, cc, project=proj)
        if proj:
            r.gpu_s = propose_s     # the proposal's cost is part of its spin-up
        reached = (None, None)
        for per in range(periods):
            for b in range(PHASE_BINS):
                _, _, st = r.advance_bin()
                if not st["finite"]:
                    raise RuntimeError("spinup: %s run not finite" % name)
            pm = period_means(r.hist)[-1]
            log_fn("spinup: %s period %d tke %.4g tau %.4g gpu %.1f s" % (name, per + 1, pm["tke"], pm["tau"],
                                                                         r.gpu_s))
            reached = time_to_reference(r.hist, ref, hold=hold)
            if reached[0] is not None:
                break
        out[name] = dict(gpu_s=reached[0], period=reached[1], periods_run=len(period_means(r.hist)),
                         gpu_s_run=r.gpu_s, periods=period_means(r.hist))
        r.close()
    return out


def summary(res, rec):
    """GPU seconds to stationary statistics, plain vs proposal; one-time cost
    of the proposal (data generation + training) and the break-even count."""
    p, q = res["plain"], res["proposal"]
    once = rec["data_gpu_s"] + rec["train_s"]
    s = dict(plain=p["gpu_s"], surrogate=q["gpu_s"],
             plain_reached=p["gpu_s"] is not None, surrogate_reached=q["gpu_s"] is not None,
             plain_lower_bound=p["gpu_s"] if p["gpu_s"] is not None else p["gpu_s_run"],
             one_time_s=once, data_gpu_s=rec["data_gpu_s"], train_s=rec["train_s"],
             tolerance=TOL)
    if q["gpu_s"] is not None:
        saved = s["plain_lower_bound"] - q["gpu_s"]
        s["ratio"] = q["gpu_s"] / s["plain_lower_bound"]
        s["break_even_runs"] = once / saved if saved > 0 else None
    return s


def load_model(cfg, path):
    d = np.load(path, allow_pickle=False)
    meta = json.loads(str(d["meta"]))
    model = SG.Proposal(cfg, channels=meta["channels"], layers=meta["layers"])
    model.load(path)
    return model, meta["record"]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["build", "eval"])
    ap.add_argument("--workload", default="spinup128")
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--spin", type=int, default=2)
    ap.add_argument("--sample", type=int, default=10)
    ap.add_argument("--per-bin", type=int, default=17, help="snapshots per phase bin while sampling")
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--periods", type=int, default=6, help="cap on the periods of each evaluation run")
    ap.add_argument("--noise", type=float, nargs="+", default=[0.1, 0.01],
                    help="noise amplitudes of the plain starts evaluated")
    ap.add_argument("--model", default=None)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    cfg = W.workload(args.workload, precision=args.precision)
    if args.mode == "build":
        model, rec = build(cfg, spin=args.spin, sample=args.sample, per_bin=args.per_bin, epochs=args.epochs,
                           noise_eval=args.noise[0], out=args.out)
    else:
        model, rec = load_model(cfg, args.model)
    result = dict(record={k: v for k, v in rec.items() if k != "train_loss"}, noise={})
    for nz in args.noise:
        res = evaluate(cfg, model, rec["reference"], periods=args.periods, noise=nz)
        result["noise"][str(nz)] = dict(eval=res, summary=summary(res, rec))
        log("spinup: noise %g: %s" % (nz, json.dumps(result["noise"][str(nz)]["summary"])))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
