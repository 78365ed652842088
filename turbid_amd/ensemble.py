"""Ensemble of independent realisations on one GPU.

    python -m turbid_amd.ensemble --members 16 --steps 20 --warmup 3

Members draw, from a seed, Nx = Ny in {64, 96, 128}, Nz in {65, 97, 129}, a
Stokes Reynolds number uniform in [400, 1000] and a settling velocity
log-uniform in [1e-4, 2e-3]; the rest is the 256^3 turbulent configuration
(same domain, forcing, boundary conditions, noise).  The member loop is the
reference's step loop (TimeStepper::Update with the member's previous CFL,
three substages, alternation toggle).

Three ways to advance them, timed side by side:
  grouped     one launch per kernel and launch shape at each position of the
              step over all members (the workgroups of every member in one
              grid, each with its own arguments), the whole ensemble step
              captured in one HIP graph (tbd_ensemble_step);
  streams     each member its own context, stream and captured step graph,
              every member's step issued before any is waited for, so members
              overlap through the hardware queues;
  sequential  one member at a time, each synchronised every step.
"""
import argparse
import copy
import json
import math
import sys
import time

import numpy as np

from . import runtime as RT
from . import workloads as W
from .roofline import step_roofline


def member_configs(n=16, seed=0, precision=1):
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n):
        nxy = int(rng.choice([64, 96, 128]))
        nz = int(rng.choice([65, 97, 129]))
        re_d = float(rng.uniform(400.0, 1000.0))
        ws = float(math.exp(rng.uniform(math.log(1e-4), math.log(2e-3))))
        cfg = W.base_config(nxy, nxy, nz, 20.0, 10.0, W.reynolds_from_stokes(re_d), W.TWO_PI / 2000.0,
                            settling=ws, schmidt=0.5, precision=precision)
        cfg.update(noise_u=0.1, noise_c=0.01, seed=1000 + m, re_delta=re_d)
        out.append(cfg)
    return out


class Ensemble:
    def __init__(self, cfgs, member_overlap=False, autotune=False, grouped=False):
        """member_overlap: let each member also run its scalar on a second
        stream.  Off by default: the members already overlap one another, and
        three streams per member over the process's few hardware queues
        serialise them (1593 vs 990 member-steps/s for the 16 members).
        The switch is the overlap bit of each member's kernel variant
        (tbd_set_variant), set per context.
        autotune: each member picks its own kernel variant (tbd_autotune,
        run alone); the overlap bit is then set back as member_overlap asks.
        grouped: also build the grouped stepper (step_grouped); it needs the
        overlap and fused-transform bits clear, which it sets."""
        self.cfgs = cfgs
        self.solvers = []
        self.steppers = []
        self.feats = []
        overlap = RT.Solver.VARIANT_BITS["overlap"]
        for cfg in cfgs:
            s = RT.Solver(cfg)
            if not member_overlap:
                s.set_variant(s.variant() & ~overlap)
            s.upload(*W.initial_state(cfg))
            self.solvers.append(s)
            self.steppers.append(RT.Stepper(cfg["dt"]))
            self.feats.append(np.zeros(50))
        self.variants = [s.autotune(3)[0] if autotune else s.variant() for s in self.solvers]
        if autotune and not member_overlap:
            for s in self.solvers:
                s.set_variant(s.variant() & ~overlap)
            self.variants = [s.variant() for s in self.solvers]
        self.group = None
        if grouped:
            fused = RT.Solver.VARIANT_BITS["fused"]
            for s in self.solvers:
                s.set_variant(s.variant() & ~(overlap | fused))
            self.variants = [s.variant() for s in self.solvers]
            self.group = RT.GroupedEnsemble(self.solvers)

    def step(self, members=None):
        """One step of the given members (all by default), concurrently."""
        idx = range(len(self.solvers)) if members is None else members
        for m in idx:
            st = self.steppers[m]
            st.update(self.feats[m][19])
            self.solvers[m].step(st.time, st.alternation)
            st.toggle()
        for m in idx:
            self.feats[m] = self.solvers[m].features()

    def step_grouped(self):
        """One step of every member through the grouped launches."""
        times, alts = [], []
        for m, st in enumerate(self.steppers):
            st.update(self.feats[m][19])
            times.append(st.time)
            alts.append(st.alternation)
        self.group.step(times, alts)
        for st in self.steppers:
            st.toggle()
        feats = self.group.features()
        for m in range(len(self.solvers)):
            self.feats[m] = feats[m]

    # variant bits tried by tune_grouped, each set or cleared on every member
    TUNE_BITS = ["rhs", "wide_yb", "wide_yf", "wide_xb", "wide_xf", "swap_yb", "swap_yf", "swap_xb", "swap_xf",
                 "wide_zgemm"]

    def tune_grouped(self, steps=5, log=None):
        """Coordinate descent over the variant bits of TUNE_BITS, each applied
        to every member at once (alike members keep merging), timing `steps`
        grouped steps after two warm-up steps that also record the variant's
        step.  The members' fields, steppers and features are restored
        afterwards (uploaded, so the wall-row history restarts as on a
        restart from files); the fastest bits are kept.  Returns (the
        members' variant bits, {tried bit name or 0: ms per step})."""
        saved = [s.download() for s in self.solvers]
        steppers = copy.deepcopy(self.steppers)
        feats = [f.copy() for f in self.feats]
        base = [s.variant() for s in self.solvers]
        flags = RT.Solver.VARIANT_BITS

        def timed(mask_on, mask_off):
            for s, b in zip(self.solvers, base):
                s.set_variant((b | mask_on) & ~mask_off)
            for _ in range(2):
                self.step_grouped()
            t = self.run_grouped(steps)
            return 1e3 * t / steps

        on, off = 0, 0
        best = timed(on, off)
        tried = {0: best}
        for name in self.TUNE_BITS:
            bit = flags[name]
            if all(b & bit for b in base):
                c_on, c_off = on & ~bit, off | bit
            else:
                c_on, c_off = on | bit, off & ~bit
            ms = timed(c_on, c_off)
            tried[name] = ms
            if log:
                log("ensemble: tune %s %s: %.3f ms per step (best %.3f)"
                    % (name, "on" if c_on & bit else "off", ms, best))
            if ms < best:
                best, on, off = ms, c_on, c_off
        for s, b, st in zip(self.solvers, base, saved):
            s.set_variant((b | on) & ~off)
            s.upload(*st)
        self.steppers = steppers
        self.feats = feats
        self.variants = [s.variant() for s in self.solvers]
        return self.variants, tried

    def run_grouped(self, nsteps):
        t0 = time.perf_counter()
        for _ in range(nsteps):
            self.step_grouped()
        return time.perf_counter() - t0

    def run_concurrent(self, nsteps):
        t0 = time.perf_counter()
        for _ in range(nsteps):
            self.step()
        return time.perf_counter() - t0

    def run_sequential(self, nsteps):
        t0 = time.perf_counter()
        for m in range(len(self.solvers)):
            for _ in range(nsteps):
                self.step([m])
        return time.perf_counter() - t0

    def finite(self):
        return all(np.isfinite(f[0]) for f in self.feats)

    def close(self):
        if self.group is not None:
            self.group.close()
        for s in self.solvers:
            s.close()


def roofline(cfgs, ms_per_step):
    """Sum over the members of the step roofline (minimal algorithmic bytes
    and flops of each member's step at MI355X peaks) against the measured
    time of one ensemble step."""
    t = sum(step_roofline(c, c.get("precision", 0) == 1, 1.0)["t_roof_ms"] for c in cfgs)
    return {"t_roof_ms": t, "ms_per_step": ms_per_step, "frac": t / ms_per_step}


def measure(cfgs, steps, warmup, legs=("grouped", "streams", "sequential"), tune=False, log=None):
    """Member-steps/s of each leg on the same member configurations (each leg
    starts from the seeded states)."""
    total = len(cfgs) * steps
    out = {"members": len(cfgs), "steps": steps}
    if "grouped" in legs:
        ens = Ensemble(cfgs, grouped=True)
        if tune:
            _, tried = ens.tune_grouped(log=log)
            out["tune_ms_per_step"] = tried
        for _ in range(warmup):
            ens.step_grouped()
        tg = ens.run_grouped(steps)
        out["grouped_steps_per_s"] = total / tg
        out["launches_per_step"] = dict(zip(("grouped", "members"), ens.group.launches()))
        out["roofline"] = roofline(cfgs, 1e3 * tg / steps)
        out["variants"] = ens.variants
        out["finite"] = ens.finite()
        ens.close()
        if log:
            log("ensemble: grouped %.1f member-steps/s (%d grouped launches per step for %d member launches), "
                "roofline %.3f" % (out["grouped_steps_per_s"], out["launches_per_step"]["grouped"],
                                   out["launches_per_step"]["members"], out["roofline"]["frac"]))
    if "streams" in legs or "sequential" in legs:
        ens = Ensemble(cfgs)
        for _ in range(warmup):
            ens.step()
        if "streams" in legs:
            out["concurrent_steps_per_s"] = total / ens.run_concurrent(steps)
        if "sequential" in legs:
            out["sequential_steps_per_s"] = total / ens.run_sequential(steps)
        out["finite"] = bool(out.get("finite", True) and ens.finite())
        ens.close()
        if log:
            log("ensemble: streams %s, sequential %s member-steps/s"
                % (out.get("concurrent_steps_per_s"), out.get("sequential_steps_per_s")))
    if "grouped_steps_per_s" in out and "sequential_steps_per_s" in out:
        out["grouped_vs_sequential"] = out["grouped_steps_per_s"] / out["sequential_steps_per_s"]
    if "grouped_steps_per_s" in out and "concurrent_steps_per_s" in out:
        out["grouped_vs_streams"] = out["grouped_steps_per_s"] / out["concurrent_steps_per_s"]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=16)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--legs", default="grouped,streams,sequential")
    ap.add_argument("--tune", action="store_true", help="tune the grouped members' variant bits together")
    args = ap.parse_args(argv)
    cfgs = member_configs(args.members, args.seed, args.precision)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    out = dict(metric="ensemble RK3 member-steps/s")
    out.update(measure(cfgs, args.steps, args.warmup, tuple(args.legs.split(",")), args.tune, log))
    out.update(precision="mixed" if args.precision == 1 else "fp64",
               grids=[[c["nx"], c["ny"], c["nz"]] for c in cfgs],
               re_delta=[round(c["re_delta"], 1) for c in cfgs],
               settling=[float("%.3g" % c["settling_velocity"]) for c in cfgs])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
