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
import json
import math
import sys
import time

import numpy as np

from . import runtime as RT
from . import workloads as W


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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=16)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--autotune", action="store_true", help="each member picks its kernel variant")
    args = ap.parse_args(argv)
    cfgs = member_configs(args.members, args.seed, args.precision)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    total = args.members * args.steps
    rates = {}
    # grouped: its own members (the grouped stepper sets their variant bits)
    ens = Ensemble(cfgs, autotune=args.autotune, grouped=True)
    for _ in range(args.warmup):
        ens.step_grouped()
    tg = ens.run_grouped(args.steps)
    rates["grouped"] = total / tg
    launches = ens.group.launches()
    variants = ens.variants
    finite_g = ens.finite()
    ens.close()
    log("ensemble: grouped %.1f member-steps/s (%d grouped launches per step for %d member launches)"
        % (rates["grouped"], launches[0], launches[1]))
    ens = Ensemble(cfgs, autotune=args.autotune)
    for _ in range(args.warmup):
        ens.step()
    tc = ens.run_concurrent(args.steps)
    ts = ens.run_sequential(args.steps)
    rates["streams"] = total / tc
    rates["sequential"] = total / ts
    log("ensemble: streams %.1f, sequential %.1f member-steps/s" % (rates["streams"], rates["sequential"]))
    out = dict(metric="ensemble RK3 member-steps/s", members=args.members, steps=args.steps,
               grouped_steps_per_s=rates["grouped"], concurrent_steps_per_s=rates["streams"],
               sequential_steps_per_s=rates["sequential"], grouped_vs_sequential=rates["grouped"] / rates["sequential"],
               grouped_vs_streams=rates["grouped"] / rates["streams"],
               launches_per_step={"grouped": launches[0], "members": launches[1]},
               precision="mixed" if args.precision == 1 else "fp64", finite=bool(ens.finite() and finite_g),
               grids=[[c["nx"], c["ny"], c["nz"]] for c in cfgs],
               re_delta=[round(c["re_delta"], 1) for c in cfgs],
               settling=[float("%.3g" % c["settling_velocity"]) for c in cfgs],
               variants=variants)
    ens.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
