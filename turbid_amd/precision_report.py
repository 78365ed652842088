"""Mixed-precision error per field against the FP64 path on the same start.

    python -m turbid_amd.precision_report --workload turb384 --steps 50

Both runs start from the same seeded state and take the reference's step
loop (TimeStepper::Update with the previous step's CFL, three substages,
alternation toggle), each with its own CFL history.  Reported: the relative
L2 difference of u, v, w, p and c after the given number of steps, and the
steps/s of both paths.  The FP64 path stands in for the FP64 reference at
sizes where the CPU restatement would take hours; it agrees with that
restatement to 1e-9 on the smaller grids of the test suite.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import runtime as RT
from . import workloads as W


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run(cfg, state, steps):
    s = RT.Solver(cfg)
    s.upload(*state)
    st = RT.Stepper(cfg["dt"])
    RT.run_steps(s, st, 1)          # first step: graph capture
    s.sync()
    t0 = time.perf_counter()
    RT.run_steps(s, st, steps - 1)
    s.sync()
    dt = time.perf_counter() - t0
    out = s.download()
    s.close()
    return out, (steps - 1) / dt


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="turb384")
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    cfg64 = W.workload(args.workload, precision=0)
    state = W.initial_state(cfg64)
    res = {}
    for prec in (0, 1):
        cfg = dict(cfg64)
        cfg["precision"] = prec
        log("precision_report: %s precision %d, %d steps" % (args.workload, prec, args.steps))
        res[prec] = run(cfg, state, args.steps)
    (u64, p64, c64), sps64 = res[0]
    (u32, p32, c32), sps32 = res[1]
    u64, u32 = u64.reshape(3, -1), u32.reshape(3, -1)
    out = dict(workload=args.workload, grid=[cfg64["nx"], cfg64["ny"], cfg64["nz"]], steps=args.steps,
               rel_l2={"u": rel(u32[0], u64[0]), "v": rel(u32[1], u64[1]), "w": rel(u32[2], u64[2]),
                       "p": rel(p32, p64), "c": rel(c32, c64)},
               steps_per_s={"fp64": sps64, "mixed": sps32})
    line = json.dumps(out)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")
    print(line, flush=True)


if __name__ == "__main__":
    main()
