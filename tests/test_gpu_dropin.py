"""The C++ drop-in host (turbid_amd/bin/channelflat) end to end on the GPU:
reference-format configuration.ini and 0/phase_NN_variable.dat in, the
step loop of FlatChannelSolver::Execute, <step>/phase_NN_variable.dat out,
compared with the FP64 CPU restatement run for the same steps."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from turbid_amd import io as TIO
from turbid_amd import workloads as W

from helpers import rel_l2, small_config

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "turbid_amd", "bin", "channelflat")


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 1e-5)])
def test_channelflat_run_matches_oracle(tmp_path, precision, tol):
    nsteps = 6
    cfg = small_config(nx=24, ny=24, nz=33)
    uvw, p, c = W.initial_state(cfg)
    ini = TIO.write_run(str(tmp_path), cfg, uvw, p, c, max_step=nsteps, write_interval=nsteps)
    env = dict(os.environ, TURBID_PRECISION=precision)
    out = subprocess.run([BIN, ini, "logcomputation"], cwd=str(tmp_path), capture_output=True, text=True,
                         env=env, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    ug, pg, cg = TIO.read_state(str(tmp_path), nsteps, cfg)
    o = O.Oracle(cfg)
    o.set_state(uvw, p, c)
    for _ in range(nsteps):
        o.step()
    uo, po, co = o.get_state()
    assert rel_l2(ug, uo) < tol
    assert rel_l2(cg, co) < tol
    log = open(tmp_path / "logcomputation").read().splitlines()
    body = [l for l in log if l.strip() and l.split()[0].isdigit()]
    assert [int(l.split()[0]) for l in body] == list(range(1, nsteps + 1))
    for ph in (0, 1):
        for what in ("phase", "variable", "boundary"):
            assert os.path.exists(tmp_path / str(nsteps) / ("phase_%02d_%s.dat" % (ph, what)))
