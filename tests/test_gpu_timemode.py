"""Time mode (output_based_on_step = 0: TimeStepper::Update2, reference
src/stepper_time.cc:120-208, with Phase::Update reassembling the solves for
the new dt, src/phase_carrier_flat.cc:250-265) and restart from a saved
index (src/variable_carrier_flat.cc:386-436, the wall-gradient history and
the alternation flag starting afresh) on the GPU against the oracle: through
the Python solver and through the C++ drop-in binary."""
import os
import subprocess

import numpy as np
import pytest

from turbid_amd import io as TIO
from turbid_amd import runtime as RT
from turbid_amd import workloads as W

from helpers import field_errors, oracle_steps, small_config

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "turbid_amd", "bin", "channelflat")
NSTEPS = 8


def _time_stepper(cfg):
    # max CFL far above the flow's: dt grows by 1.05 every step, and the
    # output time 5.5 dt0 falls inside the run, where dt is cut to land on it
    dt0 = cfg["dt"]
    return RT.TimeModeStepper(dt0, max_time=1e3, output_interval=5.5 * dt0, max_cfl=50.0)


@pytest.mark.parametrize("precision,tol", [(0, 1e-9), (1, 1e-5)])
def test_varying_dt_matches_oracle(precision, tol):
    cfg = small_config(nx=24, ny=24, nz=33, precision=precision)
    cfg["output_on_step"] = 0
    state = W.initial_state(cfg)
    g = RT.Solver(cfg)
    g.upload(*state)
    st = _time_stepper(cfg)
    dts = []
    feat = np.zeros(50)
    for _ in range(NSTEPS):
        st.update(feat[19])
        dts.append(st.dt)
        g.set_dt(st.dt)
        g.step(st.time, st.alternation)
        st.toggle()
        feat = g.features()
    got = g.download()
    g.close()
    # dt changed on every step (growth, then the cut onto the output time)
    assert len(set(dts)) == NSTEPS, dts
    assert st.output_index == 1
    ref = oracle_steps(dict(cfg, precision=0), state, NSTEPS, stepper=_time_stepper(cfg))
    errs = field_errors(got, ref)
    print("time mode, precision %d, dt %s: %s" % (precision, ["%.4g" % x for x in dts], errs))
    assert max(errs.values()) < tol, errs


def _run(tmp_path, *args, precision="fp64"):
    env = dict(os.environ, TURBID_PRECISION=precision)
    out = subprocess.run([BIN, *args], cwd=str(tmp_path), capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    return out


def _log_rows(path):
    return [l.split() for l in open(path).read().splitlines() if l.strip() and l.split()[0].isdigit()]


def test_dropin_time_mode(tmp_path):
    cfg = small_config(nx=24, ny=24, nz=33)
    cfg["output_on_step"] = 0
    state = W.initial_state(cfg)
    dt0 = cfg["dt"]
    # two output times (indices 1 and 2) at 3 dt0 and 6 dt0 of simulated time
    ini = TIO.write_run(str(tmp_path), cfg, *state, time_mode=True, max_time=6 * dt0, output_interval=3 * dt0,
                        start_time_index=0, max_cfl=50.0)
    _run(tmp_path, ini, "logcomputation")
    # the oracle through the same Update2 loop, its state kept at every output
    from oracle import oracle as O
    o = O.Oracle(cfg)
    o.set_state(*state)
    st = RT.TimeModeStepper(dt0, max_time=6 * dt0, output_interval=3 * dt0, max_cfl=50.0)
    feat = np.zeros(50)
    saved, times = {}, []
    while st.forward:
        st.update(feat[19])
        o.set_dt(st.dt)
        for s in range(3):
            o.substage(s, st.time, st.alternation)
        st.toggle()
        feat = o.features()
        times.append(st.time)
        if st.output:
            saved[st.output_index] = o.get_state()
    o.close()
    assert sorted(saved) == [1, 2]
    rows = _log_rows(tmp_path / "logcomputation")
    assert len(rows) == len(times)
    assert max(abs(float(r[1]) - t) for r, t in zip(rows, times)) < 1e-9
    for index in (1, 2):
        errs = field_errors(TIO.read_state(str(tmp_path), index, cfg), saved[index])
        print("drop-in time mode, output %d (%d steps in all): %s" % (index, len(times), errs))
        assert max(errs.values()) < 1e-9, errs


def test_dropin_restart_from_saved_step(tmp_path):
    # run 0 -> 6 writing every 3 steps, then restart from 3/ with
    # task_start_step = 3: the restarted steps 4..6 start with a zero wall
    # pressure-gradient history and the alternation flag false, as the
    # reference's load does, so they equal the oracle started from the saved
    # state with a fresh stepper at step 3
    cfg = small_config(nx=24, ny=24, nz=33)
    state = W.initial_state(cfg)
    first = tmp_path / "first"
    ini = TIO.write_run(str(first), cfg, *state, max_step=6, write_interval=3)
    _run(first, ini, "logcomputation")
    saved = TIO.read_state(str(first), 3, cfg)
    second = tmp_path / "second"
    os.makedirs(second / "3")
    for f in os.listdir(first / "3"):
        with open(first / "3" / f, "rb") as a, open(second / "3" / f, "wb") as b:
            b.write(a.read())
    ini2 = str(second / "configuration.ini")
    TIO.write_ini(ini2, cfg, max_step=6, write_interval=3, start_step=3)
    _run(second, ini2, "logcomputation")
    rows = _log_rows(second / "logcomputation")
    assert [int(r[0]) for r in rows] == [4, 5, 6]
    assert abs(float(rows[0][1]) - 4 * cfg["dt"]) < 1e-10   # the log prints 10 digits
    got = TIO.read_state(str(second), 6, cfg)
    ref = oracle_steps(cfg, tuple(np.asarray(x) for x in saved), 3, stepper=RT.Stepper(cfg["dt"], start_step=3))
    errs = field_errors(got, ref)
    print("restart from step 3: %s" % errs)
    assert max(errs.values()) < 1e-9, errs
    # and the restarted run is not the continuous one (history and flag reset)
    cont = TIO.read_state(str(first), 6, cfg)
    assert field_errors(got, cont)["u"] > 1e-9
