"""Ensemble members advanced concurrently (own streams and step graphs)
reproduce the same members advanced alone, bit for bit, and the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from turbid_amd import ensemble as E
from turbid_amd import runtime as RT
from turbid_amd import workloads as W

from helpers import rel_l2

pytestmark = pytest.mark.gpu


def _small_members():
    cfgs = []
    for m, (n, nz, re_d, ws) in enumerate([(24, 33, 400.0, 1e-3), (36, 17, 700.0, 2e-4), (16, 65, 1000.0, 5e-4)]):
        c = W.base_config(n, n, nz, 20.0, 10.0, W.reynolds_from_stokes(re_d), W.TWO_PI / 2000.0, settling=ws,
                          schmidt=0.5, precision=m % 2)
        c.update(noise_u=0.1, noise_c=0.01, seed=50 + m)
        cfgs.append(c)
    return cfgs


def test_concurrent_members_match_alone():
    cfgs = _small_members()
    ens = E.Ensemble(cfgs)
    for _ in range(4):
        ens.step()
    got = [s.download() for s in ens.solvers]
    ens.close()
    for cfg, (ug, pg, cg) in zip(cfgs, got):
        s = RT.Solver(cfg)
        s.upload(*W.initial_state(cfg))
        RT.run_steps(s, RT.Stepper(cfg["dt"]), 4)
        ua, pa, ca = s.download()
        s.close()
        assert np.array_equal(ug, ua)
        assert np.array_equal(cg, ca)


def test_member_matches_oracle():
    cfg = _small_members()[1]
    ens = E.Ensemble([cfg])
    for _ in range(3):
        ens.step()
    ug, pg, cg = ens.solvers[0].download()
    ens.close()
    o = O.Oracle(cfg)
    uvw, p, c = W.initial_state(cfg)
    o.set_state(uvw, p, c)
    st = RT.Stepper(cfg["dt"])
    feat = np.zeros(50)
    for _ in range(3):
        st.update(feat[19])
        for s in range(3):
            o.substage(s, st.time, st.alternation)
        st.toggle()
        feat = o.features()
    uo, po, co = o.get_state()
    tol = 1e-5 if cfg["precision"] == 1 else 1e-9
    assert rel_l2(ug, uo) < tol
    assert rel_l2(cg, co) < tol
