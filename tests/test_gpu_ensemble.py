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


def _alone(cfg, nsteps):
    s = RT.Solver(cfg)
    s.upload(*W.initial_state(cfg))
    RT.run_steps(s, RT.Stepper(cfg["dt"]), nsteps)
    out = s.download()
    s.close()
    return out


def _grouped_members():
    # grids of the register transforms' lengths (36 is not one of them)
    cfgs = []
    for m, (n, nz, re_d, ws) in enumerate([(24, 33, 400.0, 1e-3), (48, 17, 700.0, 2e-4), (16, 65, 1000.0, 5e-4),
                                           (24, 33, 900.0, 1e-4)]):
        c = W.base_config(n, n, nz, 20.0, 10.0, W.reynolds_from_stokes(re_d), W.TWO_PI / 2000.0, settling=ws,
                          schmidt=0.5, precision=m % 2)
        c.update(noise_u=0.1, noise_c=0.01, seed=60 + m)
        cfgs.append(c)
    return cfgs


def test_grouped_members_match_alone():
    # FP64 and mixed members of three grids in one ensemble: their launch
    # sequences differ, so the merge groups some positions and not others;
    # 5 steps replay both recorded (alternation, wall-row slot) programs
    cfgs = _grouped_members()
    ens = E.Ensemble(cfgs, grouped=True)
    for _ in range(5):
        ens.step_grouped()
    g, m = ens.group.launches()
    got = [s.download() for s in ens.solvers]
    feats = [f.copy() for f in ens.feats]
    gathered = ens.group.features()
    for k, s in enumerate(ens.solvers):
        assert np.array_equal(gathered[k], s.features())
    ens.close()
    print("grouped launches per step %d, member launches %d" % (g, m))
    assert 0 < g < m
    for cfg, (ug, pg, cg), fg in zip(cfgs, got, feats):
        ua, pa, ca = _alone(cfg, 5)
        assert np.array_equal(ug, ua)
        assert np.array_equal(pg, pa)
        assert np.array_equal(cg, ca)
        assert np.all(np.isfinite(fg[:20]))


def test_grouped_ensemble_of_the_sixteen_members_matches_streams():
    # the sixteen members of the ensemble workload (64..128 squared by
    # 65..129, mixed precision): grouped and stream-concurrent runs agree bit
    # for bit, and alike members merge position by position
    cfgs = E.member_configs(16, seed=0, precision=1)
    ens = E.Ensemble(cfgs, grouped=True)
    for _ in range(3):
        ens.step_grouped()
    g, m = ens.group.launches()
    got = [s.download() for s in ens.solvers]
    ens.close()
    print("grouped launches per step %d, member launches %d" % (g, m))
    assert 3 * g < m
    ref = E.Ensemble(cfgs)
    for _ in range(3):
        ref.step()
    want = [s.download() for s in ref.solvers]
    ref.close()
    for (ug, pg, cg), (ua, pa, ca) in zip(got, want):
        assert np.array_equal(ug, ua)
        assert np.array_equal(cg, ca)


def test_grouped_step_refuses_an_overlapping_member():
    cfg = _grouped_members()[0]
    s = RT.Solver(cfg)
    s.upload(*W.initial_state(cfg))
    s.set_variant(s.variant() | RT.Solver.VARIANT_BITS["overlap"])
    grp = RT.GroupedEnsemble([s])
    with pytest.raises(RT.TurbidError, match="overlap"):
        grp.step([cfg["dt"]], [0])
    # the fused x transforms are refused the same way, before anything runs
    s.set_variant((s.variant() & ~RT.Solver.VARIANT_BITS["overlap"]) | RT.Solver.VARIANT_BITS["fused"])
    with pytest.raises(RT.TurbidError, match="fused"):
        grp.step([cfg["dt"]], [0])
    grp.close()
    s.close()


def test_grouped_tuning_restores_the_members():
    cfgs = _grouped_members()
    ens = E.Ensemble(cfgs, grouped=True)
    ens.step_grouped()
    before = [s.download() for s in ens.solvers]
    times = [st.time for st in ens.steppers]
    bits, tried = ens.tune_grouped(steps=2)
    assert set(tried) == {0, *E.Ensemble.TUNE_BITS}
    assert all(np.isfinite(v) and v > 0 for v in tried.values())
    after = [s.download() for s in ens.solvers]
    assert [st.time for st in ens.steppers] == times
    for (ub, pb, cb), (ua, pa, ca) in zip(before, after):
        assert np.array_equal(ub, ua) and np.array_equal(cb, ca)
    assert bits == [s.variant() for s in ens.solvers]
    ens.step_grouped()
    assert ens.finite()
    ens.close()


def test_grouped_members_of_different_physics_match_alone():
    # a carrier-only member, a member whose scalar drives the carrier
    # (buoyancy: no scalar overlap, other kernels) and one with erosion: their
    # launch sequences differ in length, so the merge is ragged
    cfgs = []
    for m, (n, nz, kw) in enumerate([(24, 33, dict(nphase=1)),
                                     (24, 33, dict(affected=1, froude=0.5)),
                                     (16, 17, dict(erosion_rate=1e-4, critical_shear=1e-5))]):
        c = W.base_config(n, n, nz, 20.0, 10.0, W.reynolds_from_stokes(600.0), W.TWO_PI / 2000.0, settling=5e-4,
                          schmidt=0.5, precision=m % 2, **kw)
        c.update(noise_u=0.1, noise_c=0.01, seed=70 + m)
        cfgs.append(c)
    ens = E.Ensemble(cfgs, grouped=True)
    for _ in range(4):
        ens.step_grouped()
    got = [s.download() for s in ens.solvers]
    ens.close()
    for cfg, (ug, pg, cg) in zip(cfgs, got):
        ua, pa, ca = _alone(cfg, 4)
        assert np.array_equal(ug, ua)
        assert np.array_equal(pg, pa)
        if cfg["nphase"] > 1:
            assert np.array_equal(cg, ca)


@pytest.mark.skip(reason="grouped fused x transforms are switched off: the one run of this test faulted the GPU "
                         "(illegal memory access in the grouped step) and no cause has been shown")
def test_grouped_members_with_fused_transforms_match_alone():
    # the fused x transforms (inverse transforms, products, forward
    # transforms in one kernel) as grouped launches
    cfgs = _grouped_members()
    fused = RT.Solver.VARIANT_BITS["fused"]
    ens = E.Ensemble(cfgs, grouped=True)
    for s in ens.solvers:
        s.set_variant(s.variant() | fused)
    for _ in range(4):
        ens.step_grouped()
    got = [s.download() for s in ens.solvers]
    ens.close()
    for cfg, (ug, pg, cg) in zip(cfgs, got):
        s = RT.Solver(cfg)
        s.upload(*W.initial_state(cfg))
        s.set_variant(s.variant() | fused)
        RT.run_steps(s, RT.Stepper(cfg["dt"]), 4)
        ua, pa, ca = s.download()
        s.close()
        assert np.array_equal(ug, ua)
        assert np.array_equal(cg, ca)
