"""Host side of the phase-statistics driver: the control start and the
per-quantity differences (no GPU)."""
import numpy as np

from turbid_amd import phasestats as P
from turbid_amd import workloads as W


def test_control_start_is_the_seeded_start_plus_a_scaled_independent_draw():
    cfg = W.workload("laminar32", precision=0)
    cfg.update(noise_u=0.1, noise_c=0.01, seed=3)
    base = P.start(cfg)
    same = P.start(cfg, 0.0)
    for a, b in zip(base, same):
        assert np.array_equal(a, b)
    eps = 1e-6
    pert = P.start(cfg, eps)
    d_uvw = pert[0] - base[0]
    d_c = pert[2] - base[2]
    # the perturbation is eps times a noise field of the start's amplitude:
    # compare with the start's own noise (start minus its noise-free form)
    quiet = W.initial_state(cfg, noise_u=0.0, noise_c=0.0)
    n_uvw = base[0] - quiet[0]
    n_c = base[2] - quiet[2]
    r_u = np.linalg.norm(d_uvw) / (eps * np.linalg.norm(n_uvw))
    r_c = np.linalg.norm(d_c) / (eps * np.linalg.norm(n_c))
    assert 0.8 < r_u < 1.25 and 0.8 < r_c < 1.25
    # an independent draw, not a rescaled copy of the start's noise
    cos = np.vdot(d_uvw.ravel(), n_uvw.ravel()).real / (np.linalg.norm(d_uvw) * np.linalg.norm(n_uvw))
    assert abs(cos) < 0.05
    assert np.array_equal(pert[1], base[1])


def test_differences_are_relative_to_the_fp64_scale_of_each_kind():
    bins, nz = 3, 5
    p64 = np.zeros((bins, 8, nz))
    p64[:, 0] = 2.0          # U
    p64[:, 3] = 4.0          # C
    p64[:, 4] = 0.5          # uu, vv, ww: k-scale 1.5
    p64[:, 5] = 0.5
    p64[:, 6] = 0.5
    p64[:, 7] = 0.1          # cc
    t64 = np.full((bins, 2), 10.0)
    p = p64.copy()
    t = t64.copy()
    p[1, 1, 2] += 0.02       # V against the U scale
    p[2, 3, 0] -= 0.4        # C against its own
    p[0, 5, 4] += 0.15       # vv against the k scale
    p[2, 7, 1] += 0.01       # cc against its own
    t[1, 0] += 1.0
    d = P.differences(p64, t64, p, t)
    assert np.isclose(d["V"], 0.01) and d["U"] == 0.0 and d["W"] == 0.0
    assert np.isclose(d["C"], 0.1)
    assert np.isclose(d["vv"], 0.1) and d["uu"] == 0.0
    assert np.isclose(d["cc"], 0.1)
    assert np.isclose(d["wall_shear"], 0.1)
