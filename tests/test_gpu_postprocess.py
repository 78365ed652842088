"""Swirling strength on the GPU (tbd_swirling_strength and the drop-in's
--postprocess mode) against the restatement (oracle.swirling_strength)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from turbid_amd import io as TIO
from turbid_amd import runtime as RT
from turbid_amd import workloads as W

from helpers import small_config

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "turbid_amd", "bin", "channelflat")


def _check(got, ref, cfg, uvw):
    # the closed-form cubic and LAPACK agree to round-off away from a double
    # root and to sqrt(eps) times the gradient scale next to one
    G = O.velocity_gradient(uvw, cfg)
    scale = np.abs(G).max()
    assert np.abs(got - ref).max() < 1e-6 * scale
    big = ref > 1e-3 * ref.max()
    assert np.abs(got - ref)[big].max() < 1e-9 * ref.max()


@pytest.mark.parametrize("precision", [0, 1])
def test_swirling_strength_matches_oracle(precision):
    cfg = small_config(nx=24, ny=24, nz=33, noise_u=0.2, precision=precision)
    uvw, p, c = W.initial_state(cfg)
    g = RT.Solver(cfg)
    g.upload(uvw, p, c)
    got = g.swirling_strength()
    g.close()
    ref = O.swirling_strength(uvw, cfg)
    _check(got, ref, cfg, uvw)


def test_postprocess_binary_writes_vortex_files(tmp_path):
    cfg = small_config(nx=24, ny=24, nz=33, noise_u=0.2)
    uvw, p, c = W.initial_state(cfg)
    ini = os.path.join(str(tmp_path), "configuration.ini")
    TIO.write_ini(ini, cfg, work_index_range=(0, 1))
    TIO.write_state(str(tmp_path), 0, cfg, uvw, p, c)
    uvw1 = uvw * 0.5
    TIO.write_state(str(tmp_path), 1, cfg, uvw1, p, c)
    out = subprocess.run([BIN, "--postprocess", ini], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    for index, u in ((0, uvw), (1, uvw1)):
        got = TIO.read_real_cube(os.path.join(str(tmp_path), str(index), "phase_00_vortex.dat"))
        assert got.shape == (cfg["nz"], cfg["ny"], cfg["nx"])
        _check(got, O.swirling_strength(u, cfg), cfg, u)
