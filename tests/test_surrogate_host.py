"""Host-side pieces of the spin-up proposal: the DNS <-> reduced-grid maps
(exact on band-limited, low-degree fields) and the input assembly."""
import math

import numpy as np

from turbid_amd import surrogate as SG
from turbid_amd import workloads as W


def _bandlimited_state(mx=22, nyc=42, nz=65, kmax=11, deg=20, seed=0):
    rng = np.random.default_rng(seed)
    z = W.cgl_grid(nz)
    spec = np.zeros((4, mx, nyc, nz), dtype=np.complex128)
    for kx in range(kmax + 1):
        for ky in range(-kmax, kmax + 1):
            iy = ky if ky >= 0 else nyc + ky
            coef = rng.standard_normal((4, deg)) + 1j * rng.standard_normal((4, deg))
            prof = np.polynomial.chebyshev.chebval(z, coef.T)       # (4, nz)
            spec[:, kx, iy, :] = prof
    # real field: Hermitian kx = 0 plane, real mean
    for j in range(1, kmax + 1):
        spec[:, 0, nyc - j, :] = np.conj(spec[:, 0, j, :])
    spec[:, 0, 0, :] = spec[:, 0, 0, :].real
    return spec[:3], spec[3:4]


def test_reduce_expand_round_trip():
    uvw, c = _bandlimited_state()
    vals = SG.reduce_state(uvw, c)
    assert vals.shape == (33, 24, 24, 4)
    u2, c2 = SG.expand_state(vals, uvw.shape[1], uvw.shape[2], uvw.shape[3])
    assert np.max(np.abs(u2 - uvw)) < 1e-9 * np.max(np.abs(uvw))
    assert np.max(np.abs(c2 - c)) < 1e-9 * np.max(np.abs(c))


def test_reduce_single_mode_closed_form():
    # u = cos(2 pi (2 x / Lx)) T_3(z): on the reduced grid the same function
    mx, nyc, nz = 22, 42, 65
    uvw = np.zeros((3, mx, nyc, nz), dtype=np.complex128)
    c = np.zeros((1, mx, nyc, nz), dtype=np.complex128)
    z = W.cgl_grid(nz)
    uvw[0, 2, 0, :] = 0.5 * np.cos(3 * np.arccos(z))     # cos = (e^{i} + e^{-i})/2, kx >= 0 stored
    vals = SG.reduce_state(uvw, c)
    x = np.arange(24) / 24.0
    zr = -np.cos(math.pi * np.arange(33) / 32)
    ref = np.cos(2 * math.pi * 2 * x)[None, None, :] * np.cos(3 * np.arccos(zr))[:, None, None]
    np.testing.assert_allclose(vals[..., 0], np.broadcast_to(ref, vals.shape[:3]), atol=1e-12)


def test_laminar_grid_and_input():
    cfg = W.workload("train128")
    lam = SG.laminar_grid(cfg, 0.7)
    zr = -np.cos(math.pi * np.arange(33) / 32)
    ref = W.stokes_profile(cfg["reynolds"], zr, 0.7, omega=1.0)
    comp = 0 if cfg["force_direction"][1] == 1 else 1
    np.testing.assert_allclose(lam[:, 3, 5, comp], ref, atol=1e-14)
    x = SG.network_input(lam, 0.7)
    assert x.shape == (33 * 24 * 24, SG.CIN)
    np.testing.assert_allclose(x[:, SG.NFIELD], math.cos(0.7))
