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


def _hist(tke, tau, gpu_per_bin=1.0):
    """Per-bin statistics of a run whose period means are tke[i], tau[i]."""
    from turbid_amd import spinup as SP
    out, t = [], 0.0
    for a, b in zip(tke, tau):
        for _ in range(SP.PHASE_BINS):
            t += gpu_per_bin
            out.append(dict(tke=a, tau=b, cmean=1.0, gpu_s=t))
    return out


def test_time_to_reference_needs_consecutive_periods_in_band():
    from turbid_amd import spinup as SP
    ref = dict(tke=1.0, tau=10.0)
    # period 2 is in band, period 3 falls out, periods 4-5 are in band
    h = _hist([0.1, 0.98, 0.5, 1.02, 0.99], [8.0, 10.1, 10.0, 9.9, 10.0])
    gpu, per = SP.time_to_reference(h, ref, hold=2)
    assert per == 4 and gpu == 4 * SP.PHASE_BINS
    # the wall shear alone out of band keeps a period out
    h = _hist([1.0, 1.0, 1.0], [10.0, 12.0, 10.0])
    assert SP.time_to_reference(h, ref, hold=2) == (None, None)
    assert SP.time_to_reference(h, ref, hold=1)[1] == 1


def test_summary_break_even():
    from turbid_amd import spinup as SP
    rec = dict(data_gpu_s=100.0, train_s=50.0)
    res = dict(plain=dict(gpu_s=40.0, gpu_s_run=60.0), proposal=dict(gpu_s=10.0, gpu_s_run=20.0))
    s = SP.summary(res, rec)
    assert s["ratio"] == 0.25 and s["break_even_runs"] == 5.0 and s["one_time_s"] == 150.0
    # a plain start that never reaches the band: its run time is a lower bound
    res["plain"]["gpu_s"] = None
    s = SP.summary(res, rec)
    assert not s["plain_reached"] and s["plain_lower_bound"] == 60.0 and s["ratio"] == 10.0 / 60.0
