"""Shared helpers for the parity tests."""
import math

import numpy as np

from turbid_amd import workloads as W


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    if den == 0:
        return np.linalg.norm(a.ravel())
    return np.linalg.norm((a - b).ravel()) / den


def small_config(nx=24, ny=24, nz=33, nphase=2, noise_u=0.05, noise_c=0.01, erosion=True,
                 top="freeslip", precision=0, re_delta=180.0, affected=0, seed=3, lx=8 * math.pi,
                 ly=4 * math.pi):
    cfg = W.base_config(nx, ny, nz, lx, ly, W.reynolds_from_stokes(re_delta), W.TWO_PI / 2000.0,
                        nphase=nphase, settling=1e-3, schmidt=1.0, precision=precision, top=top,
                        affected=affected, froude=1.0)
    if erosion:
        cfg.update(erosion_rate=2e-4, critical_shear_stress=1e-5)
    cfg.update(noise_u=noise_u, noise_c=noise_c, seed=seed)
    return cfg


def spectral_split(state, cfg):
    """Flat RealCplxCube arrays -> (nv, mx, nyc, nz)."""
    _, nyc, mx = W.dims(cfg["nx"], cfg["ny"], cfg["nz"])
    return state.reshape(-1, mx, nyc, cfg["nz"])
