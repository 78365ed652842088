"""Workload definitions and seeded initial states for the oscillatory
bottom-boundary-layer runs.

Each workload is a plain dict with the fields of the ini sections the solver
reads ([Domain], [Space Stepper], [Time Stepper], [Phase Carrier],
[Phase Scalar]; reference src/configure.h:63-117).  Initial states are built
in physical space from a counter-based seeded generator, multiplied by wall
windows, and forward transformed with numpy's FFT under the solver's
truncation convention (kx < mx, ky in FFT order, 1/(Nx*Ny) scale), so the
kx = 0 plane is Hermitian.  The mean mode carries the analytic Stokes layer.

Domain and Reynolds-number mapping: a domain quoted as a x b x 1 becomes
lengths (2a, 2b, 2) because the solver's wall-normal extent is [-1, 1]
(reference src/domain.cc:55-58); a Reynolds number Re_d based on the Stokes
thickness maps to the solver's Re = Re_d^2 * omega / 2 (the re800 example:
src .../oscillatory_flow_re800/result/configuration.ini:96-97).
"""
import math

import numpy as np

TWO_PI = 2.0 * math.pi

# vertical_type_* = (a_bottom, b_bottom, a_top, b_top, top value of the mean mode)
NOSLIP_FREESLIP = (1.0, 0.0, 0.0, 1.0, 0.0)
NOSLIP_NOSLIP = (1.0, 0.0, 1.0, 0.0, 0.0)
NEUMANN_NEUMANN = (0.0, 1.0, 0.0, 1.0, 0.0)


def dims(nx, ny, nz):
    nxc, nyc = nx // 3 * 2, ny // 3 * 2
    return nxc, nyc, nxc // 2 + 1


def base_config(nx, ny, nz, lx, ly, re, dt, nphase=2, omega=1.0, settling=1e-3,
                schmidt=1.0, erosion_rate=0.0, critical_shear=1.0, precision=0,
                top="freeslip", affected=0, froude=1.0, specific_gravity=2.65):
    top_uv = NOSLIP_FREESLIP if top == "freeslip" else NOSLIP_NOSLIP
    return dict(
        nx=nx, ny=ny, nz=nz, nphase=nphase, lx=lx, ly=ly, lz=2.0,
        slope=0.0, slope_direction=0, mapping=0, mapping_coefficient=0.0,
        dt=dt, output_on_step=1,
        force_type=2, force_direction=(0, 1), force_amplitude=(0.0, omega, 0.0, 0.0),
        reynolds=re, froude=froude, averaged_concentration=1.0,
        specific_gravity=specific_gravity, affected_by_scalar_phase=affected,
        vt_u1=top_uv, vt_u2=top_uv, vt_u3=NOSLIP_NOSLIP, vt_p=NEUMANN_NEUMANN,
        schmidt=schmidt, settling_velocity=settling, erosion_rate=erosion_rate,
        critical_shear_stress=critical_shear, scalar_bc=1, precision=precision,
    )


def reynolds_from_stokes(re_delta, omega=1.0):
    return re_delta * re_delta * omega / 2.0


def workload(name, precision=0):
    """Named workloads (grid, physics) used by tests and the benchmark."""
    dt = TWO_PI / 2000.0
    if name == "laminar32":       # 32x32x65 plain FP64 laminar case
        c = base_config(32, 32, 65, 8 * math.pi, 4 * math.pi, reynolds_from_stokes(180.0), dt,
                        settling=1e-3, schmidt=1.0, precision=precision)
        c.update(noise_u=1e-4, noise_c=0.0, seed=0)
        return c
    if name == "turb256":         # 256x256x257 turbulent BBL with fine sediment
        c = base_config(256, 256, 257, 20.0, 10.0, reynolds_from_stokes(800.0), dt,
                        settling=5e-4, schmidt=0.5, precision=precision)
        c.update(noise_u=0.1, noise_c=0.01, seed=1)
        return c
    if name == "turb384":         # config-1 physics at 384x384x257
        c = base_config(384, 384, 257, 20.0, 10.0, reynolds_from_stokes(800.0), dt,
                        settling=5e-4, schmidt=0.5, precision=precision)
        c.update(noise_u=0.1, noise_c=0.01, seed=1)
        return c
    if name == "spinup128":       # box and wave of the reference's Re_delta = 800 example
        # (examples/oscillatory_flow_re800/result/configuration.ini:41,96-97: box 2 x 2 x 2,
        # omega = 0.075) at Re_delta = 1000 (Re = 37500): on 128x128x129 the layer stays
        # turbulent at 1000 and relaminarises at 800.  dt = T/12000 keeps u dt/dx near 0.5;
        # zero-flux walls for the sediment (its mass is conserved)
        omega = 0.0535237234845831 * 2 * math.pi / 10.0 / 0.448399297311834
        c = base_config(128, 128, 129, 2.0, 2.0, reynolds_from_stokes(1000.0, omega), 2 * math.pi / omega / 12000.0,
                        omega=omega, settling=5e-4, schmidt=0.5, precision=precision)
        c.update(noise_u=0.1, noise_c=0.01, seed=1, scalar_bc=0)
        return c
    if name == "train128":        # 128x128x129 surrogate-training resolution
        c = base_config(128, 128, 129, 20.0, 10.0, reynolds_from_stokes(800.0), dt,
                        settling=5e-4, schmidt=0.5, precision=precision)
        c.update(noise_u=0.1, noise_c=0.01, seed=1)
        return c
    raise KeyError(name)


# ---------------------------------------------------------------------------
# analytic laminar solutions
# ---------------------------------------------------------------------------
def _cosh_ratio(a, b):
    """cosh(a)/cosh(b) for Re(b) > 0, Re(a) <= Re(b), without overflow."""
    return (np.exp(a - b) + np.exp(-a - b)) / (1.0 + np.exp(-2.0 * b))


def stokes_profile(re, z, t, omega=1.0, top="freeslip"):
    """Laminar oscillatory flow under forcing omega*cos(omega t):
    u = Re(U(z) e^{i omega t}), i omega U = omega + U''/re.
    freeslip: U(-1)=0, U'(1)=0; noslip: U(+-1)=0 (tmtools/get_loflow.m:17-19)."""
    z = np.asarray(z, dtype=np.float64)
    k = (1.0 + 1.0j) * math.sqrt(re * omega / 2.0)
    if top == "freeslip":
        U = 1j * (_cosh_ratio(k * (z - 1.0), 2.0 * k) - 1.0)
    else:
        U = 1j * (_cosh_ratio(k * z, k) - 1.0)
    return np.real(U * np.exp(1j * omega * t))


def cgl_grid(nz):
    """Chebyshev-Gauss-Lobatto points, bottom first (src/domain.cc:306)."""
    return np.cos((nz - np.arange(nz) - 1) * math.pi / (nz - 1))


# ---------------------------------------------------------------------------
# seeded initial states
# ---------------------------------------------------------------------------
def _normal(seed, field, shape):
    ss = np.random.SeedSequence([int(seed), int(field)])
    return np.random.Generator(np.random.PCG64(ss)).standard_normal(shape)


def forward_truncated(phys):
    """Physical (nv, nz, ny, nx) -> spectral (nv, mx, nyc, nz) with the
    solver's truncation and scale, Hermitian on the kx = 0 plane."""
    nv, nz, ny, nx = phys.shape
    _, nyc, mx = dims(nx, ny, nz)
    F = np.fft.rfft2(phys, axes=(-2, -1)) / (nx * ny)
    yf = [iy if iy < nyc // 2 else iy + ny - nyc for iy in range(nyc)]
    spec = F[:, :, yf, :mx].transpose(0, 3, 2, 1).copy()
    # exact Hermitian symmetry on kx = 0 for the retained +-ky pairs
    for iy in range(1, nyc // 2):
        jy = nyc - iy
        a = spec[:, 0, iy, :]
        b = spec[:, 0, jy, :]
        h = 0.5 * (a + np.conj(b))
        spec[:, 0, iy, :] = h
        spec[:, 0, jy, :] = np.conj(h)
    spec[:, 0, 0, :] = spec[:, 0, 0, :].real
    if nyc % 2 == 0 and nyc >= 2:
        spec[:, :, nyc // 2, :] = 0.0    # the unpaired ky = -nyc/2 row
    return spec


def initial_state(cfg, seed=None, noise_u=None, noise_c=None, time=0.0):
    """Stokes layer + windowed Gaussian noise; c = 1 + noise.
    Returns uvw (3, mx, nyc, nz), p (1, mx, nyc, nz), c (1, mx, nyc, nz)."""
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    seed = cfg.get("seed", 0) if seed is None else seed
    noise_u = cfg.get("noise_u", 0.0) if noise_u is None else noise_u
    noise_c = cfg.get("noise_c", 0.0) if noise_c is None else noise_c
    z = cgl_grid(nz)
    top = "freeslip" if tuple(cfg["vt_u1"][:4]) == NOSLIP_FREESLIP[:4] else "noslip"
    win_uv = (1.0 + z) if top == "freeslip" else (1.0 - z * z)
    win_w = 1.0 - z * z
    phys = np.zeros((3, nz, ny, nx))
    if noise_u:
        for d in range(3):
            win = win_w if d == 2 else win_uv
            phys[d] = noise_u * _normal(seed, d, (nz, ny, nx)) * win[:, None, None]
        phys[:, 0] = 0.0
        phys[2, -1] = 0.0
    uvw = forward_truncated(phys)
    omega = cfg["force_amplitude"][1]
    uvw[0, 0, 0, :] += stokes_profile(cfg["reynolds"], z, time, omega, top)
    uvw[0, 0, 0, 0] = 0.0
    p = np.zeros((1,) + uvw.shape[1:], dtype=np.complex128)
    cphys = np.ones((1, nz, ny, nx))
    if noise_c:
        cphys[0] += noise_c * _normal(seed, 3, (nz, ny, nx))
    c = forward_truncated(cphys)
    c[0, 0, 0, :] = c[0, 0, 0, :].real
    return uvw, p, c


def to_flat(a):
    """(nv, mx, nyc, nz) -> flat RealCplxCube order (z fastest)."""
    return np.ascontiguousarray(a).reshape(-1)


def from_flat(a, nv, mx, nyc, nz):
    return np.asarray(a).reshape(nv, mx, nyc, nz)
