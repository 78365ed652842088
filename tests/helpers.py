"""Shared helpers for the parity tests."""
import math

import numpy as np

from turbid_amd import workloads as W


CAPTURE = None   # pytest's capture manager (set by conftest.py)


def say(msg):
    """Print a progress line to the real stdout even while pytest captures."""
    if CAPTURE is None:
        print(msg, flush=True)
        return
    with CAPTURE.global_and_fixture_disabled():
        print(msg, flush=True)


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


def manufactured_flow(cfg, amp=0.2, phase=(0.3, 0.7)):
    """A divergence-free, band-limited velocity with closed-form u.grad(u).

    u = curl(A) with A = (a1 h, a2 h, 0):
        u = -a2 h',  v = a1 h',  w = (d_x a2 - d_y a1) h,
    a1 = amp cos(al x + be y + f1), a2 = amp sin(al x) cos(be y + f2),
    al = 2 pi / lx, be = 2 pi / ly, h = (1 - z^2)^2 (1 + z/2): every component
    and h' vanish at both walls, the field is exactly divergence-free on the
    collocation grid (D1 is exact on polynomials of degree < nz), and its
    products stay inside the retained Fourier modes (|kx|, |ky| <= 2).

    Returns (phys (3, nz, ny, nx), nonlinear (3, nz, ny, nx)) on the grid
    x = lx/nx*i, y = ly/ny*j, z = CGL (bottom first), with
    nonlinear_i = u_j d_j u_i from the analytic derivatives."""
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    al, be = 2 * math.pi / cfg["lx"], 2 * math.pi / cfg["ly"]
    f1, f2 = phase
    x = (cfg["lx"] / nx * np.arange(nx))[None, None, :]
    y = (cfg["ly"] / ny * np.arange(ny))[None, :, None]
    z = W.cgl_grid(nz)
    hp = np.polynomial.Polynomial([1.0, 0.5]) * np.polynomial.Polynomial([1.0, 0.0, -1.0]) ** 2
    h, h1, h2 = (q(z)[:, None, None] for q in (hp, hp.deriv(1), hp.deriv(2)))
    t1 = al * x + be * y + f1
    t2 = be * y + f2
    a1, a1x, a1y = amp * np.cos(t1), -amp * al * np.sin(t1), -amp * be * np.sin(t1)
    a1xy, a1yy = -amp * al * be * np.cos(t1), -amp * be * be * np.cos(t1)
    a2 = amp * np.sin(al * x) * np.cos(t2)
    a2x = amp * al * np.cos(al * x) * np.cos(t2)
    a2y = -amp * be * np.sin(al * x) * np.sin(t2)
    a2xx = -amp * al * al * np.sin(al * x) * np.cos(t2)
    a2xy = -amp * al * be * np.cos(al * x) * np.sin(t2)
    full = np.broadcast_to(np.zeros((nz, ny, nx)), (nz, ny, nx))
    u, ux, uy, uz = -a2 * h1, -a2x * h1, -a2y * h1, -a2 * h2
    v, vx, vy, vz = a1 * h1, a1x * h1, a1y * h1, a1 * h2
    g = a2x - a1y
    w, wx, wy, wz = g * h, (a2xx - a1xy) * h, (a2xy - a1yy) * h, g * h1
    vel = [u + full, v + full, w + full]
    grads = [(ux, uy, uz), (vx, vy, vz), (wx, wy, wz)]
    nl = np.stack([vel[0] * gx + vel[1] * gy + vel[2] * gz for gx, gy, gz in grads])
    return np.stack(vel), nl


def oracle_steps(cfg, state, nsteps, nthreads=0, progress=None, stepper=None):
    """The oracle through FlatChannelSolver::Execute's loop (TimeStepper::Update
    with the previous step's CFL feature, three substages, toggle) for nsteps
    steps from `state`; returns (uvw, p, c).  progress(n) is called after
    every step when given.  stepper: a runtime.Stepper to use instead of a
    fresh step-mode one (a TimeModeStepper also sets the oracle's dt each
    step, as Phase::Update does)."""
    from oracle import oracle as O
    from turbid_amd import runtime as RT
    o = O.Oracle(cfg, nthreads=nthreads)
    o.set_state(*state)
    st = RT.Stepper(cfg["dt"]) if stepper is None else stepper
    timed = isinstance(st, RT.TimeModeStepper)
    feat = np.zeros(50)
    for n in range(nsteps):
        if not st.forward:
            break
        st.update(feat[19])
        if timed:
            o.set_dt(st.dt)
        for s in range(3):
            o.substage(s, st.time, st.alternation)
        st.toggle()
        feat = o.features()
        if progress:
            progress(n + 1)
    out = o.get_state()
    o.close()
    return out


def field_errors(got, ref):
    """Relative L2 per field (u, v, w, p, c) of two (uvw, p, c) states."""
    ug, pg, cg = got
    uo, po, co = ref
    ug, uo = np.asarray(ug).reshape(3, -1), np.asarray(uo).reshape(3, -1)
    errs = {"u": rel_l2(ug[0], uo[0]), "v": rel_l2(ug[1], uo[1]), "w": rel_l2(ug[2], uo[2]), "p": rel_l2(pg, po)}
    if co is not None:
        errs["c"] = rel_l2(cg, co)
    return errs
