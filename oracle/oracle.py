"""ctypes binding of the FP64 CPU restatement (oracle/turbid_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``, always as the checker or the
baseline, never as the measured or shipped path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libturbid_oracle.so")

CONFIG_FIELDS = [
    ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
    ("nphase", ctypes.c_int),
    ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
    ("slope", ctypes.c_double), ("slope_direction", ctypes.c_int),
    ("mapping", ctypes.c_int), ("mapping_coefficient", ctypes.c_double),
    ("dt", ctypes.c_double), ("output_on_step", ctypes.c_int),
    ("force_type", ctypes.c_int), ("force_direction", ctypes.c_int * 2),
    ("force_amplitude", ctypes.c_double * 4),
    ("reynolds", ctypes.c_double), ("froude", ctypes.c_double),
    ("averaged_concentration", ctypes.c_double), ("specific_gravity", ctypes.c_double),
    ("affected_by_scalar_phase", ctypes.c_int),
    ("vt_u1", ctypes.c_double * 5), ("vt_u2", ctypes.c_double * 5),
    ("vt_u3", ctypes.c_double * 5), ("vt_p", ctypes.c_double * 5),
    ("schmidt", ctypes.c_double), ("settling_velocity", ctypes.c_double),
    ("erosion_rate", ctypes.c_double), ("critical_shear_stress", ctypes.c_double),
    ("scalar_bc", ctypes.c_int), ("precision", ctypes.c_int),
]


class TboConfig(ctypes.Structure):
    _fields_ = CONFIG_FIELDS


def fill_config(struct_type, cfg):
    """Fill a ctypes config struct from a dict of field -> value."""
    s = struct_type()
    for name, ctype in struct_type._fields_:
        if name not in cfg:
            continue
        v = cfg[name]
        if hasattr(ctype, "_length_"):
            arr = getattr(s, name)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(s, name, v)
    return s


def build(force=False):
    """Compile the restatement with its Makefile (g++, OpenMP).  make runs on
    every call (it is incremental), so an edited source is never tested
    through a stale library."""
    cmd = ["make", "-s", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        # rebuilt when missing or older than any of its sources, so an edit of
        # the restatement is never tested through a stale library
        srcs = [os.path.join(_HERE, f) for f in os.listdir(_HERE) if f.endswith((".cpp", ".h"))]
        if not os.path.exists(_LIB_PATH) or any(os.path.getmtime(f) > os.path.getmtime(_LIB_PATH) for f in srcs):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.tbo_create.restype = ctypes.c_void_p
        L.tbo_create.argtypes = [ctypes.POINTER(TboConfig), ctypes.c_int]
        L.tbo_destroy.argtypes = [ctypes.c_void_p]
        L.tbo_last_error.restype = ctypes.c_char_p
        L.tbo_set_state.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.tbo_get_state.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.tbo_substage.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.tbo_step.argtypes = [ctypes.c_void_p]
        L.tbo_set_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.tbo_stepper.argtypes = [ctypes.c_void_p, ip, dp, dp, ip]
        L.tbo_features.argtypes = [ctypes.c_void_p, dp]
        L.tbo_get_physical_velocity.argtypes = [ctypes.c_void_p, dp]
        L.tbo_nonlinear.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.tbo_get_history.argtypes = [ctypes.c_void_p, dp, dp]
        L.tbo_get_operators.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.tbo_get_decomposition.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp, dp]
        L.tbo_dft_forward.argtypes = [ctypes.c_int] * 4 + [dp, dp]
        L.tbo_dft_backward.argtypes = [ctypes.c_int] * 4 + [dp, dp]
        L.tbo_chebyshev.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp, dp]
        L.tbo_eig.argtypes = [ctypes.c_int, dp, dp, dp]
        L.tbo_positivity_filter.argtypes = [ctypes.c_int] * 3 + [dp]
        L.tbo_helmholtz.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, dp, ctypes.c_int, dp, dp]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def dims(nx, ny, nz):
    nxc, nyc = nx // 3 * 2, ny // 3 * 2
    mx = nxc // 2 + 1
    return nxc, nyc, mx


def dft_forward(phys):
    """phys: (nv, nz, ny, nx) float64 -> spectral (nv, mx, nyc, nz) complex128."""
    nv, nz, ny, nx = phys.shape
    _, nyc, mx = dims(nx, ny, nz)
    src = np.ascontiguousarray(phys, dtype=np.float64)
    out = np.zeros((nv, mx, nyc, nz), dtype=np.complex128)
    rc = lib().tbo_dft_forward(nx, ny, nz, nv, _p(src), _p(out.view(np.float64)))
    assert rc == 0
    return out


def dft_backward(spec, nx, ny):
    """spec: (nv, mx, nyc, nz) complex128 -> phys (nv, nz, ny, nx)."""
    nv, mx, nyc, nz = spec.shape
    src = np.ascontiguousarray(spec, dtype=np.complex128)
    out = np.zeros((nv, nz, ny, nx), dtype=np.float64)
    rc = lib().tbo_dft_backward(nx, ny, nz, nv, _p(src.view(np.float64)), _p(out))
    assert rc == 0
    return out


def chebyshev(nz, mapping=0, coef=0.0):
    g = np.zeros(nz)
    d1 = np.zeros((nz, nz), order="F")
    d2 = np.zeros((nz, nz), order="F")
    lib().tbo_chebyshev(nz, mapping, coef, _p(g), _p(d1), _p(d2))
    return g, d1, d2


def eig(a):
    a = np.asfortranarray(a, dtype=np.float64)
    n = a.shape[0]
    lam = np.zeros(n)
    v = np.zeros((n, n), order="F")
    rc = lib().tbo_eig(n, _p(a), _p(lam), _p(v))
    if rc != 0:
        raise RuntimeError("oracle eig failed (%d)" % rc)
    return lam, v


def positivity_filter(phys):
    """The scalar's positivity filter on physical planes phys[nz, ny, nx]
    (returns a filtered copy)."""
    out = np.ascontiguousarray(phys, dtype=np.float64).copy()
    nz, ny, nx = out.shape
    rc = lib().tbo_positivity_filter(nx, ny, nz, _p(out))
    assert rc == 0
    return out


def wavenumbers(cfg):
    """kx[mx], ky[nyc] with the reference's +ky sign at iy = nyc/2
    (src/stepper_space.cc:181-213)."""
    _, nyc, mx = dims(cfg["nx"], cfg["ny"], cfg["nz"])
    kx = 2 * np.pi / cfg["lx"] * np.arange(mx)
    ky1 = 2 * np.pi / cfg["ly"]
    ky = np.array([ky1 * iy if iy <= nyc // 2 else -ky1 * (nyc - iy) for iy in range(nyc)])
    return kx, ky


def velocity_gradient(uvw, cfg):
    """Physical G[i, j] = d u_i / d x_j, shape (3, 3, nz, ny, nx), from the
    spectral velocity (3, mx, nyc, nz): SpaceStepper::Gradient (i kx, i ky,
    D1 along z; src/stepper_space.cc:373-389) then DftBackward
    (postprocessor_channel_flat.cc:203-208)."""
    kx, ky = wavenumbers(cfg)
    _, d1, _ = chebyshev(cfg["nz"], cfg.get("mapping", 0), cfg.get("mapping_coefficient", 0.0))
    d1 = np.asarray(d1)
    g = np.empty((9,) + uvw.shape[1:], dtype=np.complex128)
    for i in range(3):
        u = uvw[i]
        g[3 * i + 0] = 1j * kx[:, None, None] * u
        g[3 * i + 1] = 1j * ky[None, :, None] * u
        g[3 * i + 2] = u @ d1.T
    phys = dft_backward(g, cfg["nx"], cfg["ny"])
    return phys.reshape(3, 3, *phys.shape[1:])


def swirling_strength_tensors(G):
    """max |Im lambda| over the eigenvalues of each 3x3 tensor G[..., 3, 3]
    (arma::eig_gen, i.e. LAPACK geev, then max(unique(abs(imag(.)))),
    postprocessor_channel_flat.cc:229-232)."""
    lam = np.linalg.eigvals(G)
    return np.abs(lam.imag).max(axis=-1)


def swirling_strength(uvw, cfg):
    """Swirling strength on the physical grid, (nz, ny, nx)."""
    G = velocity_gradient(uvw, cfg)
    return swirling_strength_tensors(np.moveaxis(G, (0, 1), (-2, -1)))


class Oracle:
    """One restated solver instance (carrier + optional scalar phase)."""

    def __init__(self, cfg, nthreads=0):
        self.cfg = dict(cfg)
        s = fill_config(TboConfig, cfg)
        self._h = lib().tbo_create(ctypes.byref(s), nthreads)
        if not self._h:
            raise RuntimeError(lib().tbo_last_error().decode())
        self.nx, self.ny, self.nz = cfg["nx"], cfg["ny"], cfg["nz"]
        self.nxc, self.nyc, self.mx = dims(self.nx, self.ny, self.nz)
        self.P = self.nyc * self.mx
        self.S = self.nz * self.P

    def close(self):
        if self._h:
            lib().tbo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, uvw, p, c=None):
        u = np.ascontiguousarray(uvw, dtype=np.complex128).reshape(-1)
        pp = np.ascontiguousarray(p, dtype=np.complex128).reshape(-1)
        cc = None if c is None else np.ascontiguousarray(c, dtype=np.complex128).reshape(-1)
        assert u.size == 3 * self.S and pp.size == self.S
        rc = lib().tbo_set_state(self._h, _p(u.view(np.float64)), _p(pp.view(np.float64)),
                                 None if cc is None else _p(cc.view(np.float64)))
        if rc != 0:
            raise RuntimeError(lib().tbo_last_error().decode())

    def get_state(self):
        u = np.zeros(3 * self.S, dtype=np.complex128)
        p = np.zeros(self.S, dtype=np.complex128)
        c = np.zeros(self.S, dtype=np.complex128) if self.cfg.get("nphase", 1) > 1 else None
        lib().tbo_get_state(self._h, _p(u.view(np.float64)), _p(p.view(np.float64)),
                            None if c is None else _p(c.view(np.float64)))
        return u, p, c

    def substage(self, stage, time, alternation):
        rc = lib().tbo_substage(self._h, stage, time, int(alternation))
        if rc != 0:
            raise RuntimeError(lib().tbo_last_error().decode())

    def step(self):
        rc = lib().tbo_step(self._h)
        if rc < 0:
            raise RuntimeError(lib().tbo_last_error().decode())
        return rc

    def set_dt(self, dt):
        """Time mode: the next substages use dt (FlatCarrierPhase::Update)."""
        lib().tbo_set_dt(self._h, float(dt))

    def stepper(self):
        st, al = ctypes.c_int(), ctypes.c_int()
        t, cfl = ctypes.c_double(), ctypes.c_double()
        lib().tbo_stepper(self._h, ctypes.byref(st), ctypes.byref(t), ctypes.byref(cfl), ctypes.byref(al))
        return st.value, t.value, cfl.value, al.value

    def features(self):
        f = np.zeros(50)
        lib().tbo_features(self._h, _p(f))
        return f

    def physical_velocity(self):
        out = np.zeros(3 * self.nx * self.ny * self.nz)
        lib().tbo_get_physical_velocity(self._h, _p(out))
        return out.reshape(3, self.nz, self.ny, self.nx)

    def nonlinear(self, alternation):
        """Spectral nonlinear term u.grad u of the current state, (3, mx, nyc,
        nz): convective form (alternation 0) or divergence form (1)."""
        out = np.zeros(3 * self.S, dtype=np.complex128)
        rc = lib().tbo_nonlinear(self._h, int(alternation), _p(out.view(np.float64)))
        if rc != 0:
            raise RuntimeError(lib().tbo_last_error().decode())
        return out.reshape(3, self.mx, self.nyc, self.nz)

    def history(self):
        """Explicit-term histories (H of u, v, w; H of c or None)."""
        h = np.zeros(3 * self.S, dtype=np.complex128)
        hc = np.zeros(self.S, dtype=np.complex128) if self.cfg.get("nphase", 1) > 1 else None
        lib().tbo_get_history(self._h, _p(h.view(np.float64)), None if hc is None else _p(hc.view(np.float64)))
        return h, hc

    def operators(self):
        n = self.nz
        g = np.zeros(n)
        d1 = np.zeros((n, n), order="F")
        d2 = np.zeros((n, n), order="F")
        lib().tbo_get_operators(self._h, _p(g), _p(d1), _p(d2))
        return g, d1, d2

    def decomposition(self, which):
        n = self.nz
        V = np.zeros((n, n), order="F")
        Vi = np.zeros((n, n), order="F")
        lam = np.zeros(n)
        fix = np.zeros(4 * n)
        rc = lib().tbo_get_decomposition(self._h, which, _p(V), _p(Vi), _p(lam), _p(fix))
        if rc != 0:
            raise RuntimeError(lib().tbo_last_error().decode())
        return V, Vi, lam, fix.reshape(4, n)

    def helmholtz(self, which, stage, mode, x, k2, col0_is_mean=True):
        """x: (ncol, nz) complex; k2: (ncol,) = -(kx^2) + -(ky^2)."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        ncol = x.shape[0]
        k2 = np.ascontiguousarray(k2, dtype=np.float64)
        y = np.zeros_like(x)
        rc = lib().tbo_helmholtz(self._h, which, stage, mode, ncol, _p(k2), int(col0_is_mean),
                                 _p(x.view(np.float64)), _p(y.view(np.float64)))
        if rc != 0:
            raise RuntimeError(lib().tbo_last_error().decode())
        return y
