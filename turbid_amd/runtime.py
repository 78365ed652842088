"""ctypes binding of libturbid_amd.so (the C ABI in include/turbid_amd.h).

The library is the only compute path: if it is missing, or no GPU is
visible, every entry point raises.  Nothing here falls back to a CPU
implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libturbid_amd.so")

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

# symbols declared in include/turbid_amd.h
EXPORTS = [
    "tbd_create", "tbd_destroy", "tbd_last_error", "tbd_upload_state", "tbd_download_state",
    "tbd_download_physical_velocity", "tbd_substage", "tbd_step", "tbd_set_dt", "tbd_features",
    "tbd_sync", "tbd_profile_enable", "tbd_profile_read", "tbd_profile_reset", "tbd_get_operators",
    "tbd_get_factorisation", "tbd_set_operators", "tbd_set_factorisation", "tbd_autotune", "tbd_get_variant",
    "tbd_set_variant", "tbd_dft_forward", "tbd_dft_forward_noalloc", "tbd_dft_backward", "tbd_zcontract", "tbd_helmholtz",
    "tbd_project", "tbd_swirling_strength", "tbd_stats_accumulate", "tbd_stats_read", "tbd_stats_reset", "tbd_sno_create", "tbd_sno_destroy", "tbd_sno_param_count", "tbd_sno_set_params",
    "tbd_sno_get_params", "tbd_sno_get_grads", "tbd_sno_forward", "tbd_sno_loss_grad", "tbd_sno_train_step",
    "tbd_bf16_gemm", "tbd_ensemble_create", "tbd_ensemble_step", "tbd_ensemble_launches", "tbd_ensemble_sync", "tbd_ensemble_features",
    "tbd_ensemble_destroy",
]

# profiler classes (one per kernel): name -> (index, work kind)
KCLASSES = {
    "ypass_bwd64": (0, "bytes"), "xpass_bwd64": (1, "bytes"), "xpass_fwd64": (2, "bytes"), "ypass_fwd64": (3, "bytes"),
    "ypass_bwd32": (4, "bytes"), "xpass_bwd32": (5, "bytes"), "xpass_fwd32": (6, "bytes"), "ypass_fwd32": (7, "bytes"),
    "xfused64": (8, "bytes"), "xfused32": (9, "bytes"), "zgemm64": (10, "flops"), "zgemm32": (11, "flops"),
    "helmholtz": (12, "flops"), "products64": (13, "bytes"), "products32": (14, "bytes"), "other": (15, None),
    "helmholtz_split": (16, "flops"), "rhs_helmholtz": (17, "flops"), "rhs_helmholtz_split": (18, "flops"),
    "pressure_chain": (19, "flops"),
}
KCLASS = {k: v[0] for k, v in KCLASSES.items()}


class TbdConfig(ctypes.Structure):
    _fields_ = CONFIG_FIELDS


def make_config(cfg):
    s = TbdConfig()
    for name, ctype in TbdConfig._fields_:
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


class TurbidError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TurbidError("libturbid_amd.so not built: run `python -m turbid_amd.build` (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.tbd_create.argtypes = [ctypes.POINTER(TbdConfig), ctypes.POINTER(vp)]
        L.tbd_destroy.argtypes = [vp]
        L.tbd_destroy.restype = None
        L.tbd_last_error.restype = ctypes.c_char_p
        L.tbd_upload_state.argtypes = [vp, dp, dp, dp]
        L.tbd_download_state.argtypes = [vp, dp, dp, dp]
        L.tbd_download_physical_velocity.argtypes = [vp, dp]
        L.tbd_substage.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.tbd_step.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.tbd_set_dt.argtypes = [vp, ctypes.c_double]
        L.tbd_features.argtypes = [vp, dp]
        L.tbd_sync.argtypes = [vp]
        L.tbd_profile_enable.argtypes = [vp, ctypes.c_int]
        L.tbd_profile_read.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), dp, dp]
        L.tbd_profile_reset.argtypes = [vp]
        L.tbd_get_operators.argtypes = [vp, dp, dp, dp]
        L.tbd_get_factorisation.argtypes = [vp, ctypes.c_int, dp, dp, dp, dp]
        L.tbd_set_operators.argtypes = [vp, dp, dp]
        L.tbd_set_factorisation.argtypes = [vp, ctypes.c_int, dp, dp, dp, dp]
        L.tbd_autotune.argtypes = [vp, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_int)]
        L.tbd_get_variant.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.tbd_set_variant.argtypes = [vp, ctypes.c_int]
        L.tbd_dft_forward.argtypes = [ctypes.c_int] * 5 + [dp, dp]
        L.tbd_dft_backward.argtypes = [ctypes.c_int] * 5 + [dp, dp]
        L.tbd_dft_forward_noalloc.argtypes = [ctypes.c_int] * 5 + [vp, vp]
        L.tbd_zcontract.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int, dp, dp, dp]
        L.tbd_helmholtz.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.tbd_project.argtypes = [vp]
        L.tbd_stats_accumulate.argtypes = [vp, ctypes.c_int]
        L.tbd_swirling_strength.argtypes = [vp, dp]
        L.tbd_stats_read.argtypes = [vp, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_longlong)]
        L.tbd_stats_reset.argtypes = [vp]
        fp = ctypes.POINTER(ctypes.c_float)
        L.tbd_sno_create.argtypes = [ctypes.c_int] * 8 + [ctypes.POINTER(vp)]
        L.tbd_sno_destroy.argtypes = [vp]
        L.tbd_sno_destroy.restype = None
        L.tbd_sno_param_count.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
        L.tbd_sno_set_params.argtypes = [vp, fp]
        L.tbd_sno_get_params.argtypes = [vp, fp]
        L.tbd_sno_get_grads.argtypes = [vp, fp]
        L.tbd_sno_forward.argtypes = [vp, fp, fp, ctypes.c_int]
        L.tbd_sno_loss_grad.argtypes = [vp, fp, fp, fp, ctypes.c_int, dp]
        L.tbd_sno_train_step.argtypes = [vp, fp, fp, fp, ctypes.c_int, ctypes.c_float, dp]
        L.tbd_bf16_gemm.argtypes = [ctypes.c_int] * 3 + [fp, fp, fp]
        L.tbd_ensemble_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
        L.tbd_ensemble_step.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int)]
        L.tbd_ensemble_launches.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.tbd_ensemble_sync.argtypes = [vp]
        L.tbd_ensemble_features.argtypes = [vp, dp]
        L.tbd_ensemble_destroy.argtypes = [vp]
        L.tbd_ensemble_destroy.restype = None
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc):
    if rc != 0:
        raise TurbidError(lib().tbd_last_error().decode())


def dims(nx, ny, nz):
    nxc, nyc = nx // 3 * 2, ny // 3 * 2
    return nxc, nyc, nxc // 2 + 1


def dft_forward(phys, fp32=False):
    """phys (nv, nz, ny, nx) -> spectral (nv, mx, nyc, nz) complex128."""
    nv, nz, ny, nx = phys.shape
    _, nyc, mx = dims(nx, ny, nz)
    src = np.ascontiguousarray(phys, dtype=np.float64)
    out = np.zeros((nv, mx, nyc, nz), dtype=np.complex128)
    _check(lib().tbd_dft_forward(nx, ny, nz, nv, int(fp32), _p(src), _p(out.view(np.float64))))
    return out


def dft_forward_noalloc(nx, ny, nz, nv, phys_ptr, spec_ptr, fp32=False):
    """tbd_dft_forward_noalloc on device buffers: phys_ptr holds nv*nz*ny*nx
    doubles, spec_ptr receives nv*mx*nyc*nz complex doubles (device
    addresses the caller allocated)."""
    _check(lib().tbd_dft_forward_noalloc(nx, ny, nz, nv, int(fp32), ctypes.c_void_p(phys_ptr),
                                         ctypes.c_void_p(spec_ptr)))


def dft_backward(spec, nx, ny, fp32=False):
    nv, mx, nyc, nz = spec.shape
    src = np.ascontiguousarray(spec, dtype=np.complex128)
    out = np.zeros((nv, nz, ny, nx))
    _check(lib().tbd_dft_backward(nx, ny, nz, nv, int(fp32), _p(src.view(np.float64)), _p(out)))
    return out


def zcontract(m, x, fp32=False):
    """y[:, c] = m @ x[:, c]; m (nz, nz) real, x (ncol, nz) complex (column-contiguous)."""
    nz = m.shape[0]
    mm = np.asfortranarray(m, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    ncol = xx.shape[0]
    y = np.zeros_like(xx)
    _check(lib().tbd_zcontract(nz, ncol, int(fp32), _p(mm), _p(xx.view(np.float64)), _p(y.view(np.float64))))
    return y


class Solver:
    """One DNS context on the GPU (carrier + optional sediment phase)."""

    def __init__(self, cfg):
        self.cfg = dict(cfg)
        s = make_config(cfg)
        h = ctypes.c_void_p()
        _check(lib().tbd_create(ctypes.byref(s), ctypes.byref(h)))
        self._h = h
        self.nx, self.ny, self.nz = cfg["nx"], cfg["ny"], cfg["nz"]
        self.nxc, self.nyc, self.mx = dims(self.nx, self.ny, self.nz)
        self.P = self.nyc * self.mx
        self.S = self.nz * self.P
        self.nphase = cfg.get("nphase", 1)

    def close(self):
        if self._h:
            lib().tbd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, uvw, p, c=None):
        u = np.ascontiguousarray(uvw, dtype=np.complex128).reshape(-1)
        pp = np.ascontiguousarray(p, dtype=np.complex128).reshape(-1)
        cc = None if c is None else np.ascontiguousarray(c, dtype=np.complex128).reshape(-1)
        _check(lib().tbd_upload_state(self._h, _p(u.view(np.float64)), _p(pp.view(np.float64)),
                                      None if cc is None else _p(cc.view(np.float64))))

    def download(self):
        u = np.zeros(3 * self.S, dtype=np.complex128)
        p = np.zeros(self.S, dtype=np.complex128)
        c = np.zeros(self.S, dtype=np.complex128) if self.nphase > 1 else None
        _check(lib().tbd_download_state(self._h, _p(u.view(np.float64)), _p(p.view(np.float64)),
                                        None if c is None else _p(c.view(np.float64))))
        return u, p, c

    def state(self):
        """Downloaded state as uvw (3, mx, nyc, nz), p (1, ...), c (1, ...) or None."""
        u, p, c = self.download()
        shp = (self.mx, self.nyc, self.nz)
        return (u.reshape((3,) + shp), p.reshape((1,) + shp), None if c is None else c.reshape((1,) + shp))

    def physical_velocity(self):
        out = np.zeros(3 * self.nx * self.ny * self.nz)
        _check(lib().tbd_download_physical_velocity(self._h, _p(out)))
        return out.reshape(3, self.nz, self.ny, self.nx)

    def substage(self, stage, time, alternation):
        _check(lib().tbd_substage(self._h, stage, time, int(alternation)))

    def step(self, time, alternation):
        _check(lib().tbd_step(self._h, time, int(alternation)))

    def set_dt(self, dt):
        _check(lib().tbd_set_dt(self._h, dt))

    def swirling_strength(self):
        """Swirling strength of the current velocity, (nz, ny, nx) physical."""
        out = np.zeros(self.nz * self.ny * self.nx)
        _check(lib().tbd_swirling_strength(self._h, _p(out)))
        return out.reshape(self.nz, self.ny, self.nx)

    def stats_accumulate(self, phase_bin):
        _check(lib().tbd_stats_accumulate(self._h, int(phase_bin)))

    def stats_read(self, phase_bin):
        """(8, nz) phase-bin mean: U, V, W, C profiles and <u'^2>, <v'^2>, <w'^2>, <c'^2>."""
        out = np.zeros(8 * self.nz)
        n = ctypes.c_longlong()
        _check(lib().tbd_stats_read(self._h, int(phase_bin), _p(out), ctypes.byref(n)))
        return out.reshape(8, self.nz) / max(1, n.value), n.value

    def stats_reset(self):
        _check(lib().tbd_stats_reset(self._h))

    def project(self):
        """Make the current velocity divergence-free (proposed initial states)."""
        _check(lib().tbd_project(self._h))

    def features(self):
        f = np.zeros(50)
        _check(lib().tbd_features(self._h, _p(f)))
        return f

    def sync(self):
        _check(lib().tbd_sync(self._h))

    def profile(self, enable=True):
        _check(lib().tbd_profile_enable(self._h, int(enable)))

    def profile_reset(self):
        _check(lib().tbd_profile_reset(self._h))

    def profile_read(self, kclass):
        """(launches, total_ms, work) of a kernel class since the last reset; work in
        bytes or flops (KCLASSES)."""
        n = ctypes.c_longlong()
        ms = ctypes.c_double()
        un = ctypes.c_double()
        _check(lib().tbd_profile_read(self._h, KCLASS.get(kclass, kclass), ctypes.byref(n), ctypes.byref(ms),
                                      ctypes.byref(un)))
        return n.value, ms.value, un.value

    def operators(self):
        n = self.nz
        g = np.zeros(n)
        d1 = np.zeros((n, n), order="F")
        d2 = np.zeros((n, n), order="F")
        _check(lib().tbd_get_operators(self._h, _p(g), _p(d1), _p(d2)))
        return g, d1, d2

    def factorisation(self, which):
        n = self.nz
        V = np.zeros((n, n), order="F")
        Vi = np.zeros((n, n), order="F")
        lam = np.zeros(n)
        fix = np.zeros(4 * n)
        _check(lib().tbd_get_factorisation(self._h, which, _p(V), _p(Vi), _p(lam), _p(fix)))
        return V, Vi, lam, fix.reshape(4, n)

    def set_operators(self, d1, d2):
        """Replace D1, D2 (nz x nz) and rebuild every factorisation from them."""
        n = self.nz
        a = np.asfortranarray(d1, dtype=np.float64)
        b = np.asfortranarray(d2, dtype=np.float64)
        assert a.shape == (n, n) and b.shape == (n, n)
        _check(lib().tbd_set_operators(self._h, _p(a), _p(b)))

    def set_factorisation(self, which, V, Vi, lam, fix):
        """Replace one factorisation (layout of factorisation(which))."""
        n = self.nz
        V = np.asfortranarray(V, dtype=np.float64)
        Vi = np.asfortranarray(Vi, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1)
        fix = np.ascontiguousarray(fix, dtype=np.float64).reshape(-1)
        assert V.shape == (n, n) and Vi.shape == (n, n) and lam.size == n and fix.size == 4 * n
        _check(lib().tbd_set_factorisation(self._h, which, _p(V), _p(Vi), _p(lam), _p(fix)))

    # kernel variants: bit 1 fused x transforms, 2 fused rhs + solve and
    # pressure chain, 4 scalar on a second stream, 8/16/32/64 y backward /
    # y forward / x backward / x forward transform kernels at twice their
    # workgroup size
    VARIANT_BITS = {"fused": 1, "rhs": 2, "overlap": 4, "wide_yb": 8, "wide_yf": 16, "wide_xb": 32,
                    "wide_xf": 64, "swap_yb": 128, "swap_yf": 256, "swap_xb": 512, "swap_xf": 1024,
                    "wide_zgemm": 2048}
    WIDE_FFT = 8 | 16 | 32 | 64
    SWAP_FFT = 128 | 256 | 512 | 1024
    N_VARIANTS = 4096

    def autotune(self, steps=5):
        """Times the kernel variants on a scratch copy and keeps the fastest;
        returns (bits kept, ms per step of each variant, NaN where not tried)."""
        ms = np.zeros(self.N_VARIANTS)
        b = ctypes.c_int()
        _check(lib().tbd_autotune(self._h, int(steps), _p(ms), ctypes.byref(b)))
        return b.value, ms

    def variant(self):
        b = ctypes.c_int()
        _check(lib().tbd_get_variant(self._h, ctypes.byref(b)))
        return b.value

    def set_variant(self, bits):
        _check(lib().tbd_set_variant(self._h, int(bits)))

    def helmholtz(self, which, stage, mode, x):
        xx = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
        assert xx.size == self.S
        y = np.zeros_like(xx)
        _check(lib().tbd_helmholtz(self._h, which, stage, mode, _p(xx.view(np.float64)), _p(y.view(np.float64))))
        return y


class GroupedEnsemble:
    """Solvers advanced together through grouped launches (tbd_ensemble_*):
    one launch per kernel and launch shape at each position of the step over
    all members, captured in one HIP graph.  The solvers stay usable one by
    one (upload, download, features) and must outlive this object."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        hs = (ctypes.c_void_p * len(self.solvers))(*[s._h for s in self.solvers])
        h = ctypes.c_void_p()
        _check(lib().tbd_ensemble_create(hs, len(self.solvers), ctypes.byref(h)))
        self._h = h

    def step(self, times, alternations):
        n = len(self.solvers)
        t = np.ascontiguousarray(times, dtype=np.float64)
        a = (ctypes.c_int * n)(*[int(x) for x in alternations])
        assert t.shape == (n,)
        _check(lib().tbd_ensemble_step(self._h, _p(t), a))

    def launches(self):
        """(grouped launches, launches of the members one by one) per step."""
        g, m = ctypes.c_int(), ctypes.c_int()
        _check(lib().tbd_ensemble_launches(self._h, ctypes.byref(g), ctypes.byref(m)))
        return g.value, m.value

    def sync(self):
        _check(lib().tbd_ensemble_sync(self._h))

    def features(self):
        """(members, 50) features after the last step (one copy for all)."""
        out = np.zeros((len(self.solvers), 50))
        _check(lib().tbd_ensemble_features(self._h, _p(out)))
        return out

    def close(self):
        if self._h:
            lib().tbd_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stepper:
    """TimeStepper mirror in step mode (reference src/stepper_time.cc:72-113,
    215-246): step counter, time += dt, CFL guard on the previous step's
    feature(19), alternation flag toggled after every step
    (src/stepper_time.h:75-84)."""

    def __init__(self, dt, max_step=10**9, max_cfl=1e30, start_step=0):
        self.dt = dt
        self.step = max(0, start_step)
        self.time = self.step * dt
        self.max_step = max_step
        self.max_cfl = max_cfl
        self.forward = self.step <= max_step
        self.alternation = False
        self.cfl = 0.0

    def update(self, maxulen):
        if not self.forward:
            return
        self.step += 1
        self.time = self.time + self.dt
        self.forward = self.step < self.max_step
        self.cfl = maxulen * self.dt
        if self.cfl > self.max_cfl:
            self.forward = False

    def toggle(self):
        self.alternation = not self.alternation


class TimeModeStepper(Stepper):
    """TimeStepper mirror in time mode (output_based_on_step = 0; reference
    src/stepper_time.cc:120-208 Update2 and :215-246 Configure): dt adapts
    to the previous step's CFL (times top**k when above top*maxcfl, times
    1.05 when below bot*maxcfl, (top, bot) = cfl_adjust_scale) and is cut
    so that the run lands on every output time; the run stops at
    max_marching_time, or when the CFL exceeds 1.1*maxcfl or dt falls below
    initial_time_step/1000."""

    def __init__(self, initial_dt, max_time, output_interval, max_cfl, adjust=(0.95, 0.9), start_time_index=0):
        super().__init__(initial_dt, max_cfl=max_cfl)
        self.step = 0
        self.initial_dt = initial_dt
        self.max_time = max_time
        self.output_interval = output_interval
        self.adjust = adjust
        self.output_index = start_time_index
        self.time = start_time_index * output_interval if start_time_index > 0 else 0.0
        self.forward = self.time <= max_time
        self.output = False

    def update(self, maxulen):
        if not self.forward:
            return
        self.step += 1
        self.time = self.time + self.dt
        self.forward = self.time < self.max_time
        self.cfl = maxulen * self.dt
        if self.cfl > 0:
            top, bot = self.adjust
            if self.cfl > top * self.max_cfl:
                self.dt = self.dt * top
                while maxulen * self.dt > top * self.max_cfl:
                    self.dt = self.dt * top
            elif self.cfl < bot * self.max_cfl:
                self.dt = self.dt * 1.05
            if (self.time + self.dt) / self.output_interval - self.output_index - 1 >= 0:
                dttmp = (self.output_index + 1) * self.output_interval - self.time
                self.dt = self.dt / 20 if dttmp < self.initial_dt / 100 else dttmp
                self.output_index += 1
                self.output = True
            else:
                self.output = False
        if self.cfl > self.max_cfl * 1.1 or self.dt < self.initial_dt / 1000:
            self.forward = False
            self.output = True


def run_steps(solver, stepper, nsteps):
    """FlatChannelSolver::Execute's loop body, nsteps times (src/solver_channel_flat.cc:199-254);
    a TimeModeStepper also hands its new dt to the solver (Phase::Update)."""
    feat = np.zeros(50)
    timed = isinstance(stepper, TimeModeStepper)
    for _ in range(nsteps):
        if not stepper.forward:
            break
        stepper.update(feat[19])
        if timed:
            solver.set_dt(stepper.dt)
        solver.step(stepper.time, stepper.alternation)
        stepper.toggle()
        feat = solver.features()
    return feat


# ---------------------------------------------------------------------------
# learned spin-up proposal network
# ---------------------------------------------------------------------------
def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def bf16_gemm(a, b):
    """C = A B on the device BF16 MFMA GEMM (operands rounded to BF16, FP32 sums)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    m, k = a.shape
    k2, n = b.shape
    assert k == k2
    c = np.zeros((m, n), dtype=np.float32)
    _check(lib().tbd_bf16_gemm(m, n, k, _fp(a), _fp(b), _fp(c)))
    return c


class SpectralOperatorNet:
    """Factorized spectral neural operator on an nx*ny Fourier x nz Chebyshev
    grid (device BF16 MFMA forward, backward and Adam)."""

    def __init__(self, nx=24, ny=24, nz=33, channels=32, layers=4, cin=7, cout=4, batch=8):
        self.shape = dict(nx=nx, ny=ny, nz=nz, channels=channels, layers=layers, cin=cin, cout=cout, batch=batch)
        self.points = nx * ny * nz
        h = ctypes.c_void_p()
        _check(lib().tbd_sno_create(nx, ny, nz, channels, layers, cin, cout, batch, ctypes.byref(h)))
        self._h = h
        n = ctypes.c_longlong()
        _check(lib().tbd_sno_param_count(self._h, ctypes.byref(n)))
        self.nparams = n.value

    def close(self):
        if self._h:
            lib().tbd_sno_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        _check(lib().tbd_sno_set_params(self._h, _fp(p)))

    def params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().tbd_sno_get_params(self._h, _fp(p)))
        return p

    def grads(self):
        g = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().tbd_sno_get_grads(self._h, _fp(g)))
        return g

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        nb = x.shape[0]
        y = np.zeros((nb, self.points, self.shape["cout"]), dtype=np.float32)
        _check(lib().tbd_sno_forward(self._h, _fp(x), _fp(y), nb))
        return y

    def loss_grad(self, x, t, w):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().tbd_sno_loss_grad(self._h, _fp(x), _fp(t), _fp(w), x.shape[0], ctypes.byref(loss)))
        return loss.value

    def train_step(self, x, t, w, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().tbd_sno_train_step(self._h, _fp(x), _fp(t), _fp(w), x.shape[0], float(lr), ctypes.byref(loss)))
        return loss.value
