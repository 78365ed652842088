"""Run directories in the reference's on-disk format.

configuration.ini (keys of reference src/configure.h:51-117) and
<index>/phase_NN_{phase,variable,boundary}.dat, the variable file being the
text line "IOControlLevelOne" followed by Armadillo binary complex cubes
(header "ARMA_CUB_BIN_FC016", "rows cols slices", raw column-major
complex<double>; tmtools/w_phase_cplx.m:26-45, w_armadillo_cplx.m:35-40).
Cube shapes: carrier uvw (Nz, nyc, 3*mx) then p (Nz, nyc, mx); scalar c
(Nz, nyc, mx).
"""
import math
import os

import numpy as np

HEADER = b"ARMA_CUB_BIN_FC016"

FORCE_NAMES = {0: "None", 1: "Constant", 2: "Oscillatory", 3: "Pulsatile", 4: "Damping"}
MAPPING_NAMES = {0: "None", 1: "TwoEndTangent", 2: "TwoEndSin", 3: "OneEndTangent"}


def _write_cube(f, a, rows, cols, slices):
    a = np.ascontiguousarray(a, dtype=np.complex128).reshape(-1)
    assert a.size == rows * cols * slices
    f.write(HEADER + b"\n")
    f.write(("%d %d %d\n" % (rows, cols, slices)).encode())
    f.write(a.tobytes())


def _read_cube(f, rows, cols, slices):
    h = f.readline().rstrip(b"\n")
    if h != HEADER:
        raise ValueError("unexpected cube header %r" % h)
    dims = tuple(int(x) for x in f.readline().split())
    if dims != (rows, cols, slices):
        raise ValueError("cube shape %s, expected %s" % (dims, (rows, cols, slices)))
    n = rows * cols * slices
    return np.frombuffer(f.read(16 * n), dtype=np.complex128, count=n).copy()


def _dims(cfg):
    nyc = cfg["ny"] // 3 * 2
    mx = (cfg["nx"] // 3 * 2) // 2 + 1
    return cfg["nz"], nyc, mx


def write_state(run_dir, index, cfg, uvw, p, c=None):
    nz, nyc, mx = _dims(cfg)
    d = os.path.join(run_dir, str(index))
    os.makedirs(d, exist_ok=True)
    for ph in ([0, 1] if c is not None else [0]):
        open(os.path.join(d, "phase_%02d_phase.dat" % ph), "wb").close()
        open(os.path.join(d, "phase_%02d_boundary.dat" % ph), "wb").close()
    with open(os.path.join(d, "phase_00_variable.dat"), "wb") as f:
        f.write(b"IOControlLevelOne\n")
        _write_cube(f, uvw, nz, nyc, 3 * mx)
        _write_cube(f, p, nz, nyc, mx)
    if c is not None:
        with open(os.path.join(d, "phase_01_variable.dat"), "wb") as f:
            f.write(b"IOControlLevelOne\n")
            _write_cube(f, c, nz, nyc, mx)


def read_state(run_dir, index, cfg, nphase=None):
    nz, nyc, mx = _dims(cfg)
    nphase = cfg.get("nphase", 1) if nphase is None else nphase
    d = os.path.join(run_dir, str(index))
    with open(os.path.join(d, "phase_00_variable.dat"), "rb") as f:
        if f.readline().rstrip(b"\n") != b"IOControlLevelOne":
            raise ValueError("bad IO level line")
        uvw = _read_cube(f, nz, nyc, 3 * mx)
        p = _read_cube(f, nz, nyc, mx)
    c = None
    if nphase > 1:
        with open(os.path.join(d, "phase_01_variable.dat"), "rb") as f:
            if f.readline().rstrip(b"\n") != b"IOControlLevelOne":
                raise ValueError("bad IO level line")
            c = _read_cube(f, nz, nyc, mx)
    return uvw, p, c


def _fmt(v):
    if isinstance(v, (list, tuple)):
        return " ".join(_fmt(x) for x in v)
    if isinstance(v, float):
        return repr(v)
    return str(v)


def write_ini(path, cfg, max_step=100, write_interval=100, start_step=0, max_cfl=0.5, print_log=0, write_log=1,
              work_index_range=None, time_mode=False, max_time=100000.0, output_interval=100.0,
              start_time_index=-1, initial_dt=None):
    """configuration.ini in the reference's layout from a workload dict.
    time_mode: output_based_on_step = 0 (adaptive dt from initial_dt, output
    every output_interval of time until max_time; TimeStepper::Update2)."""
    scalar = "ErosionDeposition" if cfg.get("scalar_bc", 1) == 1 else "ZeroFlux"
    lines = [
        "[System]",
        "print_log_when_running    = %d" % print_log,
        "write_log_when_running    = %d" % write_log,
        "total_phase_number        = %d" % cfg.get("nphase", 1),
        "mpi_grid_point_number     = 1 1",
        "thread_number_openmp      = 1",
        "[PreProcessor]",
        "[PostProcessor]",
        "[Domain]",
        "applied_domain_type       = Fixed",
        "dimensionless_length      = %s" % _fmt([float(cfg["lx"]), float(cfg["ly"]), float(cfg["lz"])]),
        "grid_point_number         = %d %d %d" % (cfg["nx"], cfg["ny"], cfg["nz"]),
        "top_bottom_slope          = %s" % _fmt(float(cfg.get("slope", 0.0))),
        "slope_direction           = %d" % cfg.get("slope_direction", 0),
        "read_profile_from_file    = 0",
        "[Space Stepper]",
        "applied_stepper_type      = Chebyshev",
        "mapping_function          = %s" % MAPPING_NAMES[cfg.get("mapping", 0)],
        "mapping_coefficient       = %s" % _fmt(float(cfg.get("mapping_coefficient", 0.0))),
        "[Time Stepper]",
        "applied_stepper_type      = RungeKutta3",
        "maximum_cfl_number        = %s" % _fmt(float(max_cfl)),
        "output_based_on_step      = %d" % (0 if time_mode else 1),
        "maximum_task_step         = %d" % max_step,
        "result_write_interval     = %d" % write_interval,
        "marching_time_step        = %s" % _fmt(float(cfg["dt"])),
        "task_start_step           = %d" % start_step,
        "max_marching_time         = %s" % _fmt(float(max_time)),
        "output_interval           = %s" % _fmt(float(output_interval)),
        "task_start_time_index     = %d" % start_time_index,
        "initial_time_step         = %s" % _fmt(float(cfg["dt"] if initial_dt is None else initial_dt)),
        "cfl_adjust_scale          = 0.95 0.9",
        "result_write_initial      = 0",
        "[Phase Carrier]",
        "applied_force_type        = %s" % FORCE_NAMES[cfg.get("force_type", 2)],
        "force_direction           = %s" % _fmt(list(cfg.get("force_direction", (0, 1)))),
        "force_amplitude           = %s" % _fmt([float(x) for x in cfg.get("force_amplitude", (0, 1, 0, 0))]),
        "reynolds_number           = %s" % _fmt(float(cfg["reynolds"])),
        "vertical_type_u1          = %s" % _fmt([float(x) for x in cfg["vt_u1"]]),
        "vertical_type_u2          = %s" % _fmt([float(x) for x in cfg["vt_u2"]]),
        "vertical_type_u3          = %s" % _fmt([float(x) for x in cfg["vt_u3"]]),
        "vertical_type_p           = %s" % _fmt([float(x) for x in cfg["vt_p"]]),
        "iteration_coefficient     = 1",
        "affected_by_scalar_phase  = %d" % cfg.get("affected_by_scalar_phase", 0),
        "froude_number             = %s" % _fmt(float(cfg.get("froude", 1.0))),
        "averaged_concentration    = %s" % _fmt(float(cfg.get("averaged_concentration", 1.0))),
        "specific_gravity          = %s" % _fmt(float(cfg.get("specific_gravity", 2.65))),
        "[Phase Scalar]",
        "schmidt_number            = %s" % _fmt(float(cfg.get("schmidt", 1.0))),
        "settling_velocity         = %s" % _fmt(float(cfg.get("settling_velocity", 0.0))),
        "vertical_boundary_type    = %s" % scalar,
        "erosion_rate              = %s" % _fmt(float(cfg.get("erosion_rate", 0.0))),
        "critical_shear_stress     = %s" % _fmt(float(cfg.get("critical_shear_stress", 1.0))),
    ]
    if work_index_range is not None:   # [PostProcessor] (configure.h:60-61)
        lines += ["[PostProcessor]", "work_index_range          = %s" % _fmt(list(work_index_range))]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_run(run_dir, cfg, uvw, p, c=None, **ini_kw):
    """A complete run directory: configuration.ini + 0/ initial state."""
    os.makedirs(run_dir, exist_ok=True)
    ini = os.path.join(run_dir, "configuration.ini")
    write_ini(ini, cfg, **ini_kw)
    write_state(run_dir, 0, cfg, uvw, p, c if cfg.get("nphase", 1) > 1 else None)
    return ini


REAL_HEADER = b"ARMA_CUB_BIN_FN008"


def read_real_cube(path):
    """An Armadillo binary real cube (RealCplxCube::SaveReal): array of shape
    (slices, cols, rows), i.e. (Nz, Ny, Nx) for a physical field."""
    with open(path, "rb") as f:
        h = f.readline().rstrip(b"\n")
        if h != REAL_HEADER:
            raise ValueError("unexpected cube header %r" % h)
        rows, cols, slices = (int(x) for x in f.readline().split())
        n = rows * cols * slices
        a = np.frombuffer(f.read(8 * n), dtype=np.float64, count=n).copy()
    return a.reshape(slices, cols, rows)
