"""Roofline of one RK3 step from its minimal algorithmic counts (MI355X
published peaks: HBM 8 TB/s, dense MFMA FP64 78.6 and FP32 157.3 TFLOP/s)."""

HBM_PEAK = 8.0e12          # B/s
MFMA_PEAK = {"fp64": 78.6e12, "fp32": 157.3e12}


def step_roofline(cfg, mixed, ms_per_step):
    """Whole-step roofline with the minimal algorithmic counts per RK3 step:
    58.5 field-transforms of w*R + 2*w*S bytes and 63 z-contractions of
    4*Nz^2*P flops (mixed: FP32 transforms, 12 FP32 and 51 FP64
    contractions); frac = T_roof / T_measured."""
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    nyc, mx = ny // 3 * 2, (nx // 3 * 2) // 2 + 1
    P = mx * nyc
    R, S = nx * ny * nz, nz * P
    w = 4 if mixed else 8
    t_fft = 58.5 * (w * R + 2 * w * S) / HBM_PEAK
    zf = 4.0 * nz * nz * P
    t_z = (51 * zf / MFMA_PEAK["fp64"] + 12 * zf / MFMA_PEAK["fp32"]) if mixed else 63 * zf / MFMA_PEAK["fp64"]
    t_roof = 1e3 * (t_fft + t_z)
    return {"t_roof_ms": t_roof, "t_fft_ms": 1e3 * t_fft, "t_z_ms": 1e3 * t_z, "ms_per_step": ms_per_step,
            "frac": t_roof / ms_per_step}
