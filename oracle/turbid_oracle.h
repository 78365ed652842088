/*
 * turbid_oracle.h -- C API of the FP64 CPU restatement of TURBID's
 * pseudospectral Fourier-Chebyshev RK3 substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library
 * (turbid_amd) never links, loads or calls it.
 *
 * Layouts follow the reference's RealCplxCube (src/cube_real_cplx.cc:26-91):
 *   spectral: complex<double> interleaved, index iz + Nz*(iy + nyc*(ix + mx*d))
 *             (z fastest, ky in FFT order, components stacked)
 *   physical: double, index x + Nx*(y + Ny*(z + Nz*d)) (x fastest)
 * with nxc = Nx/3*2, nyc = Ny/3*2, mx = nxc/2+1 (src/base.cc:233-258).
 */
#ifndef TURBID_ORACLE_H_
#define TURBID_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Plain mirror of the ini sections [Domain] [Space Stepper] [Time Stepper]
 * [Phase Carrier] [Phase Scalar] (src/configure.h:63-117). */
typedef struct tbo_config {
  int nx, ny, nz;                 /* grid_point_number */
  int nphase;                     /* total_phase_number (1 or 2) */
  double lx, ly, lz;              /* dimensionless_length (lz must be 2) */
  double slope;                   /* top_bottom_slope */
  int slope_direction;            /* 0 none, 1 x, 2 y */
  int mapping;                    /* 0 None 1 TwoEndTangent 2 TwoEndSin 3 OneEndTangent */
  double mapping_coefficient;
  double dt;                      /* marching_time_step / initial_time_step */
  int output_on_step;             /* output_based_on_step */
  int force_type;                 /* 0 None 1 Constant 2 Oscillatory 3 Pulsatile 4 Damping */
  int force_direction[2];
  double force_amplitude[4];
  double reynolds, froude, averaged_concentration, specific_gravity;
  int affected_by_scalar_phase;
  double vt_u1[5], vt_u2[5], vt_u3[5], vt_p[5];
  double schmidt, settling_velocity, erosion_rate, critical_shear_stress;
  int scalar_bc;                  /* 0 ZeroFlux, 1 ErosionDeposition */
  int precision;                  /* ignored by the oracle (always FP64) */
} tbo_config;

typedef struct tbo_ctx tbo_ctx;

/* Solver lifetime.  Returns NULL on error (message via tbo_last_error). */
tbo_ctx* tbo_create(const tbo_config* cfg, int nthreads);
void tbo_destroy(tbo_ctx* ctx);
const char* tbo_last_error(void);

/* State in RealCplxCube order.  uvw: S*3 complex, p: S complex, c: S complex
 * (S = Nz*nyc*mx).  c may be NULL when nphase == 1.  set_state also performs
 * the Configure-time work (physical velocity, zero history). */
int tbo_set_state(tbo_ctx* ctx, const double* uvw, const double* p, const double* c);
int tbo_get_state(const tbo_ctx* ctx, double* uvw, double* p, double* c);

/* One RK3 substage of every phase (solver_channel_flat.cc:223-234) at the
 * time stepper's current time.  alternation = TimeStepper alternation flag. */
int tbo_substage(tbo_ctx* ctx, int stage, double time, int alternation);
/* TimeStepper::Update (step mode) + 3 substages + alternation toggle.
 * Returns 1 while the stepper is forward, 0 when it stopped, <0 on error. */
int tbo_step(tbo_ctx* ctx);
void tbo_set_dt(tbo_ctx* ctx, double dt);
void tbo_stepper(const tbo_ctx* ctx, int* step, double* time, double* cfl, int* alternation);
void tbo_features(const tbo_ctx* ctx, double* out50);

/* Physical velocity (uvwnonlinear, R*3 doubles) of the carrier. */
int tbo_get_physical_velocity(const tbo_ctx* ctx, double* uvw_phys);
/* FlatCarrierPhase::NonlinearVelocity of the current state in the
 * convective (alternation 0) or divergence (1) form: adv = S*3 complex
 * (phase_carrier_flat.cc:716-813).  The state is left unchanged. */
int tbo_nonlinear(tbo_ctx* ctx, int alternation, double* adv);
/* Explicit-term histories H of the carrier (S*3 complex) and of the scalar
 * (S complex; ignored when nphase == 1) after the last substage. */
int tbo_get_history(const tbo_ctx* ctx, double* huvw, double* hc);

/* Operator tables built by the restatement (for pinning). */
int tbo_get_operators(const tbo_ctx* ctx, double* gridz, double* d1, double* d2);
/* which: 0 u1, 1 u2, 2 u3, 3 p, 4 scalar.  V, Vinv: Nz*Nz col-major (bordered),
 * lambda: Nz, fix: 4*Nz (fwd bottom, fwd top, bwd bottom, bwd top). */
int tbo_get_decomposition(const tbo_ctx* ctx, int which, double* V, double* Vinv,
                          double* lambda, double* fix);

/* ---- stand-alone kernels of the restatement ---- */
/* DftForward / DftBackward on nv components (cube_real_cplx.cc:525-555). */
int tbo_dft_forward(int nx, int ny, int nz, int nv, const double* phys, double* spec);
int tbo_dft_backward(int nx, int ny, int nz, int nv, const double* spec, double* phys);
/* CGL grid + D1, D2 (stepper_space_chebyshev.cc:101-228), col-major. */
int tbo_chebyshev(int nz, int mapping, double coef, double* gridz, double* d1, double* d2);
/* The scalar's positivity filter, in place, on nz physical planes of nx*ny
 * values (x fastest). */
int tbo_positivity_filter(int nx, int ny, int nz, double* phys);
/* Real eigen-decomposition of a real n*n col-major matrix with real spectrum:
 * lambda (n), V (n*n col-major, unit columns).  Returns 0, or <0 if complex. */
int tbo_eig(int n, const double* a, double* lambda, double* v);
/* Helmholtz solve of one decomposition (which as above) on ncol spectral
 * columns X (Nz*ncol complex) with per-column (kx2+ky2 multipliers) given as
 * k2[ncol] = -(kx^2) + -(ky^2), stage selects sigma; mode 0: velocity/scalar
 * (forward fix), 1: pressure (boundary rows zeroed).  col0_is_mean gates the
 * pressure null mode on column 0. */
int tbo_helmholtz(const tbo_ctx* ctx, int which, int stage, int mode, int ncol,
                  const double* k2, int col0_is_mean, const double* x, double* y);

#ifdef __cplusplus
}
#endif

#endif  /* TURBID_ORACLE_H_ */
