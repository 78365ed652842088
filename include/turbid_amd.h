/*
 * turbid_amd.h -- C ABI of the MI355X pseudospectral Fourier-Chebyshev RK3
 * substep (TURBID FlatChannelSolver hot path) on one GPU.
 *
 * The reference has no FFI: its boundary is the C++ virtual interface
 * Phase::{Configure, Predict(stage), Correct(stage), Update, Save, Load}
 * (reference src/phase.h:60-73) driven by FlatChannelSolver::Execute
 * (src/solver_channel_flat.cc:199-293), with the two phases sharing static
 * variables (src/variable.h:70-85).  One tbd_ctx therefore owns the carrier
 * AND the scalar phase on the device.  The reference-side binding is the
 * C++ host in turbid_amd/csrc/host/ (GPU-backed FlatChannelSolver), which
 * calls only the functions below.
 *
 * Conventions (mirroring the reference's error behaviour at a C boundary):
 *   - every int-returning call returns 0 on success and a negative code on
 *     failure; the message ("Class::Method > Message" style) is available
 *     from tbd_last_error() (thread-local);
 *   - host pointers are borrowed for the duration of the call;
 *   - spectral arrays use the RealCplxCube order (src/cube_real_cplx.cc:26-91):
 *     interleaved complex double, index iz + Nz*(iy + nyc*(ix + mx*d)),
 *     nxc = Nx/3*2, nyc = Ny/3*2, mx = nxc/2+1 (src/base.cc:233-258);
 *   - physical arrays: double, index x + Nx*(y + Ny*(z + Nz*d));
 *   - calls on one context are synchronous with respect to the host unless
 *     stated; one host thread per context.
 */
#ifndef TURBID_AMD_H_
#define TURBID_AMD_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Plain mirror of the ini keys the path reads: [Domain], [Space Stepper],
 * [Time Stepper], [Phase Carrier], [Phase Scalar] (src/configure.h:63-117).
 * Replaces the Configuration::Load(IniFile&) structs of
 * src/phase_carrier_flat.cc:27-61, src/phase_scalar_flat.cc:27-54,
 * src/domain.cc:20-35, src/stepper_time.cc:20-41. */
typedef struct tbd_config {
  int nx, ny, nz;                 /* grid_point_number */
  int nphase;                     /* total_phase_number: 1 carrier, 2 + sediment */
  double lx, ly, lz;              /* dimensionless_length (lz must be 2) */
  double slope;                   /* top_bottom_slope */
  int slope_direction;            /* 0 none, 1 x, 2 y */
  int mapping;                    /* 0 None 1 TwoEndTangent 2 TwoEndSin 3 OneEndTangent */
  double mapping_coefficient;
  double dt;                      /* marching_time_step (step mode) / initial_time_step */
  int output_on_step;             /* output_based_on_step */
  int force_type;                 /* 0 None 1 Constant 2 Oscillatory 3 Pulsatile 4 Damping */
  int force_direction[2];
  double force_amplitude[4];
  double reynolds, froude, averaged_concentration, specific_gravity;
  int affected_by_scalar_phase;
  double vt_u1[5], vt_u2[5], vt_u3[5], vt_p[5];   /* vertical_type_* */
  double schmidt, settling_velocity, erosion_rate, critical_shear_stress;
  int scalar_bc;                  /* 0 ZeroFlux, 1 ErosionDeposition */
  int precision;                  /* 0 FP64 throughout; 1 mixed (FP32 transforms,
                                     nonlinear products and their D1 contractions) */
} tbd_config;

typedef struct tbd_ctx tbd_ctx;

/* Builds one DNS on the current device: operators, factorisations, device
 * buffers.  Replaces the constructors + Configure of ChebyshevSpaceStepper
 * (src/stepper_space_chebyshev.cc:101-250), CmDz2DecompositionCplx
 * (src/cmdz2_decomposition.cc:19-227) and both phases
 * (src/phase_carrier_flat.cc:178-246, src/phase_scalar_flat.cc:176-221). */
int tbd_create(const tbd_config* cfg, tbd_ctx** out);
void tbd_destroy(tbd_ctx* ctx);
const char* tbd_last_error(void);

/* State transfer.  Replaces Variable::Load/Save's cube I/O
 * (src/variable_carrier_flat.cc:254-377, src/variable_scalar_flat.cc:59-212):
 * uvw: 3*S complex, p: S complex, c: S complex (NULL when nphase == 1).
 * Upload also performs the Configure-time work (physical velocity, zero
 * H and wall-gradient history, src/boundary_carrier_flat.cc:24-54, 255-269). */
int tbd_upload_state(tbd_ctx* ctx, const double* uvw, const double* p, const double* c);
int tbd_download_state(tbd_ctx* ctx, double* uvw, double* p, double* c);
/* Physical carrier velocity (uvwnonlinear), 3*Nx*Ny*Nz doubles. */
int tbd_download_physical_velocity(tbd_ctx* ctx, double* uvw);

/* One RK3 substage: FlatCarrierPhase::Predict + Correct then
 * FlatScalarPhase::Predict + Correct (src/solver_channel_flat.cc:223-234).
 * time = TimeStepper time after Update; alternation = TimeStepper flag. */
int tbd_substage(tbd_ctx* ctx, int stage, double time, int alternation);
/* The three substages of one step (src/solver_channel_flat.cc:223-250). */
int tbd_step(tbd_ctx* ctx, double time, int alternation);
/* Time mode: new dt (TimeStepper::Update2 + Phase::Update reassembly,
 * src/stepper_time.cc:120-208, src/phase_carrier_flat.cc:250-265). */
int tbd_set_dt(tbd_ctx* ctx, double dt);
/* Phase::featureglobal_ after the last stage 2 (src/phase_carrier_flat.cc:633-703). */
int tbd_features(tbd_ctx* ctx, double* out50);
int tbd_sync(tbd_ctx* ctx);

/* Kernel timing with HIP events on the context's stream (while enabled, a
 * step runs its substages eagerly instead of replaying its graph).  One
 * class per kernel:
 *   0..3  y-pass backward, x-pass backward, x-pass forward, y-pass forward (FP64)
 *   4..7  the same passes in FP32
 *   8, 9  fused x transforms + nonlinear products (FP64, FP32)
 *   10, 11 wall-normal contractions (FP64, FP32)
 *   12    Helmholtz / pressure solves (FP64)
 *   13, 14 pointwise nonlinear products (FP64, FP32)
 *   15    everything else
 *   16    Helmholtz / pressure solves through the even/odd form of the
 *         factorisation (FP64; fields with the same boundary type at both walls)
 *   17, 18 right-hand side + boundary rows + Helmholtz solve in one kernel
 *         (full / even/odd factorisation; FP64; D2 through its even/odd split)
 *   19    divergence + pressure solve + projection in one kernel (FP64)
 * work: algorithmic work of the timed launches -- bytes read plus written
 * (classes 0..9, 13, 14) or the flops executed: 4*nz^2*P per contraction of
 * one component with the full operator, 4*((M+1)^2 + (M+1)*M)*P (nz = 2M+1)
 * through the even/odd split (classes 10, 11); a Helmholtz solve is two
 * contractions (class 12; class 16: 4*(2(M+1)^2 + M^2 + (M+1)*M)*P); classes
 * 17, 18 add the split D2 contraction to their solve, class 19 two split D1
 * contractions to an even/odd solve. */
int tbd_profile_enable(tbd_ctx* ctx, int enable);
int tbd_profile_read(tbd_ctx* ctx, int kclass, long long* launches, double* total_ms, double* work);
int tbd_profile_reset(tbd_ctx* ctx);

/* Operators built by the context (pinned against closed forms in tests):
 * gridz[nz], d1/d2 nz*nz column-major (stepper_space_chebyshev.cc:122-228). */
int tbd_get_operators(tbd_ctx* ctx, double* gridz, double* d1, double* d2);
/* which: 0 u1, 1 u2, 2 u3, 3 p, 4 scalar; V~, V~^-1 nz*nz column-major,
 * lambda[nz], fix[4*nz] (cmdz2_decomposition.cc:47-139). */
int tbd_get_factorisation(tbd_ctx* ctx, int which, double* V, double* Vinv, double* lambda, double* fix);
/* Operators supplied by the caller instead (the host side of the reference
 * builds them in ChebyshevSpaceStepper::Configure, stepper_space_chebyshev.cc:
 * 101-228, and CmDz2DecompositionCplx::Construct, cmdz2_decomposition.cc:
 * 19-297).  tbd_set_operators replaces D1 and D2 (nz*nz column-major) and
 * rebuilds every factorisation from them with the configured boundary rows;
 * tbd_set_factorisation replaces one factorisation (layout as in
 * tbd_get_factorisation; the eigenvectors may come in any order and scaling,
 * the even/odd form is rebuilt from them).  Both wait for the device,
 * replace the device copies and drop the captured steps; call them between
 * steps, from the thread that owns the context. */
int tbd_set_operators(tbd_ctx* ctx, const double* d1, const double* d2);
int tbd_set_factorisation(tbd_ctx* ctx, int which, const double* V, const double* Vinv, const double* lambda,
                          const double* fix);

/* Kernel variants of a step (no counterpart in the reference; every variant is
 * tested against the same oracle).  Bits: 1 x transforms fused around the
 * nonlinear products, 2 right-hand side + boundary rows + solve and the
 * divergence + pressure solve + projection chain each as one kernel (needs
 * the even/odd D1, D2), 4 scalar stages on a second stream, 8 / 16 / 32 / 64
 * the y backward / y forward / x backward / x forward transform kernels with
 * twice their default workgroup size, 128 / 256 / 512 / 1024 the same passes
 * with the two factors of their length in the other order (radix split
 * N2 x N1; lengths with two distinct factors, e.g. 24, 96, 128, 384, else the
 * bit stays clear), 2048 the plain contractions (gradients, D1 of the
 * nonlinear terms) on panels of twice the column groups where the LDS allows.
 * A switch set in the environment when the context was
 * created (TBD_FUSED, TBD_RHS_UNFUSED, TBD_NO_OVERLAP; TBD_FFT_WIDE for the
 * four widths) is pinned.  tbd_autotune times every combination of bits 1, 2,
 * 4, then each width bit, each radix-order bit and the panel-width bit in turn
 * from the best so far, for `steps` captured steps (after three warm-up steps) on a scratch
 * context with the same configuration, operators, dt and current state (so it
 * needs the device memory of a second context; this context's state is not
 * changed) and keeps the fastest; ms_per_step[4096] (optional) receives each
 * variant's time, NaN where not tried, and *bits (optional) the variant
 * kept. */
int tbd_autotune(tbd_ctx* ctx, int steps, double* ms_per_step, int* bits);
int tbd_get_variant(tbd_ctx* ctx, int* bits);
int tbd_set_variant(tbd_ctx* ctx, int bits);

/* Ensemble of independent contexts advanced together (the reference runs one
 * FlatChannelSolver per realisation, solver_channel_flat.cc:195-293 being the
 * step loop of each; it has no batched form).  The members (at most 32, not
 * owned, created with tbd_create and kept alive while the ensemble exists)
 * may differ in grid, Reynolds number, settling velocity and the rest of
 * their configuration.  tbd_ensemble_step advances every member by one RK3
 * step (member m at time[m] with alternation[m], as tbd_step): each member's
 * step is recorded once per (alternation, wall-row slot, dt, previous dt,
 * variant bits) of all members, the members' launches are merged into one launch per kernel
 * and launch shape at each position of the step (the workgroups of all
 * members in one grid, each with its own arguments), and the merged
 * sequence is captured into one HIP graph and replayed.  It is ordered after
 * the work already issued on the members and before what is issued on them
 * afterwards (uploads, downloads, tbd_features).  Members must run with the
 * overlap and fused-transform variant bits clear (tbd_set_variant) and the
 * register transforms; otherwise the step fails with an error.
 * tbd_ensemble_launches: grouped launches per step and the launches the
 * members would issue one by one, of the last recorded step.
 * tbd_ensemble_features: every member's 50 features (tbd_features) after the
 * last step, member m at out + 50 m, waiting for the ensemble's work. */
typedef struct tbd_ens tbd_ens;
int tbd_ensemble_create(tbd_ctx* const* members, int n, tbd_ens** out);
int tbd_ensemble_step(tbd_ens* e, const double* time, const int* alternation);
int tbd_ensemble_launches(tbd_ens* e, int* grouped, int* member);
int tbd_ensemble_features(tbd_ens* e, double* out);
int tbd_ensemble_sync(tbd_ens* e);
void tbd_ensemble_destroy(tbd_ens* e);

/* Stand-alone device kernels on host buffers (parity tests of single pieces).
 * DftForward / DftBackward of nv components (src/cube_real_cplx.cc:525-555);
 * fp32 != 0 runs the FP32 transform on FP64 data. */
int tbd_dft_forward(int nx, int ny, int nz, int nv, int fp32, const double* phys, double* spec);
int tbd_dft_backward(int nx, int ny, int nz, int nv, int fp32, const double* spec, double* phys);
/* No counterpart in the reference.  tbd_dft_forward for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): phys (nv*Nx*Ny*Nz doubles) and spec
 * (nv*Nz*P interleaved complex doubles, RealCplxCube order) are DEVICE
 * pointers of the caller.  The twiddles, the transform intermediate and the
 * FP32 staging live in one fixed device array of the library, 64 MiB per GPU,
 * held from the library's first use on that GPU until the process ends; the
 * z-planes are transformed in chunks that fit it.  A size whose single
 * z-plane of nv components does not fit is rejected (error, nothing
 * launched).  Runs on the null stream and returns when it has finished.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from several host threads) race. */
int tbd_dft_forward_noalloc(int nx, int ny, int nz, int nv, int fp32, const double* phys, double* spec);
/* Y = M X for ncol complex columns of length nz, M nz*nz column-major
 * (RealCplxCube::Multiply(cmat, result), src/cube_real_cplx.cc:461-471). */
int tbd_zcontract(int nz, long long ncol, int fp32, const double* m, const double* x, double* y);
/* Helmholtz solve with the context's factorisation `which` at `stage` on
 * its own fields layout: x, y each S complex (one component). mode 0 forward
 * fix, 1 pressure (src/phase_carrier_flat.cc:861-994). */
int tbd_helmholtz(tbd_ctx* ctx, int which, int stage, int mode, const double* x, double* y);

/* Swirling strength of the context's current velocity
 * (FlatChannelPostprocessor::Execute, src/postprocessor_channel_flat.cc:
 * 203-236): Gradient of u, v, w, DftBackward, and per physical point
 * max |Im lambda| over the eigenvalues of the 3x3 velocity-gradient tensor.
 * out: nx*ny*nz doubles, x fastest (the real cube phase_00_vortex.dat holds). */
int tbd_swirling_strength(tbd_ctx* ctx, double* out);

/* Phase-binned statistics on the device (reference log and postprocessing:
 * solver_channel_flat.cc:372-414 writes the mean-mode wall gradients each
 * step; the MATLAB companions average profiles by phase).  accumulate adds
 * the current state's profiles to bin `bin` (0..63): per z, rows
 * 0..3 mean-mode u, v, w, c; rows 4..7 plane means of u'^2, v'^2, w'^2, c'^2
 * (sums over the non-mean spectral columns, conjugate half included).
 * read returns the 8*nz sums of a bin and the number of states added. */
int tbd_stats_accumulate(tbd_ctx* ctx, int bin);
int tbd_stats_read(tbd_ctx* ctx, int bin, double* out, long long* count);
int tbd_stats_reset(tbd_ctx* ctx);

/* Divergence-free projection of the context's velocity, for states that do
 * not come out of a time step (a proposed initial state): the pressure-type
 * solve of FlatCarrierPhase::Correct (src/phase_carrier_flat.cc:586-620)
 * with a unit step, u -= grad phi, mean w := 0, physical velocity refreshed.
 * The wall pressure-gradient history is cleared, as after a load. */
int tbd_project(tbd_ctx* ctx);

/* ---------------------------------------------------------------------------
 * Learned spin-up proposal (no counterpart in the reference): a factorized
 * spectral neural operator on an nx*ny Fourier x nz Chebyshev grid with
 * `channels` channels and `layers` residual layers, cin input and cout output
 * fields per point, trained and evaluated with BF16 MFMA kernels (FP32
 * accumulation).  Tensors are float, [sample][z][y][x][field]; parameters are
 * one float block (tbd_sno_param_count floats) in the order lift (cin*C, C),
 * per layer (W C*C, b C, Rx (nx/2+1)*2*C*C, Ry (ny/2+1)*2*C*C, Rz nz*C*C),
 * projection (C*cout, cout).
 * ------------------------------------------------------------------------- */
typedef struct tbd_sno tbd_sno;
int tbd_sno_create(int nx, int ny, int nz, int channels, int layers, int cin, int cout, int batch,
                   tbd_sno** out);
void tbd_sno_destroy(tbd_sno* net);
int tbd_sno_param_count(tbd_sno* net, long long* count);
int tbd_sno_set_params(tbd_sno* net, const float* params);
int tbd_sno_get_params(tbd_sno* net, float* params);
int tbd_sno_get_grads(tbd_sno* net, float* grads);
/* y[nb][points][cout] = network(x[nb][points][cin]), nb <= batch */
int tbd_sno_forward(tbd_sno* net, const float* x, float* y, int nb);
/* loss = mean over samples, points and outputs of w[o] (y - t)^2, and its
 * gradient (tbd_sno_get_grads); train_step also applies one Adam update */
int tbd_sno_loss_grad(tbd_sno* net, const float* x, const float* t, const float* w, int nb, double* loss);
int tbd_sno_train_step(tbd_sno* net, const float* x, const float* t, const float* w, int nb, float lr,
                       double* loss);
/* the BF16 GEMM on host buffers: C (m x n) = A (m x k) B (k x n), row-major */
int tbd_bf16_gemm(int m, int n, int k, const float* a, const float* b, float* c);

#ifdef __cplusplus
}
#endif

#endif /* TURBID_AMD_H_ */
