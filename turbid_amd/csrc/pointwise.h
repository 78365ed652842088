// Elementwise and reduction kernels of the substage.
#pragma once

#include "common.h"

namespace tbd {

// convective products adv_i = u.grad(u_i) (phase_carrier_flat.cc:792-803):
// uvw[3] and g[9] (du/dx,du/dy,du/dz, dv/.., dw/..) physical, stride R
template <typename T>
void products_conv(const T* uvw, const T* g, T* adv, long long R, hipStream_t st);
// divergence-form products (phase_carrier_flat.cc:725-744): the six distinct
// u_i u_j in order uu, uv, uw, vv, vw, ww
template <typename T>
void products_div(const T* uvw, T* prod6, long long R, hipStream_t st);
// scalar settling-flux products with the advecting velocity u + shift
// (phase_scalar_flat.cc:256-279, 566-571, 587-590)
// conv: adv = (U dc0 + V dc1) + W dc2 ; div: out_i = c * U_i
template <typename T>
void scalar_products_conv(const T* uvw, T su, T sv, T sw, const T* dc, T* adv, long long R, hipStream_t st);
template <typename T, typename TC>
void scalar_products_div(const T* uvw, T su, T sv, T sw, const TC* c, T* out3, long long R, hipStream_t st);

// H = c1 H - dt adv, + buoyancy (K c sin, -K c cos) + dt*force on the mean
// column (phase_carrier_flat.cc:294-504).  nf fields of stride S.
struct HUpd {
  double c1 = 0, dt = 0;
  int buoy = 0;          // 1 if the scalar drives the carrier
  double K = 0;          // dt*Cbar/Fr^2*(sg-1)
  double sn = 0, cs = 0; int slope_dir = 0;
  double f1 = 0, f2 = 0; // mean-mode force (u, v)
  const double* fdev = nullptr;   // if set: the force is read from fdev[0], fdev[1]
  int nz = 0;
};
// dst[i] = src[i] for i < n (n <= 64), then dst[10 + j] = wall[src6[j]], j < 6
void collect_features(double* dst, const double* src, int n, const double* wall, const int* src6, hipStream_t st);
// device copy of `bytes` (a multiple of 4) as a kernel
void copy_words(void* dst, const void* src, size_t bytes, hipStream_t st);
// dst[i][0] = v[2 i], dst[i][1] = v[2 i + 1], i < n (n <= 32): the forcing
// pairs of an ensemble's members, one launch
void set_pairs(double* const* dst, const double* v, int n, hipStream_t st);
// dst[m * count + i] = src[m][i], m < n (n <= 32), i < count
void gather_rows(const double* const* src, int n, int count, double* dst, hipStream_t st);
// dst[0] = a, dst[1] = b (values that a captured step graph reads at replay)
void set_pair(double* dst, double a, double b, hipStream_t st);
template <typename TA>
void h_update(double2* h, const TA* adv, const double2* c, const HUpd& u, int nf, long long S, hipStream_t st);

// Carrier boundary rows (boundary_carrier_flat.cc:185-246) and the mean-mode
// top values (phase_carrier_flat.cc:542-547) written into rhs[3].
struct BcArgs {
  int nz = 0;
  long long P = 0, S = 0;
  double k = 0, s = 0;
  const double2* m1b = nullptr; const double2* m1t = nullptr;
  const double2* m2b = nullptr; const double2* m2t = nullptr;
  const double2* n1t = nullptr; const double2* n2t = nullptr;   // open-top tangential rows
  int open_top = 0;
  double top0 = 0, top1 = 0, top2 = 0;
};
void bc_carrier(double2* rhs, const BcArgs& a, hipStream_t st);

// Scalar rows: zero flux rows, then the erosion/deposition bottom row
// (boundary_scalar_flat.cc:151-152, phase_scalar_flat.cc:349-374).
void bc_scalar(double2* rhs, int nz, long long P, const double2* ehat, const double2* c, double ws, double cbar,
               double re, double sc, int erosion, hipStream_t st);
// erosion map on one plane: val = sqrt(a^2+b^2)/Re; val < tc -> 0 else M(val/tc - 1)
template <typename T>
void erosion_map(const T* dudz, const T* dvdz, double* e, long long n, double re, double tc, double M,
                 hipStream_t st);

// Bartnicki positivity filter, 5 iterations per z-plane (phase_scalar_flat.cc:427-475)
// touched (optional, device int): set to 1 when any plane is changed
// gate (optional, device int): the filter does nothing when *gate == 0
void bartnicki_filter(double* c, int nx, int ny, int nz, int* touched, hipStream_t st, const int* gate = nullptr);

// *need (device int) = 1 when some x[i] may be negative: min x <= 1e-4 max x
// (or max x <= 0), else 0.  part: 1024 doubles of scratch.  The filter above
// leaves a field with no negative value unchanged, so a check on an FP32
// transform of c (whose error is far below 1e-4 of its largest value)
// decides whether the exact FP64 filter has anything to do.
template <typename T>
void negativity_check(const T* x, long long n, double* part, int* need, hipStream_t st);

// What DftForward(DftBackward(s)) does to a truncated spectrum: the
// halfcomplex inverse keeps the real part of the kx = 0 sequences, so the
// kx = 0 plane becomes Hermitian in ky, s(ky) <- (s(ky) + conj s(-ky))/2, with
// s(+nyc/2) = 0 (not retained); every other column is reproduced.  Applied
// only when *gate == 0.
void kx0_round_trip(double2* s, int nyc, int nz, const int* gate, hipStream_t st);

// Diagnostics (phase_carrier_flat.cc:639-671): feat[0..8] max |u|,|v|,|w| over
// all / bottom / top planes, feat[19] max(|u|/dx + |v|/dy + |w|/dz(z)).
// part: kDiagMaxBlocks * 10 doubles of scratch for the per-block maxima.
static constexpr int kDiagMaxBlocks = 2048;
template <typename T>
void diagnostics(const T* uvw, long long R, int nx, int ny, int nz, double glx, double gly, const double* glz,
                 double* feat, double* part, hipStream_t st);

// feat[21] = trapz(gridz, Re c(:,0)) / 2 (phase_scalar_flat.cc:485-486)
void trapz_mean(const double2* c, const double* gridz, int nz, double* out, hipStream_t st);

// Phase statistics of one state, added into acc[kPhaseStats][nz]:
//   0..2  mean-mode profiles of u, v, w (real part of column kx = ky = 0)
//   3     mean-mode profile of c (zero if c is null)
//   4..6  <u'u'>, <v'v'>, <w'w'>: plane means of the squared fluctuations,
//         sum over the non-mean columns of |u_hat|^2 (kx > 0 counted twice,
//         the conjugate half of the plane)
//   7     <c'c'>
static constexpr int kPhaseStats = 8;
void phase_stats(const double2* uvw, const double2* c, long long S, int nz, int nyc, long long P, double* acc,
                 hipStream_t st);

// Swirling strength per point (FlatChannelPostprocessor::Execute,
// postprocessor_channel_flat.cc:212-236): g holds the nine physical gradient
// fields at g + (3i + j) R, G_ij = d u_i / d x_j; out = max |Im lambda| over
// the eigenvalues of G, i.e. |Im| of the complex pair when there is one, 0
// when all three are real.  The eigenvalues come from the characteristic
// cubic in closed form (Cardano, the root of larger magnitude taken without
// cancellation).
void swirling_strength(const double* g, long long R, double* out, hipStream_t st);

// small copies: dst[i] = Re(src[idx[i]]) for the log features
void gather_real(const double2* src, const long long* idx, int n, double* dst, hipStream_t st);
void fill_zero(void* p, size_t bytes, hipStream_t st);
// *out = (a * Re src[0]) * b
void scaled_real(const double2* src, double a, double b, double* out, hipStream_t st);
void copy_d2(double2* dst, const double2* src, long long n, hipStream_t st);

}  // namespace tbd
