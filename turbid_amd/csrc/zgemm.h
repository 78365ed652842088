// Wall-normal contractions Y = M X over every spectral column (reference
// RealCplxCube::Multiply, src/cube_real_cplx.cc:461-513) on MFMA, with the
// pointwise work around them fused into the epilogue.  M is a real nz*nz
// operator (D1, D2, V~, V~^-1) stored row-major with a padded leading
// dimension; X, Y are complex spectral fields (z fastest, one column per
// (kx,ky)).
#pragma once

#include <vector>

#include "common.h"

namespace tbd {

enum ZMode {
  Z_PLAIN = 0,    // y_f = M x_f
  Z_GRAD = 1,     // y_{3f+0,1,2} = (i kx x_f, i ky x_f, M x_f)          (SpaceStepper::Gradient)
  Z_RHS = 2,      // y_f = (s2 h_f + s1 x_f) - ((M x_f - kx^2 x_f) - ky^2 x_f)   (Laplacian + rhs)
  Z_DIV = 3,      // y = q ((M x + i kx e0) + i ky e1)                     (Divergence, pressure rhs)
  Z_CORRECT = 4,  // y_f = e_f - cf g_f, g = (i kx x, i ky x, M x); mean w := 0; wall rows saved
  Z_GRADTMP = 5,  // y_f = (i kx e0_f + i ky e1_f) + M x_f                  (GradientTmp summed)
  Z_GRADYZ = 6,   // y_f = i ky x_f, y_{3+f} = M x_f (f < 3): Z_GRAD without the x
                  // derivative, which the fused x transform takes from x_f's x-pencil
};

struct ZArgs {
  int mode = Z_PLAIN;
  int nz = 0;            // rows of M and of each column
  int lda = 0;           // leading dimension of the row-major operator
  // 0: op is the nz x nz operator.  +1 / -1: op is a centro-symmetric /
  // centro-antisymmetric operator A in its even/odd split (pack_parity),
  // applied as y_i = s_i + d_i, y_{nz-1-i} = +-(s_i - d_i), i <= M = (nz-1)/2
  int parity = 0;
  int P = 0;             // columns per field
  int nyc = 0;           // spectral y size (column -> (ix, iy))
  const void* op = nullptr;       // operator (T*), shared by all fields of the launch
  const void* x = nullptr;  long long xs = 0;   // field f at x + f*xs (complex elements)
  // when xf[0] is set: field f reads x, e0, e1 at xf[f], e0f[f], e1f[f]
  // (fields whose operands are not evenly strided)
  const void* xf[3] = {nullptr, nullptr, nullptr};
  const void* e0f[3] = {nullptr, nullptr, nullptr};
  const void* e1f[3] = {nullptr, nullptr, nullptr};
  void* y = nullptr;        long long ys = 0;
  const void* e0 = nullptr; const void* e1 = nullptr; const void* e2 = nullptr;
  long long es = 0;                             // stride between fields of e0/e1
  const void* h = nullptr;  long long hs = 0;
  const double* kx = nullptr;    // per spectral x index
  const double* ky = nullptr;    // per spectral y index
  double s1 = 0, s2 = 0, q = 0, cf = 0;
  // Z_CORRECT wall rows: [3P] bottom/top of (i kx p, i ky p, dp/dz), and
  // [2P] bottom/top of (i kx dp/dz, i ky dp/dz) when open_top
  double2* wall_bot = nullptr; double2* wall_top = nullptr;
  double2* wall2_bot = nullptr; double2* wall2_top = nullptr;
  int open_top = 0;
  // panel width: the column groups of 16 a workgroup takes are multiplied
  // by this where the LDS and the workgroup's tail-row pass allow it (an
  // autotuner choice; 1 keeps the width make_panel picks)
  int ncg_mult = 1;
};

// Even/odd split of a centro-(anti)symmetric nz x nz operator A (nz = 2M+1,
// row-major with leading dimension lda): rows 0..M hold P_ij = A_ij +
// A_i,nz-1-j (j < M) and P_iM = A_iM; rows M+1..2M+1 hold Q_ij = A_ij -
// A_i,nz-1-j (j < M).  Leading dimension of the result: ldp = ceil4(M+1).
// With e_j, o_j the even and odd halves of a column, s = P e and d = Q o.
template <typename T>
std::vector<T> pack_parity(const double* A, int nz, int lda, int* ldp);

// T: arithmetic/operator type; TX: input complex type; TY: output complex type.
// nfield fields are processed in one launch (blockIdx.y).
template <typename T, typename TX, typename TY>
void zgemm(const ZArgs& a, int nfield, hipStream_t st);

// Fused Helmholtz solve (reference FlatCarrierPhase::Helmholtz,
// phase_carrier_flat.cc:861-994, FlatScalarPhase::Helmholtz,
// phase_scalar_flat.cc:635-667) for up to three fields with their own
// factorisation.  FP64.
struct HField {
  const double* vinv = nullptr;   // row-major, lda
  const double* vnorm = nullptr;  // helmholtz(): V~ with the backward fixes folded in (fold_backward_fixes)
  const double* lambda = nullptr; // nz
  const double* fix = nullptr;    // 4*nz
  double sigma = 0;               // stage coefficient (0 for pressure)
  int nullrow = -1;               // gated row on the mean column (sigma == 0 only)
  const double2* x = nullptr;     // right-hand side
  double2* y = nullptr;           // solution (may alias x)
  const double* pfix = nullptr;   // helmholtz_split only: backward-fix weights on the split column
};

struct HArgs {
  int nz = 0, lda = 0, P = 0, nyc = 0;
  int mode = 0;                   // 0: forward fix; 1: boundary rows zeroed (pressure)
  const double* kx = nullptr;
  const double* ky = nullptr;
  HField f[3];
};

void helmholtz(const HArgs& a, int nfield, hipStream_t st);

// The same solve for one field through the even/odd form of its
// factorisation (Factorisation::pM > 0): vinv, vnorm, lambda, nullrow are the
// split pinv, pnorm, plam, pnull (lda = pld), fix the natural 4*nz fix array,
// pfix the split backward-fix weights.  Half the MFMA work of helmholtz().
void helmholtz_split(const HArgs& a, hipStream_t st);
// flops of one helmholtz_split field (both contractions, complex columns)
double helmholtz_split_flops(int nz, long long P);

// Right-hand side, boundary rows and Helmholtz solve in one kernel, for up to
// three fields with the same kind of factorisation:
//   rhs = s2 h + s1 x - ((D2 x - kx^2 x) - ky^2 x)   (Predict / Correct rhs,
//   phase_carrier_flat.cc:517-536, phase_scalar_flat.cc:319-329), rows 0 and
//   nz-1 replaced by the boundary rows (carrier: extrapolated wall pressure
//   gradient, boundary_carrier_flat.cc:185-246; scalar: zero flux and the
//   erosion/deposition bottom row, phase_scalar_flat.cc:340-384), then
//   y = the Helmholtz solve of rhs (forward fix, V~^-1, 1/(lambda - k^2 + sigma),
//   V~, sequential backward fixes; the full form's V~ has the backward fixes
//   folded into its rows 0 and nz-1).
// D2 is applied through its even/odd split (nz odd).  Plain factorisations
// take V~^-1 with its columns permuted to the order (0..M, nz-1..M+1)
// (permute_columns); split ones take their even/odd form.  y may alias x.
struct RHField {
  const double2* x = nullptr;     // field (D2 operand, s1 term)
  double2* h = nullptr;           // explicit-term accumulator (updated in place when RHArgs::hupd)
  const void* adv = nullptr;      // nonlinear term (RHArgs::hupd): float2 if RHArgs::adv32, else double2
  double2* y = nullptr;           // solution
  int comp = 0;                   // carrier component (0..2) for the wall-row arrays
  const double* vinv = nullptr;   // permuted V~^-1, or pinv
  const double* vnorm = nullptr;  // V~ with the backward fixes folded in, or pnorm
  const double* lambda = nullptr; // lambda, or plam
  const double* fix = nullptr;    // natural 4*nz fix array
  const double* pfix = nullptr;   // split only
  double sigma = 0;
};
struct RHArgs {
  int nz = 0, P = 0, nyc = 0;
  const double* kx = nullptr;
  const double* ky = nullptr;
  const double* d2p = nullptr;    // D2 in its even/odd split (pack_parity), leading dimension ldp
  int ldp = 0;
  double s1 = 0, s2 = 0;
  int split = 0;                  // factorisations in their even/odd form
  int lda = 0;                    // leading dimension of vinv/vnorm
  int bc = 0;                     // 0 carrier rows, 1 scalar rows
  // carrier rows: k ((1+s) m1 - s m2), open-top tangential rows from n1/n2,
  // mean column top = top[comp]
  double k = 0, s = 0;
  const double2* m1b = nullptr; const double2* m1t = nullptr;
  const double2* m2b = nullptr; const double2* m2t = nullptr;
  const double2* n1t = nullptr; const double2* n2t = nullptr;
  int open_top = 0;
  double top[3] = {0, 0, 0};
  // scalar rows: bottom -(E - ws c_0 cbar) Re Sc / cbar 1e7 when erosion
  const double2* ehat = nullptr;
  double ws = 0, cbar = 1, re = 0, sc = 0;
  int erosion = 0;
  // explicit-term update folded in (hupd != 0), the arithmetic of h_update:
  // h <- c1 h - dt adv; buoyancy from the scalar (w: -K c cs; the slope
  // component: +K c sn); mean column: + dt f1 (u), + dt f2 (v), the force read
  // from fdev when set.  h is written back before it enters the rhs.
  int hupd = 0, adv32 = 0;
  double c1 = 0, dt = 0;
  int buoy = 0, slope_dir = 0, force = 0;
  double K = 0, sn = 0, cs = 0, f1 = 0, f2 = 0;
  const double* fdev = nullptr;
  const double2* cb = nullptr;    // scalar state (buoyancy source)
  RHField f[3];
};
void rhs_helmholtz(const RHArgs& a, int nfield, hipStream_t st);
// flops of one rhs_helmholtz field
double rhs_helmholtz_flops(int nz, long long P, bool split);
// V~^-1 (row-major, nz x lda) with its columns in the order 0..M, nz-1..M+1
std::vector<double> permute_columns(const std::vector<double>& rowmajor, int nz, int lda);

// Divergence, pressure solve and projection in one kernel (FlatCarrierPhase::
// Correct, phase_carrier_flat.cc:586-620, boundary_carrier_flat.cc:131-170):
//   rhs_p = q ((D1 sw + i kx su) + i ky sv), rows 0 and nz-1 zeroed;
//   p     = the pressure solve of rhs_p through the even/odd form of its
//           factorisation (mean column gated at nullrow), written to p;
//   u = su - cf i kx p, v = sv - cf i ky p, w = sw - cf D1 p (mean w := 0),
//   the wall rows of (i kx p, i ky p, D1 p) (and of (i kx, i ky) D1 p when
//   open_top) saved as Z_CORRECT does.
// D1 through its even/odd split (pack_parity, centro-antisymmetric).  s and y
// may alias (the projection in place).
struct PArgs {
  int nz = 0, P = 0, nyc = 0;
  const double* kx = nullptr;
  const double* ky = nullptr;
  const double* d1p = nullptr;    // D1 split, leading dimension ldp
  int ldp = 0;
  const double* pinv = nullptr;   // pressure factorisation, even/odd form (leading dimension lda)
  const double* pnorm = nullptr;
  const double* plam = nullptr;
  const double* pfix = nullptr;
  const double* fix = nullptr;    // natural 4*nz fix array
  int lda = 0, nullrow = -1;
  double q = 0, cf = 0;
  const double2* s = nullptr;     // su, sv, sw at s + f*S
  double2* p = nullptr;
  double2* y = nullptr;           // u, v, w at y + f*S
  long long S = 0;
  double2* wall_bot = nullptr; double2* wall_top = nullptr;
  double2* wall2_bot = nullptr; double2* wall2_top = nullptr;
  int open_top = 0;
};
void pressure_chain(const PArgs& a, hipStream_t st);
double pressure_chain_flops(int nz, long long P);

// Rows 0 and nz-1 of M x for every column (thin contraction: wall values of
// dU/dz, dV/dz used by the erosion map and the log features).
void zrows_wall(const double* op_rowmajor, int lda, int nz, const double2* x, long long P,
                double2* bottom, double2* top, hipStream_t st);

// out[2f] / out[2f+1] = Re (M x_f)(row 0 / row nz-1) on the mean column of
// field f = 0..2 (fields stride S): the wall mean-mode gradients of the log
// (phase_carrier_flat.cc:686-691).
void wall_mean_rows(const double* op_rowmajor, int lda, int nz, const double2* x, long long S, double* out,
                    hipStream_t st);

}  // namespace tbd
