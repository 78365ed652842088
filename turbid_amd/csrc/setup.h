// Host-side setup of the wall-normal operators: Chebyshev-Gauss-Lobatto grid,
// D1/D2 (reference src/stepper_space_chebyshev.cc:101-228) and the bordered
// eigen-factorisation used by every Helmholtz/Poisson solve
// (reference src/cmdz2_decomposition.cc:19-297).  Matrices are column-major
// n*n unless stated otherwise.
#pragma once

#include <vector>

namespace tbd {

struct Chebyshev {
  int nz = 0;
  std::vector<double> gridz;     // CGL points, bottom (z=-1) first
  std::vector<double> gridzmap;  // mapped grid (grid lengths)
  std::vector<double> d1, d2;    // column-major
};

// mapping: 0 None, 1 TwoEndTangent, 2 TwoEndSin, 3 OneEndTangent
Chebyshev make_chebyshev(int nz, int mapping, double coef);

// Eigen-decomposition of a real matrix whose spectrum is real.  Returns 0 on
// success, <0 if the QR iteration failed or the spectrum is complex.
int real_eigen(int n, const double* a, std::vector<double>& lambda, std::vector<double>& vec);
// Inverse by LU with partial pivoting; false if singular.
bool invert(int n, const double* a, double* ainv);

// CmDz2DecompositionCplx equivalent for one boundary-condition set.
struct Factorisation {
  int nz = 0;
  double vt[5] = {0, 0, 0, 0, 0};   // (a_bottom, b_bottom, a_top, b_top, top value)
  std::vector<double> vnorm, vinv;  // bordered V~, V~^-1 (column-major nz*nz)
  std::vector<double> lambda;       // nz, rows 0 and nz-1 unused
  std::vector<double> fix;          // 4*nz: forward bottom, forward top, backward bottom, backward top
  int nullrow = -1;                 // row whose reciprocal is zeroed on the mean column when sigma == 0

  // Even/odd form, present (pM > 0) when nz = 2M+1 and the interior operator
  // is centro-symmetric (the same boundary type at both walls on an
  // unmapped grid), so every interior eigenvector is even or odd under
  // i -> nz-1-i.  Column entries are split as s_i = x_i + x_{nz-1-i},
  // s_M = x_M (i < M) at split index i, d_i = x_i - x_{nz-1-i} at M+1+i.
  // Eigen-coordinates: split index 0 the even wall pair, 1..M the even
  // eigenvectors, M+1 the odd wall pair, M+2..2M the odd eigenvectors.
  //   pinv  (2M+2) x pld row-major: rows 0..M act on s (K = M+1), rows
  //         M+1..2M on d (K = M): the eigen-coordinates of the column.
  //   pnorm (2M+2) x pld row-major: rows 0..M (a) from the even
  //         coordinates, M+1..2M+1 (b) from the odd ones; the column is
  //         y_i = a_i + b_i, y_{nz-1-i} = a_i - b_i (i <= M, b_M = 0).
  //   plam  eigenvalue per split index;  pnull  split index of nullrow.
  //   pfix  backward-fix weights on (a, b): bottom at [0, nz), top at
  //         [nz, 2nz), each as (w_i + w_{nz-1-i}, w_M) then (w_i - w_{nz-1-i}).
  int pM = 0, pld = 0, pnull = -1;
  std::vector<double> pinv, pnorm, plam, pfix;
};

// Builds the factorisation from D1, D2 and the boundary rows a*f + b*df/dz.
Factorisation make_factorisation(const Chebyshev& ch, const double* vt);

// A factorisation supplied by the caller (the arrays of CmDz2DecompositionCplx:
// bordered V~, V~^-1 column-major nz*nz, lambda[nz], fix[4*nz]) for the
// boundary rows vt.  The null row follows the same rule as
// make_factorisation; the even/odd form is rebuilt from the interior blocks
// (operator V lambda V^-1, eigenvectors V) when they allow one.
Factorisation factorisation_from(int nz, const double* vt, const double* V, const double* Vinv,
                                 const double* lambda, const double* fix);

// V~ (column-major nz*nz) with the backward fixes folded into its rows 0 and
// nz-1 (cmdz2_decomposition.cc:54-80 builds them; phase_carrier_flat.cc:
// 984-985 and phase_scalar_flat.cc:659-660 apply them to Y = V~ Z, the top
// after the bottom): row 0 becomes
// wb V~ and row nz-1 becomes wt V~ + wt_0 (wb V~ - V~_0), wb, wt = fix rows
// 2 and 3, so that one contraction with it gives the fixed Y.
std::vector<double> fold_backward_fixes(const std::vector<double>& vnorm, const std::vector<double>& fix, int nz);

}  // namespace tbd
