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
};

// Builds the factorisation from D1, D2 and the boundary rows a*f + b*df/dz.
Factorisation make_factorisation(const Chebyshev& ch, const double* vt);

}  // namespace tbd
