/*
 * turbid_oracle.cpp -- FP64 CPU restatement of TURBID's pseudospectral
 * Fourier-Chebyshev RK3 substep (carrier + sediment scalar).
 *
 * TEST INFRASTRUCTURE ONLY: the correctness oracle and the timed CPU
 * baseline.  The product (turbid_amd) never links or calls this code.
 *
 * Every function cites the reference file:line it restates.  The reference's
 * third-party arithmetic (P3DFFT/FFTW transforms, Armadillo/BLAS products,
 * LAPACK eig_gen/inv) is restated from its published definitions:
 *   - DFT conventions: FFTW forward exp(-2 pi i jk/n), backward exp(+), both
 *     unnormalised; c2r uses only the real part of X_0 (halfcomplex format);
 *   - P3DFFT truncation/padding: zip:build/reorder.F90:180-186, 226-228, 304-433;
 *   - eig_gen: balancing, Hessenberg reduction and Francis double-shift QR
 *     for the eigenvalues, inverse iteration for the eigenvectors (any
 *     scaling/order: V (R o V^-1 .) is invariant under both).
 * Parallelism: OpenMP over z-planes (transforms) and spectral columns
 * (z-contractions), as the reference's library threading does.
 */
#include "turbid_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::complex<double> cx;
typedef std::vector<double> dvec;
typedef std::vector<cx> cvec;

thread_local std::string g_err;

const double kPi = 3.14159265358979323846264338327950288;
const double kCompareSmall = 10E-12;          // common.h:42
const double kBoundaryCoefScale = 10E+6;      // common.h:55

// ---------------------------------------------------------------------------
// Mixed-radix Stockham FFT (complex, unnormalised).  sign=-1 forward.
// ---------------------------------------------------------------------------
struct FftPlan {
  int n = 0;
  std::vector<int> radices;
  cvec tw;  // tw[t] = exp(-2 pi i t / n)
};

FftPlan MakePlan(int n) {
  FftPlan p;
  p.n = n;
  int m = n;
  while (m % 4 == 0) { p.radices.push_back(4); m /= 4; }
  while (m % 2 == 0) { p.radices.push_back(2); m /= 2; }
  for (int f = 3; m > 1; f += 2) {
    while (m % f == 0) { p.radices.push_back(f); m /= f; }
  }
  p.tw.resize(n);
  for (int t = 0; t < n; ++t) {
    const double a = -2.0 * kPi * (double)t / (double)n;
    p.tw[t] = cx(std::cos(a), std::sin(a));
  }
  return p;
}

// x: input/output (n), work: scratch (n)
void FftExec(const FftPlan& p, cx* x, cx* work, int sign) {
  const int n = p.n;
  if (n == 1) return;
  cx* in = x;
  cx* out = work;
  int ns = 1;
  cx v[64];
  cx w[64];
  for (int R : p.radices) {
    const int m = n / R;
    const int step = n / (ns * R);
    // roots of unity of order R
    cx root[64];
    for (int q = 0; q < R; ++q) {
      root[q] = p.tw[(long long)q * (n / R) % n];
      if (sign > 0) root[q] = std::conj(root[q]);
    }
    for (int j = 0; j < m; ++j) {
      const int k = j % ns;
      for (int r = 0; r < R; ++r) {
        cx t = p.tw[(long long)r * k * step % n];
        if (sign > 0) t = std::conj(t);
        v[r] = in[j + r * m] * t;
      }
      for (int q = 0; q < R; ++q) {
        cx s = 0.0;
        for (int r = 0; r < R; ++r) s += v[r] * root[(r * q) % R];
        w[q] = s;
      }
      const int idxd = (j / ns) * ns * R + k;
      for (int r = 0; r < R; ++r) out[idxd + r * ns] = w[r];
    }
    std::swap(in, out);
    ns *= R;
  }
  if (in != x) std::memcpy(x, in, sizeof(cx) * n);
}

// ---------------------------------------------------------------------------
// Plane transforms with P3DFFT's op "ffn"/"nff" truncation semantics.
// ---------------------------------------------------------------------------
struct PlaneFft {
  int nx, ny, nz, nxc, nyc, mx;
  FftPlan px, py;
  PlaneFft(int nx_, int ny_, int nz_) : nx(nx_), ny(ny_), nz(nz_) {
    nxc = nx / 3 * 2;            // base.cc:233
    nyc = ny / 3 * 2;            // base.cc:234
    mx = nxc / 2 + 1;            // base.cc:256 (nxhpc, setup.F90:113-115)
    px = MakePlan(nx);
    py = MakePlan(ny);
  }
  // spectral y index iy (0..nyc-1) -> full FFT index (reorder.F90:304-358)
  int YFull(int iy) const { return iy < nyc / 2 ? iy : iy + ny - nyc; }

  // RealCplxCube::DftForward (cube_real_cplx.cc:525-539): per z-plane, x r2c,
  // keep kx < mx; y c2c, keep nyc modes; scale 1/(Nx*Ny) (:536).
  void Forward(const double* phys, cx* spec, int nv) const {
    const long long R = (long long)nx * ny * nz;
    const long long S = (long long)nz * nyc * mx;
    const double scale = 1.0 / (double)((long long)nx * ny);
#pragma omp parallel
    {
      cvec row(nx), work(std::max(nx, ny)), col(ny), tmat((size_t)mx * ny);
#pragma omp for collapse(2) schedule(static)
      for (int d = 0; d < nv; ++d) {
        for (int z = 0; z < nz; ++z) {
          const double* pl = phys + d * R + (long long)z * nx * ny;
          // x r2c of two rows at once: z = a + i b
          for (int y = 0; y < ny; y += 2) {
            const double* a = pl + (long long)y * nx;
            const double* b = (y + 1 < ny) ? pl + (long long)(y + 1) * nx : nullptr;
            for (int x = 0; x < nx; ++x) row[x] = cx(a[x], b ? b[x] : 0.0);
            FftExec(px, row.data(), work.data(), -1);
            for (int k = 0; k < mx; ++k) {
              const cx zk = row[k];
              const cx zn = std::conj(row[(nx - k) % nx]);
              tmat[(size_t)k * ny + y] = 0.5 * (zk + zn);
              if (b) tmat[(size_t)k * ny + y + 1] = cx(0.0, -0.5) * (zk - zn);
            }
          }
          for (int k = 0; k < mx; ++k) {
            for (int y = 0; y < ny; ++y) col[y] = tmat[(size_t)k * ny + y];
            FftExec(py, col.data(), work.data(), -1);
            for (int iy = 0; iy < nyc; ++iy) {
              spec[d * S + z + (long long)nz * (iy + (long long)nyc * k)] = col[YFull(iy)] * scale;
            }
          }
        }
      }
    }
  }

  // RealCplxCube::DftBackward (cube_real_cplx.cc:542-555): y zero-pad
  // (reorder.F90:180-186) + y c2c^-1, x zero-pad (:226-228) + x c2r,
  // unnormalised (btran.F90:118-243).
  void Backward(const cx* spec, double* phys, int nv) const {
    const long long R = (long long)nx * ny * nz;
    const long long S = (long long)nz * nyc * mx;
#pragma omp parallel
    {
      cvec row(nx), work(std::max(nx, ny)), col(ny), tmat((size_t)mx * ny);
#pragma omp for collapse(2) schedule(static)
      for (int d = 0; d < nv; ++d) {
        for (int z = 0; z < nz; ++z) {
          for (int k = 0; k < mx; ++k) {
            std::fill(col.begin(), col.end(), cx(0.0, 0.0));
            for (int iy = 0; iy < nyc; ++iy) {
              col[YFull(iy)] = spec[d * S + z + (long long)nz * (iy + (long long)nyc * k)];
            }
            FftExec(py, col.data(), work.data(), +1);
            for (int y = 0; y < ny; ++y) tmat[(size_t)k * ny + y] = col[y];
          }
          double* pl = phys + d * R + (long long)z * nx * ny;
          for (int y = 0; y < ny; y += 2) {
            const bool two = (y + 1 < ny);
            std::fill(row.begin(), row.end(), cx(0.0, 0.0));
            for (int k = 0; k < mx; ++k) {
              cx a = tmat[(size_t)k * ny + y];
              cx b = two ? tmat[(size_t)k * ny + y + 1] : cx(0.0, 0.0);
              if (k == 0 || 2 * k == nx) {  // halfcomplex: imaginary part not stored
                a = cx(a.real(), 0.0);
                b = cx(b.real(), 0.0);
              }
              row[k] += a + cx(0.0, 1.0) * b;
              if (k > 0 && 2 * k != nx) {
                row[nx - k] += std::conj(a) + cx(0.0, 1.0) * std::conj(b);
              }
            }
            FftExec(px, row.data(), work.data(), +1);
            double* a = pl + (long long)y * nx;
            for (int x = 0; x < nx; ++x) a[x] = row[x].real();
            if (two) {
              double* b = pl + (long long)(y + 1) * nx;
              for (int x = 0; x < nx; ++x) b[x] = row[x].imag();
            }
          }
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Dense linear algebra for the setup (eig_gen / inv replacements).
// Column-major n*n: A(i,j) = a[i + n*j].
// ---------------------------------------------------------------------------
// LU with partial pivoting, in place.  Returns false if singular.
bool LuFactor(int n, double* a, int* piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double mx = std::fabs(a[k + (size_t)n * k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = std::fabs(a[i + (size_t)n * k]);
      if (v > mx) { mx = v; p = i; }
    }
    piv[k] = p;
    if (mx == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(a[k + (size_t)n * j], a[p + (size_t)n * j]);
    }
    const double inv = 1.0 / a[k + (size_t)n * k];
    for (int i = k + 1; i < n; ++i) a[i + (size_t)n * k] *= inv;
    for (int j = k + 1; j < n; ++j) {
      const double akj = a[k + (size_t)n * j];
      if (akj == 0.0) continue;
      double* cj = a + (size_t)n * j;
      const double* ck = a + (size_t)n * k;
      for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * akj;
    }
  }
  return true;
}

void LuSolve(int n, const double* lu, const int* piv, double* b) {
  for (int k = 0; k < n; ++k) {
    if (piv[k] != k) std::swap(b[k], b[piv[k]]);
  }
  for (int j = 0; j < n; ++j) {
    const double bj = b[j];
    if (bj == 0.0) continue;
    for (int i = j + 1; i < n; ++i) b[i] -= lu[i + (size_t)n * j] * bj;
  }
  for (int j = n - 1; j >= 0; --j) {
    b[j] /= lu[j + (size_t)n * j];
    const double bj = b[j];
    for (int i = 0; i < j; ++i) b[i] -= lu[i + (size_t)n * j] * bj;
  }
}

// inv(A) by LU (armadillo inv, cmdz2_decomposition.cc:110)
bool Inverse(int n, const double* a, double* ainv) {
  dvec lu(a, a + (size_t)n * n);
  std::vector<int> piv(n);
  if (!LuFactor(n, lu.data(), piv.data())) return false;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j) {
    dvec e(n, 0.0);
    e[j] = 1.0;
    LuSolve(n, lu.data(), piv.data(), e.data());
    std::memcpy(ainv + (size_t)n * j, e.data(), sizeof(double) * n);
  }
  return true;
}

// Eigenvalues of a real matrix: balance, Householder Hessenberg reduction,
// Francis double-shift QR.  Returns false on non-convergence.
bool EigenValues(int n, const double* a_in, dvec& wr, dvec& wi) {
  dvec a(a_in, a_in + (size_t)n * n);
  auto A = [&](int i, int j) -> double& { return a[i + (size_t)n * j]; };
  // Parlett-Reinsch balancing by powers of 2
  const double radix = 2.0;
  bool done = false;
  while (!done) {
    done = true;
    for (int i = 0; i < n; ++i) {
      double r = 0.0, c = 0.0;
      for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        c += std::fabs(A(j, i));
        r += std::fabs(A(i, j));
      }
      if (c != 0.0 && r != 0.0) {
        double g = r / radix, f = 1.0;
        const double s = c + r;
        while (c < g) { f *= radix; c *= radix * radix; }
        g = r * radix;
        while (c > g) { f /= radix; c /= radix * radix; }
        if ((c + r) / f < 0.95 * s) {
          done = false;
          g = 1.0 / f;
          for (int j = 0; j < n; ++j) A(i, j) *= g;
          for (int j = 0; j < n; ++j) A(j, i) *= f;
        }
      }
    }
  }
  // Householder reduction to upper Hessenberg form
  dvec v(n);
  for (int k = 0; k < n - 2; ++k) {
    double nrm = 0.0;
    for (int i = k + 1; i < n; ++i) nrm += A(i, k) * A(i, k);
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) continue;
    const double alpha = A(k + 1, k) > 0 ? -nrm : nrm;
    double vn2 = 0.0;
    for (int i = k + 1; i < n; ++i) { v[i] = A(i, k); }
    v[k + 1] -= alpha;
    for (int i = k + 1; i < n; ++i) vn2 += v[i] * v[i];
    if (vn2 == 0.0) continue;
    const double beta = 2.0 / vn2;
    for (int j = k; j < n; ++j) {
      double s = 0.0;
      for (int i = k + 1; i < n; ++i) s += v[i] * A(i, j);
      s *= beta;
      for (int i = k + 1; i < n; ++i) A(i, j) -= s * v[i];
    }
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += A(i, j) * v[j];
      s *= beta;
      for (int j = k + 1; j < n; ++j) A(i, j) -= s * v[j];
    }
    for (int i = k + 2; i < n; ++i) A(i, k) = 0.0;
  }
  // Francis double-shift QR on the Hessenberg matrix
  wr.assign(n, 0.0);
  wi.assign(n, 0.0);
  double anorm = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = std::max(i - 1, 0); j < n; ++j) anorm += std::fabs(A(i, j));
  int nn = n - 1;
  double t = 0.0;
  while (nn >= 0) {
    int its = 0;
    for (;;) {
      int l;
      for (l = nn; l >= 1; --l) {
        double s = std::fabs(A(l - 1, l - 1)) + std::fabs(A(l, l));
        if (s == 0.0) s = anorm;
        if (std::fabs(A(l, l - 1)) + s == s) { A(l, l - 1) = 0.0; break; }
      }
      double x = A(nn, nn);
      if (l == nn) {
        wr[nn] = x + t;
        wi[nn] = 0.0;
        --nn;
        break;
      }
      double y = A(nn - 1, nn - 1);
      double w = A(nn, nn - 1) * A(nn - 1, nn);
      if (l == nn - 1) {
        const double p = 0.5 * (y - x);
        const double q = p * p + w;
        double z = std::sqrt(std::fabs(q));
        x += t;
        if (q >= 0.0) {
          z = p + (p >= 0 ? z : -z);
          wr[nn - 1] = wr[nn] = x + z;
          if (z != 0.0) wr[nn] = x - w / z;
          wi[nn - 1] = wi[nn] = 0.0;
        } else {
          wr[nn - 1] = wr[nn] = x + p;
          wi[nn - 1] = -z;
          wi[nn] = z;
        }
        nn -= 2;
        break;
      }
      if (its == 60) return false;
      if (its == 10 || its == 20) {
        t += x;
        for (int i = 0; i <= nn; ++i) A(i, i) -= x;
        const double s = std::fabs(A(nn, nn - 1)) + std::fabs(A(nn - 1, nn - 2));
        y = x = 0.75 * s;
        w = -0.4375 * s * s;
      }
      ++its;
      int m;
      double p = 0, q = 0, r = 0, z = 0;
      for (m = nn - 2; m >= l; --m) {
        z = A(m, m);
        r = x - z;
        double s = y - z;
        p = (r * s - w) / A(m + 1, m) + A(m, m + 1);
        q = A(m + 1, m + 1) - z - r - s;
        r = A(m + 2, m + 1);
        s = std::fabs(p) + std::fabs(q) + std::fabs(r);
        p /= s; q /= s; r /= s;
        if (m == l) break;
        const double u = std::fabs(A(m, m - 1)) * (std::fabs(q) + std::fabs(r));
        const double vv = std::fabs(p) * (std::fabs(A(m - 1, m - 1)) + std::fabs(z) + std::fabs(A(m + 1, m + 1)));
        if (u + vv == vv) break;
      }
      for (int i = m + 2; i <= nn; ++i) {
        A(i, i - 2) = 0.0;
        if (i != m + 2) A(i, i - 3) = 0.0;
      }
      for (int k = m; k <= nn - 1; ++k) {
        if (k != m) {
          p = A(k, k - 1);
          q = A(k + 1, k - 1);
          r = 0.0;
          if (k != nn - 1) r = A(k + 2, k - 1);
          x = std::fabs(p) + std::fabs(q) + std::fabs(r);
          if (x != 0.0) { p /= x; q /= x; r /= x; }
        }
        const double sq = std::sqrt(p * p + q * q + r * r);
        const double s = p >= 0 ? sq : -sq;
        if (s != 0.0) {
          if (k == m) {
            if (l != m) A(k, k - 1) = -A(k, k - 1);
          } else {
            A(k, k - 1) = -s * x;
          }
          p += s;
          x = p / s; y = q / s; z = r / s;
          q /= p; r /= p;
          for (int j = k; j <= nn; ++j) {
            p = A(k, j) + q * A(k + 1, j);
            if (k != nn - 1) { p += r * A(k + 2, j); A(k + 2, j) -= p * z; }
            A(k + 1, j) -= p * y;
            A(k, j) -= p * x;
          }
          const int mmin = nn < k + 3 ? nn : k + 3;
          for (int i = l; i <= mmin; ++i) {
            p = x * A(i, k) + y * A(i, k + 1);
            if (k != nn - 1) { p += z * A(i, k + 2); A(i, k + 2) -= p * r; }
            A(i, k + 1) -= p * q;
            A(i, k) -= p;
          }
        }
      }
    }
  }
  return true;
}

// Real eigen-decomposition: eigenvalues by QR, eigenvectors by inverse
// iteration on the original matrix (unit 2-norm, largest entry positive).
// Returns 0, -1 on QR failure, -2 on complex spectrum, -3 singular shift.
int EigenDecompose(int n, const double* a, dvec& lambda, dvec& vec) {
  dvec wr, wi;
  if (!EigenValues(n, a, wr, wi)) return -1;
  double anorm = 0.0;
  for (size_t i = 0; i < (size_t)n * n; ++i) anorm = std::max(anorm, std::fabs(a[i]));
  for (int i = 0; i < n; ++i) {
    if (std::fabs(wi[i]) > 1e-8 * std::max(1.0, std::fabs(wr[i]))) return -2;
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) { return wr[x] > wr[y]; });
  lambda.resize(n);
  for (int i = 0; i < n; ++i) lambda[i] = wr[order[i]];
  vec.assign((size_t)n * n, 0.0);
  int status = 0;
#pragma omp parallel for schedule(dynamic)
  for (int e = 0; e < n; ++e) {
    dvec b((size_t)n * n);
    std::vector<int> piv(n);
    const double lam = lambda[e];
    double shift = lam + (std::fabs(lam) + anorm) * 1e-13;
    std::memcpy(b.data(), a, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i) b[i + (size_t)n * i] -= shift;
    if (!LuFactor(n, b.data(), piv.data())) {
#pragma omp atomic write
      status = -3;
      continue;
    }
    dvec x(n);
    for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.01 * std::sin(1.0 + i);
    for (int it = 0; it < 3; ++it) {
      LuSolve(n, b.data(), piv.data(), x.data());
      double nrm = 0.0;
      for (int i = 0; i < n; ++i) nrm += x[i] * x[i];
      nrm = std::sqrt(nrm);
      for (int i = 0; i < n; ++i) x[i] /= nrm;
    }
    int imax = 0;
    for (int i = 1; i < n; ++i) if (std::fabs(x[i]) > std::fabs(x[imax])) imax = i;
    const double sg = x[imax] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) vec[i + (size_t)n * e] = sg * x[i];
  }
  return status;
}

// ---------------------------------------------------------------------------
// Chebyshev operators (stepper_space_chebyshev.cc:101-228, domain.cc:306).
// ---------------------------------------------------------------------------
void Chebyshev(int nz, int mapping, double coef, dvec& gridz, dvec& d1, dvec& d2) {
  gridz.resize(nz);
  for (int id = 0; id < nz; ++id) gridz[id] = std::cos((nz - id - 1) * kPi / (nz - 1));  // domain.cc:306
  d1.assign((size_t)nz * nz, 0.0);
  auto D = [&](int i, int j) -> double& { return d1[i + (size_t)nz * j]; };
  for (int iz = 1; iz < nz; ++iz)                                            // :122-126
    for (int ix = 0; ix < iz; ++ix)
      D(iz, ix) = std::pow(-1.0, iz + ix) / (gridz[iz] - gridz[ix]);
  for (int i = 0; i < nz; ++i) D(i, 0) *= 0.5;                              // :128
  for (int j = 0; j < nz; ++j) D(nz - 1, j) *= 2.0;                         // :129
  for (int iz = 0; iz < nz - 1; ++iz)                                        // :130-134
    for (int ix = iz + 1; ix < nz; ++ix) D(iz, ix) = -D(nz - iz - 1, nz - ix - 1);
  for (int iz = 0; iz < nz; ++iz) {                                          // :139-145
    D(iz, iz) = 0.0;
    for (int ix = 0; ix < nz; ++ix) {
      if (ix == iz) continue;
      D(iz, iz) = D(iz, iz) - D(iz, ix);
    }
  }
  if (mapping != 0) {                                                        // :150-186
    dvec scale(nz);
    for (int i = 0; i < nz; ++i) {
      const double z = gridz[i];
      if (mapping == 1) {
        const double t = std::tan(kPi / 2 * z);
        scale[i] = 1 / coef * (1 + std::pow(coef * t, 2)) / (1 + std::pow(t, 2));
      } else if (mapping == 2) {
        scale[i] = 1 / coef * std::cos(std::asin(coef * z)) * std::asin(coef);
      } else {
        const double t = std::tan(kPi / 4 * (z - 1));
        scale[i] = 1 / coef * (1 + std::pow(coef * t, 2)) / (1 + std::pow(t, 2));
      }
    }
    for (int j = 0; j < nz; ++j)
      for (int i = 0; i < nz; ++i) D(i, j) *= scale[i];
  }
  d2.assign((size_t)nz * nz, 0.0);                                          // :221
#pragma omp parallel for schedule(static)
  for (int j = 0; j < nz; ++j)
    for (int k = 0; k < nz; ++k) {
      const double b = d1[k + (size_t)nz * j];
      for (int i = 0; i < nz; ++i) d2[i + (size_t)nz * j] += d1[i + (size_t)nz * k] * b;
    }
  for (int iz = 0; iz < nz; ++iz) {                                          // :222-228
    double& dd = d2[iz + (size_t)nz * iz];
    dd = 0.0;
    for (int ix = 0; ix < nz; ++ix) {
      if (ix == iz) continue;
      dd = dd - d2[iz + (size_t)nz * ix];
    }
  }
}

// Mapped grid used for the z grid lengths (stepper_space_chebyshev.cc:155-208)
dvec MappedGrid(const dvec& gridz, int mapping, double coef) {
  const int nz = (int)gridz.size();
  dvec g(nz);
  for (int i = 0; i < nz; ++i) {
    const double z = gridz[i];
    if (mapping == 1) g[i] = 2 / kPi * std::atan(coef * std::tan(kPi / 2 * z));
    else if (mapping == 2) g[i] = std::asin(coef * z) / std::asin(coef);
    else if (mapping == 3) g[i] = 4 / kPi * std::atan(coef * std::tan(kPi / 4 * (z - 1))) + 1;
    else g[i] = z;
  }
  return g;
}

// ---------------------------------------------------------------------------
// CmDz2DecompositionCplx (cmdz2_decomposition.cc:19-297)
// ---------------------------------------------------------------------------
// Bartnicki positivity filter on physical z-planes (phase_scalar_flat.cc:
// 419-478): five passes; each pass sums per plane the positive values and
// their count, the zeros and the negative values; a plane whose total is
// negative is zeroed, otherwise the positive points share the negative mass
// equally (v += sum_neg / n_pos) and every other point is set to 0.
void PositivityFilter(double* phys, long long plane, int nplanes) {
  for (int it = 0; it < 5; ++it) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nplanes; ++z) {
      double* pl = phys + z * plane;
      double m[6] = {0, 0, 0, 0, 0, 0};
      for (long long e = 0; e < plane; ++e) {
        const double v = pl[e];
        if (v > 0) { m[0] += v; m[1] += 1; }
        else if (v < 0) { m[4] += v; m[5] += 1; }
        else { m[2] += v; m[3] += 1; }
      }
      if (m[0] + m[2] + m[4] < 0) {
        for (long long e = 0; e < plane; ++e) pl[e] = 0;
        continue;
      }
      for (long long e = 0; e < plane; ++e) {
        if (pl[e] > 0) pl[e] += m[4] / m[1];
        else pl[e] = 0;
      }
    }
  }
}

struct Decomposition {
  int n = 0;
  dvec vt;              // zboundarytype (5)
  dvec vnorm, vinv;     // bordered V~, V~^-1 (n*n col-major)
  dvec lambda;          // n (rows 0 and n-1 unused)
  dvec fbf, ftf, fbb, fbt;
  double sigma[3] = {0, 0, 0};

  int Construct(const dvec& d1, const dvec& d2, const double* zbt, const double* stepcoef) {
    const int N = n;
    vt.assign(zbt, zbt + 5);
    dvec c(d2);                                               // :33
    auto C = [&](int i, int j) -> double& { return c[i + (size_t)N * j]; };
    for (int j = 0; j < N; ++j) {                             // :47-48
      C(0, j) = zbt[1] * d1[0 + (size_t)N * j];
      C(N - 1, j) = zbt[3] * d1[(N - 1) + (size_t)N * j];
    }
    C(0, 0) += zbt[0];                                        // :49
    C(N - 1, N - 1) += zbt[2];                                // :50
    const double a11 = C(0, 0), ann = C(N - 1, N - 1), a1n = C(0, N - 1), an1 = C(N - 1, 0);
    const double den = 1.0 / (an1 * a1n - a11 * ann);         // :58
    fbf.assign(N, 0.0); ftf.assign(N, 0.0); fbb.assign(N, 0.0); fbt.assign(N, 0.0);
    for (int ix = 1; ix < N - 1; ++ix) {                      // :69-76
      fbb[ix] = (ann * C(0, ix) - a1n * C(N - 1, ix)) * den;
      fbt[ix] = (a11 * C(N - 1, ix) - an1 * C(0, ix)) * den;
      fbf[ix] = (ann * C(ix, 0) - an1 * C(ix, N - 1)) * den;
      ftf[ix] = (a11 * C(ix, N - 1) - a1n * C(ix, 0)) * den;
    }
    fbb[0] = -1 * ann * den;                                  // :77-80
    fbb[N - 1] = +1 * a1n * den;
    fbt[0] = +1 * an1 * den;
    fbt[N - 1] = -1 * a11 * den;
    const int m = N - 2;                                      // :85-93
    dvec fwd((size_t)m * m);
    for (int iz = 1; iz < N - 1; ++iz)
      for (int ix = 1; ix < N - 1; ++ix)
        fwd[(ix - 1) + (size_t)m * (iz - 1)] = C(ix, iz) + C(ix, 0) * fbb[iz] + C(ix, N - 1) * fbt[iz];
    dvec lam, vec;                                            // :97-100 (eig_gen, imag dropped)
    const int st = EigenDecompose(m, fwd.data(), lam, vec);
    if (st != 0) return st;
    dvec vi((size_t)m * m);
    if (!Inverse(m, vec.data(), vi.data())) return -4;        // :110
    vnorm.assign((size_t)N * N, 0.0);                         // :103-112
    vinv.assign((size_t)N * N, 0.0);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        vnorm[(i + 1) + (size_t)N * (j + 1)] = vec[i + (size_t)m * j];
        vinv[(i + 1) + (size_t)N * (j + 1)] = vi[i + (size_t)m * j];
      }
    vnorm[0] = vinv[0] = 1.0;
    vnorm[(N - 1) + (size_t)N * (N - 1)] = vinv[(N - 1) + (size_t)N * (N - 1)] = 1.0;
    lambda.assign(N, 0.0);                                    // :138-139
    for (int i = 0; i < m; ++i) lambda[i + 1] = lam[i];
    for (int s = 0; s < 3; ++s) sigma[s] = stepcoef[s];
    return 0;
  }

  // Reciprocal R(i, col) (Assemble, cmdz2_decomposition.cc:248-291).  The
  // pressure null mode (sigma==0, column of the mean mode) is gated on the
  // row whose eigenvalue has the smallest magnitude when that magnitude is at
  // round-off level relative to the spectrum (the reference's absolute
  // 1e-11 gate, :281-285, applied at the precision the eigensolver resolves).
  int NullRow() const {
    int best = -1;
    double bmin = 1e300, bmax = 0.0;
    for (int i = 1; i < n - 1; ++i) {
      bmax = std::max(bmax, std::fabs(lambda[i]));
      if (std::fabs(lambda[i]) < bmin) { bmin = std::fabs(lambda[i]); best = i; }
    }
    if (bmin < kCompareSmall || bmin < 1e-9 * bmax) return best;
    return -1;
  }
};

// ---------------------------------------------------------------------------
// The solver state (variable.h:70-85, variable_carrier_flat.cc:60-77,
// variable_scalar_flat.cc:30-34) and the substage.
// ---------------------------------------------------------------------------
struct Oracle {
  tbo_config cfg;
  int Nx, Ny, Nz, nxc, nyc, mx, P;
  long long S, R;
  PlaneFft* fft = nullptr;
  dvec gridz, d1, d2;
  dvec kx, ky;                 // multipliers per spectral x / y index (quirk: +ky at iy=nyc/2)
  dvec glx_, gly_;             // grid lengths (domain.cc:309-311, 318-319)
  dvec glz;                    // per z (domain.cc:320-332)
  Decomposition dc[5];         // u1, u2, u3, p, scalar
  bool opentop = false;
  double coef[3][3];           // stepper_time_rk3.cc:29-31, coef[id][stage]

  // carrier (phase 0)
  cvec uvw, p, huvw, suvw, lapuvw, rhsuvw, rhsp, divs, advuvw;
  cvec dudxyz, dvdxyz, dwdxyz, dpdxyz, dudxyztmp, dvdxyztmp, dwdxyztmp, ptmp0;
  dvec uvwnl;                  // uvwnonlinear (physical)
  dvec dud_r, dvd_r, dwd_r, dudt_r, dvdt_r, dwdt_r, adv_r;  // physical parts
  cvec dp1m1top, dp1m1bot, dp1m2top, dp1m2bot, dp2m1top, dp2m1bot, dp2m2top, dp2m2bot;
  double olddt = 0;
  // scalar (phase 1)
  cvec c, hphi, lapphi, rhsphi, advphi, uvw1, ptmp1, dudxyz1, dvdxyz1;
  dvec uvwnl1, ptmp1_r, dudxyz1_r, dvdxyz1_r, advphi_r;
  // stepper
  int step = 0;
  double time = 0, dt = 0, cfl = 0;
  bool forward = true, alternation = false;
  double featlocal[50], featglobal[50];

  explicit Oracle(const tbo_config& c_) : cfg(c_) {
    Nx = cfg.nx; Ny = cfg.ny; Nz = cfg.nz;
    if (Nx < 2 || Ny < 2 || Nz < 4) throw std::runtime_error("Oracle > grid too small");
    if (std::fabs(cfg.lz - 2) > kCompareSmall)                // domain.cc:55-58
      throw std::runtime_error("Domain(mpi_t,dcp_t,cfg_t) > Dimensionless z-length Should be Equal to Two!");
    fft = new PlaneFft(Nx, Ny, Nz);
    nxc = fft->nxc; nyc = fft->nyc; mx = fft->mx;
    P = nyc * mx;
    S = (long long)Nz * P;
    R = (long long)Nx * Ny * Nz;
    dt = cfg.dt;
    const double c1[3] = {0.0 / 1.0, -5.0 / 9.0, -153.0 / 128.0};
    const double c2[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
    const double c3[3] = {1.0 / 6.0, 5.0 / 24.0, 1.0 / 8.0};
    for (int s = 0; s < 3; ++s) { coef[0][s] = c1[s]; coef[1][s] = c2[s]; coef[2][s] = c3[s]; }
    std::memset(featlocal, 0, sizeof(featlocal));
    std::memset(featglobal, 0, sizeof(featglobal));
    Configure();
  }
  ~Oracle() { delete fft; }

  // SpaceStepper::Configure multipliers (stepper_space.cc:181-213)
  void Multipliers() {
    const double kx1 = 2 * kPi / cfg.lx, ky1 = 2 * kPi / cfg.ly;
    kx.resize(mx);
    ky.resize(nyc);
    for (int ix = 0; ix < mx; ++ix) kx[ix] = kx1 * (ix - 0);
    for (int iy = 0; iy < nyc; ++iy) ky[iy] = (iy <= nyc / 2) ? +ky1 * (iy - 0) : -ky1 * (nyc - iy);
  }
  double KX(long long col) const { return kx[(col % P) / nyc]; }
  double KY(long long col) const { return ky[(col % P) % nyc]; }
  double K2(long long col) const {   // cmdx2 + cmdy2 (stepper_space.cc:204,209)
    const double a = KX(col), b = KY(col);
    return (-std::pow(a, 2)) + (-std::pow(b, 2));
  }

  void Configure() {
    Chebyshev(Nz, cfg.mapping, cfg.mapping_coefficient, gridz, d1, d2);
    Multipliers();
    // grid lengths (domain.cc:304-332, stepper_space_chebyshev.cc:191-209)
    dvec gx(Nx), gy(Ny);
    for (int id = 0; id < Nx; ++id) gx[id] = cfg.lx / Nx * id;
    for (int id = 0; id < Ny; ++id) gy[id] = cfg.ly / Ny * id;
    double mxl = 1e300, myl = 1e300;
    for (int i = 0; i + 1 < Nx; ++i) mxl = std::min(mxl, gx[i + 1] - gx[i]);
    for (int i = 0; i + 1 < Ny; ++i) myl = std::min(myl, gy[i + 1] - gy[i]);
    glx_.assign(1, mxl);
    gly_.assign(1, myl);
    const dvec g = MappedGrid(gridz, cfg.mapping, cfg.mapping_coefficient);
    glz.resize(Nz);
    for (int id = 0; id < Nz; ++id) {
      if (id == 0) glz[id] = (g[1] - g[0]) / 2;
      else if (id == Nz - 1) glz[id] = (g[Nz - 1] - g[Nz - 2]) / 2;
      else glz[id] = std::min(g[id + 1] - g[id], g[id] - g[id - 1]);
    }
    // FlatCarrierPhase::Configure (phase_carrier_flat.cc:186-204)
    double sc[3];
    for (int s = 0; s < 3; ++s) sc[s] = -cfg.reynolds / dt / coef[2][s];
    for (int k = 0; k < 5; ++k) dc[k].n = Nz;
    int st = dc[0].Construct(d1, d2, cfg.vt_u1, sc);
    if (st == 0) {
      if (std::equal(cfg.vt_u1, cfg.vt_u1 + 4, cfg.vt_u2)) { dc[1] = dc[0]; dc[1].vt.assign(cfg.vt_u2, cfg.vt_u2 + 5); }
      else st = dc[1].Construct(d1, d2, cfg.vt_u2, sc);
    }
    if (st == 0) st = dc[2].Construct(d1, d2, cfg.vt_u3, sc);
    double zero[3] = {0, 0, 0};
    if (st == 0) st = dc[3].Construct(d1, d2, cfg.vt_p, zero);
    if (st == 0 && cfg.nphase > 1) {   // FlatScalarPhase::Configure (phase_scalar_flat.cc:184-200)
      double ss[3];
      for (int s = 0; s < 3; ++s) ss[s] = -cfg.reynolds * cfg.schmidt / dt / coef[2][s];
      double bc[5] = {cfg.reynolds * cfg.schmidt * cfg.settling_velocity, 1,
                      cfg.reynolds * cfg.schmidt * cfg.settling_velocity, 1, 0};
      for (int i = 0; i < 4; ++i) bc[i] = bc[i] * kBoundaryCoefScale;
      st = dc[4].Construct(d1, d2, bc, ss);
    }
    if (st != 0) throw std::runtime_error("CmDz2DecompositionCplx::Construct > eigen-decomposition failed");
    // Boundary::SetFlagOpenTop (boundary.cc:59-71)
    opentop = std::fabs(cfg.vt_u1[2] - 0) <= kCompareSmall && std::fabs(cfg.vt_u1[3] - 1) <= kCompareSmall;
    // allocate
    auto Z = [](cvec& v, long long n) { v.assign(n, cx(0.0, 0.0)); };
    Z(uvw, 3 * S); Z(p, S); Z(huvw, 3 * S); Z(suvw, 3 * S); Z(lapuvw, 3 * S); Z(rhsuvw, 3 * S);
    Z(rhsp, S); Z(divs, S); Z(advuvw, 3 * S);
    Z(dudxyz, 3 * S); Z(dvdxyz, 3 * S); Z(dwdxyz, 3 * S); Z(dpdxyz, 3 * S);
    Z(dudxyztmp, 3 * S); Z(dvdxyztmp, 3 * S); Z(dwdxyztmp, 3 * S); Z(ptmp0, S);
    uvwnl.assign(3 * R, 0.0);
    dud_r.assign(3 * R, 0.0); dvd_r.assign(3 * R, 0.0); dwd_r.assign(3 * R, 0.0);
    dudt_r.assign(3 * R, 0.0); dvdt_r.assign(3 * R, 0.0); dwdt_r.assign(3 * R, 0.0);
    adv_r.assign(3 * R, 0.0);
    const long long rowlen = 3LL * P;   // boundary_carrier_flat.cc:36-51
    Z(dp1m1top, rowlen); Z(dp1m1bot, rowlen); Z(dp1m2top, rowlen); Z(dp1m2bot, rowlen);
    Z(dp2m1top, rowlen); Z(dp2m1bot, rowlen); Z(dp2m2top, rowlen); Z(dp2m2bot, rowlen);
    olddt = dt;                         // boundary_carrier_flat.cc:265
    if (cfg.nphase > 1) {
      Z(c, S); Z(hphi, S); Z(lapphi, S); Z(rhsphi, S); Z(advphi, S); Z(uvw1, 3 * S);
      Z(ptmp1, S); Z(dudxyz1, 3 * S); Z(dvdxyz1, 3 * S);
      uvwnl1.assign(3 * R, 0.0); ptmp1_r.assign(R, 0.0); dudxyz1_r.assign(3 * R, 0.0);
      dvdxyz1_r.assign(3 * R, 0.0); advphi_r.assign(R, 0.0);
    }
    // TimeStepper::Configure (stepper_time.cc:221-228), step mode from start 0
    step = 0; time = 0.0; forward = true; alternation = false;
  }

  // Loaded state + the Configure-time transforms (phase_carrier_flat.cc:233-234,
  // variable_carrier_flat.cc:427-429)
  void SetState(const cx* u, const cx* pp, const cx* cc) {
    std::copy(u, u + 3 * S, uvw.begin());
    std::copy(pp, pp + S, p.begin());
    if (cfg.nphase > 1 && cc) std::copy(cc, cc + S, c.begin());
    fft->Backward(uvw.data(), uvwnl.data(), 3);
    for (auto* v : {&dp1m1top, &dp1m1bot, &dp1m2top, &dp1m2bot, &dp2m1top, &dp2m1bot, &dp2m2top, &dp2m2bot})
      std::fill(v->begin(), v->end(), cx(0.0, 0.0));
    std::fill(huvw.begin(), huvw.end(), cx(0.0, 0.0));
    if (cfg.nphase > 1) std::fill(hphi.begin(), hphi.end(), cx(0.0, 0.0));
    olddt = dt;
  }

  // --- z-contractions (RealCplxCube::Multiply(cmat,result), cube_real_cplx.cc:461-471)
  // y[:, col] = M * x[:, col] for ncol columns (M real Nz*Nz col-major)
  void ZMul(const dvec& M, const cx* x, cx* y, long long ncol) const {
    const int N = Nz;
#pragma omp parallel for schedule(static)
    for (long long col = 0; col < ncol; ++col) {
      const cx* xc = x + col * N;
      cx acc[1024];
      cx* yc = y + col * N;
      cx* a = (N <= 1024) ? acc : yc;
      for (int i = 0; i < N; ++i) a[i] = 0.0;
      for (int j = 0; j < N; ++j) {
        const double xr = xc[j].real(), xi = xc[j].imag();
        const double* mj = M.data() + (size_t)N * j;
        for (int i = 0; i < N; ++i) a[i] += cx(mj[i] * xr, mj[i] * xi);
      }
      if (a != yc) for (int i = 0; i < N; ++i) yc[i] = a[i];
    }
  }
  // multipliers: i*kx, i*ky as complex products (stepper_space.cc:265-296)
  static cx MulI(double k, cx v) { return cx(0.0, k) * v; }

  // SpaceStepper::Gradient(target, ddxyz, dim) (stepper_space.cc:373-380)
  void Gradient(const cx* target, cx* out) const {
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < S; ++e) {
      const long long col = e / Nz;
      out[e] = MulI(KX(col), target[e]);
      out[S + e] = MulI(KY(col), target[e]);
    }
    ZMul(d1, target, out + 2 * S, P);
  }
  // SpaceStepper::GradientTmp (stepper_space.cc:401-408)
  void GradientTmp(const cx* target, cx* out) const {
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < S; ++e) {
      const long long col = e / Nz;
      out[e] = MulI(KX(col), target[e]);
      out[S + e] = MulI(KY(col), target[S + e]);
    }
    ZMul(d1, target + 2 * S, out + 2 * S, P);
  }

  // Helmholtz core: Y = V~ (R o (V~^-1 X)), then sequential backward fixes.
  // mode 0: forward fix in place on X; mode 1: X rows 0,N-1 zeroed
  // (phase_carrier_flat.cc:867-987, phase_scalar_flat.cc:640-661).
  void Helmholtz(const Decomposition& d, int stage, int mode, cx* x, cx* y, long long ncol,
                 const double* k2 = nullptr, bool col0mean = true) const {
    const int N = Nz;
    const int nullrow = (d.sigma[stage] == 0.0 || std::fabs(d.sigma[stage]) < kCompareSmall) ? d.NullRow() : -1;
#pragma omp parallel for schedule(static)
    for (long long col = 0; col < ncol; ++col) {
      cx* xc = x + col * N;
      cx* yc = y + col * N;
      if (mode == 0) {
        for (int iz = 1; iz < N - 1; ++iz) xc[iz] += d.fbf[iz] * xc[0] + d.ftf[iz] * xc[N - 1];
      } else {
        xc[0] = cx(0.0, 0.0);
        xc[N - 1] = cx(0.0, 0.0);
      }
      cvec z(N, cx(0.0, 0.0));
      for (int j = 0; j < N; ++j) {
        const cx v = xc[j];
        const double* mj = d.vinv.data() + (size_t)N * j;
        for (int i = 0; i < N; ++i) z[i] += cx(mj[i] * v.real(), mj[i] * v.imag());
      }
      const double kk = k2 ? k2[col] : K2(col);
      for (int i = 1; i < N - 1; ++i) {
        double r = 1.0 / (kk + d.lambda[i] + d.sigma[stage]);
        if (i == nullrow && col0mean && (col % P) == 0) r = 0.0;
        z[i] *= r;
      }
      for (int i = 0; i < N; ++i) yc[i] = 0.0;
      for (int j = 0; j < N; ++j) {
        const cx v = z[j];
        const double* mj = d.vnorm.data() + (size_t)N * j;
        for (int i = 0; i < N; ++i) yc[i] += cx(mj[i] * v.real(), mj[i] * v.imag());
      }
      cx s0 = 0.0;                                              // Multiply(cmat,int) :483-488
      for (int j = 0; j < N; ++j) s0 += d.fbb[j] * yc[j];
      yc[0] = s0;
      cx s1 = 0.0;
      for (int j = 0; j < N; ++j) s1 += d.fbt[j] * yc[j];
      yc[N - 1] = s1;
    }
  }

  // ---------------------------------------------------------------------------
  // FlatCarrierPhase::NonlinearVelocity (phase_carrier_flat.cc:716-813)
  void NonlinearVelocity(int stage) {
    (void)stage;
    const cx* u = uvw.data();
    if (alternation) {
      const double* U = uvwnl.data();
#pragma omp parallel for schedule(static)
      for (long long e = 0; e < R; ++e) {                      // :725-744
        const double a = U[e], b = U[R + e], w = U[2 * R + e];
        dudt_r[e] = a * a; dudt_r[R + e] = a * b; dudt_r[2 * R + e] = a * w;
        dvdt_r[e] = b * a; dvdt_r[R + e] = b * b; dvdt_r[2 * R + e] = b * w;
        dwdt_r[e] = w * a; dwdt_r[R + e] = w * b; dwdt_r[2 * R + e] = w * w;
      }
      fft->Forward(dudt_r.data(), dudxyztmp.data(), 3);       // :747-749
      fft->Forward(dvdt_r.data(), dvdxyztmp.data(), 3);
      fft->Forward(dwdt_r.data(), dwdxyztmp.data(), 3);
      GradientTmp(dudxyztmp.data(), dudxyz.data());             // :752-754
      GradientTmp(dvdxyztmp.data(), dvdxyz.data());
      GradientTmp(dwdxyztmp.data(), dwdxyz.data());
#pragma omp parallel for schedule(static)
      for (long long e = 0; e < S; ++e) {                      // :757-770
        advuvw[e] = dudxyz[e] + dudxyz[S + e] + dudxyz[2 * S + e];
        advuvw[S + e] = dvdxyz[e] + dvdxyz[S + e] + dvdxyz[2 * S + e];
        advuvw[2 * S + e] = dwdxyz[e] + dwdxyz[S + e] + dwdxyz[2 * S + e];
      }
      Gradient(u, dudxyz.data());                               // :772-774
      Gradient(u + S, dvdxyz.data());
      Gradient(u + 2 * S, dwdxyz.data());
      if (cfg.nphase > 1) {                                     // :775-779
        fft->Backward(dudxyz.data(), dud_r.data(), 3);
        fft->Backward(dvdxyz.data(), dvd_r.data(), 3);
        fft->Backward(dwdxyz.data(), dwd_r.data(), 3);
      }
    } else {
      Gradient(u, dudxyz.data());                               // :784-786
      Gradient(u + S, dvdxyz.data());
      Gradient(u + 2 * S, dwdxyz.data());
      fft->Backward(dudxyz.data(), dud_r.data(), 3);            // :787-789
      fft->Backward(dvdxyz.data(), dvd_r.data(), 3);
      fft->Backward(dwdxyz.data(), dwd_r.data(), 3);
      const double* U = uvwnl.data();
#pragma omp parallel for schedule(static)
      for (long long e = 0; e < R; ++e) {                      // :792-803
        const double a = U[e], b = U[R + e], w = U[2 * R + e];
        adv_r[e] = a * dud_r[e] + b * dud_r[R + e] + w * dud_r[2 * R + e];
        adv_r[R + e] = a * dvd_r[e] + b * dvd_r[R + e] + w * dvd_r[2 * R + e];
        adv_r[2 * R + e] = a * dwd_r[e] + b * dwd_r[R + e] + w * dwd_r[2 * R + e];
      }
      fft->Forward(adv_r.data(), advuvw.data(), 3);             // :806
    }
  }

  // Driven force on the mean mode (phase_carrier_flat.cc:360-502)
  void Force(double& f1, double& f2) const {
    f1 = 0; f2 = 0;
    const double t = time;
    const double tm = cfg.output_on_step ? (t - dt / 2) : (t + dt / 2);
    const int d0 = cfg.force_direction[0], d1v = cfg.force_direction[1];
    const double* a = cfg.force_amplitude;
    switch (cfg.force_type) {
      case 0: break;
      case 1:
        if (d0 == 1) f1 = a[0]; else if (d0 == 2) f2 = a[0];
        break;
      case 2: {
        const double f = a[1];
        if (d1v == 1) f1 = f * std::cos(f * tm); else if (d1v == 2) f2 = f * std::cos(f * tm);
      } break;
      case 3: {
        if (d0 == 1) f1 = a[0]; else if (d0 == 2) f2 = a[0];
        const double f = a[1];
        if (d1v == 1) f1 = f1 + f * std::cos(f * tm); else if (d1v == 2) f2 = f2 + f * std::cos(f * tm);
      } break;
      case 4: {
        if (d0 == 1) f1 = a[0]; else if (d0 == 2) f2 = a[0];
        const double f = a[1];
        const double T = a[0], t0 = a[1], last = a[2], tar = a[3];
        const double damp = -std::log(tar) / (last + 0.25) / T;
        const double g = (f * std::cos(f * tm) - damp * std::sin(f * tm)) * std::exp(-damp * (t - t0));
        if (d1v == 1) f1 = f1 + g; else if (d1v == 2) f2 = f2 + g;
      } break;
      default: throw std::logic_error("FlatCarrierPhase::Predict() > Not Supported Force Type!");
    }
  }

  // FlatCarrierBoundary::Apply (boundary_carrier_flat.cc:185-246)
  void BoundaryApply(int stage) {
    const double c3m0 = coef[2][stage], c3m1 = coef[2][(stage + 2) % 3], c3m2 = coef[2][(stage + 1) % 3];
    double xm0 = 0, xm1 = 0, xm2 = 0;
    switch (stage) {
      case 0: xm0 = 2 * (0 + c3m0) * dt; xm1 = 0; xm2 = 2 * (0 - c3m1) * olddt; break;
      case 1: xm0 = 2 * (c3m0 + c3m1 + 0) * dt; xm1 = 2 * (0 + c3m1 + 0) * dt; xm2 = 0; break;
      case 2: xm0 = 2 * (c3m0 + c3m1 + c3m2) * dt; xm1 = 2 * (0 + c3m1 + c3m2) * dt; xm2 = 2 * (0 + +0 + c3m2) * dt; break;
    }
    const double s = (xm0 - xm1) / (xm1 - xm2);
    const double k = 2 * dt * c3m0;
    const int N = Nz;
    for (long long j = 0; j < 3LL * P; ++j) {
      rhsuvw[j * N + (N - 1)] = k * ((1 + s) * dp1m1top[j] + (0 - s) * dp1m2top[j]);
      rhsuvw[j * N + 0] = k * ((1 + s) * dp1m1bot[j] + (0 - s) * dp1m2bot[j]);
    }
    if (opentop) {
      for (long long j = 0; j < 2LL * P; ++j)
        rhsuvw[j * N + (N - 1)] = k * ((1 + s) * dp2m1top[j] + (0 - s) * dp2m2top[j]);
    }
  }

  // FlatCarrierPhase::Predict (phase_carrier_flat.cc:272-559)
  void CarrierPredict(int stage) {
    const double c1 = coef[0][stage], c2 = coef[1][stage], c3 = coef[2][stage];
    NonlinearVelocity(stage);
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < 3 * S; ++e) huvw[e] = c1 * huvw[e] - dt * advuvw[e];   // :294-296
    if (cfg.nphase > 1 && cfg.affected_by_scalar_phase) {                           // :325-356
      const double K = dt * cfg.averaged_concentration / std::pow(cfg.froude, 2) * (cfg.specific_gravity - 1);
      const double sn = std::sin(cfg.slope), cs = std::cos(cfg.slope);
      for (long long e = 0; e < S; ++e) ptmp0[e] = c[e];
      if (cfg.slope_direction == 1) for (long long e = 0; e < S; ++e) huvw[e] = huvw[e] + K * ptmp0[e] * sn;
      else if (cfg.slope_direction == 2) for (long long e = 0; e < S; ++e) huvw[S + e] = huvw[S + e] + K * ptmp0[e] * sn;
      for (long long e = 0; e < S; ++e) huvw[2 * S + e] = huvw[2 * S + e] - K * ptmp0[e] * cs;
    }
    double f1, f2;                                                                  // :360-504
    Force(f1, f2);
    for (int iz = 0; iz < Nz; ++iz) {
      huvw[iz] += dt * f1;
      huvw[S + iz] += dt * f2;
    }
    for (int d = 0; d < 3; ++d) ZMul(d2, uvw.data() + d * S, lapuvw.data() + d * S, P);   // :517-528
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < 3 * S; ++e) {
      const double k2x = -std::pow(KX((e % S) / Nz), 2), k2y = -std::pow(KY((e % S) / Nz), 2);
      lapuvw[e] = lapuvw[e] + k2x * uvw[e] + k2y * uvw[e];
    }
    const double redtc3 = -cfg.reynolds / dt / c3;                                  // :532-536
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < 3 * S; ++e) rhsuvw[e] = redtc3 * c2 * huvw[e] + redtc3 * uvw[e] - lapuvw[e];
    BoundaryApply(stage);                                                           // :541
    rhsuvw[Nz - 1] = cfg.vt_u1[4];                                                  // :542-547
    rhsuvw[S + Nz - 1] = cfg.vt_u2[4];
    rhsuvw[2 * S + Nz - 1] = cfg.vt_u3[4];
    for (int d = 0; d < 3; ++d)                                                     // :554, :867-931
      Helmholtz(dc[d], stage, 0, rhsuvw.data() + d * S, suvw.data() + d * S, P);
  }

  // FlatCarrierPhase::Correct (phase_carrier_flat.cc:566-709)
  void CarrierCorrect(int stage) {
    const double c3 = coef[2][stage];
    const int N = Nz;
    ZMul(d1, suvw.data() + 2 * S, divs.data(), P);                                 // :586 (stepper_space.cc:389-396)
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < S; ++e) {
      const long long col = e / Nz;
      divs[e] = divs[e] + MulI(KX(col), suvw[e]) + MulI(KY(col), suvw[S + e]);
    }
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < S; ++e) rhsp[e] = 0.5 / dt / c3 * divs[e];            // :589
    Helmholtz(dc[3], stage, 1, rhsp.data(), p.data(), P);                           // :594, :955-987
    Gradient(p.data(), dpdxyz.data());                                              // :597
    if (opentop) {                                                                  // :601-604
      for (long long e = 0; e < S; ++e) ptmp0[e] = dpdxyz[2 * S + e];
      Gradient(ptmp0.data(), dudxyztmp.data());
    }
    if (stage == 2) olddt = dt;                                                     // boundary_carrier_flat.cc:140-165
    dp1m2top = dp1m1top; dp2m2top = dp2m1top; dp1m2bot = dp1m1bot; dp2m2bot = dp2m1bot;
    for (long long j = 0; j < 3LL * P; ++j) {
      dp1m1bot[j] = dpdxyz[j * N + 0];
      dp1m1top[j] = dpdxyz[j * N + (N - 1)];
    }
    if (opentop) {
      for (long long j = 0; j < 3LL * P; ++j) {
        dp2m1bot[j] = dudxyztmp[j * N + 0];
        dp2m1top[j] = dudxyztmp[j * N + (N - 1)];
      }
    }
    const double coeffix = 2 * dt * c3;                                             // :613-614
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < 3 * S; ++e) uvw[e] = suvw[e] - coeffix * dpdxyz[e];
    for (int iz = 0; iz < Nz; ++iz) uvw[2 * S + iz] = cx(0.0, 0.0);                 // :618-620
    fft->Backward(uvw.data(), uvwnl.data(), 3);                                     // :625-626
    if (stage == 2) {                                                               // :633-703
      std::memset(featglobal, 0, sizeof(featglobal));
      const double* U = uvwnl.data();
      const long long plane = (long long)Nx * Ny;
      double m[9] = {0};
      double mcfl = 0.0;
      for (int d = 0; d < 3; ++d) {
        for (long long e = 0; e < R; ++e) m[d] = std::max(m[d], std::fabs(U[d * R + e]));
        for (long long e = 0; e < plane; ++e) {
          m[3 + d] = std::max(m[3 + d], std::fabs(U[d * R + e]));
          m[6 + d] = std::max(m[6 + d], std::fabs(U[d * R + (long long)(Nz - 1) * plane + e]));
        }
      }
      for (long long e = 0; e < R; ++e) {
        const int z = (int)(e / plane);
        const double v = std::fabs(U[e]) / glx_[0] + std::fabs(U[R + e]) / gly_[0] + std::fabs(U[2 * R + e]) / glz[z];
        mcfl = std::max(mcfl, v);
      }
      for (int i = 0; i < 9; ++i) featlocal[i] = m[i];
      featlocal[19] = mcfl;
      for (int i = 0; i < 50; ++i) featglobal[i] = featlocal[i];   // MPI_Allreduce(MAX) on one rank
      featglobal[10] = dudxyz[2 * S + 0].real();
      featglobal[11] = dvdxyz[2 * S + 0].real();
      featglobal[12] = dwdxyz[2 * S + 0].real();
      featglobal[13] = dudxyz[2 * S + N - 1].real();
      featglobal[14] = dvdxyz[2 * S + N - 1].real();
      featglobal[15] = dwdxyz[2 * S + N - 1].real();
      for (int i : {21, 22, 23, 24, 25, 28, 29}) featglobal[i] = featlocal[i];
    }
  }

  // FlatScalarPhase::Predict (phase_scalar_flat.cc:245-284)
  void ScalarPredict(int stage) {
    (void)stage;
    uvw1 = uvw;
    if (cfg.slope_direction == 1) for (int iz = 0; iz < Nz; ++iz) uvw1[iz] = uvw1[iz] + cfg.settling_velocity * std::sin(cfg.slope);
    else if (cfg.slope_direction == 2) for (int iz = 0; iz < Nz; ++iz) uvw1[S + iz] = uvw1[S + iz] + cfg.settling_velocity * std::sin(cfg.slope);
    for (int iz = 0; iz < Nz; ++iz) uvw1[2 * S + iz] = uvw1[2 * S + iz] - cfg.settling_velocity * std::cos(cfg.slope);
    fft->Backward(uvw1.data(), uvwnl1.data(), 3);
  }

  // FlatScalarPhase::NonlinearPressure (phase_scalar_flat.cc:552-599)
  void ScalarNonlinear() {
    const double* U = uvwnl1.data();
    if (alternation) {
      ptmp1 = c;
      fft->Backward(ptmp1.data(), ptmp1_r.data(), 1);
#pragma omp parallel for schedule(static)
      for (long long e = 0; e < R; ++e) {
        dvdxyz1_r[e] = ptmp1_r[e] * U[e];
        dvdxyz1_r[R + e] = ptmp1_r[e] * U[R + e];
        dvdxyz1_r[2 * R + e] = ptmp1_r[e] * U[2 * R + e];
      }
      fft->Forward(dvdxyz1_r.data(), dvdxyz1.data(), 3);
      GradientTmp(dvdxyz1.data(), dudxyz1.data());
      for (long long e = 0; e < S; ++e) advphi[e] = dudxyz1[e] + dudxyz1[S + e] + dudxyz1[2 * S + e];
    } else {
      Gradient(c.data(), dudxyz1.data());
      fft->Backward(dudxyz1.data(), dudxyz1_r.data(), 3);
#pragma omp parallel for schedule(static)
      for (long long e = 0; e < R; ++e)
        advphi_r[e] = U[e] * dudxyz1_r[e] + U[R + e] * dudxyz1_r[R + e] + U[2 * R + e] * dudxyz1_r[2 * R + e];
      fft->Forward(advphi_r.data(), advphi.data(), 1);
    }
  }

  // FlatScalarPhase::Correct (phase_scalar_flat.cc:291-533)
  void ScalarCorrect(int stage) {
    const double c1 = coef[0][stage], c2 = coef[1][stage], c3 = coef[2][stage];
    const int N = Nz;
    ScalarNonlinear();
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < S; ++e) hphi[e] = c1 * hphi[e] - dt * advphi[e];      // :313-315
    ZMul(d2, c.data(), lapphi.data(), P);                                           // :319-322
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < S; ++e) {
      const long long col = e / Nz;
      lapphi[e] = lapphi[e] + (-std::pow(KX(col), 2)) * c[e] + (-std::pow(KY(col), 2)) * c[e];
    }
    const double redtc3 = -cfg.reynolds * cfg.schmidt / dt / c3;                    // :325-329
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < S; ++e) rhsphi[e] = redtc3 * c2 * hphi[e] + redtc3 * c[e] - lapphi[e];
    for (long long col = 0; col < P; ++col) {                                       // boundary_scalar_flat.cc:151-152
      rhsphi[col * N + 0] = cx(0, 0) * kBoundaryCoefScale;
      rhsphi[col * N + N - 1] = cx(0, 0) * kBoundaryCoefScale;
    }
    if (cfg.scalar_bc == 1) {                                                       // :340-384
      const double tc = cfg.critical_shear_stress, M = cfg.erosion_rate;
#pragma omp parallel for schedule(static)
      for (long long e = 0; e < R; ++e) {
        const double a = dud_r[2 * R + e], b = dvd_r[2 * R + e];
        double val = std::sqrt(a * a + b * b) / cfg.reynolds;
        if (val < tc) val = 0; else val = M * (val / tc - 1);
        ptmp1_r[e] = val;
      }
      fft->Forward(ptmp1_r.data(), ptmp1.data(), 1);
      const double ws = cfg.settling_velocity * std::cos(cfg.slope);
      for (long long col = 0; col < P; ++col) {
        cx v = ptmp1[col * N] - ws * c[col * N] * cfg.averaged_concentration;
        v = -1.0 * v * cfg.reynolds * cfg.schmidt / cfg.averaged_concentration * kBoundaryCoefScale;
        rhsphi[col * N] = v;
      }
      featlocal[28] = ptmp1[0].real();
      featlocal[29] = (ws * c[0] * cfg.averaged_concentration).real();
    }
    Helmholtz(dc[4], stage, 0, rhsphi.data(), c.data(), P);                         // :399
    // Bartnicki positivity filter (:419-478)
    fft->Backward(c.data(), ptmp1_r.data(), 1);
    PositivityFilter(ptmp1_r.data(), (long long)Nx * Ny, Nz);
    fft->Forward(ptmp1_r.data(), c.data(), 1);
    featlocal[21] = Trapz(c.data()) / 2;                                            // :485-486
  }
  // arma::trapz(gridz, real(col0))
  double Trapz(const cx* col) const {
    double s = 0.0;
    for (int i = 0; i + 1 < Nz; ++i) s += (gridz[i + 1] - gridz[i]) * (col[i].real() + col[i + 1].real()) / 2;
    return s;
  }

  void Substage(int stage) {
    CarrierPredict(stage);
    CarrierCorrect(stage);
    if (cfg.nphase > 1) {
      ScalarPredict(stage);
      ScalarCorrect(stage);
    }
  }

  // TimeStepper::Update (stepper_time.cc:72-113) + the stage loop + toggle
  // (solver_channel_flat.cc:199-254), step mode.
  int Step() {
    if (!forward) return 0;
    step += 1;
    time = time + dt;
    cfl = featglobal[19] * dt;
    for (int s = 0; s < 3; ++s) Substage(s);
    alternation = !alternation;
    return 1;
  }
};

}  // namespace

struct tbo_ctx {
  Oracle* o;
};

extern "C" {

const char* tbo_last_error(void) { return g_err.c_str(); }

tbo_ctx* tbo_create(const tbo_config* cfg, int nthreads) {
  try {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    tbo_ctx* c = new tbo_ctx;
    c->o = new Oracle(*cfg);
    return c;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void tbo_destroy(tbo_ctx* ctx) {
  if (!ctx) return;
  delete ctx->o;
  delete ctx;
}

int tbo_set_state(tbo_ctx* ctx, const double* uvw, const double* p, const double* c) {
  try {
    ctx->o->SetState(reinterpret_cast<const cx*>(uvw), reinterpret_cast<const cx*>(p),
                     reinterpret_cast<const cx*>(c));
    return 0;
  } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

int tbo_get_state(const tbo_ctx* ctx, double* uvw, double* p, double* c) {
  const Oracle* o = ctx->o;
  if (uvw) std::memcpy(uvw, o->uvw.data(), sizeof(cx) * 3 * o->S);
  if (p) std::memcpy(p, o->p.data(), sizeof(cx) * o->S);
  if (c && o->cfg.nphase > 1) std::memcpy(c, o->c.data(), sizeof(cx) * o->S);
  return 0;
}

int tbo_substage(tbo_ctx* ctx, int stage, double time, int alternation) {
  try {
    ctx->o->time = time;
    ctx->o->alternation = alternation != 0;
    ctx->o->Substage(stage);
    return 0;
  } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

int tbo_step(tbo_ctx* ctx) {
  try { return ctx->o->Step(); } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

// Time mode: the new dt, and Phase::Update's reassembly of the solves for it
// (FlatCarrierPhase::Update, phase_carrier_flat.cc:250-265: u1, u2, u3 with
// -Re/(dt c3), the pressure unchanged; FlatScalarPhase::Update,
// phase_scalar_flat.cc:222-234: -Re Sc/(dt c3)).  Assemble rebuilds only the
// reciprocals, which this restatement forms from sigma at each solve.
void tbo_set_dt(tbo_ctx* ctx, double dt) {
  Oracle& o = *ctx->o;
  o.dt = dt;
  for (int s = 0; s < 3; ++s) {
    const double sv = -o.cfg.reynolds / dt / o.coef[2][s];
    for (int k = 0; k < 3; ++k) o.dc[k].sigma[s] = sv;
    if (o.cfg.nphase > 1) o.dc[4].sigma[s] = -o.cfg.reynolds * o.cfg.schmidt / dt / o.coef[2][s];
  }
}

void tbo_stepper(const tbo_ctx* ctx, int* step, double* time, double* cfl, int* alternation) {
  if (step) *step = ctx->o->step;
  if (time) *time = ctx->o->time;
  if (cfl) *cfl = ctx->o->cfl;
  if (alternation) *alternation = ctx->o->alternation ? 1 : 0;
}

void tbo_features(const tbo_ctx* ctx, double* out50) {
  std::memcpy(out50, ctx->o->featglobal, sizeof(double) * 50);
}

int tbo_get_physical_velocity(const tbo_ctx* ctx, double* out) {
  std::memcpy(out, ctx->o->uvwnl.data(), sizeof(double) * 3 * ctx->o->R);
  return 0;
}

int tbo_nonlinear(tbo_ctx* ctx, int alternation, double* adv) {
  try {
    Oracle* o = ctx->o;
    const bool keep = o->alternation;
    o->alternation = alternation != 0;
    o->NonlinearVelocity(0);
    o->alternation = keep;
    std::memcpy(adv, o->advuvw.data(), sizeof(cx) * 3 * o->S);
    return 0;
  } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

int tbo_get_history(const tbo_ctx* ctx, double* huvw, double* hc) {
  const Oracle* o = ctx->o;
  if (huvw) std::memcpy(huvw, o->huvw.data(), sizeof(cx) * 3 * o->S);
  if (hc && o->cfg.nphase > 1) std::memcpy(hc, o->hphi.data(), sizeof(cx) * o->S);
  return 0;
}

int tbo_get_operators(const tbo_ctx* ctx, double* gridz, double* d1, double* d2) {
  const Oracle* o = ctx->o;
  if (gridz) std::memcpy(gridz, o->gridz.data(), sizeof(double) * o->Nz);
  if (d1) std::memcpy(d1, o->d1.data(), sizeof(double) * o->Nz * o->Nz);
  if (d2) std::memcpy(d2, o->d2.data(), sizeof(double) * o->Nz * o->Nz);
  return 0;
}

int tbo_get_decomposition(const tbo_ctx* ctx, int which, double* V, double* Vinv, double* lambda, double* fix) {
  const Oracle* o = ctx->o;
  if (which < 0 || which > 4 || (which == 4 && o->cfg.nphase < 2)) { g_err = "bad decomposition index"; return -1; }
  const Decomposition& d = o->dc[which];
  const size_t N = o->Nz;
  if (V) std::memcpy(V, d.vnorm.data(), sizeof(double) * N * N);
  if (Vinv) std::memcpy(Vinv, d.vinv.data(), sizeof(double) * N * N);
  if (lambda) std::memcpy(lambda, d.lambda.data(), sizeof(double) * N);
  if (fix) {
    std::memcpy(fix, d.fbf.data(), sizeof(double) * N);
    std::memcpy(fix + N, d.ftf.data(), sizeof(double) * N);
    std::memcpy(fix + 2 * N, d.fbb.data(), sizeof(double) * N);
    std::memcpy(fix + 3 * N, d.fbt.data(), sizeof(double) * N);
  }
  return 0;
}

int tbo_dft_forward(int nx, int ny, int nz, int nv, const double* phys, double* spec) {
  try {
    PlaneFft f(nx, ny, nz);
    f.Forward(phys, reinterpret_cast<cx*>(spec), nv);
    return 0;
  } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

int tbo_dft_backward(int nx, int ny, int nz, int nv, const double* spec, double* phys) {
  try {
    PlaneFft f(nx, ny, nz);
    f.Backward(reinterpret_cast<const cx*>(spec), phys, nv);
    return 0;
  } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

int tbo_chebyshev(int nz, int mapping, double coef, double* gridz, double* d1, double* d2) {
  dvec g, a, b;
  Chebyshev(nz, mapping, coef, g, a, b);
  if (gridz) std::memcpy(gridz, g.data(), sizeof(double) * nz);
  if (d1) std::memcpy(d1, a.data(), sizeof(double) * nz * nz);
  if (d2) std::memcpy(d2, b.data(), sizeof(double) * nz * nz);
  return 0;
}

int tbo_positivity_filter(int nx, int ny, int nz, double* phys) {
  if (nx <= 0 || ny <= 0 || nz <= 0) { g_err = "bad plane shape"; return -1; }
  PositivityFilter(phys, (long long)nx * ny, nz);
  return 0;
}

int tbo_eig(int n, const double* a, double* lambda, double* v) {
  dvec lam, vec;
  const int st = EigenDecompose(n, a, lam, vec);
  if (st != 0) { g_err = "eigen-decomposition failed"; return st; }
  std::memcpy(lambda, lam.data(), sizeof(double) * n);
  std::memcpy(v, vec.data(), sizeof(double) * n * n);
  return 0;
}

int tbo_helmholtz(const tbo_ctx* ctx, int which, int stage, int mode, int ncol, const double* k2,
                  int col0_is_mean, const double* x, double* y) {
  const Oracle* o = ctx->o;
  if (which < 0 || which > 4 || (which == 4 && o->cfg.nphase < 2)) { g_err = "bad decomposition index"; return -1; }
  cvec xx(reinterpret_cast<const cx*>(x), reinterpret_cast<const cx*>(x) + (size_t)ncol * o->Nz);
  o->Helmholtz(o->dc[which], stage, mode, xx.data(), reinterpret_cast<cx*>(y), ncol, k2, col0_is_mean != 0);
  return 0;
}

}  // extern "C"
