// Host setup: Chebyshev operators and the bordered eigen-factorisation.
//
// The eigen-decomposition replaces the reference's LAPACK call (eig_gen,
// cmdz2_decomposition.cc:99; ROCm ships no nonsymmetric eigensolver): the
// matrix is balanced, reduced to Hessenberg form with accumulated Householder
// reflections, its eigenvalues found by Francis double-shift QR, and each
// eigenvector by inverse iteration on the Hessenberg matrix followed by the
// back-transformation.  The factorisation V (R o V^-1 .) is invariant under
// eigenvector scaling and ordering, so those are free.
#include "setup.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "common.h"

namespace tbd {

static const double kPi = 3.14159265358979323846264338327950288;

Chebyshev make_chebyshev(int nz, int mapping, double coef) {
  Chebyshev ch;
  ch.nz = nz;
  ch.gridz.resize(nz);
  for (int i = 0; i < nz; ++i) ch.gridz[i] = std::cos((nz - i - 1) * kPi / (nz - 1));  // domain.cc:306
  std::vector<double>& D = ch.d1;
  D.assign((size_t)nz * nz, 0.0);
  auto at = [&](int i, int j) -> double& { return D[i + (size_t)nz * j]; };
  const std::vector<double>& z = ch.gridz;
  // lower triangle (-1)^(i+j)/(z_i - z_j); column 0 halved, last row doubled;
  // upper triangle by centro-antisymmetry; diagonal = -(sum of the row)
  for (int i = 1; i < nz; ++i)
    for (int j = 0; j < i; ++j) at(i, j) = std::pow(-1.0, i + j) / (z[i] - z[j]);
  for (int i = 0; i < nz; ++i) at(i, 0) *= 0.5;
  for (int j = 0; j < nz; ++j) at(nz - 1, j) *= 2.0;
  for (int i = 0; i < nz - 1; ++i)
    for (int j = i + 1; j < nz; ++j) at(i, j) = -at(nz - i - 1, nz - j - 1);
  for (int i = 0; i < nz; ++i) {
    double s = 0.0;
    for (int j = 0; j < nz; ++j) {
      if (j == i) continue;
      s = s - at(i, j);
    }
    at(i, i) = s;
  }
  ch.gridzmap = z;
  if (mapping != 0) {   // stretched grids: D1 <- diag(dxi/dz) D1 (stepper_space_chebyshev.cc:150-186)
    std::vector<double> sc(nz);
    for (int i = 0; i < nz; ++i) {
      const double zi = z[i];
      if (mapping == 1) {
        const double t = std::tan(kPi / 2 * zi);
        ch.gridzmap[i] = 2 / kPi * std::atan(coef * t);
        sc[i] = 1 / coef * (1 + std::pow(coef * t, 2)) / (1 + std::pow(t, 2));
      } else if (mapping == 2) {
        ch.gridzmap[i] = std::asin(coef * zi) / std::asin(coef);
        sc[i] = 1 / coef * std::cos(std::asin(coef * zi)) * std::asin(coef);
      } else if (mapping == 3) {
        const double t = std::tan(kPi / 4 * (zi - 1));
        ch.gridzmap[i] = 4 / kPi * std::atan(coef * t) + 1;
        sc[i] = 1 / coef * (1 + std::pow(coef * t, 2)) / (1 + std::pow(t, 2));
      } else {
        throw Error("ChebyshevSpaceStepper::Configure > Unsupported Mapping Function!");
      }
    }
    for (int j = 0; j < nz; ++j)
      for (int i = 0; i < nz; ++i) at(i, j) *= sc[i];
  }
  // D2 = D1*D1 with the diagonal re-derived from the row sums
  ch.d2.assign((size_t)nz * nz, 0.0);
  for (int j = 0; j < nz; ++j)
    for (int k = 0; k < nz; ++k) {
      const double b = D[k + (size_t)nz * j];
      const double* ak = &D[(size_t)nz * k];
      double* cj = &ch.d2[(size_t)nz * j];
      for (int i = 0; i < nz; ++i) cj[i] += ak[i] * b;
    }
  for (int i = 0; i < nz; ++i) {
    double s = 0.0;
    for (int j = 0; j < nz; ++j) {
      if (j == i) continue;
      s = s - ch.d2[i + (size_t)nz * j];
    }
    ch.d2[i + (size_t)nz * i] = s;
  }
  return ch;
}

// ---------------------------------------------------------------------------
// dense helpers
// ---------------------------------------------------------------------------
bool invert(int n, const double* a, double* ainv) {
  // Gauss-Jordan with partial pivoting on [A | I]
  std::vector<double> m((size_t)n * n);
  std::memcpy(m.data(), a, sizeof(double) * n * n);
  std::vector<double> inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i + (size_t)n * i] = 1.0;
  auto M = [&](int i, int j) -> double& { return m[i + (size_t)n * j]; };
  auto I = [&](int i, int j) -> double& { return inv[i + (size_t)n * j]; };
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(M(i, k)) > std::fabs(M(p, k))) p = i;
    if (M(p, k) == 0.0) return false;
    if (p != k)
      for (int j = 0; j < n; ++j) { std::swap(M(k, j), M(p, j)); std::swap(I(k, j), I(p, j)); }
    const double piv = 1.0 / M(k, k);
    for (int j = 0; j < n; ++j) { M(k, j) *= piv; I(k, j) *= piv; }
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      const double f = M(i, k);
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) { M(i, j) -= f * M(k, j); I(i, j) -= f * I(k, j); }
    }
  }
  std::memcpy(ainv, inv.data(), sizeof(double) * n * n);
  return true;
}

namespace {

// Francis double-shift QR on an upper Hessenberg matrix h (col-major),
// eigenvalues only.
bool hessenberg_qr(int n, std::vector<double> h, std::vector<double>& wr, std::vector<double>& wi) {
  auto H = [&](int i, int j) -> double& { return h[i + (size_t)n * j]; };
  wr.assign(n, 0.0);
  wi.assign(n, 0.0);
  double norm = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = std::max(0, i - 1); j < n; ++j) norm += std::fabs(H(i, j));
  int hi = n - 1;
  double shift_acc = 0.0;
  int iter = 0;
  while (hi >= 0) {
    // locate the active block [lo, hi]
    int lo = hi;
    while (lo > 0) {
      double s = std::fabs(H(lo - 1, lo - 1)) + std::fabs(H(lo, lo));
      if (s == 0.0) s = norm;
      if (std::fabs(H(lo, lo - 1)) + s == s) { H(lo, lo - 1) = 0.0; break; }
      --lo;
    }
    double x = H(hi, hi);
    if (lo == hi) {              // 1x1 block
      wr[hi] = x + shift_acc;
      wi[hi] = 0.0;
      --hi;
      iter = 0;
      continue;
    }
    double y = H(hi - 1, hi - 1);
    double w = H(hi, hi - 1) * H(hi - 1, hi);
    if (lo == hi - 1) {          // 2x2 block
      const double p = 0.5 * (y - x);
      const double q = p * p + w;
      const double zr = std::sqrt(std::fabs(q));
      x += shift_acc;
      if (q >= 0.0) {
        const double z = p + (p >= 0.0 ? zr : -zr);
        wr[hi - 1] = x + z;
        wr[hi] = (z != 0.0) ? x - w / z : x + z;
        wi[hi - 1] = wi[hi] = 0.0;
      } else {
        wr[hi - 1] = wr[hi] = x + p;
        wi[hi - 1] = zr;
        wi[hi] = -zr;
      }
      hi -= 2;
      iter = 0;
      continue;
    }
    if (iter >= 80) return false;
    if (iter == 10 || iter == 20 || iter == 40) {   // exceptional shifts
      shift_acc += x;
      for (int i = 0; i <= hi; ++i) H(i, i) -= x;
      const double s = std::fabs(H(hi, hi - 1)) + std::fabs(H(hi - 1, hi - 2));
      x = y = 0.75 * s;
      w = -0.4375 * s * s;
    }
    ++iter;
    // find two consecutive small subdiagonals
    int m = hi - 2;
    double p = 0, q = 0, r = 0;
    for (; m >= lo; --m) {
      const double z = H(m, m);
      const double rr = x - z;
      const double ss = y - z;
      p = (rr * ss - w) / H(m + 1, m) + H(m, m + 1);
      q = H(m + 1, m + 1) - z - rr - ss;
      r = H(m + 2, m + 1);
      const double s = std::fabs(p) + std::fabs(q) + std::fabs(r);
      p /= s; q /= s; r /= s;
      if (m == lo) break;
      const double u = std::fabs(H(m, m - 1)) * (std::fabs(q) + std::fabs(r));
      const double v = std::fabs(p) * (std::fabs(H(m - 1, m - 1)) + std::fabs(z) + std::fabs(H(m + 1, m + 1)));
      if (u + v == v) break;
    }
    for (int i = m + 2; i <= hi; ++i) {
      H(i, i - 2) = 0.0;
      if (i != m + 2) H(i, i - 3) = 0.0;
    }
    // double-shift QR sweep (bulge chase) on rows/columns lo..hi
    for (int k = m; k <= hi - 1; ++k) {
      const bool last = (k == hi - 1);
      if (k != m) {
        p = H(k, k - 1);
        q = H(k + 1, k - 1);
        r = last ? 0.0 : H(k + 2, k - 1);
        x = std::fabs(p) + std::fabs(q) + std::fabs(r);
        if (x == 0.0) continue;
        p /= x; q /= x; r /= x;
      }
      double s = std::sqrt(p * p + q * q + r * r);
      if (p < 0) s = -s;
      if (s == 0.0) continue;
      if (k == m) {
        if (lo != m) H(k, k - 1) = -H(k, k - 1);
      } else {
        H(k, k - 1) = -s * x;
      }
      p += s;
      x = p / s;
      y = q / s;
      const double z = r / s;
      q /= p;
      r /= p;
      for (int j = k; j <= hi; ++j) {
        double t = H(k, j) + q * H(k + 1, j);
        if (!last) { t += r * H(k + 2, j); H(k + 2, j) -= t * z; }
        H(k + 1, j) -= t * y;
        H(k, j) -= t * x;
      }
      const int imax = std::min(hi, k + 3);
      for (int i = lo; i <= imax; ++i) {
        double t = x * H(i, k) + y * H(i, k + 1);
        if (!last) { t += z * H(i, k + 2); H(i, k + 2) -= t * r; }
        H(i, k + 1) -= t * q;
        H(i, k) -= t;
      }
    }
  }
  return true;
}

}  // namespace

int real_eigen(int n, const double* a, std::vector<double>& lambda, std::vector<double>& vec) {
  std::vector<double> A(a, a + (size_t)n * n);
  auto at = [&](int i, int j) -> double& { return A[i + (size_t)n * j]; };
  // 1. balance by powers of two: A <- D^-1 A D
  std::vector<double> scale(n, 1.0);
  for (bool changed = true; changed;) {
    changed = false;
    for (int i = 0; i < n; ++i) {
      double c = 0.0, r = 0.0;
      for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        c += std::fabs(at(j, i));
        r += std::fabs(at(i, j));
      }
      if (c == 0.0 || r == 0.0) continue;
      double f = 1.0, g = r / 2.0;
      const double s0 = c + r;
      while (c < g) { f *= 2.0; c *= 4.0; }
      g = r * 2.0;
      while (c > g) { f /= 2.0; c /= 4.0; }
      if ((c + r) / f < 0.95 * s0) {
        changed = true;
        scale[i] *= f;
        for (int j = 0; j < n; ++j) at(i, j) /= f;
        for (int j = 0; j < n; ++j) at(j, i) *= f;
      }
    }
  }
  // 2. Householder reduction H = Q^T A Q, Q accumulated
  std::vector<double> Q((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Q[i + (size_t)n * i] = 1.0;
  std::vector<double> v(n);
  for (int k = 0; k + 2 < n; ++k) {
    double nrm = 0.0;
    for (int i = k + 1; i < n; ++i) nrm += at(i, k) * at(i, k);
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) continue;
    const double alpha = at(k + 1, k) > 0 ? -nrm : nrm;
    std::fill(v.begin(), v.end(), 0.0);
    for (int i = k + 1; i < n; ++i) v[i] = at(i, k);
    v[k + 1] -= alpha;
    double vv = 0.0;
    for (int i = k + 1; i < n; ++i) vv += v[i] * v[i];
    if (vv == 0.0) continue;
    const double beta = 2.0 / vv;
    for (int j = 0; j < n; ++j) {           // A <- (I - b v v^T) A
      double s = 0.0;
      for (int i = k + 1; i < n; ++i) s += v[i] * at(i, j);
      s *= beta;
      for (int i = k + 1; i < n; ++i) at(i, j) -= s * v[i];
    }
    for (int i = 0; i < n; ++i) {           // A <- A (I - b v v^T), Q <- Q (I - b v v^T)
      double s = 0.0, sq = 0.0;
      for (int j = k + 1; j < n; ++j) { s += at(i, j) * v[j]; sq += Q[i + (size_t)n * j] * v[j]; }
      s *= beta;
      sq *= beta;
      for (int j = k + 1; j < n; ++j) { at(i, j) -= s * v[j]; Q[i + (size_t)n * j] -= sq * v[j]; }
    }
    for (int i = k + 2; i < n; ++i) at(i, k) = 0.0;
  }
  // 3. eigenvalues
  std::vector<double> wr, wi;
  if (!hessenberg_qr(n, A, wr, wi)) return -1;
  double hnorm = 0.0;
  for (size_t i = 0; i < (size_t)n * n; ++i) hnorm = std::max(hnorm, std::fabs(A[i]));
  for (int i = 0; i < n; ++i)
    if (std::fabs(wi[i]) > 1e-8 * std::max(1.0, std::fabs(wr[i]))) return -2;
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) { return wr[x] > wr[y]; });
  lambda.resize(n);
  for (int i = 0; i < n; ++i) lambda[i] = wr[order[i]];
  // 4. eigenvectors: inverse iteration on H - mu I (Hessenberg LU with
  //    row interchanges), then x = D Q y, unit 2-norm, largest entry positive
  vec.assign((size_t)n * n, 0.0);
  std::vector<double> U((size_t)n * n), y(n), x(n), lmul(n);
  std::vector<char> swp(n);
  for (int e = 0; e < n; ++e) {
    const double mu = lambda[e] + (std::fabs(lambda[e]) + hnorm) * 1e-13;
    U = A;
    auto u = [&](int i, int j) -> double& { return U[i + (size_t)n * j]; };
    for (int i = 0; i < n; ++i) u(i, i) -= mu;
    for (int k = 0; k + 1 < n; ++k) {
      swp[k] = 0;
      if (std::fabs(u(k + 1, k)) > std::fabs(u(k, k))) {
        swp[k] = 1;
        for (int j = k; j < n; ++j) std::swap(u(k, j), u(k + 1, j));
      }
      const double piv = u(k, k) != 0.0 ? u(k, k) : 1e-300;
      const double l = u(k + 1, k) / piv;
      lmul[k] = l;
      u(k + 1, k) = 0.0;
      for (int j = k + 1; j < n; ++j) u(k + 1, j) -= l * u(k, j);
    }
    if (u(n - 1, n - 1) == 0.0) u(n - 1, n - 1) = 1e-300;
    for (int i = 0; i < n; ++i) y[i] = 1.0;
    for (int it = 0; it < 3; ++it) {
      if (it > 0) {   // forward elimination on the right-hand side
        for (int k = 0; k + 1 < n; ++k) {
          if (swp[k]) std::swap(y[k], y[k + 1]);
          y[k + 1] -= lmul[k] * y[k];
        }
      }
      for (int i = n - 1; i >= 0; --i) {
        double s = y[i];
        for (int j = i + 1; j < n; ++j) s -= u(i, j) * y[j];
        y[i] = s / u(i, i);
      }
      double nrm = 0.0;
      for (int i = 0; i < n; ++i) nrm += y[i] * y[i];
      nrm = std::sqrt(nrm);
      for (int i = 0; i < n; ++i) y[i] /= nrm;
    }
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += Q[i + (size_t)n * j] * y[j];
      x[i] = s * scale[i];
    }
    double nrm = 0.0;
    int imax = 0;
    for (int i = 0; i < n; ++i) {
      nrm += x[i] * x[i];
      if (std::fabs(x[i]) > std::fabs(x[imax])) imax = i;
    }
    nrm = std::sqrt(nrm);
    const double sg = x[imax] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) vec[i + (size_t)n * e] = sg * x[i] / nrm;
  }
  return 0;
}

// Even/odd form of a factorisation (the layout is described with
// Factorisation).  The eigenvectors are symmetrised, grouped by parity, and
// the two half-size blocks inverted separately; nothing is built when the
// operator is not centro-symmetric or an eigenvector is not clearly even or
// odd (a near-degenerate pair of opposite parity).
// optol: tolerance of the operator's centro-symmetry check, relative to its
// largest entry (looser for an operator reassembled from supplied V, lambda).
static void make_split(Factorisation& f, const std::vector<double>& fwd, const std::vector<double>& lam,
                       const std::vector<double>& vec, double optol = 1e-12) {
  const int N = f.nz;
  if (N % 2 == 0 || N < 7) return;
  const int m = N - 2, M = (N - 1) / 2;
  double mx = 0.0, dv = 0.0;
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      mx = std::max(mx, std::fabs(fwd[i + (size_t)m * j]));
      dv = std::max(dv, std::fabs(fwd[i + (size_t)m * j] - fwd[(m - 1 - i) + (size_t)m * (m - 1 - j)]));
    }
  if (!(dv <= optol * mx)) return;
  std::vector<int> ev, od;
  for (int k = 0; k < m; ++k) {
    const double* v = &vec[(size_t)m * k];
    double se = 0.0, so = 0.0, nn = 0.0;
    for (int i = 0; i < m; ++i) {
      se += (v[i] + v[m - 1 - i]) * (v[i] + v[m - 1 - i]);
      so += (v[i] - v[m - 1 - i]) * (v[i] - v[m - 1 - i]);
      nn += v[i] * v[i];
    }
    if (so <= 1e-16 * nn) ev.push_back(k);
    else if (se <= 1e-16 * nn) od.push_back(k);
    else return;
  }
  if ((int)ev.size() != M || (int)od.size() != M - 1) return;
  // half blocks on interior split indices 1..M (even) and 1..M-1 (odd), column-major
  const int me = M, mo = M - 1;
  std::vector<double> re((size_t)me * me), ro((size_t)mo * mo), rei(re.size()), roi(ro.size());
  for (int c = 0; c < me; ++c) {
    const double* v = &vec[(size_t)m * ev[c]];
    for (int r = 0; r < M - 1; ++r) re[r + (size_t)me * c] = 0.5 * (v[r] + v[m - 1 - r]);
    re[(M - 1) + (size_t)me * c] = v[M - 1];
  }
  for (int c = 0; c < mo; ++c) {
    const double* v = &vec[(size_t)m * od[c]];
    for (int r = 0; r < M - 1; ++r) ro[r + (size_t)mo * c] = 0.5 * (v[r] - v[m - 1 - r]);
  }
  if (!invert(me, re.data(), rei.data()) || !invert(mo, ro.data(), roi.data())) return;
  const int ld = (M + 1 + 3) / 4 * 4;
  f.pinv.assign((size_t)(2 * M + 2) * ld, 0.0);
  f.pnorm.assign((size_t)(2 * M + 2) * ld, 0.0);
  auto PI = [&](int r, int c) -> double& { return f.pinv[(size_t)r * ld + c]; };
  auto PN = [&](int r, int c) -> double& { return f.pnorm[(size_t)r * ld + c]; };
  // eigen-coordinates from (s, d): the even part of x is s_i / 2 (s_M at i = M)
  PI(0, 0) = 1.0;
  for (int k = 0; k < me; ++k) {
    for (int i = 1; i < M; ++i) PI(1 + k, i) = 0.5 * rei[k + (size_t)me * (i - 1)];
    PI(1 + k, M) = rei[k + (size_t)me * (M - 1)];
  }
  PI(M + 1, 0) = 1.0;
  for (int k = 0; k < mo; ++k)
    for (int i = 1; i < M; ++i) PI(M + 2 + k, i) = 0.5 * roi[k + (size_t)mo * (i - 1)];
  // columns from the coordinates: the wall pair gives y_0 = (s_0 + d_0)/2,
  // y_{nz-1} = (s_0 - d_0)/2
  PN(0, 0) = 0.5;
  for (int i = 1; i <= M; ++i)
    for (int k = 0; k < me; ++k) PN(i, 1 + k) = re[(i - 1) + (size_t)me * k];
  PN(M + 1, 0) = 0.5;
  for (int i = 1; i < M; ++i)
    for (int k = 0; k < mo; ++k) PN(M + 1 + i, 1 + k) = ro[(i - 1) + (size_t)mo * k];
  f.plam.assign(N, 0.0);
  for (int k = 0; k < me; ++k) f.plam[1 + k] = lam[ev[k]];
  for (int k = 0; k < mo; ++k) f.plam[M + 2 + k] = lam[od[k]];
  f.pnull = -1;
  if (f.nullrow >= 1) {
    const int j = f.nullrow - 1;
    for (int k = 0; k < me; ++k)
      if (ev[k] == j) f.pnull = 1 + k;
    for (int k = 0; k < mo; ++k)
      if (od[k] == j) f.pnull = M + 2 + k;
  }
  f.pfix.assign(2 * (size_t)N, 0.0);
  for (int t = 0; t < 2; ++t) {
    const double* w = &f.fix[(2 + t) * (size_t)N];
    double* o = &f.pfix[t * (size_t)N];
    for (int i = 0; i < M; ++i) {
      o[i] = w[i] + w[N - 1 - i];
      o[M + 1 + i] = w[i] - w[N - 1 - i];
    }
    o[M] = w[M];
  }
  f.pM = M;
  f.pld = ld;
}

Factorisation make_factorisation(const Chebyshev& ch, const double* vt) {
  const int N = ch.nz;
  Factorisation f;
  f.nz = N;
  for (int i = 0; i < 5; ++i) f.vt[i] = vt[i];
  // bordered operator: rows 0 and N-1 become a*f + b*D1 f (cmdz2_decomposition.cc:47-50)
  std::vector<double> c = ch.d2;
  auto C = [&](int i, int j) -> double& { return c[i + (size_t)N * j]; };
  for (int j = 0; j < N; ++j) {
    C(0, j) = vt[1] * ch.d1[0 + (size_t)N * j];
    C(N - 1, j) = vt[3] * ch.d1[(N - 1) + (size_t)N * j];
  }
  C(0, 0) += vt[0];
  C(N - 1, N - 1) += vt[2];
  const double a11 = C(0, 0), ann = C(N - 1, N - 1), a1n = C(0, N - 1), an1 = C(N - 1, 0);
  const double den = 1.0 / (an1 * a1n - a11 * ann);
  f.fix.assign(4 * (size_t)N, 0.0);
  double* fbf = &f.fix[0];
  double* ftf = &f.fix[N];
  double* fbb = &f.fix[2 * N];
  double* fbt = &f.fix[3 * N];
  for (int i = 1; i < N - 1; ++i) {   // :69-76
    fbb[i] = (ann * C(0, i) - a1n * C(N - 1, i)) * den;
    fbt[i] = (a11 * C(N - 1, i) - an1 * C(0, i)) * den;
    fbf[i] = (ann * C(i, 0) - an1 * C(i, N - 1)) * den;
    ftf[i] = (a11 * C(i, N - 1) - a1n * C(i, 0)) * den;
  }
  fbb[0] = -1 * ann * den;            // :77-80
  fbb[N - 1] = +1 * a1n * den;
  fbt[0] = +1 * an1 * den;
  fbt[N - 1] = -1 * a11 * den;
  // interior operator with the boundary unknowns eliminated (:85-93)
  const int m = N - 2;
  std::vector<double> fwd((size_t)m * m);
  for (int j = 1; j < N - 1; ++j)
    for (int i = 1; i < N - 1; ++i)
      fwd[(i - 1) + (size_t)m * (j - 1)] = C(i, j) + C(i, 0) * fbb[j] + C(i, N - 1) * fbt[j];
  std::vector<double> lam, vec;
  const int st = real_eigen(m, fwd.data(), lam, vec);
  if (st != 0)
    throw Error("CmDz2DecompositionCplx::Construct > eigen-decomposition failed (" + std::to_string(st) + ")");
  std::vector<double> vi((size_t)m * m);
  if (!invert(m, vec.data(), vi.data())) throw Error("CmDz2DecompositionCplx::Construct > singular eigenvectors");
  f.vnorm.assign((size_t)N * N, 0.0);
  f.vinv.assign((size_t)N * N, 0.0);
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      f.vnorm[(i + 1) + (size_t)N * (j + 1)] = vec[i + (size_t)m * j];
      f.vinv[(i + 1) + (size_t)N * (j + 1)] = vi[i + (size_t)m * j];
    }
  f.vnorm[0] = f.vinv[0] = 1.0;
  f.vnorm[(N - 1) + (size_t)N * (N - 1)] = f.vinv[(N - 1) + (size_t)N * (N - 1)] = 1.0;
  f.lambda.assign(N, 0.0);
  for (int i = 0; i < m; ++i) f.lambda[i + 1] = lam[i];
  // null mode of a pure Neumann problem: the eigenvalue of smallest magnitude
  // when it sits at round-off level of the spectrum (the reference gates
  // |lambda| < 1e-11 on the mean column, :273-291)
  double bmin = 1e300, bmax = 0.0;
  int best = -1;
  for (int i = 1; i < N - 1; ++i) {
    bmax = std::max(bmax, std::fabs(f.lambda[i]));
    if (std::fabs(f.lambda[i]) < bmin) { bmin = std::fabs(f.lambda[i]); best = i; }
  }
  f.nullrow = (bmin < 10E-12 || bmin < 1e-9 * bmax) ? best : -1;
  make_split(f, fwd, lam, vec);
  return f;
}

std::vector<double> fold_backward_fixes(const std::vector<double>& vnorm, const std::vector<double>& fix, int nz) {
  TBD_CHECK(vnorm.size() == (size_t)nz * nz && fix.size() == (size_t)4 * nz,
            "fold_backward_fixes > bad factorisation size");
  const double* wb = fix.data() + 2 * (size_t)nz;
  const double* wt = fix.data() + 3 * (size_t)nz;
  std::vector<double> out(vnorm);
  for (int j = 0; j < nz; ++j) {
    const double* col = vnorm.data() + (size_t)nz * j;
    long double b = 0.0L, t = 0.0L;
    for (int i = 0; i < nz; ++i) {
      b += (long double)wb[i] * col[i];
      t += (long double)wt[i] * col[i];
    }
    t += (long double)wt[0] * (b - (long double)col[0]);
    out[(size_t)nz * j] = (double)b;
    out[(size_t)nz * j + nz - 1] = (double)t;
  }
  return out;
}

Factorisation factorisation_from(int nz, const double* vt, const double* V, const double* Vinv,
                                 const double* lambda, const double* fix) {
  const int N = nz, m = N - 2;
  if (N < 4) throw Error("CmDz2DecompositionCplx::Construct > nz too small");
  Factorisation f;
  f.nz = N;
  for (int i = 0; i < 5; ++i) f.vt[i] = vt[i];
  f.vnorm.assign(V, V + (size_t)N * N);
  f.vinv.assign(Vinv, Vinv + (size_t)N * N);
  f.lambda.assign(lambda, lambda + N);
  f.fix.assign(fix, fix + 4 * (size_t)N);
  for (double v : f.vnorm)
    if (!std::isfinite(v)) throw Error("CmDz2DecompositionCplx::Construct > V is not finite");
  for (double v : f.vinv)
    if (!std::isfinite(v)) throw Error("CmDz2DecompositionCplx::Construct > V^-1 is not finite");
  for (double v : f.lambda)
    if (!std::isfinite(v)) throw Error("CmDz2DecompositionCplx::Construct > lambda is not finite");
  for (double v : f.fix)
    if (!std::isfinite(v)) throw Error("CmDz2DecompositionCplx::Construct > fix vectors are not finite");
  double bmin = 1e300, bmax = 0.0;
  int best = -1;
  for (int i = 1; i < N - 1; ++i) {
    bmax = std::max(bmax, std::fabs(f.lambda[i]));
    if (std::fabs(f.lambda[i]) < bmin) { bmin = std::fabs(f.lambda[i]); best = i; }
  }
  f.nullrow = (bmin < 10E-12 || bmin < 1e-9 * bmax) ? best : -1;
  // interior eigenvectors, eigenvalues and operator V lambda V^-1
  std::vector<double> vec((size_t)m * m), lam(m), fwd((size_t)m * m, 0.0);
  for (int j = 0; j < m; ++j) {
    lam[j] = f.lambda[j + 1];
    for (int i = 0; i < m; ++i) vec[i + (size_t)m * j] = f.vnorm[(i + 1) + (size_t)N * (j + 1)];
  }
  for (int j = 0; j < m; ++j)
    for (int k = 0; k < m; ++k) {
      const double w = lam[k] * f.vinv[(k + 1) + (size_t)N * (j + 1)];
      if (w == 0.0) continue;
      for (int i = 0; i < m; ++i) fwd[i + (size_t)m * j] += vec[i + (size_t)m * k] * w;
    }
  // the reassembled operator carries the rounding of V lambda V^-1; the
  // parity test of every eigenvector is what admits the split
  make_split(f, fwd, lam, vec, 1e-9);
  return f;
}

}  // namespace tbd
