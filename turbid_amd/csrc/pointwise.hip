// Elementwise and reduction kernels of the substage (see pointwise.h).
#include "pointwise.h"
#include "group.h"

#include <algorithm>

namespace tbd {

Recorder*& Recorder::active() {
  thread_local Recorder* r = nullptr;
  return r;
}

static constexpr int kNT = 256;

static inline dim3 grid1d(long long n, int per = 1) {
  long long b = (n + (long long)kNT * per - 1) / ((long long)kNT * per);
  if (b > 65536LL * 8) b = 65536LL * 8;
  return dim3((unsigned)b);
}

#define GRID_STRIDE(i, n) \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)
// the same over the member's own grid (group.h)
#define MEMBER_STRIDE(i, n) \
  for (long long i = (long long)b.x * blockDim.x + threadIdx.x; i < (n); i += (long long)b.gx * blockDim.x)

template <typename T>
struct PConvA {
  const T* U;
  const T* g;
  T* adv;
  long long R;
};
template <typename T, bool G>
__global__ void k_products_conv(PConvA<T> a0, GroupP grp) {
  Blk b;
  const PConvA<T>& A = member<G>(a0, grp, b);
  const T* __restrict__ U = A.U;
  const T* __restrict__ g = A.g;
  T* __restrict__ adv = A.adv;
  const long long R = A.R;
  MEMBER_STRIDE(e, R) {
    const T a = U[e], b = U[R + e], w = U[2 * R + e];
    adv[e] = a * g[e] + b * g[R + e] + w * g[2 * R + e];
    adv[R + e] = a * g[3 * R + e] + b * g[4 * R + e] + w * g[5 * R + e];
    adv[2 * R + e] = a * g[6 * R + e] + b * g[7 * R + e] + w * g[8 * R + e];
  }
}
template <typename T>
void products_conv(const T* uvw, const T* g, T* adv, long long R, hipStream_t st) {
  glaunch(k_products_conv<T, false>, k_products_conv<T, true>, grid1d(R), dim3(kNT), 0, st, PConvA<T>{uvw, g, adv, R});
}

template <typename T>
struct PDivA {
  const T* U;
  T* o;
  long long R;
};
template <typename T, bool G>
__global__ void k_products_div(PDivA<T> a0, GroupP grp) {
  Blk b;
  const PDivA<T>& A = member<G>(a0, grp, b);
  const T* __restrict__ U = A.U;
  T* __restrict__ o = A.o;
  const long long R = A.R;
  MEMBER_STRIDE(e, R) {
    const T a = U[e], b = U[R + e], w = U[2 * R + e];
    o[e] = a * a;
    o[R + e] = a * b;
    o[2 * R + e] = a * w;
    o[3 * R + e] = b * b;
    o[4 * R + e] = b * w;
    o[5 * R + e] = w * w;
  }
}
template <typename T>
void products_div(const T* uvw, T* prod6, long long R, hipStream_t st) {
  glaunch(k_products_div<T, false>, k_products_div<T, true>, grid1d(R), dim3(kNT), 0, st, PDivA<T>{uvw, prod6, R});
}

template <typename T>
struct SConvA {
  const T* U;
  T su, sv, sw;
  const T* dc;
  T* adv;
  long long R;
};
template <typename T, bool G>
__global__ void k_scalar_conv(SConvA<T> a0, GroupP grp) {
  Blk b;
  const SConvA<T>& A = member<G>(a0, grp, b);
  const T* __restrict__ U = A.U;
  const T* __restrict__ dc = A.dc;
  T* __restrict__ adv = A.adv;
  const T su = A.su, sv = A.sv, sw = A.sw;
  const long long R = A.R;
  MEMBER_STRIDE(e, R) {
    adv[e] = (U[e] + su) * dc[e] + (U[R + e] + sv) * dc[R + e] + (U[2 * R + e] + sw) * dc[2 * R + e];
  }
}
template <typename T>
void scalar_products_conv(const T* uvw, T su, T sv, T sw, const T* dc, T* adv, long long R, hipStream_t st) {
  glaunch(k_scalar_conv<T, false>, k_scalar_conv<T, true>, grid1d(R), dim3(kNT), 0, st,
          SConvA<T>{uvw, su, sv, sw, dc, adv, R});
}

template <typename T, typename TC>
struct SDivA {
  const T* U;
  T su, sv, sw;
  const TC* c;
  T* o;
  long long R;
};
template <typename T, typename TC, bool G>
__global__ void k_scalar_div(SDivA<T, TC> a0, GroupP grp) {
  Blk b;
  const SDivA<T, TC>& A = member<G>(a0, grp, b);
  const T* __restrict__ U = A.U;
  const TC* __restrict__ c = A.c;
  T* __restrict__ o = A.o;
  const T su = A.su, sv = A.sv, sw = A.sw;
  const long long R = A.R;
  MEMBER_STRIDE(e, R) {
    const T cv = (T)c[e];
    o[e] = cv * (U[e] + su);
    o[R + e] = cv * (U[R + e] + sv);
    o[2 * R + e] = cv * (U[2 * R + e] + sw);
  }
}
template <typename T, typename TC>
void scalar_products_div(const T* uvw, T su, T sv, T sw, const TC* c, T* out3, long long R, hipStream_t st) {
  glaunch(k_scalar_div<T, TC, false>, k_scalar_div<T, TC, true>, grid1d(R), dim3(kNT), 0, st,
          SDivA<T, TC>{uvw, su, sv, sw, c, out3, R});
}

template <typename TA>
struct HUpdA {
  double2* h;
  const TA* adv;
  const double2* c;
  HUpd u;
  int nf;
  long long S;
};
template <typename TA, bool G>
__global__ void k_h_update(HUpdA<TA> a0, GroupP grp) {
  Blk b;
  const HUpdA<TA>& A = member<G>(a0, grp, b);
  double2* __restrict__ h = A.h;
  const TA* __restrict__ adv = A.adv;
  const double2* __restrict__ c = A.c;
  const HUpd& u = A.u;
  const int nf = A.nf;
  const long long S = A.S;
  MEMBER_STRIDE(i, nf * S) {
    const int f = (int)(i / S);
    const long long e = i - (long long)f * S;
    double2 hv = h[i];
    const TA a = adv[i];
    hv.x = u.c1 * hv.x - u.dt * (double)a.x;
    hv.y = u.c1 * hv.y - u.dt * (double)a.y;
    if (u.buoy) {
      const double2 cv = c[e];
      if (f == 2) {
        hv.x = hv.x - u.K * cv.x * u.cs;
        hv.y = hv.y - u.K * cv.y * u.cs;
      } else if ((f == 0 && u.slope_dir == 1) || (f == 1 && u.slope_dir == 2)) {
        hv.x = hv.x + u.K * cv.x * u.sn;
        hv.y = hv.y + u.K * cv.y * u.sn;
      }
    }
    if (e < u.nz) {                 // mean column (kx = ky = 0), all z
      const double f1 = u.fdev ? u.fdev[0] : u.f1, f2 = u.fdev ? u.fdev[1] : u.f2;
      if (f == 0) hv.x += u.dt * f1;
      else if (f == 1) hv.x += u.dt * f2;
    }
    h[i] = hv;
  }
}
template <typename TA>
void h_update(double2* h, const TA* adv, const double2* c, const HUpd& u, int nf, long long S, hipStream_t st) {
  glaunch(k_h_update<TA, false>, k_h_update<TA, true>, grid1d(nf * S), dim3(kNT), 0, st,
          HUpdA<TA>{h, adv, c, u, nf, S});
}

// lambda^3 - I1 lambda^2 + I2 lambda - I3 = 0; with lambda = t + I1/3:
// t^3 + p t + q = 0, p = I2 - I1^2/3, q = -2 I1^3/27 + I1 I2/3 - I3.  When
// D = (q/2)^2 + (p/3)^3 > 0 the roots are t1 = A + B and -(A + B)/2 +- i
// (sqrt 3 / 2)(A - B) with A = cbrt(-q/2 + sqrt D), B = -p / (3 A).
__device__ __forceinline__ double swirl3(const double a[9]) {
  const double i1 = a[0] + a[4] + a[8];
  const double i2 = a[0] * a[4] - a[1] * a[3] + a[0] * a[8] - a[2] * a[6] + a[4] * a[8] - a[5] * a[7];
  const double i3 = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                    a[2] * (a[3] * a[7] - a[4] * a[6]);
  const double p = i2 - i1 * i1 / 3.0;
  const double q = -2.0 * i1 * i1 * i1 / 27.0 + i1 * i2 / 3.0 - i3;
  const double disc = 0.25 * q * q + p * p * p / 27.0;
  if (!(disc > 0.0)) return 0.0;
  const double sq = sqrt(disc);
  // the term without cancellation: sign(-q) (|q|/2 + sqrt D)
  const double m = q > 0.0 ? -(0.5 * q + sq) : (-0.5 * q + sq);
  const double A = cbrt(m);
  const double B = A != 0.0 ? -p / (3.0 * A) : 0.0;
  return 0.8660254037844386 * fabs(A - B);
}

__global__ void k_swirl(const double* __restrict__ g, long long R, double* __restrict__ out) {
  GRID_STRIDE(i, R) {
    double a[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = g[k * R + i];
    out[i] = swirl3(a);
  }
}
void swirling_strength(const double* g, long long R, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_swirl, grid1d(R), dim3(kNT), 0, st, g, R, out);
  TBD_HIP(hipGetLastError());
}

__global__ void k_set_pair(double* dst, double a, double b) {
  if (threadIdx.x == 0) {
    dst[0] = a;
    dst[1] = b;
  }
}
void set_pair(double* dst, double a, double b, hipStream_t st) {
  hipLaunchKernelGGL(k_set_pair, dim3(1), dim3(64), 0, st, dst, a, b);
  TBD_HIP(hipGetLastError());
}

struct BcCA {
  double2* rhs;
  BcArgs a;
};
struct PairsA {
  double* dst[kMaxMembers];
  double v[2 * kMaxMembers];
  int n;
};
__global__ void k_set_pairs(PairsA a) {
  const int i = threadIdx.x;
  if (i < a.n) {
    a.dst[i][0] = a.v[2 * i];
    a.dst[i][1] = a.v[2 * i + 1];
  }
}
void set_pairs(double* const* dst, const double* v, int n, hipStream_t st) {
  TBD_CHECK(n >= 1 && n <= kMaxMembers, "set_pairs > between 1 and kMaxMembers pairs");
  PairsA a;
  for (int i = 0; i < n; ++i) {
    a.dst[i] = dst[i];
    a.v[2 * i] = v[2 * i];
    a.v[2 * i + 1] = v[2 * i + 1];
  }
  a.n = n;
  hipLaunchKernelGGL(k_set_pairs, dim3(1), dim3(64), 0, st, a);
  TBD_HIP(hipGetLastError());
}

struct GatherA {
  const double* src[kMaxMembers];
  double* dst;
  int count;
};
__global__ void k_gather_rows(GatherA a) {
  const int m = blockIdx.x;
  for (int i = threadIdx.x; i < a.count; i += blockDim.x) a.dst[(long long)m * a.count + i] = a.src[m][i];
}
void gather_rows(const double* const* src, int n, int count, double* dst, hipStream_t st) {
  TBD_CHECK(n >= 1 && n <= kMaxMembers, "gather_rows > between 1 and kMaxMembers rows");
  GatherA a;
  for (int i = 0; i < n; ++i) a.src[i] = src[i];
  a.dst = dst;
  a.count = count;
  hipLaunchKernelGGL(k_gather_rows, dim3(n), dim3(64), 0, st, a);
  TBD_HIP(hipGetLastError());
}

template <bool G>
__global__ void k_bc_carrier(BcCA a0, GroupP grp) {
  Blk b;
  const BcCA& A = member<G>(a0, grp, b);
  double2* __restrict__ rhs = A.rhs;
  const BcArgs& a = A.a;
  MEMBER_STRIDE(j, 3 * a.P) {
    const int f = (int)(j / a.P);
    const long long col = j - (long long)f * a.P;
    const double s = a.s, k = a.k;
    const double2 m1t = a.m1t[j], m2t = a.m2t[j], m1b = a.m1b[j], m2b = a.m2b[j];
    double2 top = make_double2(k * ((1 + s) * m1t.x + (0 - s) * m2t.x), k * ((1 + s) * m1t.y + (0 - s) * m2t.y));
    const double2 bot = make_double2(k * ((1 + s) * m1b.x + (0 - s) * m2b.x), k * ((1 + s) * m1b.y + (0 - s) * m2b.y));
    if (a.open_top && f < 2) {
      const double2 n1 = a.n1t[j], n2 = a.n2t[j];
      top = make_double2(k * ((1 + s) * n1.x + (0 - s) * n2.x), k * ((1 + s) * n1.y + (0 - s) * n2.y));
    }
    if (col == 0) top = make_double2(f == 0 ? a.top0 : (f == 1 ? a.top1 : a.top2), 0.0);
    double2* r = rhs + (long long)f * a.S + col * a.nz;
    r[0] = bot;
    r[a.nz - 1] = top;
  }
}
void bc_carrier(double2* rhs, const BcArgs& a, hipStream_t st) {
  glaunch(k_bc_carrier<false>, k_bc_carrier<true>, grid1d(3 * a.P), dim3(kNT), 0, st, BcCA{rhs, a});
}

struct BcSA {
  double2* rhs;
  int nz;
  long long P;
  const double2* ehat;
  const double2* c;
  double ws, cbar, re, sc;
  int erosion;
};
template <bool G>
__global__ void k_bc_scalar(BcSA a0, GroupP grp) {
  Blk b;
  const BcSA& A = member<G>(a0, grp, b);
  double2* __restrict__ rhs = A.rhs;
  const double2* __restrict__ ehat = A.ehat;
  const double2* __restrict__ c = A.c;
  const int nz = A.nz, erosion = A.erosion;
  const long long P = A.P;
  const double ws = A.ws, cbar = A.cbar, re = A.re, sc = A.sc;
  MEMBER_STRIDE(col, P) {
    double2* r = rhs + col * nz;
    double2 bv = make_double2(0.0, 0.0);
    if (erosion) {
      const double2 e = ehat[col], cv = c[col * nz];
      double2 v = make_double2(e.x - ws * cv.x * cbar, e.y - ws * cv.y * cbar);
      const double scale = 10E+6;   // BOUNDARY_COEF_SCALE (common.h:55)
      v = make_double2(-1.0 * v.x * re * sc / cbar * scale, -1.0 * v.y * re * sc / cbar * scale);
      bv = v;
    }
    r[0] = bv;
    r[nz - 1] = make_double2(0.0, 0.0);
  }
}
void bc_scalar(double2* rhs, int nz, long long P, const double2* ehat, const double2* c, double ws, double cbar,
               double re, double sc, int erosion, hipStream_t st) {
  glaunch(k_bc_scalar<false>, k_bc_scalar<true>, grid1d(P), dim3(kNT), 0, st,
          BcSA{rhs, nz, P, ehat, c, ws, cbar, re, sc, erosion});
}

template <typename T>
struct ErosA {
  const T* a;
  const T* b;
  double* e;
  long long n;
  double re, tc, M;
};
template <typename T, bool G>
__global__ void k_erosion(ErosA<T> a0, GroupP grp) {
  Blk b;
  const ErosA<T>& A = member<G>(a0, grp, b);
  const T* __restrict__ pa = A.a;
  const T* __restrict__ pb = A.b;
  double* __restrict__ e = A.e;
  const long long n = A.n;
  const double re = A.re, tc = A.tc, M = A.M;
  MEMBER_STRIDE(i, n) {
    const double x = (double)pa[i], y = (double)pb[i];
    double val = sqrt(x * x + y * y) / re;
    if (val < tc) val = 0;
    else val = M * (val / tc - 1);
    e[i] = val;
  }
}
template <typename T>
void erosion_map(const T* dudz, const T* dvdz, double* e, long long n, double re, double tc, double M,
                 hipStream_t st) {
  glaunch(k_erosion<T, false>, k_erosion<T, true>, grid1d(n), dim3(kNT), 0, st,
          ErosA<T>{dudz, dvdz, e, n, re, tc, M});
}

// One workgroup per z-plane; the plane statistics are FP64 sums.  A plane
// with no negative entries is left untouched after the first reduction (each
// further pass would add 0 to the positive entries and zero the zeros).
struct FiltA {
  double* c;
  long long plane;
  int* touched;
  const int* gate;
};
template <bool G>
__global__ __launch_bounds__(1024) void k_filter(FiltA a0, GroupP grp) {
  Blk b;
  const FiltA& A = member<G>(a0, grp, b);
  const long long plane = A.plane;
  int* touched = A.touched;
  if (A.gate && *A.gate == 0) return;
  __shared__ double red[4][16];
  double* pl = A.c + (long long)b.x * plane;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int it = 0; it < 5; ++it) {
    double pos = 0, npos = 0, neg = 0, nneg = 0;
    for (long long e = threadIdx.x; e < plane; e += blockDim.x) {
      const double v = pl[e];
      if (v > 0) { pos += v; npos += 1; }
      else if (v < 0) { neg += v; nneg += 1; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      pos += __shfl_xor(pos, o);
      npos += __shfl_xor(npos, o);
      neg += __shfl_xor(neg, o);
      nneg += __shfl_xor(nneg, o);
    }
    if (lane == 0) { red[0][wave] = pos; red[1][wave] = npos; red[2][wave] = neg; red[3][wave] = nneg; }
    __syncthreads();
    double P = 0, NP = 0, N = 0, NN = 0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; ++w) { P += red[0][w]; NP += red[1][w]; N += red[2][w]; NN += red[3][w]; }
    __syncthreads();
    if (NN == 0) break;                 // uniform across the block
    if (touched && threadIdx.x == 0) *touched = 1;
    if (P + 0.0 + N < 0) {
      for (long long e = threadIdx.x; e < plane; e += blockDim.x) pl[e] = 0;
    } else {
      const double add = N / NP;
      for (long long e = threadIdx.x; e < plane; e += blockDim.x) {
        const double v = pl[e];
        pl[e] = v > 0 ? v + add : 0.0;
      }
    }
    __syncthreads();
  }
}
void bartnicki_filter(double* c, int nx, int ny, int nz, int* touched, hipStream_t st, const int* gate) {
  glaunch(k_filter<false>, k_filter<true>, dim3(nz), dim3(1024), 0, st, FiltA{c, (long long)nx * ny, touched, gate});
}

// Smallest and largest value of x: per-block partials, then one block that
// sets *need = 1 when the smallest is not above kNegMargin times the largest
// (or the largest is not positive), else 0.
static constexpr int kMinMaxBlocks = 512;
static constexpr double kNegMargin = 1e-4;
template <typename T>
struct MmA {
  const T* x;
  long long n;
  double* part;
};
template <typename T, bool G>
__global__ __launch_bounds__(256) void k_minmax_part(MmA<T> a0, GroupP grp) {
  Blk bk;
  const MmA<T>& A = member<G>(a0, grp, bk);
  const T* __restrict__ x = A.x;
  const long long n = A.n;
  double* part = A.part;
  __shared__ double lo[4], hi[4];
  double a = 1e300, b = -1e300;
  // four independent loads per step of the grid-stride loop, then the rest
  const long long stride = (long long)bk.gx * 256;
  long long i = bk.x * 256LL + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const T v0 = x[i], v1 = x[i + stride], v2 = x[i + 2 * stride], v3 = x[i + 3 * stride];
    a = fmin(fmin(a, (double)v0), fmin((double)v1, fmin((double)v2, (double)v3)));
    b = fmax(fmax(b, (double)v0), fmax((double)v1, fmax((double)v2, (double)v3)));
  }
  for (; i < n; i += stride) {
    const double v = (double)x[i];
    a = fmin(a, v);
    b = fmax(b, v);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a = fmin(a, __shfl_xor(a, o));
    b = fmax(b, __shfl_xor(b, o));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { lo[wave] = a; hi[wave] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) { a = fmin(a, lo[w]); b = fmax(b, hi[w]); }
    part[2 * bk.x] = fmin(a, lo[0]);
    part[2 * bk.x + 1] = fmax(b, hi[0]);
  }
}
struct MmFA {
  const double* part;
  int nb;
  int* need;
};
template <bool G>
__global__ __launch_bounds__(64) void k_minmax_final(MmFA a0, GroupP grp) {
  Blk bk;
  const MmFA& A = member<G>(a0, grp, bk);
  const double* __restrict__ part = A.part;
  const int nb = A.nb;
  int* need = A.need;
  double a = 1e300, b = -1e300;
  for (int i = threadIdx.x; i < nb; i += 64) {
    a = fmin(a, part[2 * i]);
    b = fmax(b, part[2 * i + 1]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a = fmin(a, __shfl_xor(a, o));
    b = fmax(b, __shfl_xor(b, o));
  }
  if (threadIdx.x == 0) *need = (b <= 0.0 || a <= kNegMargin * b) ? 1 : 0;
}
template <typename T>
void negativity_check(const T* x, long long n, double* part, int* need, hipStream_t st) {
  const int nb = (int)std::min<long long>(kMinMaxBlocks, std::max<long long>(1, (n + 255) / 256));
  glaunch(k_minmax_part<T, false>, k_minmax_part<T, true>, dim3(nb), dim3(256), 0, st, MmA<T>{x, n, part});
  glaunch(k_minmax_final<false>, k_minmax_final<true>, dim3(1), dim3(64), 0, st, MmFA{part, nb, need});
}
template void negativity_check<float>(const float*, long long, double*, int*, hipStream_t);
template void negativity_check<double>(const double*, long long, double*, int*, hipStream_t);

struct Kx0A {
  double2* s;
  int nyc, nz;
  const int* gate;
};
template <bool G>
__global__ void k_kx0_round_trip(Kx0A a0, GroupP grp) {
  Blk bk;
  const Kx0A& A = member<G>(a0, grp, bk);
  if (A.gate && *A.gate != 0) return;
  double2* __restrict__ s = A.s;
  const int nyc = A.nyc, nz = A.nz;
  const int half = nyc / 2;
  for (int i = bk.x * blockDim.x + threadIdx.x; i < (half + 1) * nz; i += bk.gx * blockDim.x) {
    const int iy = i / nz, z = i - iy * nz;
    double2* a = s + (long long)iy * nz + z;
    if (iy == 0) {
      a->y = 0.0;
    } else if (iy == half) {          // ky = -nyc/2: its partner +nyc/2 is not retained
      a->x = 0.5 * a->x;
      a->y = 0.5 * a->y;
    } else {
      double2* b = s + (long long)(nyc - iy) * nz + z;
      const double2 av = *a, bv = *b;
      const double hx = 0.5 * (av.x + bv.x), hy = 0.5 * (av.y - bv.y);
      *a = make_double2(hx, hy);
      *b = make_double2(hx, -hy);
    }
  }
}
void kx0_round_trip(double2* s, int nyc, int nz, const int* gate, hipStream_t st) {
  glaunch(k_kx0_round_trip<false>, k_kx0_round_trip<true>, dim3(((nyc / 2 + 1) * nz + 255) / 256), dim3(256), 0, st,
          Kx0A{s, nyc, nz, gate});
}

// Max-norm and CFL diagnostics: per-block partial maxima (z planes strided
// over blockIdx.y, plane points over blockIdx.x), then one block reduces the
// partials.  No atomics, so the result does not depend on scheduling.
static constexpr int kDiagVals = 10;
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
template <typename T>
struct DiagA {
  const T* U;
  long long R;
  int plane, nz;
  double glx, gly;
  const double* glz;
  double* part;
};
template <typename T, bool G>
__global__ __launch_bounds__(kNT) void k_diag_part(DiagA<T> a0, GroupP grp) {
  Blk bk;
  const DiagA<T>& A = member<G>(a0, grp, bk);
  const T* __restrict__ U = A.U;
  const long long R = A.R;
  const int plane = A.plane, nz = A.nz;
  const double glx = A.glx, gly = A.gly;
  const double* __restrict__ glz = A.glz;
  double* __restrict__ part = A.part;
  double m[kDiagVals] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int z = bk.y; z < nz; z += bk.gy) {
    const double gz = glz[z];
    const T* u = U + (long long)z * plane;
    double a0 = 0, a1 = 0, a2 = 0, cf = 0;
    for (int e = bk.x * blockDim.x + threadIdx.x; e < plane; e += bk.gx * blockDim.x) {
      const double a = fabs((double)u[e]), b = fabs((double)u[R + e]), w = fabs((double)u[2 * R + e]);
      a0 = fmax(a0, a);
      a1 = fmax(a1, b);
      a2 = fmax(a2, w);
      cf = fmax(cf, a / glx + b / gly + w / gz);
    }
    m[0] = fmax(m[0], a0);
    m[1] = fmax(m[1], a1);
    m[2] = fmax(m[2], a2);
    m[9] = fmax(m[9], cf);
    if (z == 0) { m[3] = fmax(m[3], a0); m[4] = fmax(m[4], a1); m[5] = fmax(m[5], a2); }
    if (z == nz - 1) { m[6] = fmax(m[6], a0); m[7] = fmax(m[7], a1); m[8] = fmax(m[8], a2); }
  }
  __shared__ double red[kDiagVals][kNT / 64];
#pragma unroll
  for (int i = 0; i < kDiagVals; ++i) {
    const double v = wave_max(m[i]);
    if ((threadIdx.x & 63) == 0) red[i][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < kDiagVals) {
    double v = 0;
    for (int w = 0; w < kNT / 64; ++w) v = fmax(v, red[threadIdx.x][w]);
    part[(long long)(bk.y * bk.gx + bk.x) * kDiagVals + threadIdx.x] = v;
  }
}
struct DiagFA {
  const double* part;
  int nblk;
  double* feat;
};
template <bool G>
__global__ __launch_bounds__(kNT) void k_diag_final(DiagFA a0, GroupP grp) {
  Blk bk;
  const DiagFA& A = member<G>(a0, grp, bk);
  const double* __restrict__ part = A.part;
  const int nblk = A.nblk;
  double* feat = A.feat;
  __shared__ double red[kDiagVals][kNT / 64];
  for (int i = 0; i < kDiagVals; ++i) {
    double v = 0;
    for (int b = threadIdx.x; b < nblk; b += kNT) v = fmax(v, part[(long long)b * kDiagVals + i]);
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[i][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < kDiagVals) {
    double v = 0;
    for (int w = 0; w < kNT / 64; ++w) v = fmax(v, red[threadIdx.x][w]);
    feat[threadIdx.x < 9 ? threadIdx.x : 19] = v;
  }
}
template <typename T>
void diagnostics(const T* uvw, long long R, int nx, int ny, int nz, double glx, double gly, const double* glz,
                 double* feat, double* part, hipStream_t st) {
  const int plane = nx * ny;
  int nbx = (plane + kNT * 8 - 1) / (kNT * 8);
  nbx = nbx < 1 ? 1 : (nbx > 32 ? 32 : nbx);
  int nby = kDiagMaxBlocks / nbx;
  nby = nby > nz ? nz : (nby < 1 ? 1 : nby);
  glaunch(k_diag_part<T, false>, k_diag_part<T, true>, dim3(nbx, nby), dim3(kNT), 0, st,
          DiagA<T>{uvw, R, plane, nz, glx, gly, glz, part});
  glaunch(k_diag_final<false>, k_diag_final<true>, dim3(1), dim3(kNT), 0, st, DiagFA{part, nbx * nby, feat});
}

// one workgroup per z: reductions over the P spectral columns of that plane
__global__ __launch_bounds__(256) void k_phase_stats(const double2* __restrict__ uvw, const double2* __restrict__ c,
                                                     long long S, int nz, int nyc, long long P, double* acc) {
  const int z = blockIdx.x;
  double s[4] = {0, 0, 0, 0};
  for (long long col = 1 + threadIdx.x; col < P; col += blockDim.x) {
    const double w = (col / nyc) > 0 ? 2.0 : 1.0;
    const long long off = col * nz + z;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const double2 v = uvw[f * S + off];
      s[f] += w * (v.x * v.x + v.y * v.y);
    }
    if (c) {
      const double2 v = c[off];
      s[3] += w * (v.x * v.x + v.y * v.y);
    }
  }
  __shared__ double red[4][4];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    double v = s[f];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[f][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int f = threadIdx.x;
    double v = 0;
    for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) v += red[f][wv];
    acc[(long long)(4 + f) * nz + z] += v;
  } else if (threadIdx.x < 8) {
    const int f = threadIdx.x - 4;
    const double m = f < 3 ? uvw[f * S + z].x : (c ? c[z].x : 0.0);
    acc[(long long)f * nz + z] += m;
  }
}
void phase_stats(const double2* uvw, const double2* c, long long S, int nz, int nyc, long long P, double* acc,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_phase_stats, dim3(nz), dim3(256), 0, st, uvw, c, S, nz, nyc, P, acc);
  TBD_HIP(hipGetLastError());
}

struct TrapzA {
  const double2* c;
  const double* z;
  int nz;
  double* out;
};
template <bool G>
__global__ void k_trapz(TrapzA a0, GroupP grp) {
  Blk bk;
  const TrapzA& A = member<G>(a0, grp, bk);
  const double2* __restrict__ c = A.c;
  const double* __restrict__ z = A.z;
  const int nz = A.nz;
  double* out = A.out;
  __shared__ double part[64];
  double s = 0.0;
  for (int i = threadIdx.x; i + 1 < nz; i += blockDim.x) s += (z[i + 1] - z[i]) * (c[i].x + c[i + 1].x) / 2;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
    *out = t / 2;
  }
}
void trapz_mean(const double2* c, const double* gridz, int nz, double* out, hipStream_t st) {
  glaunch(k_trapz<false>, k_trapz<true>, dim3(1), dim3(256), 0, st, TrapzA{c, gridz, nz, out});
}

__global__ void k_gather_real(const double2* __restrict__ src, const long long* __restrict__ idx, int n,
                              double* __restrict__ dst) {
  const int i = threadIdx.x;
  if (i < n) dst[i] = src[idx[i]].x;
}
void gather_real(const double2* src, const long long* idx, int n, double* dst, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_real, dim3(1), dim3(64), 0, st, src, idx, n, dst);
  TBD_HIP(hipGetLastError());
}

template <typename W>
struct ZeroA {
  W* p;
  long long n;
};
template <typename W, bool G>
__global__ void k_zero(ZeroA<W> a0, GroupP grp) {
  Blk b;
  const ZeroA<W>& A = member<G>(a0, grp, b);
  W* p = A.p;
  MEMBER_STRIDE(i, A.n) p[i] = (W)0;
}
void fill_zero(void* p, size_t bytes, hipStream_t st) {
  typedef unsigned long long u64;
  typedef unsigned char u8;
  if (bytes % 8 == 0) {
    glaunch(k_zero<u64, false>, k_zero<u64, true>, grid1d(bytes / 8), dim3(kNT), 0, st,
            ZeroA<u64>{(u64*)p, (long long)(bytes / 8)});
  } else {
    glaunch(k_zero<u8, false>, k_zero<u8, true>, grid1d(bytes), dim3(kNT), 0, st, ZeroA<u8>{(u8*)p, (long long)bytes});
  }
}

struct ScRA {
  const double2* src;
  double a, b;
  double* out;
};
template <bool G>
__global__ void k_scaled_real(ScRA a0, GroupP grp) {
  Blk bk;
  const ScRA& A = member<G>(a0, grp, bk);
  if (threadIdx.x == 0) *A.out = (A.a * A.src[0].x) * A.b;
}
void scaled_real(const double2* src, double a, double b, double* out, hipStream_t st) {
  glaunch(k_scaled_real<false>, k_scaled_real<true>, dim3(1), dim3(64), 0, st, ScRA{src, a, b, out});
}

// dst[i] = src[i] for i < n, then dst[10 + j] = wall[src6[j]] for j < 6 (the
// step's features: the diagnostics' 50 values and the mean-mode wall gradients)
struct FeatA {
  double* dst;
  const double* src;
  int n;
  const double* wall;
  int src6[6];
};
template <bool G>
__global__ void k_features(FeatA a0, GroupP grp) {
  Blk bk;
  const FeatA& A = member<G>(a0, grp, bk);
  const int i = threadIdx.x;
  if (i < A.n) A.dst[i] = A.src[i];
  __syncthreads();
  if (i < 6) A.dst[10 + i] = A.wall[A.src6[i]];
}
void collect_features(double* dst, const double* src, int n, const double* wall, const int* src6, hipStream_t st) {
  TBD_CHECK(n >= 16 && n <= 64, "collect_features > n must be in 16..64");
  FeatA a;
  a.dst = dst;
  a.src = src;
  a.n = n;
  a.wall = wall;
  for (int j = 0; j < 6; ++j) a.src6[j] = src6[j];
  glaunch(k_features<false>, k_features<true>, dim3(1), dim3(64), 0, st, a);
}

// dst = src, n bytes (a multiple of 4)
struct CopyA {
  unsigned int* dst;
  const unsigned int* src;
  long long n;
};
template <bool G>
__global__ void k_copy4(CopyA a0, GroupP grp) {
  Blk b;
  const CopyA& A = member<G>(a0, grp, b);
  MEMBER_STRIDE(i, A.n) A.dst[i] = A.src[i];
}
void copy_words(void* dst, const void* src, size_t bytes, hipStream_t st) {
  TBD_CHECK(bytes % 4 == 0, "copy_words > size must be a multiple of 4 bytes");
  glaunch(k_copy4<false>, k_copy4<true>, grid1d(bytes / 4), dim3(kNT), 0, st,
          CopyA{(unsigned int*)dst, (const unsigned int*)src, (long long)(bytes / 4)});
}

__global__ void k_copy_d2(double2* __restrict__ d, const double2* __restrict__ s, long long n) {
  GRID_STRIDE(i, n) d[i] = s[i];
}
void copy_d2(double2* dst, const double2* src, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_d2, grid1d(n), dim3(kNT), 0, st, dst, src, n);
  TBD_HIP(hipGetLastError());
}

template void products_conv<double>(const double*, const double*, double*, long long, hipStream_t);
template void products_conv<float>(const float*, const float*, float*, long long, hipStream_t);
template void products_div<double>(const double*, double*, long long, hipStream_t);
template void products_div<float>(const float*, float*, long long, hipStream_t);
template void scalar_products_conv<double>(const double*, double, double, double, const double*, double*, long long,
                                           hipStream_t);
template void scalar_products_conv<float>(const float*, float, float, float, const float*, float*, long long,
                                          hipStream_t);
template void scalar_products_div<double, double>(const double*, double, double, double, const double*, double*,
                                                  long long, hipStream_t);
template void scalar_products_div<float, float>(const float*, float, float, float, const float*, float*, long long,
                                                hipStream_t);
template void scalar_products_div<float, double>(const float*, float, float, float, const double*, float*, long long,
                                                 hipStream_t);
template void h_update<double2>(double2*, const double2*, const double2*, const HUpd&, int, long long, hipStream_t);
template void h_update<float2>(double2*, const float2*, const double2*, const HUpd&, int, long long, hipStream_t);
template void erosion_map<double>(const double*, const double*, double*, long long, double, double, double,
                                  hipStream_t);
template void erosion_map<float>(const float*, const float*, double*, long long, double, double, double, hipStream_t);
template void diagnostics<double>(const double*, long long, int, int, int, double, double, const double*, double*, double*,
                                  hipStream_t);
template void diagnostics<float>(const float*, long long, int, int, int, double, double, const double*, double*, double*,
                                 hipStream_t);

}  // namespace tbd
