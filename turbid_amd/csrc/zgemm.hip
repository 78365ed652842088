// MFMA z-contractions (see zgemm.h).
//
// A workgroup (4 waves) stages a panel of 16 complex columns (all nz rows)
// in LDS, then each wave sweeps 16-row tiles of the operator with
// v_mfma_{f64,f32}_16x16x4: the A operand is 16 operator rows x 4 k, the B
// operand 4 k x 16 columns of the panel, with separate accumulators for the
// real and imaginary parts (the operator is real).  Within a block of 16 k
// each lane group g = lane/16 owns k = kb + 4g .. kb + 4g + 3, so operator
// rows and panel columns are read as contiguous runs.  Rows and k beyond the
// last full 16 (nz = 16*T + r, r = 1 for the CGL sizes 65/97/129/257) are
// finished on the vector ALUs.
#include "zgemm.h"

namespace tbd {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
// Accumulator layout of the 16x16x4 instructions (lane l, g = l/16, result
// register r, column l%16), measured on gfx950: the f32 variant holds row
// 4g + r, the f64 variant holds row g + 4r.  Operands: A[l%16][l/16] and
// B[l/16][l%16] for both.
template <typename T> struct AccT;
template <> struct AccT<double> {
  typedef d4v type;
  static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};
template <> struct AccT<float> {
  typedef f4v type;
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

__device__ __forceinline__ d4v mfma16(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma16(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

static constexpr int kNC = 16;    // complex columns per panel
static constexpr int kNT = 256;   // threads per workgroup (4 waves)

__host__ __device__ inline int panel_ld(int nz) { return nz | 1; }

// Y(row, c) = sum_k op[row][k] X(k, c) for all rows of the panel; emit(row, c, value)
template <typename T, class Emit>
__device__ __forceinline__ void gemm_panel(const T* __restrict__ op, int lda, int nz, const C2<T>* xs, int ldx,
                                           Emit emit) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int MT = nz / 16;
  const int KM = MT * 16;
  const int g = lane >> 4;
  const int cl = lane & 15;
  for (int mt = wave; mt < MT; mt += kNT / 64) {
    typename AccT<T>::type are = {0, 0, 0, 0};
    typename AccT<T>::type aim = {0, 0, 0, 0};
    const T* arow = op + (long long)(mt * 16 + cl) * lda + 4 * g;
    const C2<T>* xc = xs + cl * ldx + 4 * g;
    for (int kb = 0; kb < KM; kb += 16) {
      const T a0 = arow[kb], a1 = arow[kb + 1], a2 = arow[kb + 2], a3 = arow[kb + 3];
      const C2<T> b0 = xc[kb], b1 = xc[kb + 1], b2 = xc[kb + 2], b3 = xc[kb + 3];
      are = mfma16(a0, b0.x, are);
      aim = mfma16(a0, b0.y, aim);
      are = mfma16(a1, b1.x, are);
      aim = mfma16(a1, b1.y, aim);
      are = mfma16(a2, b2.x, are);
      aim = mfma16(a2, b2.y, aim);
      are = mfma16(a3, b3.x, are);
      aim = mfma16(a3, b3.y, aim);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = mt * 16 + AccT<T>::row(g, r);
      T vr = are[r], vi = aim[r];
      const T* orow = op + (long long)row * lda;
      for (int k = KM; k < nz; ++k) {
        const T m = orow[k];
        const C2<T> xv = xs[cl * ldx + k];
        vr += m * xv.x;
        vi += m * xv.y;
      }
      emit(row, cl, mk<T>(vr, vi));
    }
  }
  // rows KM .. nz-1 on the vector ALUs
  for (int idx = threadIdx.x; idx < (nz - KM) * kNC; idx += kNT) {
    const int rr = idx / kNC, c = idx - rr * kNC;
    const int row = KM + rr;
    const T* orow = op + (long long)row * lda;
    T vr = 0, vi = 0;
    for (int k = 0; k < nz; ++k) {
      const T m = orow[k];
      const C2<T> xv = xs[c * ldx + k];
      vr += m * xv.x;
      vi += m * xv.y;
    }
    emit(row, c, mk<T>(vr, vi));
  }
}

template <typename TY, typename T>
__device__ __forceinline__ void st(TY* p, C2<T> v) {
  TY o;
  o.x = v.x;
  o.y = v.y;
  *p = o;
}
template <typename T, typename TX>
__device__ __forceinline__ C2<T> ld(const TX* p) {
  const TX v = *p;
  return mk<T>((T)v.x, (T)v.y);
}

template <typename T, typename TX, typename TY>
__device__ __forceinline__ void zepilogue(const ZArgs& a, int f, long long col, int row, C2<T> v, C2<T> xv) {
  const int nz = a.nz;
  const long long off = col * nz + row;
  TY* y = (TY*)a.y;
  const T kx = (T)a.kx[col / a.nyc];
  const T ky = (T)a.ky[col % a.nyc];
  switch (a.mode) {
    case Z_PLAIN:
      st<TY, T>(y + f * a.ys + off, v);
      break;
    case Z_GRAD:
      st<TY, T>(y + (3 * f + 0) * a.ys + off, cmuli<T>(kx, xv));
      st<TY, T>(y + (3 * f + 1) * a.ys + off, cmuli<T>(ky, xv));
      st<TY, T>(y + (3 * f + 2) * a.ys + off, v);
      break;
    case Z_RHS: {
      const C2<T> h = ld<T, TX>((const TX*)a.h + f * a.hs + off);
      const T k2x = -(kx * kx), k2y = -(ky * ky);
      const C2<T> lap = cadd<T>(cadd<T>(v, cscale<T>(k2x, xv)), cscale<T>(k2y, xv));
      const C2<T> r = csub<T>(cadd<T>(cscale<T>((T)a.s2, h), cscale<T>((T)a.s1, xv)), lap);
      st<TY, T>(y + f * a.ys + off, r);
    } break;
    case Z_DIV: {
      const C2<T> su = ld<T, TX>((const TX*)a.e0 + off);
      const C2<T> sv = ld<T, TX>((const TX*)a.e1 + off);
      const C2<T> dv = cadd<T>(cadd<T>(v, cmuli<T>(kx, su)), cmuli<T>(ky, sv));
      st<TY, T>(y + off, cscale<T>((T)a.q, dv));
    } break;
    case Z_CORRECT: {
      const C2<T> gx = cmuli<T>(kx, xv), gy = cmuli<T>(ky, xv);
      const T cf = (T)a.cf;
      const C2<T> su = ld<T, TX>((const TX*)a.e0 + off);
      const C2<T> sv = ld<T, TX>((const TX*)a.e0 + a.es + off);
      const C2<T> sw = ld<T, TX>((const TX*)a.e0 + 2 * a.es + off);
      st<TY, T>(y + off, csub<T>(su, cscale<T>(cf, gx)));
      st<TY, T>(y + a.ys + off, csub<T>(sv, cscale<T>(cf, gy)));
      C2<T> w = csub<T>(sw, cscale<T>(cf, v));
      if (col == 0) w = mk<T>((T)0, (T)0);
      st<TY, T>(y + 2 * a.ys + off, w);
      if (row == 0 || row == nz - 1) {
        double2* wb = (row == 0) ? a.wall_bot : a.wall_top;
        const long long P = a.P;
        wb[col] = make_double2(gx.x, gx.y);
        wb[P + col] = make_double2(gy.x, gy.y);
        wb[2 * P + col] = make_double2(v.x, v.y);
        if (a.open_top) {
          double2* w2 = (row == 0) ? a.wall2_bot : a.wall2_top;
          const C2<T> hx = cmuli<T>(kx, v), hy = cmuli<T>(ky, v);
          w2[col] = make_double2(hx.x, hx.y);
          w2[P + col] = make_double2(hy.x, hy.y);
        }
      }
    } break;
    case Z_GRADTMP: {
      const C2<T> f0 = ld<T, TX>((const TX*)a.e0 + f * a.es + off);
      const C2<T> f1 = ld<T, TX>((const TX*)a.e1 + f * a.es + off);
      st<TY, T>(y + f * a.ys + off, cadd<T>(cadd<T>(cmuli<T>(kx, f0), cmuli<T>(ky, f1)), v));
    } break;
    default:
      break;
  }
}

template <typename T, typename TX, typename TY>
__global__ __launch_bounds__(kNT) void k_zgemm(ZArgs a) {
  extern __shared__ char smem_raw[];
  C2<T>* xs = reinterpret_cast<C2<T>*>(smem_raw);
  const int nz = a.nz;
  const int ldx = panel_ld(nz);
  const int f = blockIdx.y;
  const long long c0 = (long long)blockIdx.x * kNC;
  const int ncol = (int)min((long long)kNC, a.P - c0);
  const TX* x = (const TX*)a.x + f * a.xs + c0 * nz;
  for (int idx = threadIdx.x; idx < kNC * nz; idx += kNT) {
    const int c = idx / nz, k = idx - c * nz;
    xs[c * ldx + k] = (c < ncol) ? ld<T, TX>(x + (long long)c * nz + k) : mk<T>((T)0, (T)0);
  }
  __syncthreads();
  gemm_panel<T>((const T*)a.op, a.lda, nz, xs, ldx, [&](int row, int c, C2<T> v) {
    if (c < ncol) zepilogue<T, TX, TY>(a, f, c0 + c, row, v, xs[c * ldx + row]);
  });
}

template <typename T, typename TX, typename TY>
void zgemm(const ZArgs& a, int nfield, hipStream_t st) {
  TBD_CHECK(a.nz >= 4 && a.lda >= a.nz && a.P > 0, "zgemm > bad shape");
  dim3 grid(ceil_div(a.P, kNC), nfield);
  const size_t lds = sizeof(C2<T>) * kNC * panel_ld(a.nz);
  hipLaunchKernelGGL((k_zgemm<T, TX, TY>), grid, dim3(kNT), lds, st, a);
  TBD_HIP(hipGetLastError());
}

template void zgemm<double, double2, double2>(const ZArgs&, int, hipStream_t);
template void zgemm<float, double2, float2>(const ZArgs&, int, hipStream_t);
template void zgemm<float, float2, float2>(const ZArgs&, int, hipStream_t);

// ---------------------------------------------------------------------------
// fused Helmholtz
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kNT) void k_helmholtz(HArgs a) {
  extern __shared__ char smem_raw[];
  const int nz = a.nz;
  const int ldx = panel_ld(nz);
  double2* bufA = reinterpret_cast<double2*>(smem_raw);
  double2* bufB = bufA + kNC * ldx;
  const int fi = blockIdx.y;
  const double* vinv = fi == 0 ? a.f[0].vinv : (fi == 1 ? a.f[1].vinv : a.f[2].vinv);
  const double* vnorm = fi == 0 ? a.f[0].vnorm : (fi == 1 ? a.f[1].vnorm : a.f[2].vnorm);
  const double* lam = fi == 0 ? a.f[0].lambda : (fi == 1 ? a.f[1].lambda : a.f[2].lambda);
  const double* fix = fi == 0 ? a.f[0].fix : (fi == 1 ? a.f[1].fix : a.f[2].fix);
  const double sigma = fi == 0 ? a.f[0].sigma : (fi == 1 ? a.f[1].sigma : a.f[2].sigma);
  const int nullrow = fi == 0 ? a.f[0].nullrow : (fi == 1 ? a.f[1].nullrow : a.f[2].nullrow);
  const double2* x = fi == 0 ? a.f[0].x : (fi == 1 ? a.f[1].x : a.f[2].x);
  double2* y = fi == 0 ? a.f[0].y : (fi == 1 ? a.f[1].y : a.f[2].y);
  const long long c0 = (long long)blockIdx.x * kNC;
  const int ncol = (int)min((long long)kNC, (long long)a.P - c0);
  for (int idx = threadIdx.x; idx < kNC * nz; idx += kNT) {
    const int c = idx / nz, k = idx - c * nz;
    bufA[c * ldx + k] = (c < ncol) ? x[(c0 + c) * nz + k] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  if (a.mode == 0) {
    // forward fix: X_i += fbf_i X_0 + ftf_i X_{N-1}, i = 1..N-2
    const double* fbf = fix;
    const double* ftf = fix + nz;
    for (int idx = threadIdx.x; idx < kNC * (nz - 2); idx += kNT) {
      const int c = idx / (nz - 2), iz = 1 + idx - c * (nz - 2);
      const double2 x0 = bufA[c * ldx], xn = bufA[c * ldx + nz - 1];
      const double2 xi = bufA[c * ldx + iz];
      const double2 t = cadd<double>(cscale<double>(fbf[iz], x0), cscale<double>(ftf[iz], xn));
      bufA[c * ldx + iz] = cadd<double>(xi, t);
    }
  } else {
    for (int c = threadIdx.x; c < kNC; c += kNT) {
      bufA[c * ldx] = make_double2(0.0, 0.0);
      bufA[c * ldx + nz - 1] = make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  // Z = R o (V~^-1 X) into bufB
  gemm_panel<double>(vinv, a.lda, nz, bufA, ldx, [&](int row, int c, double2 v) {
    double r = 1.0;
    if (row > 0 && row < nz - 1) {
      const long long col = c0 + c;
      const long long cc = col < a.P ? col : 0;
      const double kx = a.kx[cc / a.nyc], ky = a.ky[cc % a.nyc];
      const double k2 = (-(kx * kx)) + (-(ky * ky));
      r = 1.0 / (((k2 + lam[row]) + sigma));
      if (row == nullrow && col == 0) r = 0.0;
    }
    bufB[c * ldx + row] = cscale<double>(r, v);
  });
  __syncthreads();
  // Y = V~ Z into bufA
  gemm_panel<double>(vnorm, a.lda, nz, bufB, ldx, [&](int row, int c, double2 v) { bufA[c * ldx + row] = v; });
  __syncthreads();
  // sequential backward fixes: row 0 from Y, then row N-1 from the updated Y
  const double* fbb = fix + 2 * nz;
  const double* fbt = fix + 3 * nz;
  const int c = threadIdx.x >> 4;      // 16 threads per column
  const int part = threadIdx.x & 15;
  for (int pass = 0; pass < 2; ++pass) {
    const double* w = pass == 0 ? fbb : fbt;
    double sr = 0.0, si = 0.0;
    for (int j = part; j < nz; j += 16) {
      const double2 v = bufA[c * ldx + j];
      sr += w[j] * v.x;
      si += w[j] * v.y;
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
      sr += __shfl_xor(sr, o, 16);
      si += __shfl_xor(si, o, 16);
    }
    __syncthreads();
    if (part == 0) bufA[c * ldx + (pass == 0 ? 0 : nz - 1)] = make_double2(sr, si);
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < kNC * nz; idx += kNT) {
    const int cc = idx / nz, k = idx - cc * nz;
    if (cc < ncol) y[(c0 + cc) * nz + k] = bufA[cc * ldx + k];
  }
}

void helmholtz(const HArgs& a, int nfield, hipStream_t st) {
  TBD_CHECK(a.nz >= 4 && a.lda >= a.nz && a.P > 0 && nfield >= 1 && nfield <= 3, "helmholtz > bad shape");
  dim3 grid(ceil_div(a.P, kNC), nfield);
  const size_t lds = sizeof(double2) * 2 * kNC * panel_ld(a.nz);
  hipLaunchKernelGGL(k_helmholtz, grid, dim3(kNT), lds, st, a);
  TBD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// wall rows of M x: one wave per column
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kNT) void k_zrows_wall(const double* __restrict__ op, int lda, int nz,
                                                   const double2* __restrict__ x, long long P,
                                                   double2* __restrict__ bot, double2* __restrict__ top) {
  const int lane = threadIdx.x & 63;
  const long long col = (long long)blockIdx.x * (kNT / 64) + (threadIdx.x >> 6);
  if (col >= P) return;
  const double* r0 = op;
  const double* r1 = op + (long long)(nz - 1) * lda;
  double br = 0, bi = 0, tr = 0, ti = 0;
  for (int k = lane; k < nz; k += 64) {
    const double2 v = x[col * nz + k];
    br += r0[k] * v.x;
    bi += r0[k] * v.y;
    tr += r1[k] * v.x;
    ti += r1[k] * v.y;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    br += __shfl_xor(br, o);
    bi += __shfl_xor(bi, o);
    tr += __shfl_xor(tr, o);
    ti += __shfl_xor(ti, o);
  }
  if (lane == 0) {
    if (bot) bot[col] = make_double2(br, bi);
    if (top) top[col] = make_double2(tr, ti);
  }
}

void zrows_wall(const double* op, int lda, int nz, const double2* x, long long P, double2* bottom, double2* top,
                hipStream_t st) {
  dim3 grid(ceil_div(P, kNT / 64));
  hipLaunchKernelGGL(k_zrows_wall, grid, dim3(kNT), 0, st, op, lda, nz, x, P, bottom, top);
  TBD_HIP(hipGetLastError());
}

// one wave per (field, wall): 3 fields x 2 rows
__global__ __launch_bounds__(384) void k_wall_mean(const double* __restrict__ op, int lda, int nz,
                                                   const double2* __restrict__ x, long long S, double* out) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;        // 0..5
  const int f = w >> 1, top = w & 1;
  const double* row = op + (long long)(top ? nz - 1 : 0) * lda;
  const double2* col = x + (long long)f * S;
  double s = 0.0;
  for (int k = lane; k < nz; k += 64) s += row[k] * col[k].x;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) out[w] = s;
}

void wall_mean_rows(const double* op, int lda, int nz, const double2* x, long long S, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_wall_mean, dim3(1), dim3(384), 0, st, op, lda, nz, x, S, out);
  TBD_HIP(hipGetLastError());
}

}  // namespace tbd
