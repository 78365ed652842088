// MFMA z-contractions (see zgemm.h).
//
// A workgroup (8 waves) stages a panel of ncol = 16*ncg complex columns (all
// nz rows) in LDS.  The work is split into units of two 16-row operator tiles
// x 16 panel columns; each wave takes units round-robin and runs
// v_mfma_{f64,f32}_16x16x4 with the A operand = 16 operator rows x 4 k and the
// B operand = 4 k x 16 panel columns, separate accumulators for the real and
// imaginary parts (the operator is real), so one B read feeds four MFMAs.
// Within a block of 16 k each lane group g = lane/16 owns k = kb + 4g ..
// kb + 4g + 3, so operator rows and panel columns are read as contiguous runs;
// the operator values of the next block are loaded while the current block's
// MFMAs issue.  Rows and k beyond the last full 16 (nz = 16*T + r, r = 1 for
// the CGL sizes 65/97/129/257) are finished on the vector ALUs.
//
// ncg is chosen per nz so that every wave has a unit (nz = 257: 1 group,
// 129: 2, 97: 3, 65: 4).  In buffered mode the results stay in registers
// until every wave has finished reading the panel, so the panel can be
// overwritten with them (the fused Helmholtz solve needs one LDS panel, not
// two, and two workgroups fit per CU).
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

template <typename T> struct V4 { T x, y, z, w; };
__device__ __forceinline__ V4<double> ld4(const double* p) {
  const double2 u = reinterpret_cast<const double2*>(p)[0];
  const double2 v = reinterpret_cast<const double2*>(p)[1];
  return {u.x, u.y, v.x, v.y};
}
__device__ __forceinline__ V4<float> ld4(const float* p) {
  const float4 u = *reinterpret_cast<const float4*>(p);
  return {u.x, u.y, u.z, u.w};
}

static constexpr int kNT = 512;           // threads per workgroup
static constexpr int kWaves = kNT / 64;
static constexpr int kMaxU = 2;           // units (and VALU tail outputs) per thread in buffered mode

struct Panel {
  int nz, MT, KM, npair, ncg, ncol, ldx;
};

__host__ __device__ inline int panel_ld(int nz) { return nz | 1; }

__host__ inline Panel make_panel(int nz, size_t elem, size_t lds_limit) {
  Panel pn;
  pn.nz = nz;
  pn.MT = nz / 16;
  pn.KM = pn.MT * 16;
  pn.npair = (pn.MT + 1) / 2;
  pn.ldx = panel_ld(nz);
  int ncg = pn.npair > 0 ? (kWaves + pn.npair - 1) / pn.npair : 4;
  if (ncg > 4) ncg = 4;
  while (ncg > 1 && (size_t)16 * ncg * pn.ldx * elem > lds_limit) --ncg;
  pn.ncg = ncg;
  pn.ncol = 16 * ncg;
  return pn;
}

// One unit: rows mt0*16 .. mt0*16+31 (second tile if TWO) x the 16 panel
// columns of group cg.  out[t][r] = row mt0*16 + 16t + AccT::row(g, r), column
// cg*16 + lane%16.
template <typename T, bool TWO>
__device__ __forceinline__ void unit_mfma(const T* __restrict__ op, int lda, const Panel& pn,
                                          const C2<T>* __restrict__ xs, int mt0, int cg, C2<T> (&out)[2][4]) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int cl = lane & 15;
  typedef typename AccT<T>::type Acc;
  Acc r0 = {0, 0, 0, 0}, i0 = {0, 0, 0, 0}, r1 = {0, 0, 0, 0}, i1 = {0, 0, 0, 0};
  const T* A0 = op + (long long)(mt0 * 16 + cl) * lda + 4 * g;
  const T* A1 = A0 + 16LL * lda;
  const C2<T>* B = xs + (cg * 16 + cl) * pn.ldx + 4 * g;
  V4<T> a0 = ld4(A0), a1 = a0;
  if (TWO) a1 = ld4(A1);
  for (int kb = 0; kb < pn.KM; kb += 16) {
    const C2<T> b0 = B[kb], b1 = B[kb + 1], b2 = B[kb + 2], b3 = B[kb + 3];
    V4<T> n0 = a0, n1 = a1;
    if (kb + 16 < pn.KM) {
      n0 = ld4(A0 + kb + 16);
      if (TWO) n1 = ld4(A1 + kb + 16);
    }
    r0 = mfma16(a0.x, b0.x, r0);
    i0 = mfma16(a0.x, b0.y, i0);
    if (TWO) { r1 = mfma16(a1.x, b0.x, r1); i1 = mfma16(a1.x, b0.y, i1); }
    r0 = mfma16(a0.y, b1.x, r0);
    i0 = mfma16(a0.y, b1.y, i0);
    if (TWO) { r1 = mfma16(a1.y, b1.x, r1); i1 = mfma16(a1.y, b1.y, i1); }
    r0 = mfma16(a0.z, b2.x, r0);
    i0 = mfma16(a0.z, b2.y, i0);
    if (TWO) { r1 = mfma16(a1.z, b2.x, r1); i1 = mfma16(a1.z, b2.y, i1); }
    r0 = mfma16(a0.w, b3.x, r0);
    i0 = mfma16(a0.w, b3.y, i0);
    if (TWO) { r1 = mfma16(a1.w, b3.x, r1); i1 = mfma16(a1.w, b3.y, i1); }
    a0 = n0;
    a1 = n1;
  }
  const C2<T>* xc = xs + (cg * 16 + cl) * pn.ldx;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = mt0 * 16 + AccT<T>::row(g, r);
    T vr = r0[r], vi = i0[r];
    T wr = r1[r], wi = i1[r];
    const T* o0 = op + (long long)row * lda;
    const T* o1 = o0 + 16LL * lda;
    for (int k = pn.KM; k < pn.nz; ++k) {
      const C2<T> xv = xc[k];
      const T m0 = o0[k];
      vr += m0 * xv.x;
      vi += m0 * xv.y;
      if (TWO) {
        const T m1 = o1[k];
        wr += m1 * xv.x;
        wi += m1 * xv.y;
      }
    }
    out[0][r] = mk<T>(vr, vi);
    out[1][r] = mk<T>(wr, wi);
  }
}

template <typename T>
__device__ __forceinline__ C2<T> tail_row(const T* __restrict__ op, int lda, const Panel& pn,
                                          const C2<T>* __restrict__ xs, int row, int c) {
  const T* orow = op + (long long)row * lda;
  const C2<T>* xc = xs + c * pn.ldx;
  T vr = 0, vi = 0;
  for (int k = 0; k < pn.nz; ++k) {
    const T m = orow[k];
    const C2<T> xv = xc[k];
    vr += m * xv.x;
    vi += m * xv.y;
  }
  return mk<T>(vr, vi);
}

// Y(row, c) = sum_k op[row][k] X(k, c) for every row and panel column c;
// emit(row, c, value).  BUFFERED > 0: every value is emitted after a barrier
// that follows the last panel read (emit may overwrite the panel); each wave
// holds at most BUFFERED units and each thread at most BUFFERED tail values.
template <typename T, int BUFFERED, class Emit>
__device__ __forceinline__ void gemm_panel(const T* __restrict__ op, int lda, const Panel& pn, const C2<T>* xs,
                                           Emit emit) {
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int cl = lane & 15;
  const int nunit = pn.npair * pn.ncg;
  const int ntail = (pn.nz - pn.KM) * pn.ncol;
  if constexpr (BUFFERED == 0) {
    for (int u = wave; u < nunit; u += kWaves) {
      const int p = u / pn.ncg, cg = u - p * pn.ncg;
      const int mt0 = 2 * p;
      C2<T> out[2][4];
      const bool two = mt0 + 1 < pn.MT;
      if (two) unit_mfma<T, true>(op, lda, pn, xs, mt0, cg, out);
      else unit_mfma<T, false>(op, lda, pn, xs, mt0, cg, out);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        emit(mt0 * 16 + AccT<T>::row(g, r), cg * 16 + cl, out[0][r]);
        if (two) emit(mt0 * 16 + 16 + AccT<T>::row(g, r), cg * 16 + cl, out[1][r]);
      }
    }
    for (int idx = threadIdx.x; idx < ntail; idx += kNT) {
      const int rr = idx / pn.ncol, c = idx - rr * pn.ncol;
      emit(pn.KM + rr, c, tail_row<T>(op, lda, pn, xs, pn.KM + rr, c));
    }
  } else {
    C2<T> buf[BUFFERED][2][4];
    C2<T> tl[BUFFERED];
#pragma unroll
    for (int j = 0; j < BUFFERED; ++j) {
      const int u = wave + j * kWaves;
      if (u < nunit) {
        const int p = u / pn.ncg, cg = u - p * pn.ncg;
        const int mt0 = 2 * p;
        if (mt0 + 1 < pn.MT) unit_mfma<T, true>(op, lda, pn, xs, mt0, cg, buf[j]);
        else unit_mfma<T, false>(op, lda, pn, xs, mt0, cg, buf[j]);
      }
    }
#pragma unroll
    for (int t = 0; t < BUFFERED; ++t) {
      const int idx = threadIdx.x + t * kNT;
      if (idx < ntail) {
        const int rr = idx / pn.ncol, c = idx - rr * pn.ncol;
        tl[t] = tail_row<T>(op, lda, pn, xs, pn.KM + rr, c);
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < BUFFERED; ++j) {
      const int u = wave + j * kWaves;
      if (u < nunit) {
        const int p = u / pn.ncg, cg = u - p * pn.ncg;
        const int mt0 = 2 * p;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          emit(mt0 * 16 + AccT<T>::row(g, r), cg * 16 + cl, buf[j][0][r]);
          if (mt0 + 1 < pn.MT) emit(mt0 * 16 + 16 + AccT<T>::row(g, r), cg * 16 + cl, buf[j][1][r]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < BUFFERED; ++t) {
      const int idx = threadIdx.x + t * kNT;
      if (idx < ntail) {
        const int rr = idx / pn.ncol, c = idx - rr * pn.ncol;
        emit(pn.KM + rr, c, tl[t]);
      }
    }
  }
}

// units per wave / tail values per thread that the buffered mode must hold
static inline int buffered_depth(const Panel& pn) {
  const int nunit = pn.npair * pn.ncg;
  const int u = (nunit + kWaves - 1) / kWaves;
  const int t = ((pn.nz - pn.KM) * pn.ncol + kNT - 1) / kNT;
  return u > t ? u : t;
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

static constexpr size_t kLdsLimit = 80 * 1024;   // two workgroups per CU

template <typename T, typename TX, typename TY>
__global__ __launch_bounds__(kNT, 4) void k_zgemm(ZArgs a, Panel pn) {
  extern __shared__ char smem_raw[];
  C2<T>* xs = reinterpret_cast<C2<T>*>(smem_raw);
  const int nz = a.nz;
  const int ldx = pn.ldx;
  const int f = blockIdx.y;
  const long long c0 = (long long)blockIdx.x * pn.ncol;
  const int ncol = (int)min((long long)pn.ncol, a.P - c0);
  const TX* x = (const TX*)a.x + f * a.xs + c0 * nz;
  for (int idx = threadIdx.x; idx < pn.ncol * nz; idx += kNT) {
    const int c = idx / nz, k = idx - c * nz;
    xs[c * ldx + k] = (c < ncol) ? ld<T, TX>(x + (long long)c * nz + k) : mk<T>((T)0, (T)0);
  }
  __syncthreads();
  gemm_panel<T, 0>((const T*)a.op, a.lda, pn, xs, [&](int row, int c, C2<T> v) {
    if (c < ncol) zepilogue<T, TX, TY>(a, f, c0 + c, row, v, xs[c * ldx + row]);
  });
}

template <typename T, typename TX, typename TY>
void zgemm(const ZArgs& a, int nfield, hipStream_t st) {
  TBD_CHECK(a.nz >= 4 && a.lda >= a.nz && a.lda % 4 == 0 && a.P > 0, "zgemm > bad shape");
  const Panel pn = make_panel(a.nz, sizeof(C2<T>), kLdsLimit);
  const size_t lds = sizeof(C2<T>) * pn.ncol * pn.ldx;
  TBD_CHECK(lds <= 160 * 1024, "zgemm > nz too large for one LDS panel");
  dim3 grid(ceil_div(a.P, pn.ncol), nfield);
  hipLaunchKernelGGL((k_zgemm<T, TX, TY>), grid, dim3(kNT), lds, st, a, pn);
  TBD_HIP(hipGetLastError());
}

template void zgemm<double, double2, double2>(const ZArgs&, int, hipStream_t);
template void zgemm<float, double2, float2>(const ZArgs&, int, hipStream_t);
template void zgemm<float, float2, float2>(const ZArgs&, int, hipStream_t);

// ---------------------------------------------------------------------------
// fused Helmholtz: one LDS panel, both contractions in buffered mode
// ---------------------------------------------------------------------------
template <int DEPTH>
__global__ __launch_bounds__(kNT, DEPTH == 1 ? 4 : 2) void k_helmholtz(HArgs a, Panel pn) {
  extern __shared__ char smem_raw[];
  const int nz = a.nz;
  const int ldx = pn.ldx;
  double2* xs = reinterpret_cast<double2*>(smem_raw);
  const int fi = blockIdx.y;
  const double* vinv = fi == 0 ? a.f[0].vinv : (fi == 1 ? a.f[1].vinv : a.f[2].vinv);
  const double* vnorm = fi == 0 ? a.f[0].vnorm : (fi == 1 ? a.f[1].vnorm : a.f[2].vnorm);
  const double* lam = fi == 0 ? a.f[0].lambda : (fi == 1 ? a.f[1].lambda : a.f[2].lambda);
  const double* fix = fi == 0 ? a.f[0].fix : (fi == 1 ? a.f[1].fix : a.f[2].fix);
  const double sigma = fi == 0 ? a.f[0].sigma : (fi == 1 ? a.f[1].sigma : a.f[2].sigma);
  const int nullrow = fi == 0 ? a.f[0].nullrow : (fi == 1 ? a.f[1].nullrow : a.f[2].nullrow);
  const double2* x = fi == 0 ? a.f[0].x : (fi == 1 ? a.f[1].x : a.f[2].x);
  double2* y = fi == 0 ? a.f[0].y : (fi == 1 ? a.f[1].y : a.f[2].y);
  const long long c0 = (long long)blockIdx.x * pn.ncol;
  const int ncol = (int)min((long long)pn.ncol, (long long)a.P - c0);
  for (int idx = threadIdx.x; idx < pn.ncol * nz; idx += kNT) {
    const int c = idx / nz, k = idx - c * nz;
    xs[c * ldx + k] = (c < ncol) ? x[(c0 + c) * nz + k] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  if (a.mode == 0) {
    // forward fix: X_i += fbf_i X_0 + ftf_i X_{N-1}, i = 1..N-2
    const double* fbf = fix;
    const double* ftf = fix + nz;
    for (int idx = threadIdx.x; idx < pn.ncol * (nz - 2); idx += kNT) {
      const int c = idx / (nz - 2), iz = 1 + idx - c * (nz - 2);
      const double2 x0 = xs[c * ldx], xn = xs[c * ldx + nz - 1];
      const double2 xi = xs[c * ldx + iz];
      const double2 t = cadd<double>(cscale<double>(fbf[iz], x0), cscale<double>(ftf[iz], xn));
      xs[c * ldx + iz] = cadd<double>(xi, t);
    }
  } else {
    for (int c = threadIdx.x; c < pn.ncol; c += kNT) {
      xs[c * ldx] = make_double2(0.0, 0.0);
      xs[c * ldx + nz - 1] = make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  // Z = R o (V~^-1 X), written over the panel
  gemm_panel<double, DEPTH>(vinv, a.lda, pn, xs, [&](int row, int c, double2 v) {
    double r = 1.0;
    if (row > 0 && row < nz - 1) {
      const long long col = c0 + c;
      const long long cc = col < a.P ? col : 0;
      const double kx = a.kx[cc / a.nyc], ky = a.ky[cc % a.nyc];
      const double k2 = (-(kx * kx)) + (-(ky * ky));
      r = 1.0 / (((k2 + lam[row]) + sigma));
      if (row == nullrow && col == 0) r = 0.0;
    }
    xs[c * ldx + row] = cscale<double>(r, v);
  });
  __syncthreads();
  // Y = V~ Z, written over the panel
  gemm_panel<double, DEPTH>(vnorm, a.lda, pn, xs, [&](int row, int c, double2 v) { xs[c * ldx + row] = v; });
  __syncthreads();
  // sequential backward fixes: row 0 from Y, then row N-1 from the updated Y
  const double* fbb = fix + 2 * nz;
  const double* fbt = fix + 3 * nz;
  const int part = threadIdx.x & 15;      // 16 threads per column
  for (int pass = 0; pass < 2; ++pass) {
    const double* w = pass == 0 ? fbb : fbt;
    for (int cb = 0; cb < pn.ncol; cb += kNT / 16) {
      const int c = cb + (threadIdx.x >> 4);
      double sr = 0.0, si = 0.0;
      if (c < pn.ncol) {
        for (int j = part; j < nz; j += 16) {
          const double2 v = xs[c * ldx + j];
          sr += w[j] * v.x;
          si += w[j] * v.y;
        }
      }
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o, 16);
        si += __shfl_xor(si, o, 16);
      }
      __syncthreads();
      if (part == 0 && c < pn.ncol) xs[c * ldx + (pass == 0 ? 0 : nz - 1)] = make_double2(sr, si);
      __syncthreads();
    }
  }
  for (int idx = threadIdx.x; idx < pn.ncol * nz; idx += kNT) {
    const int cc = idx / nz, k = idx - cc * nz;
    if (cc < ncol) y[(c0 + cc) * nz + k] = xs[cc * ldx + k];
  }
}

void helmholtz(const HArgs& a, int nfield, hipStream_t st) {
  TBD_CHECK(a.nz >= 4 && a.lda >= a.nz && a.lda % 4 == 0 && a.P > 0 && nfield >= 1 && nfield <= 3,
            "helmholtz > bad shape");
  const Panel pn = make_panel(a.nz, sizeof(double2), kLdsLimit);
  const int depth = buffered_depth(pn);
  TBD_CHECK(depth <= kMaxU, "helmholtz > nz too large for the register-buffered contraction");
  const size_t lds = sizeof(double2) * pn.ncol * pn.ldx;
  TBD_CHECK(lds <= 160 * 1024, "helmholtz > nz too large for one LDS panel");
  dim3 grid(ceil_div(a.P, pn.ncol), nfield);
  if (depth <= 1)
    hipLaunchKernelGGL(k_helmholtz<1>, grid, dim3(kNT), lds, st, a, pn);
  else
    hipLaunchKernelGGL(k_helmholtz<2>, grid, dim3(kNT), lds, st, a, pn);
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
