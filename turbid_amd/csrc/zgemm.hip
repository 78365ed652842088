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
#include "group.h"

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
static constexpr int kSplitMaxCol = 64;   // columns of a panel at most (per-column LDS tables)

// A contraction segment: output rows [0, rows) of the operator block stored
// at op + opofs*lda, contracted with panel entries [kofs, kofs + K) of every
// column, results emitted at panel index rbase + row.  A plain operator is
// one segment; an operator applied through its even/odd split is two.
struct Seg {
  int rows, K, kofs, opofs, rbase;
  int MT, RM, KM, npair;   // 16-row tiles, rows covered by them, k covered by full blocks, tile pairs
};
struct Panel {
  int nz;        // column length in HBM
  int nseg;
  Seg sg[2];
  int ncg, ncol, ldx;
  int nunit, ntail;
  int parity;    // 0 plain, +1 centro-symmetric, -1 centro-antisymmetric operator
};

__host__ inline Seg make_seg(int rows, int K, int kofs, int opofs, int rbase) {
  Seg g;
  g.rows = rows; g.K = K; g.kofs = kofs; g.opofs = opofs; g.rbase = rbase;
  g.MT = rows / 16;
  g.RM = g.MT * 16;
  g.KM = (K / 16) * 16;
  g.npair = (g.MT + 1) / 2;
  return g;
}

// parity != 0 needs nz odd: the operator is stored as the (M+1) x (M+1)
// even block followed by the (M+1) x M odd block (M = (nz-1)/2)
__host__ inline Panel make_panel(int nz, size_t elem, size_t lds_limit, int parity = 0, int ncg_mult = 1) {
  Panel pn;
  pn.nz = nz;
  pn.parity = parity;
  if (parity == 0) {
    pn.nseg = 1;
    pn.sg[0] = make_seg(nz, nz, 0, 0, 0);
    pn.sg[1] = pn.sg[0];
    pn.ldx = nz | 1;
  } else {
    const int M = (nz - 1) / 2;
    pn.nseg = 2;
    pn.sg[0] = make_seg(M + 1, M + 1, 0, 0, 0);
    pn.sg[1] = make_seg(M + 1, M, M + 1, M + 1, M + 1);
    pn.ldx = (nz + 1) | 1;    // indices 0 .. 2M+1
  }
  int pairs = 0, tailrows = 0;
  for (int i = 0; i < pn.nseg; ++i) {
    pairs += pn.sg[i].npair;
    tailrows += pn.sg[i].rows - pn.sg[i].RM;
  }
  int ncg = pairs > 0 ? (kWaves + pairs - 1) / pairs : 1;
  if (ncg > 4) ncg = 4;
  ncg *= ncg_mult;
  if (ncg > kSplitMaxCol / 16) ncg = kSplitMaxCol / 16;
  while (ncg > 1 && (size_t)16 * ncg * pn.ldx * elem > lds_limit) --ncg;
  // keep the rows past the last full tile within what one pass of the
  // workgroup (or the register-buffered mode) holds
  while (ncg > 1 && tailrows * 16 * ncg > kNT) --ncg;
  pn.ncg = ncg;
  pn.ncol = 16 * ncg;
  pn.nunit = pairs * ncg;
  pn.ntail = tailrows * pn.ncol;
  return pn;
}

// One unit of segment sg: rows mt0*16 .. mt0*16+31 (second tile if TWO) x the
// 16 panel columns of group cg.  out[t][r] = segment row mt0*16 + 16t +
// AccT::row(g, r), column cg*16 + lane%16.
template <typename T, bool TWO>
__device__ __forceinline__ void unit_mfma(const T* __restrict__ op, int lda, const Seg& sg, int ldx,
                                          const C2<T>* __restrict__ xs, int mt0, int cg, C2<T> (&out)[2][4]) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int cl = lane & 15;
  typedef typename AccT<T>::type Acc;
  Acc r0 = {0, 0, 0, 0}, i0 = {0, 0, 0, 0}, r1 = {0, 0, 0, 0}, i1 = {0, 0, 0, 0};
  const T* ob = op + (long long)sg.opofs * lda;
  const T* A0 = ob + (long long)(mt0 * 16 + cl) * lda + 4 * g;
  const T* A1 = A0 + 16LL * lda;
  const C2<T>* xc = xs + (cg * 16 + cl) * ldx + sg.kofs;
  const C2<T>* B = xc + 4 * g;
  // 16 k of MFMAs on operand set (p, q); B read from the panel
  auto block = [&](const V4<T>& p, const V4<T>& q, int kb) {
    const C2<T> b0 = B[kb], b1 = B[kb + 1], b2 = B[kb + 2], b3 = B[kb + 3];
    r0 = mfma16(p.x, b0.x, r0);
    i0 = mfma16(p.x, b0.y, i0);
    if (TWO) { r1 = mfma16(q.x, b0.x, r1); i1 = mfma16(q.x, b0.y, i1); }
    r0 = mfma16(p.y, b1.x, r0);
    i0 = mfma16(p.y, b1.y, i0);
    if (TWO) { r1 = mfma16(q.y, b1.x, r1); i1 = mfma16(q.y, b1.y, i1); }
    r0 = mfma16(p.z, b2.x, r0);
    i0 = mfma16(p.z, b2.y, i0);
    if (TWO) { r1 = mfma16(q.z, b2.x, r1); i1 = mfma16(q.z, b2.y, i1); }
    r0 = mfma16(p.w, b3.x, r0);
    i0 = mfma16(p.w, b3.y, i0);
    if (TWO) { r1 = mfma16(q.w, b3.x, r1); i1 = mfma16(q.w, b3.y, i1); }
  };
  // operand of the first k step past the full blocks (k = KM + g; zero beyond
  // K), fetched before the main loop so that its latency is hidden
  const bool kt = sg.KM + g < sg.K;
  T at0 = kt ? ob[(long long)(mt0 * 16 + cl) * lda + sg.KM + g] : T(0);
  T at1 = (TWO && kt) ? ob[(long long)(mt0 * 16 + 16 + cl) * lda + sg.KM + g] : T(0);
  // two operand register sets used alternately: the loads of one set are in
  // flight while the other set's MFMAs issue (no register copies)
  // Blocks in pairs with no branch between a prefetch and its use (the load
  // of the next pair's first block clamped to the last block: a redundant
  // in-bounds load at the end); the scheduling barriers keep each prefetch
  // ahead of the block before its use.  An odd last block runs on its own.
  const int kl = sg.KM - 16;
  if (sg.KM > 0) {
    V4<T> pa = ld4(A0), qa = pa, pb = pa, qb = pa;
    if (TWO) qa = ld4(A1);
    int kb = 0;
    for (; kb + 32 <= sg.KM; kb += 32) {
      pb = ld4(A0 + kb + 16);
      if (TWO) qb = ld4(A1 + kb + 16);
      __builtin_amdgcn_sched_barrier(0);
      block(pa, qa, kb);
      const int k2 = min(kb + 32, kl);
      pa = ld4(A0 + k2);
      if (TWO) qa = ld4(A1 + k2);
      __builtin_amdgcn_sched_barrier(0);
      block(pb, qb, kb + 16);
    }
    if (kb < sg.KM) block(pa, qa, kb);
  }
  // k beyond the last full block (one k for the CGL sizes), four k per MFMA
  // step with the operand zero past K
  for (int kk = sg.KM; kk < sg.K; kk += 4) {
    const bool in = kk + g < sg.K;
    if (kk > sg.KM) {
      at0 = in ? ob[(long long)(mt0 * 16 + cl) * lda + kk + g] : T(0);
      if (TWO) at1 = in ? ob[(long long)(mt0 * 16 + 16 + cl) * lda + kk + g] : T(0);
    }
    const C2<T> bt = in ? xc[kk + g] : mk<T>(0, 0);
    r0 = mfma16(at0, bt.x, r0);
    i0 = mfma16(at0, bt.y, i0);
    if (TWO) { r1 = mfma16(at1, bt.x, r1); i1 = mfma16(at1, bt.y, i1); }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    out[0][r] = mk<T>(r0[r], i0[r]);
    out[1][r] = mk<T>(r1[r], i1[r]);
  }
}

// Rows beyond the last full tile of each segment on the vector ALUs: each
// output is a dot product of length K split over a group of G lanes
// (coalesced operand reads) and reduced with cross-lane shuffles.
__host__ __device__ inline int tail_group(int ntail) {
  int g = 64;
  while (g > 1 && g * ntail > kNT) g >>= 1;
  return g;
}
// tail output idx -> (segment, row within the segment, column)
__device__ __forceinline__ void tail_index(const Panel& pn, int idx, int& si, int& row, int& c) {
  const int n0 = (pn.sg[0].rows - pn.sg[0].RM) * pn.ncol;
  si = idx < n0 ? 0 : 1;
  const int li = si == 0 ? idx : idx - n0;
  const int rr = li / pn.ncol;
  c = li - rr * pn.ncol;
  row = pn.sg[si].RM + rr;
}
template <typename T>
__device__ __forceinline__ C2<T> tail_dot(const T* __restrict__ op, int lda, const Panel& pn,
                                          const C2<T>* __restrict__ xs, int idx, int G, int lig, bool active) {
  T vr = 0, vi = 0;
  if (active) {
    int si, row, c;
    tail_index(pn, idx, si, row, c);
    const Seg& sg = pn.sg[si];
    const T* orow = op + (long long)(sg.opofs + row) * lda;
    const C2<T>* xc = xs + c * pn.ldx + sg.kofs;
    int k = lig;
    // four operand loads in flight at a time
    for (; k + 3 * G < sg.K; k += 4 * G) {
      const T m0 = orow[k], m1 = orow[k + G], m2 = orow[k + 2 * G], m3 = orow[k + 3 * G];
      const C2<T> x0 = xc[k], x1 = xc[k + G], x2 = xc[k + 2 * G], x3 = xc[k + 3 * G];
      vr += m0 * x0.x; vi += m0 * x0.y;
      vr += m1 * x1.x; vi += m1 * x1.y;
      vr += m2 * x2.x; vi += m2 * x2.y;
      vr += m3 * x3.x; vi += m3 * x3.y;
    }
    for (; k < sg.K; k += G) {
      const T m = orow[k];
      const C2<T> xv = xc[k];
      vr += m * xv.x;
      vi += m * xv.y;
    }
  }
  for (int o = G >> 1; o > 0; o >>= 1) {
    vr += __shfl_xor(vr, o, G);
    vi += __shfl_xor(vi, o, G);
  }
  return mk<T>(vr, vi);
}

// unit u -> (segment, first tile, column group)
__device__ __forceinline__ void unit_index(const Panel& pn, int u, int& si, int& mt0, int& cg) {
  const int n0 = pn.sg[0].npair * pn.ncg;
  si = u < n0 ? 0 : 1;
  const int lu = si == 0 ? u : u - n0;
  const int p = lu / pn.ncg;
  cg = lu - p * pn.ncg;
  mt0 = 2 * p;
}

// Y(row, c) = sum_k op[row][k] X(k, c) for every segment row and panel column
// c; emit(panel index, c, value).  BUFFERED > 0: every value is emitted after
// a barrier that follows the last panel read (emit may overwrite the panel);
// each wave holds at most BUFFERED units and each thread at most BUFFERED
// tail values.
template <typename T, int BUFFERED, class Emit>
__device__ __forceinline__ void gemm_panel(const T* __restrict__ op, int lda, const Panel& pn, const C2<T>* xs,
                                           Emit emit) {
  // wave-uniform (a scalar register): the unit loops branch on the scalar unit
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int cl = lane & 15;
  const int nunit = pn.nunit;
  const int ntail = pn.ntail;
  const int G = tail_group(ntail);
  const int per = kNT / G;                       // tail outputs per pass
  const int lig = threadIdx.x & (G - 1);
  auto emit_unit = [&](int si, int mt0, int cg, const C2<T> (&out)[2][4]) {
    const Seg& sg = pn.sg[si];
    const bool two = mt0 + 1 < sg.MT;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      emit(sg.rbase + mt0 * 16 + AccT<T>::row(g, r), cg * 16 + cl, out[0][r]);
      if (two) emit(sg.rbase + mt0 * 16 + 16 + AccT<T>::row(g, r), cg * 16 + cl, out[1][r]);
    }
  };
  auto run_unit = [&](int si, int mt0, int cg, C2<T> (&out)[2][4]) {
    const Seg& sg = pn.sg[si];
    if (mt0 + 1 < sg.MT) unit_mfma<T, true>(op, lda, sg, pn.ldx, xs, mt0, cg, out);
    else unit_mfma<T, false>(op, lda, sg, pn.ldx, xs, mt0, cg, out);
  };
  auto emit_tail = [&](int idx, C2<T> v) {
    int si, row, c;
    tail_index(pn, idx, si, row, c);
    emit(pn.sg[si].rbase + row, c, v);
  };
  if constexpr (BUFFERED == 0) {
    for (int u = wave; u < nunit; u += kWaves) {
      int si, mt0, cg;
      unit_index(pn, u, si, mt0, cg);
      C2<T> out[2][4];
      run_unit(si, mt0, cg, out);
      emit_unit(si, mt0, cg, out);
    }
    for (int base = 0; base < ntail; base += per) {
      const int idx = base + threadIdx.x / G;
      const C2<T> v = tail_dot<T>(op, lda, pn, xs, idx, G, lig, idx < ntail);
      if (lig == 0 && idx < ntail) emit_tail(idx, v);
    }
  } else {
    C2<T> buf[BUFFERED][2][4];
    C2<T> tl[BUFFERED];
    // the tail rows wait on operand loads: waves 4..7 compute them before
    // their units, waves 0..3 after (wave w runs on SIMD w % 4), so that the
    // MFMAs of the other wave of the workgroup on the SIMD hide the wait
    auto tails = [&] {
#pragma unroll
      for (int t = 0; t < BUFFERED; ++t) {
        const int idx = t * per + threadIdx.x / G;
        if (t * per < ntail) tl[t] = tail_dot<T>(op, lda, pn, xs, idx, G, lig, idx < ntail);
      }
    };
    if (wave & 4) tails();
#pragma unroll
    for (int j = 0; j < BUFFERED; ++j) {
      const int u = wave + j * kWaves;
      if (u < nunit) {
        int si, mt0, cg;
        unit_index(pn, u, si, mt0, cg);
        run_unit(si, mt0, cg, buf[j]);
      }
    }
    if (!(wave & 4)) tails();
    __syncthreads();
#pragma unroll
    for (int j = 0; j < BUFFERED; ++j) {
      const int u = wave + j * kWaves;
      if (u < nunit) {
        int si, mt0, cg;
        unit_index(pn, u, si, mt0, cg);
        emit_unit(si, mt0, cg, buf[j]);
      }
    }
#pragma unroll
    for (int t = 0; t < BUFFERED; ++t) {
      const int idx = t * per + threadIdx.x / G;
      if (lig == 0 && idx < ntail) emit_tail(idx, tl[t]);
    }
  }
}

// units per wave / tail values per thread that the buffered mode must hold
static inline int buffered_depth(const Panel& pn) {
  const int u = (pn.nunit + kWaves - 1) / kWaves;
  const int per = kNT / tail_group(pn.ntail);
  const int t = (pn.ntail + per - 1) / per;
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

// Runs load(idx) then store(idx, value) for idx = threadIdx.x + j*kNT < n,
// NB indices at a time: all NB loads of a batch are issued before its first
// store, so a thread keeps NB global reads in flight instead of one.
template <int NB, typename V, class Load, class Store>
__device__ __forceinline__ void batched(int n, Load load, Store store) {
  for (int base = threadIdx.x; base < n; base += NB * kNT) {
    V v[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int idx = base + j * kNT;
      if (idx < n) v[j] = load(idx);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int idx = base + j * kNT;
      if (idx < n) store(idx, v[j]);
    }
  }
}

// Operand bases of field f (evenly strided, or per field when xf[0] is set).
__device__ __forceinline__ const void* pick3(const void* const* p, int f) {
  return f == 0 ? p[0] : (f == 1 ? p[1] : p[2]);
}
template <typename TX>
__device__ __forceinline__ const TX* zx(const ZArgs& a, int f) {
  return a.xf[0] ? (const TX*)pick3(a.xf, f) : (const TX*)a.x + f * a.xs;
}
template <typename TX>
__device__ __forceinline__ const TX* ze0(const ZArgs& a, int f) {
  return a.xf[0] ? (const TX*)pick3(a.e0f, f) : (const TX*)a.e0 + f * a.es;
}
template <typename TX>
__device__ __forceinline__ const TX* ze1(const ZArgs& a, int f) {
  return a.xf[0] ? (const TX*)pick3(a.e1f, f) : (const TX*)a.e1 + f * a.es;
}

// The global inputs an epilogue mode reads besides the contraction result.
template <typename T> struct EpiIn { C2<T> x, e0, e1, e2; };

template <typename T, typename TX>
__device__ __forceinline__ EpiIn<T> epi_load(const ZArgs& a, int f, long long off) {
  EpiIn<T> in;
  in.x = in.e0 = in.e1 = in.e2 = mk<T>((T)0, (T)0);
  const TX* x = zx<TX>(a, f);
  switch (a.mode) {
    case Z_GRAD:
    case Z_GRADYZ:
      in.x = ld<T, TX>(x + off);
      break;
    case Z_RHS:
      in.x = ld<T, TX>(x + off);
      in.e0 = ld<T, TX>((const TX*)a.h + f * a.hs + off);
      break;
    case Z_DIV:
      in.e0 = ld<T, TX>((const TX*)a.e0 + off);
      in.e1 = ld<T, TX>((const TX*)a.e1 + off);
      break;
    case Z_CORRECT:
      in.x = ld<T, TX>(x + off);
      in.e0 = ld<T, TX>((const TX*)a.e0 + off);
      in.e1 = ld<T, TX>((const TX*)a.e0 + a.es + off);
      in.e2 = ld<T, TX>((const TX*)a.e0 + 2 * a.es + off);
      break;
    case Z_GRADTMP:
      in.e0 = ld<T, TX>(ze0<TX>(a, f) + off);
      in.e1 = ld<T, TX>(ze1<TX>(a, f) + off);
      break;
    default:
      break;
  }
  return in;
}

template <typename T, typename TX, typename TY>
__device__ __forceinline__ void zepilogue(const ZArgs& a, int f, long long col, int row, C2<T> v,
                                          const EpiIn<T>& in) {
  const int nz = a.nz;
  const long long off = col * nz + row;
  TY* y = (TY*)a.y;
  const int ci = (int)col, ix = ci / a.nyc;
  const T kx = (T)a.kx[ix];
  const T ky = (T)a.ky[ci - ix * a.nyc];
  const C2<T> xv = in.x;
  switch (a.mode) {
    case Z_PLAIN:
      st<TY, T>(y + f * a.ys + off, v);
      break;
    case Z_GRAD:
      st<TY, T>(y + (3 * f + 0) * a.ys + off, cmuli<T>(kx, xv));
      st<TY, T>(y + (3 * f + 1) * a.ys + off, cmuli<T>(ky, xv));
      st<TY, T>(y + (3 * f + 2) * a.ys + off, v);
      break;
    case Z_GRADYZ:
      st<TY, T>(y + f * a.ys + off, cmuli<T>(ky, xv));
      st<TY, T>(y + (3 + f) * a.ys + off, v);
      break;
    case Z_RHS: {
      const C2<T> h = in.e0;
      const T k2x = -(kx * kx), k2y = -(ky * ky);
      const C2<T> lap = cadd<T>(cadd<T>(v, cscale<T>(k2x, xv)), cscale<T>(k2y, xv));
      const C2<T> r = csub<T>(cadd<T>(cscale<T>((T)a.s2, h), cscale<T>((T)a.s1, xv)), lap);
      st<TY, T>(y + f * a.ys + off, r);
    } break;
    case Z_DIV: {
      const C2<T> su = in.e0;
      const C2<T> sv = in.e1;
      const C2<T> dv = cadd<T>(cadd<T>(v, cmuli<T>(kx, su)), cmuli<T>(ky, sv));
      st<TY, T>(y + off, cscale<T>((T)a.q, dv));
    } break;
    case Z_CORRECT: {
      const C2<T> gx = cmuli<T>(kx, xv), gy = cmuli<T>(ky, xv);
      const T cf = (T)a.cf;
      const C2<T> su = in.e0;
      const C2<T> sv = in.e1;
      const C2<T> sw = in.e2;
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
      const C2<T> f0 = in.e0;
      const C2<T> f1 = in.e1;
      st<TY, T>(y + f * a.ys + off, cadd<T>(cadd<T>(cmuli<T>(kx, f0), cmuli<T>(ky, f1)), v));
    } break;
    default:
      break;
  }
}

static constexpr size_t kLdsLimit = 80 * 1024;   // two workgroups per CU

// (kx, ky) of spectral column col (columns past P, the padding of the last
// panel, read column 0); 32-bit index arithmetic (P < 2^31)
__device__ __forceinline__ double2 kcol(const double* kx, const double* ky, long long col, long long P, int nyc) {
  const int ci = col < P ? (int)col : 0;
  const int ix = ci / nyc;
  return make_double2(kx[ix], ky[ci - ix * nyc]);
}

// Panel load, buffered contraction written over the panel, then the epilogue
// as one coalesced pass (row fastest) that re-reads the input column where the
// mode needs it.
struct ZL {
  ZArgs a;
  Panel pn;
};
template <typename T, typename TX, typename TY, int DEPTH, bool G>
__global__ __launch_bounds__(kNT, sizeof(T) == 4 ? (DEPTH == 1 ? 8 : 4) : (DEPTH == 1 ? 4 : 2)) void k_zgemm(ZL l0, GroupP grp) {
  Blk bk;
  const ZL& L = member<G>(l0, grp, bk);
  const ZArgs& a = L.a;
  const Panel& pn = L.pn;
  extern __shared__ char smem_raw[];
  C2<T>* xs = reinterpret_cast<C2<T>*>(smem_raw);
  const int nz = a.nz;
  const int ldx = pn.ldx;
  const int f = bk.y;
  const long long c0 = (long long)bk.x * pn.ncol;
  const int ncol = (int)min((long long)pn.ncol, a.P - c0);
  const TX* x = zx<TX>(a, f) + c0 * nz;
  const C2<T> zero = mk<T>((T)0, (T)0);
  // Z_GRAD (Z_GRADYZ) on the even/odd split: the pointwise outputs (i kx x,
  // i ky x; i ky x) are written from the load pass, so the epilogue neither
  // re-reads x nor writes them
  const bool yz = a.mode == Z_GRADYZ;
  const bool grad_early = (a.mode == Z_GRAD || yz) && pn.parity != 0;
  __shared__ double2 kv[kSplitMaxCol];   // (kx, ky) of each panel column
  if (grad_early) {
    for (int c = threadIdx.x; c < pn.ncol; c += kNT) kv[c] = kcol(a.kx, a.ky, c0 + c, a.P, a.nyc);
    __syncthreads();
  }
  if (pn.parity == 0) {
    batched<8, C2<T>>(pn.ncol * nz, [&](int idx) {
      const int c = idx / nz;
      return (c < ncol) ? ld<T, TX>(x + idx) : zero;
    }, [&](int idx, C2<T> v) {
      const int c = idx / nz, k = idx - c * nz;
      xs[c * ldx + k] = v;
    });
  } else {
    // even part e_j = (x_j + x_{N-1-j})/2 at j <= M (e_M = x_M), odd part
    // o_j = (x_j - x_{N-1-j})/2 at M+1+j, j < M
    const int M = (nz - 1) / 2;
    struct LoHi { C2<T> lo, hi; };
    batched<5, LoHi>(pn.ncol * (M + 1), [&](int idx) {
      const int c = idx / (M + 1), j = idx - c * (M + 1);
      LoHi r = {zero, zero};
      if (c < ncol) {
        r.lo = ld<T, TX>(x + (long long)c * nz + j);
        r.hi = ld<T, TX>(x + (long long)c * nz + nz - 1 - j);
      }
      return r;
    }, [&](int idx, LoHi r) {
      const int c = idx / (M + 1), j = idx - c * (M + 1);
      if (grad_early && c < ncol) {
        TY* y = (TY*)a.y;
        const T kx = (T)kv[c].x, ky = (T)kv[c].y;
        const long long off = (c0 + c) * nz + j;
        TY* yx = y + (3 * f + 0) * a.ys;
        TY* yy = y + (yz ? f : 3 * f + 1) * a.ys;
        if (!yz) st<TY, T>(yx + off, cmuli<T>(kx, r.lo));
        st<TY, T>(yy + off, cmuli<T>(ky, r.lo));
        if (j < M) {
          const long long offh = (c0 + c) * nz + nz - 1 - j;
          if (!yz) st<TY, T>(yx + offh, cmuli<T>(kx, r.hi));
          st<TY, T>(yy + offh, cmuli<T>(ky, r.hi));
        }
      }
      if (j < M) {
        xs[c * ldx + j] = mk<T>((T)0.5 * (r.lo.x + r.hi.x), (T)0.5 * (r.lo.y + r.hi.y));
        xs[c * ldx + M + 1 + j] = mk<T>((T)0.5 * (r.lo.x - r.hi.x), (T)0.5 * (r.lo.y - r.hi.y));
      } else {
        xs[c * ldx + j] = r.lo;
      }
    });
  }
  __syncthreads();
  gemm_panel<T, DEPTH>((const T*)a.op, a.lda, pn, xs, [&](int row, int c, C2<T> v) { xs[c * ldx + row] = v; });
  __syncthreads();
  const int M = (nz - 1) / 2;
  if (grad_early) {
    TY* y = (TY*)a.y + (yz ? 3 + f : 3 * f + 2) * a.ys + c0 * nz;
    for (int idx = threadIdx.x; idx < ncol * nz; idx += kNT) {
      const int c = idx / nz, row = idx - c * nz;
      // y_i = s_i + d_i, y_{N-1-i} = -(s_i - d_i) (D1 centro-antisymmetric), i <= M
      const bool top = row > M;
      const int i = top ? nz - 1 - row : row;
      const C2<T> sv = xs[c * ldx + i], dv = xs[c * ldx + M + 1 + i];
      st<TY, T>(y + idx, top ? (pn.parity > 0 ? csub<T>(sv, dv) : csub<T>(dv, sv)) : cadd<T>(sv, dv));
    }
    return;
  }
  batched<sizeof(T) == 4 ? 2 : 4, EpiIn<T>>(ncol * nz, [&](int idx) {
    return epi_load<T, TX>(a, f, c0 * nz + idx);
  }, [&](int idx, const EpiIn<T>& in) {
    const int c = idx / nz, row = idx - c * nz;
    C2<T> v;
    if (pn.parity == 0) {
      v = xs[c * ldx + row];
    } else {
      // y_i = s_i + d_i, y_{N-1-i} = parity (s_i - d_i), i <= M
      const bool top = row > M;
      const int i = top ? nz - 1 - row : row;
      const C2<T> sv = xs[c * ldx + i], dv = xs[c * ldx + M + 1 + i];
      v = top ? (pn.parity > 0 ? csub<T>(sv, dv) : csub<T>(dv, sv)) : cadd<T>(sv, dv);
    }
    zepilogue<T, TX, TY>(a, f, c0 + c, row, v, in);
  });
}

template <typename T, typename TX, typename TY>
void zgemm(const ZArgs& a, int nfield, hipStream_t st) {
  TBD_CHECK(a.nz >= 4 && a.lda % 4 == 0 && a.P > 0, "zgemm > bad shape");
  TBD_CHECK(a.mode != Z_GRADYZ || nfield <= 3, "zgemm > Z_GRADYZ takes at most 3 fields");
  TBD_CHECK(a.parity == 0 ? a.lda >= a.nz : (a.nz % 2 == 1 && a.lda >= (a.nz + 1) / 2),
            "zgemm > operator layout does not match the column length");
  const Panel pn = make_panel(a.nz, sizeof(C2<T>), kLdsLimit, a.parity, a.ncg_mult);
  const size_t lds = sizeof(C2<T>) * pn.ncol * pn.ldx;
  TBD_CHECK(pn.ncol <= kSplitMaxCol, "zgemm > too many columns per panel");
  TBD_CHECK(lds + sizeof(double2) * kSplitMaxCol <= 160 * 1024, "zgemm > nz too large for one LDS panel");
  const int depth = buffered_depth(pn);
  TBD_CHECK(depth <= kMaxU, "zgemm > nz too large for the register-buffered contraction");
  dim3 grid(ceil_div(a.P, pn.ncol), nfield);
  if (depth <= 1)
    glaunch(k_zgemm<T, TX, TY, 1, false>, k_zgemm<T, TX, TY, 1, true>, grid, dim3(kNT), lds, st, ZL{a, pn});
  else
    glaunch(k_zgemm<T, TX, TY, 2, false>, k_zgemm<T, TX, TY, 2, true>, grid, dim3(kNT), lds, st, ZL{a, pn});
}

template void zgemm<double, double2, double2>(const ZArgs&, int, hipStream_t);
template void zgemm<float, double2, float2>(const ZArgs&, int, hipStream_t);
template void zgemm<float, float2, float2>(const ZArgs&, int, hipStream_t);

// ---------------------------------------------------------------------------
// fused Helmholtz: one LDS panel, both contractions in buffered mode
// ---------------------------------------------------------------------------
struct HL {
  HArgs a;
  Panel pn;
};
template <int DEPTH, bool G>
__global__ __launch_bounds__(kNT, DEPTH == 1 ? 4 : 2) void k_helmholtz(HL l0, GroupP grp) {
  Blk bk;
  const HL& L = member<G>(l0, grp, bk);
  const HArgs& a = L.a;
  const Panel& pn = L.pn;
  extern __shared__ char smem_raw[];
  const int nz = a.nz;
  const int ldx = pn.ldx;
  double2* xs = reinterpret_cast<double2*>(smem_raw);
  const int fi = bk.y;
  const double* vinv = fi == 0 ? a.f[0].vinv : (fi == 1 ? a.f[1].vinv : a.f[2].vinv);
  const double* vnorm = fi == 0 ? a.f[0].vnorm : (fi == 1 ? a.f[1].vnorm : a.f[2].vnorm);
  const double* lam = fi == 0 ? a.f[0].lambda : (fi == 1 ? a.f[1].lambda : a.f[2].lambda);
  const double* fix = fi == 0 ? a.f[0].fix : (fi == 1 ? a.f[1].fix : a.f[2].fix);
  const double sigma = fi == 0 ? a.f[0].sigma : (fi == 1 ? a.f[1].sigma : a.f[2].sigma);
  const int nullrow = fi == 0 ? a.f[0].nullrow : (fi == 1 ? a.f[1].nullrow : a.f[2].nullrow);
  const double2* x = fi == 0 ? a.f[0].x : (fi == 1 ? a.f[1].x : a.f[2].x);
  double2* y = fi == 0 ? a.f[0].y : (fi == 1 ? a.f[1].y : a.f[2].y);
  const long long c0 = (long long)bk.x * pn.ncol;
  const int ncol = (int)min((long long)pn.ncol, (long long)a.P - c0);
  __shared__ double2 kv[kSplitMaxCol];   // (kx, ky) of each panel column
  for (int c = threadIdx.x; c < pn.ncol; c += kNT) kv[c] = kcol(a.kx, a.ky, c0 + c, a.P, a.nyc);
  batched<8, double2>(pn.ncol * nz, [&](int idx) {
    return (idx / nz < ncol) ? x[c0 * nz + idx] : make_double2(0.0, 0.0);
  }, [&](int idx, double2 v) {
    const int c = idx / nz, k = idx - c * nz;
    xs[c * ldx + k] = v;
  });
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
      const double kx = kv[c].x, ky = kv[c].y;
      const double k2 = (-(kx * kx)) + (-(ky * ky));
      r = 1.0 / (((k2 + lam[row]) + sigma));
      if (row == nullrow && col == 0) r = 0.0;
    }
    xs[c * ldx + row] = cscale<double>(r, v);
  });
  __syncthreads();
  // Y = V~ Z, written over the panel; rows 0 and nz-1 of this V~ carry the
  // backward fixes (the bottom, then the top after it), so Y is final
  gemm_panel<double, DEPTH>(vnorm, a.lda, pn, xs, [&](int row, int c, double2 v) { xs[c * ldx + row] = v; });
  __syncthreads();
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
  TBD_CHECK(pn.ncol <= kSplitMaxCol, "helmholtz > too many columns per panel");
  TBD_CHECK(lds + sizeof(double2) * kSplitMaxCol <= 160 * 1024, "helmholtz > nz too large for one LDS panel");
  dim3 grid(ceil_div(a.P, pn.ncol), nfield);
  if (depth <= 1)
    glaunch(k_helmholtz<1, false>, k_helmholtz<1, true>, grid, dim3(kNT), lds, st, HL{a, pn});
  else
    glaunch(k_helmholtz<2, false>, k_helmholtz<2, true>, grid, dim3(kNT), lds, st, HL{a, pn});
}

// ---------------------------------------------------------------------------
// Helmholtz through the even/odd form: the column is loaded with the forward
// fix applied and split into s (indices 0..M) and d (M+1..2M); the
// block-diagonal pinv gives the eigen-coordinates, scaled in registers; the
// split V~ gives a (0..M) and b (M+1..2M+1); the backward fixes are dot
// products over (a, b) and the column is written as y_i = a_i + b_i,
// y_{nz-1-i} = a_i - b_i.
// ---------------------------------------------------------------------------

__host__ inline void make_split_panels(int nz, size_t lds_limit, Panel& p1, Panel& p2) {
  const int M = (nz - 1) / 2;
  p2 = make_panel(nz, sizeof(double2), lds_limit, 1);
  p1 = p2;
  p1.sg[1] = make_seg(M, M, M + 1, M + 1, M + 1);
  const int pairs = p1.sg[0].npair + p1.sg[1].npair;
  const int tailrows = (p1.sg[0].rows - p1.sg[0].RM) + (p1.sg[1].rows - p1.sg[1].RM);
  p1.nunit = pairs * p1.ncg;
  p1.ntail = tailrows * p1.ncol;
}

// Backward fixes of every panel column of the even/odd form, 16 to 64 lanes
// per column: t0 = wb . Y and tN = wt . Y + fbt0 (t0 - Y_0) (the top fix sees
// row 0 already replaced); Y is the split (a, b) of the panel column, so
// Y_0 = a_0 + b_0 and wb, wt are the weights folded onto (a, b).  (The full
// form has these rows folded into its V~ instead.)
__device__ __forceinline__ void backward_fixes(const double2* xs, int ldx, int ncolp, int nz, int M,
                                               const double* wb, const double* wt, double fbt0,
                                               double2 (*tfix)[kSplitMaxCol]) {
  // L lanes per column (every thread busy for ncolp <= kNT / 16), four
  // weight pairs in flight per lane
  int L = 64;
  while (L > 16 && L * ncolp > kNT) L >>= 1;
  const int part = threadIdx.x & (L - 1);
  for (int cb = 0; cb < ncolp; cb += kNT / L) {
    const int c = cb + threadIdx.x / L;
    double br = 0.0, bi = 0.0, tr = 0.0, ti = 0.0;
    if (c < ncolp) {
      const double2* xc = xs + c * ldx;
      int j = part;
      for (; j + 3 * L < nz; j += 4 * L) {
        double b[4], t[4];
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          b[u] = wb[j + u * L];
          t[u] = wt[j + u * L];
          v[u] = xc[j + u * L];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          br += b[u] * v[u].x;
          bi += b[u] * v[u].y;
          tr += t[u] * v[u].x;
          ti += t[u] * v[u].y;
        }
      }
      for (; j < nz; j += L) {
        const double2 v = xc[j];
        br += wb[j] * v.x;
        bi += wb[j] * v.y;
        tr += wt[j] * v.x;
        ti += wt[j] * v.y;
      }
    }
    for (int o = L >> 1; o > 0; o >>= 1) {
      br += __shfl_xor(br, o, L);
      bi += __shfl_xor(bi, o, L);
      tr += __shfl_xor(tr, o, L);
      ti += __shfl_xor(ti, o, L);
    }
    if (part == 0 && c < ncolp) {
      const double2 y0 = cadd<double>(xs[c * ldx], xs[c * ldx + M + 1]);
      tfix[0][c] = make_double2(br, bi);
      tfix[1][c] = make_double2(tr + fbt0 * (br - y0.x), ti + fbt0 * (bi - y0.y));
    }
  }
}

// Writes the panel's ncol columns to y (column stride nz): the full form from
// the panel in natural order (its V~ carries the backward fixes in rows 0 and
// nz-1); the even/odd form with rows 0 and nz-1 from tfix and y_i = a_i + b_i,
// y_{nz-1-i} = a_i - b_i from the panel.
template <bool SPLIT>
__device__ __forceinline__ void store_columns(const double2* xs, int ldx, int ncol, int nz, int M,
                                              const double2 (*tfix)[kSplitMaxCol], double2* y) {
  for (int idx = threadIdx.x; idx < ncol * nz; idx += kNT) {
    const int c = idx / nz, k = idx - c * nz;
    double2 v;
    if (!SPLIT) {
      v = xs[c * ldx + k];
    } else if (k == 0) {
      v = tfix[0][c];
    } else if (k == nz - 1) {
      v = tfix[1][c];
    } else if (k <= M) {
      v = cadd<double>(xs[c * ldx + k], xs[c * ldx + M + 1 + k]);
    } else {
      const int i = nz - 1 - k;
      v = csub<double>(xs[c * ldx + i], xs[c * ldx + M + 1 + i]);
    }
    y[idx] = v;
  }
}

struct HSL {
  HArgs a;
  Panel p1, p2;
};
template <int DEPTH, bool G>
__global__ __launch_bounds__(kNT, DEPTH == 1 ? 4 : 2) void k_helmholtz_split(HSL l0, GroupP grp) {
  Blk bk;
  const HSL& L = member<G>(l0, grp, bk);
  const HArgs& a = L.a;
  const Panel& p1 = L.p1;
  const Panel& p2 = L.p2;
  extern __shared__ char smem_raw[];
  __shared__ double2 tfix[2][kSplitMaxCol];
  const int nz = a.nz;
  const int M = (nz - 1) / 2;
  const int ldx = p1.ldx;
  double2* xs = reinterpret_cast<double2*>(smem_raw);
  const double* fix = a.f[0].fix;
  const double* lam = a.f[0].lambda;
  const double sigma = a.f[0].sigma;
  const int nullrow = a.f[0].nullrow;
  const long long c0 = (long long)bk.x * p1.ncol;
  const int ncol = (int)min((long long)p1.ncol, (long long)a.P - c0);
  __shared__ double2 kv[kSplitMaxCol];   // (kx, ky) of each panel column
  for (int c = threadIdx.x; c < p1.ncol; c += kNT) kv[c] = kcol(a.kx, a.ky, c0 + c, a.P, a.nyc);
  const double2* x = a.f[0].x + c0 * nz;
  struct Quad { double2 lo, hi, x0, xn; };
  const double2 zero = make_double2(0.0, 0.0);
  batched<5, Quad>(p1.ncol * (M + 1), [&](int idx) {
    const int c = idx / (M + 1), j = idx - c * (M + 1);
    Quad q = {zero, zero, zero, zero};
    if (c < ncol) {
      const double2* xc = x + (long long)c * nz;
      q.lo = xc[j];
      q.hi = xc[nz - 1 - j];
      if (a.mode == 0 && j > 0) {
        q.x0 = xc[0];
        q.xn = xc[nz - 1];
      }
    }
    return q;
  }, [&](int idx, Quad q) {
    const int c = idx / (M + 1), j = idx - c * (M + 1);
    double2 lo = q.lo, hi = q.hi;
    if (a.mode == 0) {
      if (j > 0) {   // X_i += fbf_i X_0 + ftf_i X_{N-1} on interior rows
        lo = cadd<double>(lo, cadd<double>(cscale<double>(fix[j], q.x0), cscale<double>(fix[nz + j], q.xn)));
        hi = cadd<double>(hi, cadd<double>(cscale<double>(fix[nz - 1 - j], q.x0),
                                           cscale<double>(fix[2 * nz - 1 - j], q.xn)));
      }
    } else if (j == 0) {
      lo = zero;
      hi = zero;
    }
    if (j < M) {
      xs[c * ldx + j] = cadd<double>(lo, hi);
      xs[c * ldx + M + 1 + j] = csub<double>(lo, hi);
    } else {
      xs[c * ldx + j] = lo;
    }
  });
  __syncthreads();
  gemm_panel<double, DEPTH>(a.f[0].vinv, a.lda, p1, xs, [&](int row, int c, double2 v) {
    double r = 1.0;
    if (row != 0 && row != M + 1 && row < nz) {   // row nz: zero row, no eigenvalue
      const long long col = c0 + c;
      const double kx = kv[c].x, ky = kv[c].y;
      const double k2 = (-(kx * kx)) + (-(ky * ky));
      r = 1.0 / (((k2 + lam[row]) + sigma));
      if (row == nullrow && col == 0) r = 0.0;
    }
    xs[c * ldx + row] = cscale<double>(r, v);
  });
  __syncthreads();
  gemm_panel<double, DEPTH>(a.f[0].vnorm, a.lda, p2, xs, [&](int row, int c, double2 v) { xs[c * ldx + row] = v; });
  __syncthreads();
  backward_fixes(xs, ldx, p1.ncol, nz, M, a.f[0].pfix, a.f[0].pfix + nz, fix[3 * nz], tfix);
  __syncthreads();
  store_columns<true>(xs, ldx, ncol, nz, M, tfix, a.f[0].y + c0 * nz);
}

void helmholtz_split(const HArgs& a, hipStream_t st) {
  TBD_CHECK(a.nz >= 7 && a.nz % 2 == 1 && a.P > 0 && a.lda >= (a.nz + 1) / 2 && a.lda % 4 == 0 && a.f[0].pfix,
            "helmholtz_split > bad shape");
  Panel p1, p2;
  make_split_panels(a.nz, kLdsLimit, p1, p2);
  TBD_CHECK(p1.ncol <= kSplitMaxCol, "helmholtz_split > too many columns per panel");
  const int depth = std::max(buffered_depth(p1), buffered_depth(p2));
  TBD_CHECK(depth <= kMaxU, "helmholtz_split > nz too large for the register-buffered contraction");
  const size_t lds = sizeof(double2) * p1.ncol * p1.ldx;
  TBD_CHECK(lds + sizeof(double2) * 3 * kSplitMaxCol <= 160 * 1024, "helmholtz_split > nz too large for one LDS panel");
  dim3 grid(ceil_div(a.P, p1.ncol), 1);
  if (depth <= 1)
    glaunch(k_helmholtz_split<1, false>, k_helmholtz_split<1, true>, grid, dim3(kNT), lds, st, HSL{a, p1, p2});
  else
    glaunch(k_helmholtz_split<2, false>, k_helmholtz_split<2, true>, grid, dim3(kNT), lds, st, HSL{a, p1, p2});
  TBD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// rhs assembly + boundary rows + Helmholtz in one kernel
// ---------------------------------------------------------------------------
__host__ inline Panel with_ncg(Panel pn, int ncg, int ldx) {
  pn.ncg = ncg;
  pn.ncol = 16 * ncg;
  pn.ldx = ldx;
  int pairs = 0, tailrows = 0;
  for (int i = 0; i < pn.nseg; ++i) {
    pairs += pn.sg[i].npair;
    tailrows += pn.sg[i].rows - pn.sg[i].RM;
  }
  pn.nunit = pairs * ncg;
  pn.ntail = tailrows * pn.ncol;
  return pn;
}

struct RHL {
  RHArgs a;
  Panel pd, p1, p2;
};
template <int DEPTH, bool SPLIT, bool G>
__global__ __launch_bounds__(kNT, DEPTH == 1 ? 4 : 2) void k_rhs_helmholtz(RHL l0, GroupP grp) {
  Blk bk;
  const RHL& L = member<G>(l0, grp, bk);
  const RHArgs& a = L.a;
  const Panel& pd = L.pd;
  const Panel& p1 = L.p1;
  const Panel& p2 = L.p2;
  extern __shared__ char smem_raw[];
  __shared__ double2 bcv[2][kSplitMaxCol];
  __shared__ double2 tfix[2][kSplitMaxCol];
  const int nz = a.nz;
  const int M = (nz - 1) / 2;
  const int ldx = pd.ldx;
  double2* xs = reinterpret_cast<double2*>(smem_raw);
  const int fi = bk.y;
  const RHField& F = fi == 0 ? a.f[0] : (fi == 1 ? a.f[1] : a.f[2]);
  const double2* x = F.x;
  double2* h = F.h;
  // eigenvalues and the two forward-fix rows staged in LDS after the panel
  // (read per element in the epilogues)
  double* lam = reinterpret_cast<double*>(xs + pd.ncol * ldx);
  double* fixs = lam + nz + 1;
  const double* fix = F.fix;
  // (nz eigenvalues; the even/odd form's last panel index, nz, is a zero row of
  // its V~^-1 that the V~ contraction does not read)
  for (int i = threadIdx.x; i < nz + 1; i += kNT) lam[i] = i < nz ? F.lambda[i] : 0.0;
  for (int i = threadIdx.x; i < 2 * nz; i += kNT) fixs[i] = fix[i];
  const double sigma = F.sigma;
  const long long c0 = (long long)bk.x * pd.ncol;
  const int ncol = (int)min((long long)pd.ncol, (long long)a.P - c0);
  __shared__ double2 kv[kSplitMaxCol];   // (kx, ky) of each panel column
  __shared__ double qv[kSplitMaxCol];    // s1 + kx^2 + ky^2 of each panel column
  for (int c = threadIdx.x; c < pd.ncol; c += kNT) {
    kv[c] = kcol(a.kx, a.ky, c0 + c, a.P, a.nyc);
    qv[c] = a.s1 + kv[c].x * kv[c].x + kv[c].y * kv[c].y;
  }
  const double2 zero = make_double2(0.0, 0.0);
  // boundary rows of every column of the panel
  if ((int)threadIdx.x < pd.ncol) {
    const int c = threadIdx.x;
    double2 bot = zero, top = zero;
    if (c < ncol) {
      const long long col = c0 + c;
      if (a.bc == 0) {
        const long long j = (long long)F.comp * a.P + col;
        const double sk = a.s, k = a.k;
        const double2 m1t = a.m1t[j], m2t = a.m2t[j], m1b = a.m1b[j], m2b = a.m2b[j];
        top = make_double2(k * ((1 + sk) * m1t.x + (0 - sk) * m2t.x), k * ((1 + sk) * m1t.y + (0 - sk) * m2t.y));
        bot = make_double2(k * ((1 + sk) * m1b.x + (0 - sk) * m2b.x), k * ((1 + sk) * m1b.y + (0 - sk) * m2b.y));
        if (a.open_top && F.comp < 2) {
          const double2 n1 = a.n1t[j], n2 = a.n2t[j];
          top = make_double2(k * ((1 + sk) * n1.x + (0 - sk) * n2.x), k * ((1 + sk) * n1.y + (0 - sk) * n2.y));
        }
        if (col == 0) top = make_double2(a.top[F.comp], 0.0);
      } else if (a.erosion) {
        const double2 e = a.ehat[col], cv = x[col * nz];
        double2 v = make_double2(e.x - a.ws * cv.x * a.cbar, e.y - a.ws * cv.y * a.cbar);
        const double scale = 10E+6;
        bot = make_double2(-1.0 * v.x * a.re * a.sc / a.cbar * scale, -1.0 * v.y * a.re * a.sc / a.cbar * scale);
      }
    }
    bcv[0][c] = bot;
    bcv[1][c] = top;
  }
  // even/odd halves of x for the D2 contraction
  struct LoHi { double2 lo, hi; };
  batched<5, LoHi>(pd.ncol * (M + 1), [&](int idx) {
    const int c = idx / (M + 1), j = idx - c * (M + 1);
    LoHi r = {zero, zero};
    if (c < ncol) {
      r.lo = x[(c0 + c) * nz + j];
      r.hi = x[(c0 + c) * nz + nz - 1 - j];
    }
    return r;
  }, [&](int idx, LoHi r) {
    const int c = idx / (M + 1), j = idx - c * (M + 1);
    if (j < M) {
      xs[c * ldx + j] = make_double2(0.5 * (r.lo.x + r.hi.x), 0.5 * (r.lo.y + r.hi.y));
      xs[c * ldx + M + 1 + j] = make_double2(0.5 * (r.lo.x - r.hi.x), 0.5 * (r.lo.y - r.hi.y));
    } else {
      xs[c * ldx + j] = r.lo;
    }
  });
  __syncthreads();
  // D2 x through the even/odd halves (s rows 0..M, d rows M+1..), each slot
  // emitted as v - q x_slot with q = s1 + kx^2 + ky^2 while the slot still
  // holds its half of x (the buffered contraction emits after its last
  // panel read, and every slot is read and written by the thread that emits
  // it).  Then for the row pair (j, nz-1-j):
  //   s1 x - ((D2 x - kx^2 x) - ky^2 x) = -(s'_j +- d'_j)
  // and the right-hand side needs no second read of x.
  gemm_panel<double, DEPTH>(a.d2p, a.ldp, pd, xs, [&](int row, int c, double2 v) {
    // the split's last index holds no half of x
    const double q = row > 2 * M ? 0.0 : qv[c];
    const double2 xv = xs[c * ldx + row];
    xs[c * ldx + row] = make_double2(v.x - q * xv.x, v.y - q * xv.y);
  });
  __syncthreads();
  // rhs per row pair (j, nz-1-j); rows 0 and nz-1 are the boundary rows, the
  // interior ones get the forward fix; written back to the two slots read
  struct In { double2 hl, hh, al, ah, cl, ch; };
  auto adv_at = [&](long long e) -> double2 {
    if (a.adv32) {
      const float2 v = reinterpret_cast<const float2*>(F.adv)[e];
      return make_double2((double)v.x, (double)v.y);
    }
    return reinterpret_cast<const double2*>(F.adv)[e];
  };
  // h <- c1 h - dt adv (+ buoyancy, + mean-column force), as h_update does
  auto hnew = [&](double2 hv, double2 av, double2 cv, long long col) {
    hv.x = a.c1 * hv.x - a.dt * av.x;
    hv.y = a.c1 * hv.y - a.dt * av.y;
    if (a.buoy) {
      if (F.comp == 2) {
        hv.x = hv.x - a.K * cv.x * a.cs;
        hv.y = hv.y - a.K * cv.y * a.cs;
      } else if ((F.comp == 0 && a.slope_dir == 1) || (F.comp == 1 && a.slope_dir == 2)) {
        hv.x = hv.x + a.K * cv.x * a.sn;
        hv.y = hv.y + a.K * cv.y * a.sn;
      }
    }
    if (a.force && col == 0) {
      const double f1 = a.fdev ? a.fdev[0] : a.f1, f2 = a.fdev ? a.fdev[1] : a.f2;
      if (F.comp == 0) hv.x += a.dt * f1;
      else if (F.comp == 1) hv.x += a.dt * f2;
    }
    return hv;
  };
  batched<2, In>(pd.ncol * (M + 1), [&](int idx) {
    const int c = idx / (M + 1), j = idx - c * (M + 1);
    In r = {zero, zero, zero, zero, zero, zero};
    if (c < ncol) {
      const long long o = (c0 + c) * nz;
      const bool upd = a.hupd != 0;
      if (j > 0 || upd) {
        r.hl = h[o + j];
        if (j < M) r.hh = h[o + nz - 1 - j];
      }
      if (upd) {
        r.al = adv_at(o + j);
        if (j < M) r.ah = adv_at(o + nz - 1 - j);
        if (a.buoy) {
          r.cl = a.cb[o + j];
          if (j < M) r.ch = a.cb[o + nz - 1 - j];
        }
      }
    }
    return r;
  }, [&](int idx, In r) {
    const int c = idx / (M + 1), j = idx - c * (M + 1);
    const double2 bot = bcv[0][c], top = bcv[1][c];
    double2 lo = bot, hi = top;
    if (a.hupd && c < ncol) {   // every row of h, boundary rows included
      const long long col = c0 + c;
      const long long o = col * nz;
      r.hl = hnew(r.hl, r.al, r.cl, col);
      h[o + j] = r.hl;
      if (j < M) {
        r.hh = hnew(r.hh, r.ah, r.ch, col);
        h[o + nz - 1 - j] = r.hh;
      }
    }
    if (j > 0) {
      const double2 sv = xs[c * ldx + j];
      const double2 dv = j < M ? xs[c * ldx + M + 1 + j] : zero;
      auto rhs = [&](double2 v, double2 hv) { return csub<double>(cscale<double>(a.s2, hv), v); };
      auto fixed = [&](int i, double2 v) {
        return cadd<double>(v, cadd<double>(cscale<double>(fixs[i], bot), cscale<double>(fixs[nz + i], top)));
      };
      lo = fixed(j, rhs(cadd<double>(sv, dv), r.hl));
      if (j < M) hi = fixed(nz - 1 - j, rhs(csub<double>(sv, dv), r.hh));
    }
    if (j == M) {
      xs[c * ldx + M] = lo;
    } else if (SPLIT) {
      xs[c * ldx + j] = cadd<double>(lo, hi);
      xs[c * ldx + M + 1 + j] = csub<double>(lo, hi);
    } else {
      xs[c * ldx + j] = lo;
      xs[c * ldx + M + 1 + j] = hi;
    }
  });
  __syncthreads();
  // Z = R o (V~^-1 X)
  const int brow = SPLIT ? M + 1 : nz - 1;
  gemm_panel<double, DEPTH>(F.vinv, a.lda, p1, xs, [&](int row, int c, double2 v) {
    double r = 1.0;
    if (row != 0 && row != brow) {
      const long long col = c0 + c;
      const double kx = kv[c].x, ky = kv[c].y;
      const double k2 = (-(kx * kx)) + (-(ky * ky));
      r = 1.0 / (((k2 + lam[row]) + sigma));
    }
    xs[c * ldx + row] = cscale<double>(r, v);
  });
  __syncthreads();
  gemm_panel<double, DEPTH>(F.vnorm, a.lda, p2, xs, [&](int row, int c, double2 v) { xs[c * ldx + row] = v; });
  __syncthreads();
  if (SPLIT) {
    backward_fixes(xs, ldx, pd.ncol, nz, M, F.pfix, F.pfix + nz, fix[3 * nz], tfix);
    __syncthreads();
  }
  store_columns<SPLIT>(xs, ldx, ncol, nz, M, tfix, F.y + c0 * nz);
}

void rhs_helmholtz(const RHArgs& a, int nfield, hipStream_t st) {
  TBD_CHECK(a.nz >= 7 && a.nz % 2 == 1 && a.P > 0 && nfield >= 1 && nfield <= 3 && a.d2p &&
                a.ldp >= (a.nz + 1) / 2 && a.ldp % 4 == 0 && a.lda % 4 == 0 &&
                a.lda >= (a.split ? (a.nz + 1) / 2 : a.nz),
            "rhs_helmholtz > bad shape");
  const Panel pd0 = make_panel(a.nz, sizeof(double2), kLdsLimit, 1);
  Panel q1, q2;
  if (a.split) {
    make_split_panels(a.nz, kLdsLimit, q1, q2);
  } else {
    q1 = make_panel(a.nz, sizeof(double2), kLdsLimit, 0);
    q2 = q1;
  }
  const int ncg = std::min(std::min(pd0.ncg, q1.ncg), q2.ncg);
  const int ldx = pd0.ldx;
  const Panel pd = with_ncg(pd0, ncg, ldx), p1 = with_ncg(q1, ncg, ldx), p2 = with_ncg(q2, ncg, ldx);
  TBD_CHECK(pd.ncol <= kSplitMaxCol, "rhs_helmholtz > too many columns per panel");
  const int depth = std::max(std::max(buffered_depth(pd), buffered_depth(p1)), buffered_depth(p2));
  TBD_CHECK(depth <= kMaxU, "rhs_helmholtz > nz too large for the register-buffered contraction");
  const size_t lds = sizeof(double2) * pd.ncol * ldx + sizeof(double) * (3 * a.nz + 1);
  TBD_CHECK(lds + sizeof(double2) * 5 * kSplitMaxCol <= 160 * 1024, "rhs_helmholtz > nz too large for one LDS panel");
  dim3 grid(ceil_div(a.P, pd.ncol), nfield);
  if (a.split) {
    if (depth <= 1)
      glaunch(k_rhs_helmholtz<1, true, false>, k_rhs_helmholtz<1, true, true>, grid, dim3(kNT), lds, st,
              RHL{a, pd, p1, p2});
    else
      glaunch(k_rhs_helmholtz<2, true, false>, k_rhs_helmholtz<2, true, true>, grid, dim3(kNT), lds, st,
              RHL{a, pd, p1, p2});
  } else {
    if (depth <= 1)
      glaunch(k_rhs_helmholtz<1, false, false>, k_rhs_helmholtz<1, false, true>, grid, dim3(kNT), lds, st,
              RHL{a, pd, p1, p2});
    else
      glaunch(k_rhs_helmholtz<2, false, false>, k_rhs_helmholtz<2, false, true>, grid, dim3(kNT), lds, st,
              RHL{a, pd, p1, p2});
  }
}

// ---------------------------------------------------------------------------
// divergence + pressure solve + projection in one kernel
// ---------------------------------------------------------------------------
struct PL {
  PArgs a;
  Panel pd, p1, p2;
};
template <int DEPTH, bool G>
__global__ __launch_bounds__(kNT, DEPTH == 1 ? 4 : 2) void k_pressure_chain(PL l0, GroupP grp) {
  Blk bk;
  const PL& L = member<G>(l0, grp, bk);
  const PArgs& a = L.a;
  const Panel& pd = L.pd;
  const Panel& p1 = L.p1;
  const Panel& p2 = L.p2;
  extern __shared__ char smem_raw[];
  __shared__ double2 tfix[2][kSplitMaxCol];
  const int nz = a.nz;
  const int M = (nz - 1) / 2;
  const int ldx = pd.ldx;
  const long long S = a.S;
  double2* xs = reinterpret_cast<double2*>(smem_raw);
  const long long c0 = (long long)bk.x * pd.ncol;
  const int ncol = (int)min((long long)pd.ncol, (long long)a.P - c0);
  __shared__ double2 kv[kSplitMaxCol];   // (kx, ky) of each panel column
  for (int c = threadIdx.x; c < pd.ncol; c += kNT) kv[c] = kcol(a.kx, a.ky, c0 + c, a.P, a.nyc);
  const double2 zero = make_double2(0.0, 0.0);
  const double2* su = a.s;
  const double2* sv = a.s + S;
  const double2* sw = a.s + 2 * S;
  struct LoHi { double2 lo, hi; };
  // even/odd halves of sw for D1
  batched<5, LoHi>(pd.ncol * (M + 1), [&](int idx) {
    const int c = idx / (M + 1), j = idx - c * (M + 1);
    LoHi r = {zero, zero};
    if (c < ncol) {
      r.lo = sw[(c0 + c) * nz + j];
      r.hi = sw[(c0 + c) * nz + nz - 1 - j];
    }
    return r;
  }, [&](int idx, LoHi r) {
    const int c = idx / (M + 1), j = idx - c * (M + 1);
    if (j < M) {
      xs[c * ldx + j] = make_double2(0.5 * (r.lo.x + r.hi.x), 0.5 * (r.lo.y + r.hi.y));
      xs[c * ldx + M + 1 + j] = make_double2(0.5 * (r.lo.x - r.hi.x), 0.5 * (r.lo.y - r.hi.y));
    } else {
      xs[c * ldx + j] = r.lo;
    }
  });
  __syncthreads();
  gemm_panel<double, DEPTH>(a.d1p, a.ldp, pd, xs, [&](int row, int c, double2 v) { xs[c * ldx + row] = v; });
  __syncthreads();
  // pressure right-hand side per row pair, boundary rows zeroed, split (s, d)
  struct UV { double2 ul, uh, vl, vh; };
  batched<3, UV>(pd.ncol * (M + 1), [&](int idx) {
    const int c = idx / (M + 1), j = idx - c * (M + 1);
    UV r = {zero, zero, zero, zero};
    if (c < ncol && j > 0) {
      const long long o = (c0 + c) * nz;
      r.ul = su[o + j];
      r.vl = sv[o + j];
      if (j < M) {
        r.uh = su[o + nz - 1 - j];
        r.vh = sv[o + nz - 1 - j];
      }
    }
    return r;
  }, [&](int idx, UV r) {
    const int c = idx / (M + 1), j = idx - c * (M + 1);
    double2 lo = zero, hi = zero;
    if (j > 0) {
      const long long col = c0 + c;
      const double kx = kv[c].x, ky = kv[c].y;
      const double2 sv1 = xs[c * ldx + j], dv1 = xs[c * ldx + M + 1 + j];
      auto div = [&](double2 v, double2 u0, double2 v0) {
        const double2 t = cadd<double>(cadd<double>(v, cmuli<double>(kx, u0)), cmuli<double>(ky, v0));
        return cscale<double>(a.q, t);
      };
      lo = div(cadd<double>(sv1, dv1), r.ul, r.vl);
      if (j < M) hi = div(csub<double>(dv1, sv1), r.uh, r.vh);
    }
    if (j == M) {
      xs[c * ldx + M] = lo;
    } else {
      xs[c * ldx + j] = cadd<double>(lo, hi);
      xs[c * ldx + M + 1 + j] = csub<double>(lo, hi);
    }
  });
  __syncthreads();
  gemm_panel<double, DEPTH>(a.pinv, a.lda, p1, xs, [&](int row, int c, double2 v) {
    double r = 1.0;
    if (row != 0 && row != M + 1 && row < nz) {   // row nz: zero row, no eigenvalue
      const long long col = c0 + c;
      const double kx = kv[c].x, ky = kv[c].y;
      const double k2 = (-(kx * kx)) + (-(ky * ky));
      r = 1.0 / (((k2 + a.plam[row]) + 0.0));
      if (row == a.nullrow && col == 0) r = 0.0;
    }
    xs[c * ldx + row] = cscale<double>(r, v);
  });
  __syncthreads();
  gemm_panel<double, DEPTH>(a.pnorm, a.lda, p2, xs, [&](int row, int c, double2 v) { xs[c * ldx + row] = v; });
  __syncthreads();
  backward_fixes(xs, ldx, pd.ncol, nz, M, a.pfix, a.pfix + nz, a.fix[3 * nz], tfix);
  __syncthreads();
  // p to HBM, and u = su - cf i kx p, v = sv - cf i ky p while p is in the
  // panel (natural row order; the wall rows of (i kx p, i ky p) saved).  The
  // panel's (a, b) are p's even/odd halves except at the wall pair, which is
  // then replaced by ((p_0 + p_N-1)/2, (p_0 - p_N-1)/2) for D1 (this pass
  // reads neither slot).
  struct UV2 { double2 u, v; };
  batched<4, UV2>(ncol * nz, [&](int idx) {
    const long long off = c0 * nz + idx;
    UV2 r = {su[off], sv[off]};
    return r;
  }, [&](int idx, UV2 r) {
    const int c = idx / nz, k = idx - c * nz;
    double2 pv;
    if (k == 0) {
      pv = tfix[0][c];
    } else if (k == nz - 1) {
      pv = tfix[1][c];
    } else if (k <= M) {
      pv = cadd<double>(xs[c * ldx + k], xs[c * ldx + M + 1 + k]);
    } else {
      const int i = nz - 1 - k;
      pv = csub<double>(xs[c * ldx + i], xs[c * ldx + M + 1 + i]);
    }
    const long long off = c0 * nz + idx;
    a.p[off] = pv;
    const double2 gx = cmuli<double>(kv[c].x, pv), gy = cmuli<double>(kv[c].y, pv);
    a.y[off] = csub<double>(r.u, cscale<double>(a.cf, gx));
    a.y[S + off] = csub<double>(r.v, cscale<double>(a.cf, gy));
    if (k == 0 || k == nz - 1) {
      double2* wb = (k == 0) ? a.wall_bot : a.wall_top;
      wb[c0 + c] = gx;
      wb[a.P + c0 + c] = gy;
    }
  });
  if ((int)threadIdx.x < pd.ncol) {
    const int c = threadIdx.x;
    const double2 t0 = tfix[0][c], tn = tfix[1][c];
    xs[c * ldx] = make_double2(0.5 * (t0.x + tn.x), 0.5 * (t0.y + tn.y));
    xs[c * ldx + M + 1] = make_double2(0.5 * (t0.x - tn.x), 0.5 * (t0.y - tn.y));
  }
  __syncthreads();
  gemm_panel<double, DEPTH>(a.d1p, a.ldp, pd, xs, [&](int row, int c, double2 v) { xs[c * ldx + row] = v; });
  __syncthreads();
  // w = sw - cf D1 p per row pair (mean w := 0), wall rows of D1 p (and of
  // (i kx, i ky) D1 p when open_top)
  struct Wv { double2 wl, wh; };
  batched<4, Wv>(pd.ncol * (M + 1), [&](int idx) {
    const int c = idx / (M + 1), j = idx - c * (M + 1);
    Wv r = {zero, zero};
    if (c < ncol) {
      const long long o = (c0 + c) * nz;
      r.wl = sw[o + j];
      if (j < M) r.wh = sw[o + nz - 1 - j];
    }
    return r;
  }, [&](int idx, Wv r) {
    const int c = idx / (M + 1), j = idx - c * (M + 1);
    if (c >= ncol) return;
    const long long col = c0 + c;
    const double kx = kv[c].x, ky = kv[c].y;
    const double2 s2 = xs[c * ldx + j], d2 = xs[c * ldx + M + 1 + j];
    auto put = [&](int row, double2 v, double2 w0) {
      const long long off = col * nz + row;
      double2 w = csub<double>(w0, cscale<double>(a.cf, v));
      if (col == 0) w = make_double2(0.0, 0.0);
      a.y[2 * S + off] = w;
      if (row == 0 || row == nz - 1) {
        double2* wb = (row == 0) ? a.wall_bot : a.wall_top;
        wb[2 * a.P + col] = v;
        if (a.open_top) {
          double2* w2 = (row == 0) ? a.wall2_bot : a.wall2_top;
          w2[col] = cmuli<double>(kx, v);
          w2[a.P + col] = cmuli<double>(ky, v);
        }
      }
    };
    put(j, cadd<double>(s2, d2), r.wl);
    if (j < M) put(nz - 1 - j, csub<double>(d2, s2), r.wh);
  });
}

void pressure_chain(const PArgs& a, hipStream_t st) {
  TBD_CHECK(a.nz >= 7 && a.nz % 2 == 1 && a.P > 0 && a.d1p && a.pinv && a.pnorm && a.pfix && a.fix &&
                a.ldp >= (a.nz + 1) / 2 && a.ldp % 4 == 0 && a.lda >= (a.nz + 1) / 2 && a.lda % 4 == 0,
            "pressure_chain > bad shape");
  const Panel pd0 = make_panel(a.nz, sizeof(double2), kLdsLimit, -1);
  Panel q1, q2;
  make_split_panels(a.nz, kLdsLimit, q1, q2);
  const int ncg = std::min(std::min(pd0.ncg, q1.ncg), q2.ncg);
  const int ldx = pd0.ldx;
  const Panel pd = with_ncg(pd0, ncg, ldx), p1 = with_ncg(q1, ncg, ldx), p2 = with_ncg(q2, ncg, ldx);
  TBD_CHECK(pd.ncol <= kSplitMaxCol, "pressure_chain > too many columns per panel");
  const int depth = std::max(std::max(buffered_depth(pd), buffered_depth(p1)), buffered_depth(p2));
  TBD_CHECK(depth <= kMaxU, "pressure_chain > nz too large for the register-buffered contraction");
  const size_t lds = sizeof(double2) * pd.ncol * ldx;
  TBD_CHECK(lds + sizeof(double2) * 3 * kSplitMaxCol <= 160 * 1024, "pressure_chain > nz too large for one LDS panel");
  dim3 grid(ceil_div(a.P, pd.ncol), 1);
  if (depth <= 1)
    glaunch(k_pressure_chain<1, false>, k_pressure_chain<1, true>, grid, dim3(kNT), lds, st, PL{a, pd, p1, p2});
  else
    glaunch(k_pressure_chain<2, false>, k_pressure_chain<2, true>, grid, dim3(kNT), lds, st, PL{a, pd, p1, p2});
}

double pressure_chain_flops(int nz, long long P) {
  const double M = (nz - 1) / 2;
  const double d1 = 4.0 * ((M + 1) * (M + 1) + (M + 1) * M) * (double)P;
  return 2 * d1 + helmholtz_split_flops(nz, P);
}

std::vector<double> permute_columns(const std::vector<double>& rm, int nz, int lda) {
  const int M = (nz - 1) / 2;
  std::vector<double> out(rm.size(), 0.0);
  for (int r = 0; r < nz; ++r)
    for (int p = 0; p < nz; ++p) {
      const int nat = p <= M ? p : 3 * M + 1 - p;
      out[(size_t)r * lda + p] = rm[(size_t)r * lda + nat];
    }
  return out;
}

double rhs_helmholtz_flops(int nz, long long P, bool split) {
  const double M = (nz - 1) / 2;
  const double d2 = 4.0 * ((M + 1) * (M + 1) + (M + 1) * M) * (double)P;
  return d2 + (split ? helmholtz_split_flops(nz, P) : 8.0 * nz * (double)nz * P);
}

double helmholtz_split_flops(int nz, long long P) {
  const double M = (nz - 1) / 2;
  // pinv: (M+1)^2 + M^2, split V~: (M+1)^2 + (M+1) M multiply-adds per real column
  return 4.0 * ((M + 1) * (M + 1) + M * M + (M + 1) * (M + 1) + (M + 1) * M) * (double)P;
}

// ---------------------------------------------------------------------------
// wall rows of M x: one wave per column
// ---------------------------------------------------------------------------
struct ZRowsA {
  const double* op;
  int lda, nz;
  const double2* x;
  long long P;
  double2* bot;
  double2* top;
};
template <bool G>
__global__ __launch_bounds__(kNT) void k_zrows_wall(ZRowsA a0, GroupP grp) {
  Blk bk;
  const ZRowsA& A = member<G>(a0, grp, bk);
  const double* __restrict__ op = A.op;
  const int lda = A.lda, nz = A.nz;
  const double2* __restrict__ x = A.x;
  const long long P = A.P;
  double2* __restrict__ bot = A.bot;
  double2* __restrict__ top = A.top;
  const int lane = threadIdx.x & 63;
  const long long col = (long long)bk.x * (kNT / 64) + (threadIdx.x >> 6);
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
  glaunch(k_zrows_wall<false>, k_zrows_wall<true>, grid, dim3(kNT), 0, st, ZRowsA{op, lda, nz, x, P, bottom, top});
}

// one wave per (field, wall): 3 fields x 2 rows
struct WallMeanA {
  const double* op;
  int lda, nz;
  const double2* x;
  long long S;
  double* out;
};
template <bool G>
__global__ __launch_bounds__(384) void k_wall_mean(WallMeanA a0, GroupP grp) {
  Blk bk;
  const WallMeanA& A = member<G>(a0, grp, bk);
  const double* __restrict__ op = A.op;
  const int lda = A.lda, nz = A.nz;
  const double2* __restrict__ x = A.x;
  const long long S = A.S;
  double* out = A.out;
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
  glaunch(k_wall_mean<false>, k_wall_mean<true>, dim3(1), dim3(384), 0, st, WallMeanA{op, lda, nz, x, S, out});
}

template <typename T>
std::vector<T> pack_parity(const double* A, int nz, int lda, int* ldp) {
  TBD_CHECK(nz % 2 == 1 && nz >= 3, "pack_parity > nz must be odd");
  const int M = (nz - 1) / 2;
  const int ld = (M + 1 + 3) / 4 * 4;
  std::vector<T> out((size_t)(2 * M + 2) * ld, (T)0);
  for (int i = 0; i <= M; ++i) {
    const double* ar = A + (size_t)i * lda;
    for (int j = 0; j < M; ++j) {
      out[(size_t)i * ld + j] = (T)(ar[j] + ar[nz - 1 - j]);
      out[(size_t)(M + 1 + i) * ld + j] = (T)(ar[j] - ar[nz - 1 - j]);
    }
    out[(size_t)i * ld + M] = (T)ar[M];
  }
  *ldp = ld;
  return out;
}
template std::vector<double> pack_parity<double>(const double*, int, int, int*);
template std::vector<float> pack_parity<float>(const double*, int, int, int*);

}  // namespace tbd
