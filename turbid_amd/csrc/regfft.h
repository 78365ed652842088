// Register-resident DFT building blocks for the plane transforms.
//
// A length N = N1*N2 transform is done in two levels (Cooley-Tukey,
// n = N2*n1 + n2, k = k1 + N1*k2):
//   level A: for each n2, a DFT of size N1 over the stride-N2 subsequence,
//            then the twiddle W_N^(n2*k1);
//   level B: for each k1, a DFT of size N2 over n2, giving X[k1 + N1*k2].
// Each level runs in the registers of one thread (one "task" holds one
// small DFT); the data moves between the levels through LDS once.  N1 and N2
// divide 48, so every in-register twiddle is a 48th root of unity and is a
// compile-time constant after unrolling.
#pragma once

#include "common.h"

namespace tbd {

// sin(2 pi j / 48), j = 0..12, correctly rounded
__host__ __device__ constexpr double qsin48(int j) {
  return j == 0 ? 0.0
       : j == 1 ? 0.1305261922200516
       : j == 2 ? 0.25881904510252074
       : j == 3 ? 0.3826834323650898
       : j == 4 ? 0.5
       : j == 5 ? 0.6087614290087207
       : j == 6 ? 0.7071067811865476
       : j == 7 ? 0.7933533402912352
       : j == 8 ? 0.8660254037844386
       : j == 9 ? 0.9238795325112867
       : j == 10 ? 0.9659258262890683
       : j == 11 ? 0.9914448613738104
       : 1.0;
}
__host__ __device__ constexpr double sin48(int j) {
  return j <= 12 ? qsin48(j) : j <= 24 ? qsin48(24 - j) : j <= 36 ? -qsin48(j - 24) : -qsin48(48 - j);
}
__host__ __device__ constexpr double cos48(int j) { return sin48((j + 12) % 48); }

// v * exp(SIGN * 2 pi i j / 48), exact for the quarter turns
template <typename T, int SIGN>
__device__ __forceinline__ C2<T> rot48(C2<T> v, int j) {
  j %= 48;
  if (j == 0) return v;
  if (j == 24) return mk<T>(-v.x, -v.y);
  if (j == 12) return SIGN > 0 ? mk<T>(-v.y, v.x) : mk<T>(v.y, -v.x);
  if (j == 36) return SIGN > 0 ? mk<T>(v.y, -v.x) : mk<T>(-v.y, v.x);
  const T c = (T)cos48(j), s = (T)(SIGN * sin48(j));
  return mk<T>(v.x * c - v.y * s, v.x * s + v.y * c);
}

// In-place DFT of M points held in registers, natural order in and out.
// SIGN -1: forward (exp(-2 pi i nk/M)), +1: inverse, unnormalised.
template <typename T, int M, int SIGN> struct RegDft;

template <typename T, int SIGN> struct RegDft<T, 1, SIGN> {
  static __device__ __forceinline__ void run(C2<T>*) {}
};
template <typename T, int SIGN> struct RegDft<T, 2, SIGN> {
  static __device__ __forceinline__ void run(C2<T>* v) {
    const C2<T> a = v[0], b = v[1];
    v[0] = cadd<T>(a, b);
    v[1] = csub<T>(a, b);
  }
};
template <typename T, int SIGN> struct RegDft<T, 3, SIGN> {
  static __device__ __forceinline__ void run(C2<T>* v) {
    const T h = (T)0.86602540378443864676;   // sqrt(3)/2
    const C2<T> sm = cadd<T>(v[1], v[2]), df = csub<T>(v[1], v[2]);
    const C2<T> c0 = mk<T>(v[0].x - (T)0.5 * sm.x, v[0].y - (T)0.5 * sm.y);
    const C2<T> id = mk<T>(-(T)SIGN * h * df.y, (T)SIGN * h * df.x);
    v[0] = cadd<T>(v[0], sm);
    v[1] = cadd<T>(c0, id);
    v[2] = csub<T>(c0, id);
  }
};
template <typename T, int SIGN> struct RegDft<T, 4, SIGN> {
  static __device__ __forceinline__ void run(C2<T>* v) {
    const C2<T> a0 = cadd<T>(v[0], v[2]), a1 = csub<T>(v[0], v[2]);
    const C2<T> a2 = cadd<T>(v[1], v[3]), a3 = csub<T>(v[1], v[3]);
    // SIGN * i * a3
    const C2<T> ia3 = SIGN > 0 ? mk<T>(-a3.y, a3.x) : mk<T>(a3.y, -a3.x);
    v[0] = cadd<T>(a0, a2);
    v[1] = cadd<T>(a1, ia3);
    v[2] = csub<T>(a0, a2);
    v[3] = csub<T>(a1, ia3);
  }
};

// M = P*Q: DFT_P over n1 for each n2, twiddle W_M^(n2*k1), DFT_Q over n2 for each k1
template <typename T, int M, int SIGN> struct RegDft {
  static constexpr int P = (M % 4 == 0) ? 4 : ((M % 2 == 0) ? 2 : 3);
  static constexpr int Q = M / P;
  static_assert(48 % M == 0, "in-register DFT lengths divide 48");
  static __device__ __forceinline__ void run(C2<T>* v) {
    C2<T> y[M];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
      C2<T> a[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) a[n1] = v[Q * n1 + n2];
      RegDft<T, P, SIGN>::run(a);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) y[k1 * Q + n2] = rot48<T, SIGN>(a[k1], (n2 * k1) * (48 / M));
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      RegDft<T, Q, SIGN>::run(y + k1 * Q);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = y[k1 * Q + k2];
    }
  }
};

// Two-level split of a transform length: N1 (level A, the larger factor)
// and N2, both dividing 48; {0, 0} when the length has no such split.
struct Split { int n1, n2; };
__host__ __device__ constexpr Split two_level(int n) {
  return n == 576 ? Split{24, 24} : n == 384 ? Split{24, 16} : n == 288 ? Split{24, 12}
       : n == 256 ? Split{16, 16} : n == 192 ? Split{16, 12} : n == 144 ? Split{12, 12}
       : n == 128 ? Split{16, 8}  : n == 96 ? Split{12, 8}   : n == 72 ? Split{12, 6}
       : n == 64 ? Split{8, 8}    : n == 48 ? Split{8, 6}    : n == 36 ? Split{6, 6}
       : n == 32 ? Split{8, 4}    : n == 24 ? Split{6, 4}    : n == 18 ? Split{6, 3}
       : n == 16 ? Split{4, 4}    : n == 12 ? Split{4, 3}    : n == 8 ? Split{4, 2}
       : n == 6 ? Split{3, 2}     : n == 4 ? Split{4, 1}     : n == 3 ? Split{3, 1}
       : n == 2 ? Split{2, 1}     : Split{0, 0};
}

// Complex LDS slots of the level-A -> level-B exchange E[s][k1][n2] for S
// sequences: rows of N2+1, sequence blocks of N1*(N2+1)+1 (odd strides).
template <int N1, int N2>
__host__ __device__ constexpr int xrow() { return N2 + 1; }
template <int N1, int N2>
__host__ __device__ constexpr int xseq() { return N1 * (N2 + 1) + 1; }

// twiddle W_N^(e) from the table exp(-2 pi i t/N), conjugated for SIGN > 0
template <typename T, int SIGN>
__device__ __forceinline__ C2<T> twiddle(const C2<T>* __restrict__ tw, int e) {
  C2<T> w = tw[e];
  if (SIGN > 0) w.y = -w.y;
  return w;
}

// level A on registers v (x[N2*n1 + n2], n1 < N1) -> E[s][k1][n2]
template <typename T, int N1, int N2, int SIGN>
__device__ __forceinline__ void level_a_store(C2<T>* v, C2<T>* E, int s, int n2, const C2<T>* __restrict__ tw) {
  RegDft<T, N1, SIGN>::run(v);
  C2<T>* e = E + s * xseq<N1, N2>() + n2;
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1)
    e[k1 * xrow<N1, N2>()] = (k1 == 0 || n2 == 0) ? v[k1] : cmul<T>(v[k1], twiddle<T, SIGN>(tw, n2 * k1));
}
// level B: w <- E[s][k1][*], DFT_N2; w[k2] = X[k1 + N1*k2]
template <typename T, int N1, int N2, int SIGN>
__device__ __forceinline__ void level_b_load(C2<T>* w, const C2<T>* E, int s, int k1) {
  const C2<T>* e = E + s * xseq<N1, N2>() + k1 * xrow<N1, N2>();
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) w[n2] = e[n2];
  RegDft<T, N2, SIGN>::run(w);
}

// The same exchange through half the LDS: one real plane E[s][k1][n2]
// carries first the real parts, then (after a barrier on both sides) the
// imaginary parts.  level_a_regs leaves the twiddled level-A outputs in v;
// level_a_put writes part 0 (real) or 1 (imaginary); level_b_get reads it
// into w; level_b_dft finishes level B.
template <typename T, int N1, int N2, int SIGN>
__device__ __forceinline__ void level_a_regs(C2<T>* v, int n2, const C2<T>* __restrict__ tw) {
  RegDft<T, N1, SIGN>::run(v);
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1)
    if (n2 != 0) v[k1] = cmul<T>(v[k1], twiddle<T, SIGN>(tw, n2 * k1));
}
//
// Slot of (s, k1, n2) in the real plane.  SW = 0: rows of N2+1, sequence
// blocks of xseq (odd strides).  For N1 = N2 = 16 the plane can instead be
// swizzled (256 words per sequence): the bank's low four bits are s XOR the
// index that runs over 16 lanes of one pass, the two bits above them
// (s + the index that runs over 4 lanes) mod 4, so that every 64-lane pass
// touches 64 distinct banks.  SW = 1: n2 in the low bits, for the y forward
// pass (put: lanes over n2 and 4 sequences; get: lanes over 16 sequences and
// 4 k1).  SW = 2: k1 in the low bits, for the y backward pass (put: lanes
// over 16 sequences and 4 n2; get: lanes over k1 and 4 sequences).
template <int N1, int N2, int SW>
__device__ __forceinline__ int xslot(int s, int k1, int n2) {
  static_assert(SW == 0 || (N1 == 16 && N2 == 16), "swizzled plane is for 16 x 16");
  if (SW == 1) return s * 256 + 64 * (k1 >> 2) + 16 * ((k1 + s) & 3) + ((s ^ n2) & 15);
  if (SW == 2) return s * 256 + 64 * (n2 >> 2) + 16 * ((n2 + s) & 3) + ((s ^ k1) & 15);
  return s * xseq<N1, N2>() + k1 * xrow<N1, N2>() + n2;
}
template <typename T, int N1, int N2, int SW = 0>
__device__ __forceinline__ void level_a_put(const C2<T>* v, T* E, int s, int n2, int part) {
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) E[xslot<N1, N2, SW>(s, k1, n2)] = part == 0 ? v[k1].x : v[k1].y;
}
template <typename T, int N1, int N2, int SW = 0>
__device__ __forceinline__ void level_b_get(C2<T>* w, const T* E, int s, int k1, int part) {
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) {
    const T u = E[xslot<N1, N2, SW>(s, k1, n2)];
    if (part == 0) w[n2].x = u;
    else w[n2].y = u;
  }
}
template <typename T, int N2, int SIGN>
__device__ __forceinline__ void level_b_dft(C2<T>* w) {
  RegDft<T, N2, SIGN>::run(w);
}

}  // namespace tbd
