// Fused x transforms + nonlinear products (see nonlinear.h).
#include "nonlinear.h"

#include "fft_device.h"
#include "group.h"
#include "regfft.h"

namespace tbd {

static constexpr int kNT = 256;
static constexpr int kPPT = 8;     // packed row-pair positions per thread
static constexpr int kFftB = 4;    // butterflies per thread per stage

struct XfArgs {
  FftLen L;
  int ny = 0, mx = 0, nz = 0, nseq = 0, nout = 0;
  long long R = 0;
  const void* tw = nullptr;
  const void* gin = nullptr;
  const void* U = nullptr;
  void* gout = nullptr;
  void* bot0 = nullptr;
  void* bot1 = nullptr;
  double su = 0, sv = 0, sw = 0;
  // conv form with the x derivatives formed here: the x-pencils of u, v, w,
  // and kx = dkx * (spectral x index)
  const void* ux = nullptr;
  double dkx = 0;
};

// inverse x transform input: zero-padded, Hermitian-completed (halfcomplex)
template <typename T>
__device__ __forceinline__ void load_c2r(C2<T>* buf, int ldn, const C2<T>* in, int N, int ny, int mx, int y0,
                                         int nseq) {
  for (int idx = threadIdx.x; idx < nseq * N; idx += kNT) {
    const int s = idx / N, x = idx - s * N;
    buf[s * ldn + x] = mk<T>((T)0, (T)0);
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < mx * nseq; idx += kNT) {
    const int k = idx / nseq, s = idx - k * nseq;
    const int y = y0 + 2 * s;
    C2<T> A = (y < ny) ? in[(long long)k * ny + y] : mk<T>((T)0, (T)0);
    C2<T> B = (y + 1 < ny) ? in[(long long)k * ny + y + 1] : mk<T>((T)0, (T)0);
    const bool selfconj = (k == 0) || (2 * k == N);
    if (selfconj) { A.y = (T)0; B.y = (T)0; }
    buf[s * ldn + k] = mk<T>(A.x - B.y, A.y + B.x);
    if (!selfconj) buf[s * ldn + N - k] = mk<T>(A.x + B.y, B.x - A.y);
  }
  __syncthreads();
}

// forward x transform output: truncate to kx < mx, unpack the two rows
template <typename T>
__device__ __forceinline__ void store_r2c(const C2<T>* buf, int ldn, C2<T>* out, int N, int ny, int mx, int y0,
                                          int nseq) {
  const int nr = 2 * nseq;
  for (int idx = threadIdx.x; idx < mx * nr; idx += kNT) {
    const int k = idx / nr, r = idx - k * nr;
    const int y = y0 + r;
    if (y >= ny) continue;
    const int s = r >> 1;
    const C2<T> zk = buf[s * ldn + k];
    const C2<T> zn = buf[s * ldn + (k == 0 ? 0 : N - k)];
    C2<T> val;
    if ((r & 1) == 0) val = mk<T>((T)0.5 * (zk.x + zn.x), (T)0.5 * (zk.y - zn.y));
    else val = mk<T>((T)0.5 * (zk.y + zn.y), (T)0.5 * (zn.x - zk.x));
    out[(long long)k * ny + y] = val;
  }
}

// packed physical rows (2s, 2s+1) of a field at this thread's positions
template <typename T>
__device__ __forceinline__ void load_rows(C2<T>* v, const T* base, int N, int ny, int y0, int nseq, T shift) {
#pragma unroll
  for (int b = 0; b < kPPT; ++b) {
    const int q = threadIdx.x + b * kNT;
    if (q < nseq * N) {
      const int s = q / N, x = q - s * N;
      const int y = y0 + 2 * s;
      const T a = (y < ny) ? base[(long long)y * N + x] + shift : (T)0;
      const T c = (y + 1 < ny) ? base[(long long)(y + 1) * N + x] + shift : (T)0;
      v[b] = mk<T>(a, c);
    }
  }
}

template <typename T, int MODE>
__global__ __launch_bounds__(kNT) void k_xfused(XfArgs a) {
  extern __shared__ char smem_raw[];
  C2<T>* buf = reinterpret_cast<C2<T>*>(smem_raw);
  const int N = a.L.n;
  const int ldn = lds_ld(N);
  C2<T>* tw = buf + a.nseq * ldn;
  const int z = blockIdx.y;
  const int nseq = a.nseq;
  const int y0 = blockIdx.x * 2 * nseq;
  const long long plane = (long long)a.mx * a.ny;            // x-pencil plane
  const long long pstride = (long long)a.nz * plane;          // x-pencil component
  const C2<T>* gin = (const C2<T>*)a.gin;
  C2<T>* gout = (C2<T>*)a.gout;
  const T* U = (const T*)a.U;
  const T shifts[3] = {(T)a.su, (T)a.sv, (T)a.sw};
  const T* Ub = U + (long long)z * a.ny * N;
  load_tw<T, kNT>(tw, (const C2<T>*)a.tw, N);
  __syncthreads();

  if (MODE == 0) {
    // adv_i = sum_j U_j g_ij
    C2<T> acc[3][kPPT];
    C2<T> uj[kPPT];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int b = 0; b < kPPT; ++b) acc[i][b] = mk<T>((T)0, (T)0);
    for (int j = 0; j < 3; ++j) {
      load_rows<T>(uj, Ub + j * a.R, N, a.ny, y0, nseq, shifts[j]);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        if (i < a.nout) {   // uniform across the block
          const int gcomp = (a.nout == 3) ? 3 * i + j : j;
          load_c2r<T>(buf, ldn, gin + gcomp * pstride + z * plane, N, a.ny, a.mx, y0, nseq);
          lds_fft<T, kNT, kFftB>(buf, ldn, nseq, a.L, tw, +1);
          if (a.bot0 && z == 0 && j == 2 && i < 2) {   // physical bottom plane of du/dz, dv/dz
            T* bot = (T*)(i == 0 ? a.bot0 : a.bot1);
            for (int idx = threadIdx.x; idx < nseq * N; idx += kNT) {
              const int s = idx / N, x = idx - s * N;
              const int y = y0 + 2 * s;
              const C2<T> v = buf[s * ldn + x];
              if (y < a.ny) bot[(long long)y * N + x] = v.x;
              if (y + 1 < a.ny) bot[(long long)(y + 1) * N + x] = v.y;
            }
          }
#pragma unroll
          for (int b = 0; b < kPPT; ++b) {
            const int q = threadIdx.x + b * kNT;
            if (q < nseq * N) {
              const int s = q / N, x = q - s * N;
              const C2<T> g = buf[s * ldn + x];
              acc[i][b].x = acc[i][b].x + uj[b].x * g.x;
              acc[i][b].y = acc[i][b].y + uj[b].y * g.y;
            }
          }
          __syncthreads();
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (i < a.nout) {
#pragma unroll
        for (int b = 0; b < kPPT; ++b) {
          const int q = threadIdx.x + b * kNT;
          if (q < nseq * N) {
            const int s = q / N, x = q - s * N;
            buf[s * ldn + x] = acc[i][b];
          }
        }
        __syncthreads();
        lds_fft<T, kNT, kFftB>(buf, ldn, nseq, a.L, tw, -1);
        store_r2c<T>(buf, ldn, gout + i * pstride + z * plane, N, a.ny, a.mx, y0, nseq);
        __syncthreads();
      }
    }
  } else if (MODE == 1) {
    // uu, uv, uw, vv, vw, ww
    C2<T> u[3][kPPT];
    for (int j = 0; j < 3; ++j) load_rows<T>(u[j], Ub + j * a.R, N, a.ny, y0, nseq, (T)0);
    constexpr int pa[6] = {0, 0, 0, 1, 1, 2};
    constexpr int pb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int o = 0; o < 6; ++o) {
#pragma unroll
      for (int b = 0; b < kPPT; ++b) {
        const int q = threadIdx.x + b * kNT;
        if (q < nseq * N) {
          const int s = q / N, x = q - s * N;
          const C2<T> ua = u[pa[o]][b], ub = u[pb[o]][b];
          buf[s * ldn + x] = mk<T>(ua.x * ub.x, ua.y * ub.y);
        }
      }
      __syncthreads();
      lds_fft<T, kNT, kFftB>(buf, ldn, nseq, a.L, tw, -1);
      store_r2c<T>(buf, ldn, gout + o * pstride + z * plane, N, a.ny, a.mx, y0, nseq);
      __syncthreads();
    }
  } else {
    // c (U_j + shift_j), j = 0..2
    C2<T> cv[kPPT];
    C2<T> uj[kPPT];
    load_c2r<T>(buf, ldn, gin + z * plane, N, a.ny, a.mx, y0, nseq);
    lds_fft<T, kNT, kFftB>(buf, ldn, nseq, a.L, tw, +1);
#pragma unroll
    for (int b = 0; b < kPPT; ++b) {
      const int q = threadIdx.x + b * kNT;
      if (q < nseq * N) {
        const int s = q / N, x = q - s * N;
        cv[b] = buf[s * ldn + x];
      }
    }
    __syncthreads();
    for (int j = 0; j < 3; ++j) {
      load_rows<T>(uj, Ub + j * a.R, N, a.ny, y0, nseq, shifts[j]);
#pragma unroll
      for (int b = 0; b < kPPT; ++b) {
        const int q = threadIdx.x + b * kNT;
        if (q < nseq * N) {
          const int s = q / N, x = q - s * N;
          buf[s * ldn + x] = mk<T>(cv[b].x * uj[b].x, cv[b].y * uj[b].y);
        }
      }
      __syncthreads();
      lds_fft<T, kNT, kFftB>(buf, ldn, nseq, a.L, tw, -1);
      store_r2c<T>(buf, ldn, gout + j * pstride + z * plane, N, a.ny, a.mx, y0, nseq);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Two-level register version (lengths N = N1*N2 whose factors divide 48).
// A workgroup takes S packed row pairs of one z-plane.  The inverse x
// transform leaves task (k1 = t % N1, s = t / N1) holding the physical points
// x = k1 + N1*k2, k2 < N2; the forward transform uses the swapped split
// (N2 first), whose level-A input is exactly that set, so products are
// formed in registers and only the transforms' own exchanges touch LDS.
// ---------------------------------------------------------------------------
template <int N1, int N2, int NT>
__host__ __device__ constexpr int xf_seqs() {
  return (NT / (N1 > N2 ? N1 : N2)) > 0 ? NT / (N1 > N2 ? N1 : N2) : 1;
}
template <int N1, int N2, int NT>
__host__ __device__ constexpr int xf_lds() {
  constexpr int S = xf_seqs<N1, N2, NT>();
  constexpr int a = S * xseq<N1, N2>(), b = S * xseq<N2, N1>(), c = S * (N1 * N2 + 1);
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

// inverse x transform of one x-pencil plane (kx < mx, y fastest) for the
// rows y0 .. y0+2S: task (k1, s) receives w[k2] = row pair at x = k1 + N1*k2.
// With dkx != 0 the plane is first multiplied by i kx, kx = dkx * index (the
// x derivative, as Gradient's i kx multiplier applies it in spectral space).
template <typename T, int N1, int N2, int NT>
__device__ __forceinline__ void xback_regs(const C2<T>* __restrict__ in, int ny, int mx, int y0, C2<T>* L,
                                           const C2<T>* __restrict__ tw, C2<T>* w, double dkx = 0.0) {
  constexpr int N = N1 * N2;
  constexpr int S = xf_seqs<N1, N2, NT>();
  const int t = threadIdx.x;
  // all of a thread's spectra are loaded before any is packed into LDS
  constexpr int NI = (S * N + NT - 1) / NT;
  C2<T> ra[NI], rb[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int idx = t + i * NT;
    const int s = idx % S, kk = idx / S;
    const int ya = y0 + 2 * s;
    const bool direct = kk < mx;
    const int kx = direct ? kk : N - kk;
    // every load reads an in-bounds element (column and rows clamped) and
    // what lies outside the transform is zeroed afterwards: the loads carry
    // no branch, so they are all in flight together
    const bool col = idx < S * N && (direct || kx < mx);
    // lanes past the last sequence (idx >= S N, when NT does not divide
    // S N) have kk > N and so kx < 0: their column is clamped too
    const long long kc = (long long)(kx < 0 ? 0 : (kx < mx ? kx : mx - 1)) * ny;
    const C2<T> A = in[kc + (ya < ny ? ya : ny - 1)];
    const C2<T> B = in[kc + (ya + 1 < ny ? ya + 1 : ny - 1)];
    ra[i] = (col && ya < ny) ? A : mk<T>((T)0, (T)0);
    rb[i] = (col && ya + 1 < ny) ? B : mk<T>((T)0, (T)0);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int idx = t + i * NT;
    if (idx < S * N) {
      const int s = idx % S, kk = idx / S;
      const bool direct = kk < mx;
      const int kx = direct ? kk : N - kk;
      C2<T> z = mk<T>((T)0, (T)0);
      if (direct || kx < mx) {
        C2<T> A = ra[i], B = rb[i];
        if (dkx != 0.0) {
          const T k = (T)(dkx * (double)kx);
          A = mk<T>(-(k * A.y), k * A.x);
          B = mk<T>(-(k * B.y), k * B.x);
        }
        if (kx == 0 || 2 * kx == N) { A.y = (T)0; B.y = (T)0; }
        z = direct ? mk<T>(A.x - B.y, A.y + B.x) : mk<T>(A.x + B.y, B.x - A.y);
      }
      L[s * (N + 1) + kk] = z;
    }
  }
  __syncthreads();
  C2<T> v[N1];
  const int n2 = t % N2, sa = t / N2;
  if (t < S * N2) {
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = L[sa * (N + 1) + N2 * n1 + n2];
  }
  __syncthreads();
  if (t < S * N2) level_a_store<T, N1, N2, +1>(v, L, sa, n2, tw);
  __syncthreads();
  if (t < S * N1) level_b_load<T, N1, N2, +1>(w, L, t / N1, t % N1);
}

// forward x transform of the row pairs held as in xback_regs, truncated to
// kx < mx and unpacked into the x-pencil plane out
template <typename T, int N1, int N2, int NT>
__device__ __forceinline__ void xfwd_regs(C2<T>* w, C2<T>* __restrict__ out, int ny, int mx, int y0, C2<T>* L,
                                          const C2<T>* __restrict__ tw) {
  constexpr int N = N1 * N2;
  constexpr int S = xf_seqs<N1, N2, NT>();
  const int t = threadIdx.x;
  __syncthreads();
  if (t < S * N1) level_a_store<T, N2, N1, -1>(w, L, t / N1, t % N1, tw);
  __syncthreads();
  C2<T> q[N1];
  const int k1b = t % N2, sb = t / N2;
  if (t < S * N2) level_b_load<T, N2, N1, -1>(q, L, sb, k1b);
  __syncthreads();
  if (t < S * N2) {
#pragma unroll
    for (int k2 = 0; k2 < N1; ++k2) L[sb * (N + 1) + k1b + N2 * k2] = q[k2];
  }
  __syncthreads();
  const int nr = 2 * S;
  for (int idx = t; idx < mx * nr; idx += NT) {
    const int kk = idx / nr, r = idx - kk * nr;
    const int y = y0 + r;
    if (y >= ny) continue;
    const int s = r >> 1;
    const C2<T> zk = L[s * (N + 1) + kk];
    const C2<T> zn = L[s * (N + 1) + (kk == 0 ? 0 : N - kk)];
    C2<T> val;
    if ((r & 1) == 0) val = mk<T>((T)0.5 * (zk.x + zn.x), (T)0.5 * (zk.y - zn.y));
    else val = mk<T>((T)0.5 * (zk.y + zn.y), (T)0.5 * (zn.x - zk.x));
    out[(long long)kk * ny + y] = val;
  }
}

// packed physical rows (ya, ya+1) of one component at x = k1 + N1*k2
template <typename T, int N1, int N2>
__device__ __forceinline__ void load_pair(C2<T>* u, const T* __restrict__ base, int N, int ny, int ya, int k1,
                                          T shift) {
  // rows past the last are read at the last row (in bounds) and zeroed: no
  // branch between the loads, which are then all in flight together
  const T* ra = base + (long long)(ya < ny ? ya : ny - 1) * N + k1;
  const T* rb = base + (long long)(ya + 1 < ny ? ya + 1 : ny - 1) * N + k1;
  T la[N2], lb[N2];
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) {
    la[k2] = ra[N1 * k2];
    lb[k2] = rb[N1 * k2];
  }
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2)
    u[k2] = mk<T>(ya < ny ? la[k2] + shift : (T)0, ya + 1 < ny ? lb[k2] + shift : (T)0);
}

// FP32 at lengths up to 256: three waves per SIMD (168 VGPRs; the conv and
// div forms at 256 spill 12 / 48 bytes per lane) instead of two
template <typename T, int N1, int N2, int NT, int MODE, bool G>
__global__ __launch_bounds__(NT, (sizeof(T) == 4 && NT == 256) ? (N1 * N2 <= 256 ? 3 : 2) : 1) void k_xfused2(XfArgs a0,
                                                                                            GroupP grp) {
  Blk bk;
  const XfArgs& a = member<G>(a0, grp, bk);
  constexpr int N = N1 * N2;
  constexpr int S = xf_seqs<N1, N2, NT>();
  __shared__ __align__(16) char raw[sizeof(C2<T>) * xf_lds<N1, N2, NT>()];
  C2<T>* L = reinterpret_cast<C2<T>*>(raw);
  const int z = bk.y;
  const int y0 = bk.x * 2 * S;
  const long long plane = (long long)a.mx * a.ny;
  const long long pstride = (long long)a.nz * plane;
  const C2<T>* gin = (const C2<T>*)a.gin;
  C2<T>* gout = (C2<T>*)a.gout;
  // the twiddles in LDS: a global twiddle load issued after this workgroup's
  // output stores would wait for those stores (one in-order memory counter),
  // which serialised the transforms of the fused kernel on store latency
  __shared__ C2<T> tw[N];
  for (int i = threadIdx.x; i < N; i += NT) tw[i] = ((const C2<T>*)a.tw)[i];
  __syncthreads();
  const T* Ub = (const T*)a.U + (long long)z * a.ny * N;
  const T shifts[3] = {(T)a.su, (T)a.sv, (T)a.sw};
  const int t = threadIdx.x;
  const bool act = t < S * N1;
  const int k1 = t % N1, s = t / N1;
  const int ya = y0 + 2 * s;
  if (MODE == 0) {
    // adv_i = sum_j (U_j + shift_j) g_ij for the output i = the grid's z
    const int i = bk.z;
    C2<T> acc[N2];
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) acc[k2] = mk<T>((T)0, (T)0);
    for (int j = 0; j < 3; ++j) {
      // g_ij from g (9 components, or 3 for the scalar), or with ux set:
      // j = 0 from ux_i times i kx, j = 1, 2 from g's components i, 3 + i
      const C2<T>* src = a.ux ? (j == 0 ? (const C2<T>*)a.ux + i * pstride : gin + (j == 1 ? i : 3 + i) * pstride)
                              : gin + ((a.nout == 3) ? 3 * i + j : j) * pstride;
      C2<T> w[N2];
      xback_regs<T, N1, N2, NT>(src + z * plane, a.ny, a.mx, y0, L, tw, w, (a.ux && j == 0) ? a.dkx : 0.0);
      if (act) {
        if (a.bot0 && z == 0 && j == 2 && i < 2) {   // physical bottom plane of du/dz, dv/dz
          T* bot = (T*)(i == 0 ? a.bot0 : a.bot1) + (long long)ya * N + k1;
#pragma unroll
          for (int k2 = 0; k2 < N2; ++k2) {
            if (ya < a.ny) bot[N1 * k2] = w[k2].x;
            if (ya + 1 < a.ny) bot[N + N1 * k2] = w[k2].y;
          }
        }
        C2<T> u[N2];
        load_pair<T, N1, N2>(u, Ub + j * a.R, N, a.ny, ya, k1, shifts[j]);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
          acc[k2].x = acc[k2].x + u[k2].x * w[k2].x;
          acc[k2].y = acc[k2].y + u[k2].y * w[k2].y;
        }
      }
    }
    xfwd_regs<T, N1, N2, NT>(acc, gout + i * pstride + z * plane, a.ny, a.mx, y0, L, tw);
  } else if (MODE == 1) {
    // uu, uv, uw, vv, vw, ww
    C2<T> u[3][N2];
    if (act)
#pragma unroll
      for (int j = 0; j < 3; ++j) load_pair<T, N1, N2>(u[j], Ub + j * a.R, N, a.ny, ya, k1, (T)0);
    constexpr int pa[6] = {0, 0, 0, 1, 1, 2};
    constexpr int pb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int o = 0; o < 6; ++o) {
      C2<T> pr[N2];
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const C2<T> ua = u[pa[o]][k2], ub = u[pb[o]][k2];
        pr[k2] = mk<T>(ua.x * ub.x, ua.y * ub.y);
      }
      xfwd_regs<T, N1, N2, NT>(pr, gout + o * pstride + z * plane, a.ny, a.mx, y0, L, tw);
    }
  } else {
    // c (U_j + shift_j), j = 0..2
    C2<T> cv[N2];
    xback_regs<T, N1, N2, NT>(gin + z * plane, a.ny, a.mx, y0, L, tw, cv);
    for (int j = 0; j < 3; ++j) {
      C2<T> pr[N2];
      if (act) {
        C2<T> u[N2];
        load_pair<T, N1, N2>(u, Ub + j * a.R, N, a.ny, ya, k1, shifts[j]);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) pr[k2] = mk<T>(cv[k2].x * u[k2].x, cv[k2].y * u[k2].y);
      }
      xfwd_regs<T, N1, N2, NT>(pr, gout + j * pstride + z * plane, a.ny, a.mx, y0, L, tw);
    }
  }
}

template <typename T> struct XfNt;
template <> struct XfNt<double> { static constexpr int v = 128; };
template <> struct XfNt<float> { static constexpr int v = 256; };

#define TBD_XF_LENGTHS(F)                                                                               \
  F(384, 24, 16) F(256, 16, 16) F(192, 16, 12) F(128, 16, 8) F(96, 12, 8) F(64, 8, 8) F(48, 8, 6)    \
  F(32, 8, 4) F(24, 6, 4) F(16, 4, 4) F(12, 4, 3) F(8, 4, 2) F(6, 3, 2) F(4, 4, 1) F(3, 3, 1) F(2, 2, 1)

template <typename T, int MODE>
static bool launch_xf2(const FftPlan& p, const XfArgs& a, hipStream_t st) {
  constexpr int NT = XfNt<T>::v;
  switch (p.d.nx) {
#define TBD_CASE(NN, A, B)                                                                    \
  case NN: {                                                                                  \
    constexpr int S = xf_seqs<A, B, NT>();                                                    \
    dim3 grid(ceil_div(p.d.ny, 2 * S), p.d.nz, MODE == 0 ? a.nout : 1);                       \
    glaunch(k_xfused2<T, A, B, NT, MODE, false>, k_xfused2<T, A, B, NT, MODE, true>, grid,    \
            dim3(NT), 0, st, a);                                                             \
    return true;                                                                              \
  }
    TBD_XF_LENGTHS(TBD_CASE)
#undef TBD_CASE
    default:
      return false;
  }
}

template <typename T> static const void* twx_of(const FftPlan& p);
template <> const void* twx_of<double>(const FftPlan& p) { return p.twx64; }
template <> const void* twx_of<float>(const FftPlan& p) { return p.twx32; }

template <typename T>
static XfArgs make_args(const FftPlan& p, long long R) {
  XfArgs a;
  a.L = p.lx;
  a.ny = p.d.ny;
  a.mx = p.d.mx;
  a.nz = p.d.nz;
  a.R = R;
  a.tw = twx_of<T>(p);
  const int n = p.d.nx;
  int nseq = (kPPT * kNT) / n;           // positions per thread <= kPPT
  nseq = std::min(nseq, 2 * kFftB * kNT / n);
  nseq = std::min(nseq, (p.d.ny + 1) / 2);
  a.nseq = std::max(1, std::min(nseq, 16));
  TBD_CHECK(a.nseq * n <= kPPT * kNT && a.nseq * n / 2 <= kFftB * kNT, "xfused > row length too large");
  return a;
}

template <typename T, int MODE>
static void launch_xf(const FftPlan& p, const XfArgs& a, hipStream_t st) {
  if (p.two_level && launch_xf2<T, MODE>(p, a, st)) {
    TBD_HIP(hipGetLastError());
    return;
  }
  TBD_CHECK(a.ux == nullptr, "xfused_conv > x derivatives from the x-pencils need the register transform length");
  const int n = p.d.nx;
  const size_t lds = sizeof(C2<T>) * ((size_t)a.nseq * lds_ld(n) + n);
  dim3 grid(ceil_div(p.d.ny, 2 * a.nseq), p.d.nz);
  hipLaunchKernelGGL((k_xfused<T, MODE>), grid, dim3(kNT), lds, st, a);
  TBD_HIP(hipGetLastError());
}

bool xfused_registers(const FftPlan& p) {
  if (!p.two_level) return false;
  switch (p.d.nx) {
#define TBD_CASE(NN, A, B) case NN:
    TBD_XF_LENGTHS(TBD_CASE)
#undef TBD_CASE
    return true;
    default:
      return false;
  }
}

template <typename T>
void xfused_conv(const FftPlan& p, const C2<T>* g, const T* U, long long R, T su, T sv, T sw, int nout, C2<T>* adv,
                 T* bot0, T* bot1, hipStream_t st, const C2<T>* ux, double dkx) {
  TBD_CHECK(nout == 1 || nout == 3, "xfused_conv > nout must be 1 or 3");
  TBD_CHECK(ux == nullptr || (nout == 3 && dkx != 0.0), "xfused_conv > ux needs nout = 3 and dkx");
  XfArgs a = make_args<T>(p, R);
  a.gin = g; a.U = U; a.gout = adv; a.nout = nout;
  a.ux = ux; a.dkx = dkx;
  a.su = su; a.sv = sv; a.sw = sw;
  a.bot0 = (nout == 3) ? bot0 : nullptr;
  a.bot1 = (nout == 3) ? bot1 : nullptr;
  if (a.bot0 && !a.bot1) a.bot0 = nullptr;
  launch_xf<T, 0>(p, a, st);
}

template <typename T>
void xfused_div6(const FftPlan& p, const T* U, long long R, C2<T>* prod, hipStream_t st) {
  XfArgs a = make_args<T>(p, R);
  a.U = U; a.gout = prod; a.nout = 6;
  launch_xf<T, 1>(p, a, st);
}

template <typename T>
void xfused_cdiv(const FftPlan& p, const C2<T>* c, const T* U, long long R, T su, T sv, T sw, C2<T>* prod,
                 hipStream_t st) {
  XfArgs a = make_args<T>(p, R);
  a.gin = c; a.U = U; a.gout = prod; a.nout = 3;
  a.su = su; a.sv = sv; a.sw = sw;
  launch_xf<T, 2>(p, a, st);
}

template void xfused_conv<double>(const FftPlan&, const double2*, const double*, long long, double, double, double,
                                  int, double2*, double*, double*, hipStream_t, const double2*, double);
template void xfused_conv<float>(const FftPlan&, const float2*, const float*, long long, float, float, float, int,
                                 float2*, float*, float*, hipStream_t, const float2*, double);
template void xfused_div6<double>(const FftPlan&, const double*, long long, double2*, hipStream_t);
template void xfused_div6<float>(const FftPlan&, const float*, long long, float2*, hipStream_t);
template void xfused_cdiv<double>(const FftPlan&, const double2*, const double*, long long, double, double, double,
                                  double2*, hipStream_t);
template void xfused_cdiv<float>(const FftPlan&, const float2*, const float*, long long, float, float, float, float2*,
                                 hipStream_t);

}  // namespace tbd
