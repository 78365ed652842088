// LDS-staged Stockham transforms for the plane DFTs (see fft.h).
//
// A workgroup owns a batch of equal-length sequences in LDS.  Each stage of
// radix R in {4, 2, 3} reads all butterfly inputs into registers, barriers,
// and writes the outputs back in self-sorting (Stockham) order, so a single
// LDS buffer suffices.  Truncation, zero padding, the real<->complex packing
// of two rows into one complex sequence, the 1/(Nx*Ny) scale and the
// transposes to and from the z-fastest spectral layout are fused into the
// global loads/stores of the passes.
#include "fft.h"
#include "fft_device.h"
#include "group.h"
#include "regfft.h"

#include <cmath>
#include <cstdlib>
#include <vector>

namespace tbd {

FftLen make_fftlen(int n) {
  FftLen L;
  L.n = n;
  int m = n;
  while (m % 4 == 0 && L.nstage < 12) { L.radix[L.nstage++] = 4; m /= 4; }
  while (m % 2 == 0 && L.nstage < 12) { L.radix[L.nstage++] = 2; m /= 2; }
  while (m % 3 == 0 && L.nstage < 12) { L.radix[L.nstage++] = 3; m /= 3; }
  TBD_CHECK(m == 1, "FFT length " + std::to_string(n) + " must be of the form 2^a 3^b");
  return L;
}

void FftPlan::create(const Dims& dims) {
  d = dims;
  const char* stockham = std::getenv("TBD_FFT_STOCKHAM");
  two_level = !(stockham && stockham[0] == '1');
  lx = make_fftlen(d.nx);
  ly = make_fftlen(d.ny);
  auto mk_tw = [](int n, double2** t64, float2** t32) {
    std::vector<double2> a(n);
    std::vector<float2> b(n);
    for (int t = 0; t < n; ++t) {
      const double ang = -2.0 * M_PI * (double)t / (double)n;
      a[t].x = std::cos(ang);
      a[t].y = std::sin(ang);
      b[t].x = (float)a[t].x;
      b[t].y = (float)a[t].y;
    }
    TBD_HIP(hipMalloc(t64, sizeof(double2) * n));
    TBD_HIP(hipMalloc(t32, sizeof(float2) * n));
    TBD_HIP(hipMemcpy(*t64, a.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    TBD_HIP(hipMemcpy(*t32, b.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  };
  mk_tw(d.nx, &twx64, &twx32);
  mk_tw(d.ny, &twy64, &twy32);
}

void FftPlan::destroy() {
  if (twx64) (void)hipFree(twx64);
  if (twy64) (void)hipFree(twy64);
  if (twx32) (void)hipFree(twx32);
  if (twy32) (void)hipFree(twy32);
  twx64 = twy64 = nullptr;
  twx32 = twy32 = nullptr;
}

template <typename T> __host__ __device__ inline const C2<T>* twx(const FftPlan& p);
template <> __host__ __device__ inline const double2* twx<double>(const FftPlan& p) { return p.twx64; }
template <> __host__ __device__ inline const float2* twx<float>(const FftPlan& p) { return p.twx32; }
template <typename T> __host__ __device__ inline const C2<T>* twy(const FftPlan& p);
template <> __host__ __device__ inline const double2* twy<double>(const FftPlan& p) { return p.twy64; }
template <> __host__ __device__ inline const float2* twy<float>(const FftPlan& p) { return p.twy32; }

// ---------------------------------------------------------------------------
// x pass, forward: rows (y0..y0+2*nseq) of plane pl -> T[pl][kx][y]
// ---------------------------------------------------------------------------
template <typename T, int NT, int MAXB>
__global__ __launch_bounds__(NT) void k_xpass_fwd(const T* __restrict__ phys, long long pstride,
                                                  C2<T>* __restrict__ tmp, FftLen L, const C2<T>* twg,
                                                  int ny, int mx, int nplanes, int nseq, const int* gate) {
  if (gate && *gate == 0) return;
  extern __shared__ char smem_raw[];
  C2<T>* buf = reinterpret_cast<C2<T>*>(smem_raw);
  const int N = L.n;
  const int ldn = lds_ld(N);
  C2<T>* tw = buf + nseq * ldn;
  const int pl = blockIdx.y;                 // d*nplanes + zl
  const int d = pl / nplanes, zl = pl - d * nplanes;
  const int y0 = blockIdx.x * 2 * nseq;
  const T* base = phys + (long long)d * pstride + (long long)zl * ny * N;
  load_tw<T, NT>(tw, twg, N);
  for (int idx = threadIdx.x; idx < nseq * N; idx += NT) {
    const int s = idx / N, x = idx - s * N;
    const int y = y0 + 2 * s;
    const T a = (y < ny) ? base[(long long)y * N + x] : (T)0;
    const T b = (y + 1 < ny) ? base[(long long)(y + 1) * N + x] : (T)0;
    buf[s * ldn + x] = mk<T>(a, b);
  }
  __syncthreads();
  lds_fft<T, NT, MAXB>(buf, ldn, nseq, L, tw, -1);
  const int nr = 2 * nseq;
  C2<T>* out = tmp + (long long)pl * mx * ny;
  for (int idx = threadIdx.x; idx < mx * nr; idx += NT) {
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

// ---------------------------------------------------------------------------
// x pass, backward: T[pl][kx][y] (kx < mx) -> rows, halfcomplex c2r
// ---------------------------------------------------------------------------
template <typename T, int NT, int MAXB>
__global__ __launch_bounds__(NT) void k_xpass_bwd(const C2<T>* __restrict__ tmp, T* __restrict__ phys,
                                                  long long pstride, FftLen L, const C2<T>* twg,
                                                  int ny, int mx, int nplanes, int nseq, const int* gate) {
  if (gate && *gate == 0) return;
  extern __shared__ char smem_raw[];
  C2<T>* buf = reinterpret_cast<C2<T>*>(smem_raw);
  const int N = L.n;
  const int ldn = lds_ld(N);
  C2<T>* tw = buf + nseq * ldn;
  const int pl = blockIdx.y;
  const int d = pl / nplanes, zl = pl - d * nplanes;
  const int y0 = blockIdx.x * 2 * nseq;
  load_tw<T, NT>(tw, twg, N);
  for (int idx = threadIdx.x; idx < nseq * N; idx += NT) {
    const int s = idx / N, x = idx - s * N;
    buf[s * ldn + x] = mk<T>((T)0, (T)0);
  }
  __syncthreads();
  const C2<T>* in = tmp + (long long)pl * mx * ny;
  for (int idx = threadIdx.x; idx < mx * nseq; idx += NT) {
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
  lds_fft<T, NT, MAXB>(buf, ldn, nseq, L, tw, +1);
  T* base = phys + (long long)d * pstride + (long long)zl * ny * N;
  for (int idx = threadIdx.x; idx < nseq * N; idx += NT) {
    const int s = idx / N, x = idx - s * N;
    const int y = y0 + 2 * s;
    const C2<T> z = buf[s * ldn + x];
    if (y < ny) base[(long long)y * N + x] = z.x;
    if (y + 1 < ny) base[(long long)(y + 1) * N + x] = z.y;
  }
}

// ---------------------------------------------------------------------------
// y pass, forward: T[d][z][kx][y] -> spec (z fastest), keep nyc, scale
// grid: (z tiles, mx, nv)
// ---------------------------------------------------------------------------
template <typename T, typename SpecT, int NT, int MAXB>
__global__ __launch_bounds__(NT) void k_ypass_fwd(const C2<T>* __restrict__ tmp, SpecT* __restrict__ spec,
                                                  long long sstride, FftLen L, const C2<T>* twg, int nz,
                                                  int mx, int nyc, int nplanes, int zoff, int zt, T scale,
                                                  const int* gate) {
  if (gate && *gate == 0) return;
  extern __shared__ char smem_raw[];
  C2<T>* buf = reinterpret_cast<C2<T>*>(smem_raw);
  const int N = L.n;
  const int ldn = lds_ld(N);
  C2<T>* tw = buf + zt * ldn;
  const int z0 = blockIdx.x * zt;
  const int k = blockIdx.y;
  const int d = blockIdx.z;
  const int nseq = min(zt, nplanes - z0);
  load_tw<T, NT>(tw, twg, N);
  for (int idx = threadIdx.x; idx < nseq * N; idx += NT) {
    const int j = idx / N, y = idx - j * N;
    buf[j * ldn + y] = tmp[(((long long)d * nplanes + z0 + j) * mx + k) * N + y];
  }
  __syncthreads();
  lds_fft<T, NT, MAXB>(buf, ldn, nseq, L, tw, -1);
  const int half = nyc / 2;
  SpecT* out = spec + (long long)d * sstride + (long long)k * nyc * nz + zoff + z0;
  for (int idx = threadIdx.x; idx < nyc * zt; idx += NT) {
    const int iy = idx / zt, j = idx - iy * zt;
    if (j >= nseq) continue;
    const int yf = iy < half ? iy : iy + N - nyc;
    const C2<T> v = buf[j * ldn + yf];
    SpecT o;
    o.x = v.x * scale;
    o.y = v.y * scale;
    out[(long long)iy * nz + j] = o;
  }
}

// y pass, backward: spec (z fastest) -> T[d][z][kx][y], zero padded
template <typename T, typename SpecT, int NT, int MAXB>
__global__ __launch_bounds__(NT) void k_ypass_bwd(const SpecT* __restrict__ spec, long long sstride,
                                                  C2<T>* __restrict__ tmp, FftLen L, const C2<T>* twg,
                                                  int nz, int mx, int nyc, int nplanes, int zoff, int zt,
                                                  const int* gate) {
  if (gate && *gate == 0) return;
  extern __shared__ char smem_raw[];
  C2<T>* buf = reinterpret_cast<C2<T>*>(smem_raw);
  const int N = L.n;
  const int ldn = lds_ld(N);
  C2<T>* tw = buf + zt * ldn;
  const int z0 = blockIdx.x * zt;
  const int k = blockIdx.y;
  const int d = blockIdx.z;
  const int nseq = min(zt, nplanes - z0);
  load_tw<T, NT>(tw, twg, N);
  for (int idx = threadIdx.x; idx < nseq * N; idx += NT) {
    const int j = idx / N, y = idx - j * N;
    buf[j * ldn + y] = mk<T>((T)0, (T)0);
  }
  __syncthreads();
  const int half = nyc / 2;
  const SpecT* in = spec + (long long)d * sstride + (long long)k * nyc * nz + zoff + z0;
  for (int idx = threadIdx.x; idx < nyc * zt; idx += NT) {
    const int iy = idx / zt, j = idx - iy * zt;
    if (j >= nseq) continue;
    const int yf = iy < half ? iy : iy + N - nyc;
    const SpecT v = in[(long long)iy * nz + j];
    buf[j * ldn + yf] = mk<T>((T)v.x, (T)v.y);
  }
  __syncthreads();
  lds_fft<T, NT, MAXB>(buf, ldn, nseq, L, tw, +1);
  for (int idx = threadIdx.x; idx < nseq * N; idx += NT) {
    const int j = idx / N, y = idx - j * N;
    tmp[(((long long)d * nplanes + z0 + j) * mx + k) * N + y] = buf[j * ldn + y];
  }
}

// ---------------------------------------------------------------------------
// Two-level register transforms (regfft.h) for lengths N = N1*N2 with both
// factors dividing 48.  A workgroup of NT threads handles S sequences; every
// level-A task holds one DFT of size N1 in registers, every level-B task one
// DFT of size N2, and the data crosses LDS once between them.  Global loads
// and stores are arranged so that consecutive lanes touch consecutive
// addresses: y for the x-pencil side, z for the spectral side, x for rows.
// ---------------------------------------------------------------------------
template <int N1, int N2, int NT>
__host__ __device__ constexpr int seqs2() {
  return (NT / (N1 > N2 ? N1 : N2)) > 0 ? NT / (N1 > N2 ? N1 : N2) : 1;
}
template <typename T, int N1, int N2, int NT>
__host__ __device__ constexpr int lds2_elems(bool with_rows) {
  return (seqs2<N1, N2, NT>() * xseq<N1, N2>() > (with_rows ? seqs2<N1, N2, NT>() * (N1 * N2 + 1) : 0))
             ? seqs2<N1, N2, NT>() * xseq<N1, N2>()
             : seqs2<N1, N2, NT>() * (N1 * N2 + 1);
}

// y pass, forward: T[d][z][kx][y] -> spectral (z fastest), keep nyc, scale
template <typename T, typename SpecT>
struct YfA {
  const C2<T>* tmp;
  SpecT* spec;
  long long sstride;
  const C2<T>* tw;
  int scol, mx, nyc, nplanes, zoff;
  T scale;
  const int* gate;
};
template <typename T, typename SpecT, int N1, int N2, int NT, bool G>
__global__ __launch_bounds__(NT) void k_ypf2(YfA<T, SpecT> a0, GroupP grp) {
  Blk bk;
  const YfA<T, SpecT>& A = member<G>(a0, grp, bk);
  if (A.gate && *A.gate == 0) return;
  const C2<T>* __restrict__ tmp = A.tmp;
  SpecT* __restrict__ spec = A.spec;
  const long long sstride = A.sstride;
  const C2<T>* __restrict__ tw = A.tw;
  const int scol = A.scol, mx = A.mx, nyc = A.nyc, nplanes = A.nplanes, zoff = A.zoff;
  const T scale = A.scale;
  constexpr int N = N1 * N2;
  constexpr int S = seqs2<N1, N2, NT>();
  // level-A -> level-B exchange through one real plane: real parts, then
  // imaginary parts (half the LDS of a complex exchange, twice the occupancy)
  __shared__ __align__(16) char raw[sizeof(T) * lds2_elems<T, N1, N2, NT>(false)];
  T* E = reinterpret_cast<T*>(raw);
  const int z0 = bk.x * S, k = bk.y, d = bk.z;
  const int ns = min(S, nplanes - z0);
  const int t = threadIdx.x;
  const int n2 = t % N2, sa = t / N2;
  const bool act_a = t < S * N2 && sa < ns;
  const int sb = t % S, k1 = t / S;
  const bool act_b = t < S * N1 && sb < ns;
  constexpr int SW = (N1 == 16 && N2 == 16) ? 1 : 0;
  C2<T> v[N1];
  if (act_a) {
    const C2<T>* src = tmp + (((long long)d * nplanes + z0 + sa) * mx + k) * N + n2;
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = src[N2 * n1];
    level_a_regs<T, N1, N2, -1>(v, n2, tw);
    level_a_put<T, N1, N2, SW>(v, E, sa, n2, 0);
  }
  __syncthreads();
  C2<T> w[N2];
  if (act_b) level_b_get<T, N1, N2, SW>(w, E, sb, k1, 0);
  __syncthreads();
  if (act_a) level_a_put<T, N1, N2, SW>(v, E, sa, n2, 1);
  __syncthreads();
  if (act_b) {
    level_b_get<T, N1, N2, SW>(w, E, sb, k1, 1);
    level_b_dft<T, N2, -1>(w);
    const int half = nyc / 2;
    SpecT* out = spec + (long long)d * sstride + (long long)k * nyc * scol + zoff + z0 + sb;
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
      const int y = k1 + N1 * k2;
      const int iy = y < half ? y : (y >= N - half ? y - N + nyc : -1);
      if (iy >= 0) {
        SpecT o;
        o.x = w[k2].x * scale;
        o.y = w[k2].y * scale;
        out[(long long)iy * scol] = o;
      }
    }
  }
}

// y pass, backward: spectral (z fastest) -> T[d][z][kx][y], zero padded
template <typename T, typename SpecT>
struct YbA {
  const SpecT* spec;
  long long sstride;
  C2<T>* tmp;
  const C2<T>* tw;
  int scol, mx, nyc, nplanes, zoff;
  const int* gate;
};
template <typename T, typename SpecT, int N1, int N2, int NT, bool G>
__global__ __launch_bounds__(NT) void k_ypb2(YbA<T, SpecT> a0, GroupP grp) {
  Blk bk;
  const YbA<T, SpecT>& A = member<G>(a0, grp, bk);
  if (A.gate && *A.gate == 0) return;
  const SpecT* __restrict__ spec = A.spec;
  const long long sstride = A.sstride;
  C2<T>* __restrict__ tmp = A.tmp;
  const C2<T>* __restrict__ tw = A.tw;
  const int scol = A.scol, mx = A.mx, nyc = A.nyc, nplanes = A.nplanes, zoff = A.zoff;
  constexpr int N = N1 * N2;
  constexpr int S = seqs2<N1, N2, NT>();
  // exchange through one real plane, real then imaginary parts (see k_ypf2)
  __shared__ __align__(16) char raw[sizeof(T) * lds2_elems<T, N1, N2, NT>(false)];
  T* E = reinterpret_cast<T*>(raw);
  const int z0 = bk.x * S, k = bk.y, d = bk.z;
  const int ns = min(S, nplanes - z0);
  const int t = threadIdx.x;
  const int sa = t % S, n2 = t / S;
  const bool act_a = t < S * N2 && sa < ns;
  const int k1 = t % N1, sb = t / N1;
  const bool act_b = t < S * N1 && sb < ns;
  constexpr int SW = (N1 == 16 && N2 == 16) ? 2 : 0;
  C2<T> v[N1];
  if (act_a) {
    const int half = nyc / 2;
    const SpecT* in = spec + (long long)d * sstride + (long long)k * nyc * scol + zoff + z0 + sa;
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
      const int y = N2 * n1 + n2;
      const int iy = y < half ? y : (y >= N - half ? y - N + nyc : -1);
      if (iy >= 0) {
        const SpecT u = in[(long long)iy * scol];
        v[n1] = mk<T>((T)u.x, (T)u.y);
      } else {
        v[n1] = mk<T>((T)0, (T)0);
      }
    }
    level_a_regs<T, N1, N2, +1>(v, n2, tw);
    level_a_put<T, N1, N2, SW>(v, E, sa, n2, 0);
  }
  __syncthreads();
  C2<T> w[N2];
  if (act_b) level_b_get<T, N1, N2, SW>(w, E, sb, k1, 0);
  __syncthreads();
  if (act_a) level_a_put<T, N1, N2, SW>(v, E, sa, n2, 1);
  __syncthreads();
  if (act_b) {
    level_b_get<T, N1, N2, SW>(w, E, sb, k1, 1);
    level_b_dft<T, N2, +1>(w);
    C2<T>* out = tmp + (((long long)d * nplanes + z0 + sb) * mx + k) * N + k1;
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) out[N1 * k2] = w[k2];
  }
}

// x pass, forward: rows (y0 .. y0+2S) of plane pl -> T[pl][kx][y]
template <typename T>
struct XfA {
  const T* phys;
  long long pstride;
  C2<T>* tmp;
  const C2<T>* tw;
  int ny, mx, nplanes;
  const int* gate;
};
template <typename T, int N1, int N2, int NT, bool G>
__global__ __launch_bounds__(NT) void k_xpf2(XfA<T> a0, GroupP grp) {
  Blk bk;
  const XfA<T>& A = member<G>(a0, grp, bk);
  if (A.gate && *A.gate == 0) return;
  const T* __restrict__ phys = A.phys;
  const long long pstride = A.pstride;
  C2<T>* __restrict__ tmp = A.tmp;
  const C2<T>* __restrict__ tw = A.tw;
  const int ny = A.ny, mx = A.mx, nplanes = A.nplanes;
  constexpr int N = N1 * N2;
  constexpr int S = seqs2<N1, N2, NT>();
  __shared__ __align__(16) char raw[sizeof(C2<T>) * lds2_elems<T, N1, N2, NT>(true)];
  C2<T>* E = reinterpret_cast<C2<T>*>(raw);
  C2<T>* Z = E;   // natural-order spectra after level B (aliases E)
  const int pl = bk.y;
  const int d = pl / nplanes, zl = pl - d * nplanes;
  const int y0 = bk.x * 2 * S;
  const T* base = phys + (long long)d * pstride + (long long)zl * ny * N;
  const int t = threadIdx.x;
  if (t < S * N2) {
    const int n2 = t % N2, s = t / N2;
    const int ya = y0 + 2 * s;
    const T* ra = base + (long long)ya * N + n2;
    C2<T> v[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
      const T a = (ya < ny) ? ra[N2 * n1] : (T)0;
      const T b = (ya + 1 < ny) ? ra[N + N2 * n1] : (T)0;
      v[n1] = mk<T>(a, b);
    }
    level_a_store<T, N1, N2, -1>(v, E, s, n2, tw);
  }
  __syncthreads();
  C2<T> w[N2];
  const int k1 = t % N1, sb = t / N1;
  if (t < S * N1) level_b_load<T, N1, N2, -1>(w, E, sb, k1);
  __syncthreads();
  if (t < S * N1) {
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) Z[sb * (N + 1) + k1 + N1 * k2] = w[k2];
  }
  __syncthreads();
  const int nr = 2 * S;
  C2<T>* out = tmp + (long long)pl * mx * ny;
  for (int idx = t; idx < mx * nr; idx += NT) {
    const int kk = idx / nr, r = idx - kk * nr;
    const int y = y0 + r;
    if (y >= ny) continue;
    const int s = r >> 1;
    const C2<T> zk = Z[s * (N + 1) + kk];
    const C2<T> zn = Z[s * (N + 1) + (kk == 0 ? 0 : N - kk)];
    C2<T> val;
    if ((r & 1) == 0) val = mk<T>((T)0.5 * (zk.x + zn.x), (T)0.5 * (zk.y - zn.y));
    else val = mk<T>((T)0.5 * (zk.y + zn.y), (T)0.5 * (zn.x - zk.x));
    out[(long long)kk * ny + y] = val;
  }
}

// x pass, backward: T[pl][kx][y] (kx < mx) -> rows, halfcomplex c2r
template <typename T>
struct XbA {
  const C2<T>* tmp;
  T* phys;
  long long pstride;
  const C2<T>* tw;
  int ny, mx, nplanes;
  const int* gate;
};
template <typename T, int N1, int N2, int NT, bool G>
__global__ __launch_bounds__(NT) void k_xpb2(XbA<T> a0, GroupP grp) {
  Blk bk;
  const XbA<T>& A = member<G>(a0, grp, bk);
  if (A.gate && *A.gate == 0) return;
  const C2<T>* __restrict__ tmp = A.tmp;
  T* __restrict__ phys = A.phys;
  const long long pstride = A.pstride;
  const C2<T>* __restrict__ tw = A.tw;
  const int ny = A.ny, mx = A.mx, nplanes = A.nplanes;
  constexpr int N = N1 * N2;
  constexpr int S = seqs2<N1, N2, NT>();
  __shared__ __align__(16) char raw[sizeof(C2<T>) * lds2_elems<T, N1, N2, NT>(true)];
  C2<T>* E = reinterpret_cast<C2<T>*>(raw);
  C2<T>* Z = E;   // packed two-row spectra (aliases E)
  const int pl = bk.y;
  const int d = pl / nplanes, zl = pl - d * nplanes;
  const int y0 = bk.x * 2 * S;
  const int t = threadIdx.x;
  const C2<T>* in = tmp + (long long)pl * mx * ny;
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
    ra[i] = mk<T>((T)0, (T)0);
    rb[i] = ra[i];
    if (idx < S * N && (direct || kx < mx)) {
      if (ya < ny) ra[i] = in[(long long)kx * ny + ya];
      if (ya + 1 < ny) rb[i] = in[(long long)kx * ny + ya + 1];
    }
  }
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
        if (kx == 0 || 2 * kx == N) { A.y = (T)0; B.y = (T)0; }
        z = direct ? mk<T>(A.x - B.y, A.y + B.x) : mk<T>(A.x + B.y, B.x - A.y);
      }
      Z[s * (N + 1) + kk] = z;
    }
  }
  __syncthreads();
  C2<T> v[N1];
  const int n2 = t % N2, sa = t / N2;
  if (t < S * N2) {
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = Z[sa * (N + 1) + N2 * n1 + n2];
  }
  __syncthreads();
  if (t < S * N2) level_a_store<T, N1, N2, +1>(v, E, sa, n2, tw);
  __syncthreads();
  if (t < S * N1) {
    const int k1 = t % N1, s = t / N1;
    C2<T> w[N2];
    level_b_load<T, N1, N2, +1>(w, E, s, k1);
    const int ya = y0 + 2 * s;
    T* ra = phys + (long long)d * pstride + ((long long)zl * ny + ya) * N + k1;
    if (ya < ny) {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) ra[N1 * k2] = w[k2].x;
    }
    if (ya + 1 < ny) {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) ra[N + N1 * k2] = w[k2].y;
    }
  }
}

template <typename T> struct Nt2;
template <> struct Nt2<double> { static constexpr int v = 128; };
template <> struct Nt2<float> { static constexpr int v = 256; };

// Supported two-level lengths; F(N, N1, N2) is expanded per length.
#define TBD_FOR_LENGTHS(F)                                                                              \
  F(384, 24, 16) F(256, 16, 16) F(192, 16, 12) F(128, 16, 8) F(96, 12, 8) F(64, 8, 8) F(48, 8, 6)    \
  F(32, 8, 4) F(24, 6, 4) F(16, 4, 4) F(12, 4, 3) F(8, 4, 2) F(6, 3, 2) F(4, 4, 1) F(3, 3, 1) F(2, 2, 1)
// The same lengths with the two factors in the other order (level A the
// smaller one), an autotuner choice per pass (FftPlan::swap_*)
#define TBD_FOR_SWAPPED(F)                                                                              \
  F(384, 16, 24) F(192, 12, 16) F(128, 8, 16) F(96, 8, 12) F(48, 6, 8) F(32, 4, 8) F(24, 4, 6)       \
  F(12, 3, 4) F(8, 2, 4) F(6, 2, 3)

bool fft_has_two_level(int n) {
  switch (n) {
#define TBD_CASE(NN, A, B) case NN: return true;
    TBD_FOR_LENGTHS(TBD_CASE)
#undef TBD_CASE
    default:
      return false;
  }
}

bool fft_has_swap(int n) {
  switch (n) {
#define TBD_CASE(NN, A, B) case NN: return true;
    TBD_FOR_SWAPPED(TBD_CASE)
#undef TBD_CASE
    default:
      return false;
  }
}

template <typename T, typename SpecT, int NT, int A, int B>
static void ypf2_launch(const FftPlan& p, const C2<T>* tmp, SpecT* spec, long long sstride, int nv, int zoff,
                        int nplanes, hipStream_t st, int sc, T scale, const int* gate) {
  constexpr int S = seqs2<A, B, NT>();
  dim3 grid(ceil_div(nplanes, S), p.d.mx, nv);
  glaunch(k_ypf2<T, SpecT, A, B, NT, false>, k_ypf2<T, SpecT, A, B, NT, true>, grid, dim3(NT), 0, st,
          YfA<T, SpecT>{tmp, spec, sstride, twy<T>(p), sc, p.d.mx, p.d.nyc, nplanes, zoff, scale, gate});
}
template <typename T, typename SpecT, int NT>
static bool ypf2_nt(const FftPlan& p, const C2<T>* tmp, SpecT* spec, long long sstride, int nv, int zoff, int nplanes,
                 hipStream_t st, int scol, const int* gate) {
  const T scale = (T)(1.0 / ((double)p.d.nx * (double)p.d.ny));
  const int sc = scol > 0 ? scol : p.d.nz;
#define TBD_CASE(NN, A, B) \
  case NN: ypf2_launch<T, SpecT, NT, A, B>(p, tmp, spec, sstride, nv, zoff, nplanes, st, sc, scale, gate); return true;
  if (p.swap_yf) switch (p.d.ny) { TBD_FOR_SWAPPED(TBD_CASE) default: break; }
  switch (p.d.ny) { TBD_FOR_LENGTHS(TBD_CASE) default: return false; }
#undef TBD_CASE
}

template <typename T, typename SpecT, int NT, int A, int B>
static void ypb2_launch(const FftPlan& p, const SpecT* spec, long long sstride, C2<T>* tmp, int nv, int zoff,
                        int nplanes, hipStream_t st, int sc, const int* gate) {
  constexpr int S = seqs2<A, B, NT>();
  dim3 grid(ceil_div(nplanes, S), p.d.mx, nv);
  glaunch(k_ypb2<T, SpecT, A, B, NT, false>, k_ypb2<T, SpecT, A, B, NT, true>, grid, dim3(NT), 0, st,
          YbA<T, SpecT>{spec, sstride, tmp, twy<T>(p), sc, p.d.mx, p.d.nyc, nplanes, zoff, gate});
}
template <typename T, typename SpecT, int NT>
static bool ypb2_nt(const FftPlan& p, const SpecT* spec, long long sstride, C2<T>* tmp, int nv, int zoff, int nplanes,
                 hipStream_t st, int scol, const int* gate) {
  const int sc = scol > 0 ? scol : p.d.nz;
#define TBD_CASE(NN, A, B) \
  case NN: ypb2_launch<T, SpecT, NT, A, B>(p, spec, sstride, tmp, nv, zoff, nplanes, st, sc, gate); return true;
  if (p.swap_yb) switch (p.d.ny) { TBD_FOR_SWAPPED(TBD_CASE) default: break; }
  switch (p.d.ny) { TBD_FOR_LENGTHS(TBD_CASE) default: return false; }
#undef TBD_CASE
}

template <typename T, int NT, int A, int B>
static void xpf2_launch(const FftPlan& p, const T* phys, long long pstride, C2<T>* tmp, int nv, int nplanes,
                        hipStream_t st, const int* gate) {
  constexpr int S = seqs2<A, B, NT>();
  dim3 grid(ceil_div(p.d.ny, 2 * S), nv * nplanes);
  glaunch(k_xpf2<T, A, B, NT, false>, k_xpf2<T, A, B, NT, true>, grid, dim3(NT), 0, st,
          XfA<T>{phys, pstride, tmp, twx<T>(p), p.d.ny, p.d.mx, nplanes, gate});
}
template <typename T, int NT>
static bool xpf2_nt(const FftPlan& p, const T* phys, long long pstride, C2<T>* tmp, int nv, int nplanes,
                 hipStream_t st, const int* gate) {
#define TBD_CASE(NN, A, B) \
  case NN: xpf2_launch<T, NT, A, B>(p, phys, pstride, tmp, nv, nplanes, st, gate); return true;
  if (p.swap_xf) switch (p.d.nx) { TBD_FOR_SWAPPED(TBD_CASE) default: break; }
  switch (p.d.nx) { TBD_FOR_LENGTHS(TBD_CASE) default: return false; }
#undef TBD_CASE
}

template <typename T, int NT, int A, int B>
static void xpb2_launch(const FftPlan& p, const C2<T>* tmp, T* phys, long long pstride, int nv, int nplanes,
                        hipStream_t st, const int* gate) {
  constexpr int S = seqs2<A, B, NT>();
  dim3 grid(ceil_div(p.d.ny, 2 * S), nv * nplanes);
  glaunch(k_xpb2<T, A, B, NT, false>, k_xpb2<T, A, B, NT, true>, grid, dim3(NT), 0, st,
          XbA<T>{tmp, phys, pstride, twx<T>(p), p.d.ny, p.d.mx, nplanes, gate});
}
template <typename T, int NT>
static bool xpb2_nt(const FftPlan& p, const C2<T>* tmp, T* phys, long long pstride, int nv, int nplanes,
                 hipStream_t st, const int* gate) {
#define TBD_CASE(NN, A, B) \
  case NN: xpb2_launch<T, NT, A, B>(p, tmp, phys, pstride, nv, nplanes, st, gate); return true;
  if (p.swap_xb) switch (p.d.nx) { TBD_FOR_SWAPPED(TBD_CASE) default: break; }
  switch (p.d.nx) { TBD_FOR_LENGTHS(TBD_CASE) default: return false; }
#undef TBD_CASE
}

// The two-level kernels at the default workgroup size (Nt2: 16 sequences of
// 256 per FP32 workgroup, 8 per FP64 one) or twice it (FftPlan::wide_*), a
// choice the autotuner makes per grid and pass.
template <typename T, typename SpecT>
static bool ypf2(const FftPlan& p, const C2<T>* tmp, SpecT* spec, long long sstride, int nv, int zoff, int nplanes,
                 hipStream_t st, int scol, const int* gate) {
  return p.wide_yf ? ypf2_nt<T, SpecT, 2 * Nt2<T>::v>(p, tmp, spec, sstride, nv, zoff, nplanes, st, scol, gate)
                : ypf2_nt<T, SpecT, Nt2<T>::v>(p, tmp, spec, sstride, nv, zoff, nplanes, st, scol, gate);
}
template <typename T, typename SpecT>
static bool ypb2(const FftPlan& p, const SpecT* spec, long long sstride, C2<T>* tmp, int nv, int zoff, int nplanes,
                 hipStream_t st, int scol, const int* gate) {
  return p.wide_yb ? ypb2_nt<T, SpecT, 2 * Nt2<T>::v>(p, spec, sstride, tmp, nv, zoff, nplanes, st, scol, gate)
                : ypb2_nt<T, SpecT, Nt2<T>::v>(p, spec, sstride, tmp, nv, zoff, nplanes, st, scol, gate);
}
template <typename T>
static bool xpf2(const FftPlan& p, const T* phys, long long pstride, C2<T>* tmp, int nv, int nplanes,
                 hipStream_t st, const int* gate) {
  return p.wide_xf ? xpf2_nt<T, 2 * Nt2<T>::v>(p, phys, pstride, tmp, nv, nplanes, st, gate)
                : xpf2_nt<T, Nt2<T>::v>(p, phys, pstride, tmp, nv, nplanes, st, gate);
}
template <typename T>
static bool xpb2(const FftPlan& p, const C2<T>* tmp, T* phys, long long pstride, int nv, int nplanes,
                 hipStream_t st, const int* gate) {
  return p.wide_xb ? xpb2_nt<T, 2 * Nt2<T>::v>(p, tmp, phys, pstride, nv, nplanes, st, gate)
                : xpb2_nt<T, Nt2<T>::v>(p, tmp, phys, pstride, nv, nplanes, st, gate);
}

// ---------------------------------------------------------------------------
// launch configuration
// ---------------------------------------------------------------------------
static constexpr int kNT = 256;
static constexpr int kMaxB = 8;
static constexpr size_t kLdsBudget = 64 * 1024;

// sequences per block such that registers (MAXB butterflies of radix >= 2)
// and the LDS budget hold them
template <typename T>
static int seq_per_block(int n) {
  int byb = (2 * kMaxB * kNT) / n;              // nseq*n/2 <= MAXB*NT
  const size_t per = sizeof(C2<T>) * (size_t)lds_ld(n);
  int bylds = (int)((kLdsBudget - sizeof(C2<T>) * n) / per);
  int s = std::min(byb, bylds);
  if (s > 32) s = 32;
  if (s < 1) s = 1;
  return s;
}

template <typename T>
static size_t lds_bytes(int n, int nseq) {
  return sizeof(C2<T>) * ((size_t)nseq * lds_ld(n) + n);
}

template <typename T>
void fft_xpass_forward(const FftPlan& p, const T* phys, long long pstride, C2<T>* tmp, int nv, int nplanes,
                       hipStream_t st, const int* gate) {
  if (p.two_level && xpf2<T>(p, phys, pstride, tmp, nv, nplanes, st, gate)) {
    TBD_HIP(hipGetLastError());
    return;
  }
  const int n = p.d.nx;
  const int nseq = std::max(1, std::min(seq_per_block<T>(n), (p.d.ny + 1) / 2));
  dim3 grid(ceil_div(p.d.ny, 2 * nseq), nv * nplanes);
  hipLaunchKernelGGL((k_xpass_fwd<T, kNT, kMaxB>), grid, dim3(kNT), lds_bytes<T>(n, nseq), st, phys,
                     pstride, tmp, p.lx, twx<T>(p), p.d.ny, p.d.mx, nplanes, nseq, gate);
  TBD_HIP(hipGetLastError());
}

template <typename T>
void fft_xpass_backward(const FftPlan& p, const C2<T>* tmp, T* phys, long long pstride, int nv, int nplanes,
                        hipStream_t st, const int* gate) {
  if (p.two_level && xpb2<T>(p, tmp, phys, pstride, nv, nplanes, st, gate)) {
    TBD_HIP(hipGetLastError());
    return;
  }
  const int n = p.d.nx;
  const int nseq = std::max(1, std::min(seq_per_block<T>(n), (p.d.ny + 1) / 2));
  dim3 grid(ceil_div(p.d.ny, 2 * nseq), nv * nplanes);
  hipLaunchKernelGGL((k_xpass_bwd<T, kNT, kMaxB>), grid, dim3(kNT), lds_bytes<T>(n, nseq), st, tmp, phys,
                     pstride, p.lx, twx<T>(p), p.d.ny, p.d.mx, nplanes, nseq, gate);
  TBD_HIP(hipGetLastError());
}

template <typename T, typename SpecT>
void fft_ypass_forward(const FftPlan& p, const C2<T>* tmp, SpecT* spec, long long sstride, int nv, int zoff,
                       int nplanes, hipStream_t st, int scol, const int* gate) {
  if (p.two_level && ypf2<T, SpecT>(p, tmp, spec, sstride, nv, zoff, nplanes, st, scol, gate)) {
    TBD_HIP(hipGetLastError());
    return;
  }
  const int n = p.d.ny;
  const int zt = std::max(1, std::min(seq_per_block<T>(n), nplanes));
  dim3 grid(ceil_div(nplanes, zt), p.d.mx, nv);
  const T scale = (T)(1.0 / ((double)p.d.nx * (double)p.d.ny));
  hipLaunchKernelGGL((k_ypass_fwd<T, SpecT, kNT, kMaxB>), grid, dim3(kNT), lds_bytes<T>(n, zt), st, tmp, spec,
                     sstride, p.ly, twy<T>(p), scol > 0 ? scol : p.d.nz, p.d.mx, p.d.nyc, nplanes, zoff, zt, scale,
                     gate);
  TBD_HIP(hipGetLastError());
}

template <typename T, typename SpecT>
void fft_ypass_backward(const FftPlan& p, const SpecT* spec, long long sstride, C2<T>* tmp, int nv, int zoff,
                        int nplanes, hipStream_t st, int scol, const int* gate) {
  if (p.two_level && ypb2<T, SpecT>(p, spec, sstride, tmp, nv, zoff, nplanes, st, scol, gate)) {
    TBD_HIP(hipGetLastError());
    return;
  }
  const int n = p.d.ny;
  const int zt = std::max(1, std::min(seq_per_block<T>(n), nplanes));
  dim3 grid(ceil_div(nplanes, zt), p.d.mx, nv);
  hipLaunchKernelGGL((k_ypass_bwd<T, SpecT, kNT, kMaxB>), grid, dim3(kNT), lds_bytes<T>(n, zt), st, spec, sstride,
                     tmp, p.ly, twy<T>(p), scol > 0 ? scol : p.d.nz, p.d.mx, p.d.nyc, nplanes, zoff, zt, gate);
  TBD_HIP(hipGetLastError());
}

template <typename T, typename SpecT>
void fft_forward(const FftPlan& p, const T* phys, long long pstride, SpecT* spec, long long sstride, int nv,
                 int zoff, int nplanes, C2<T>* tmp, hipStream_t st, int scol, const int* gate) {
  fft_xpass_forward<T>(p, phys, pstride, tmp, nv, nplanes, st, gate);
  fft_ypass_forward<T, SpecT>(p, tmp, spec, sstride, nv, zoff, nplanes, st, scol, gate);
}

template <typename T, typename SpecT>
void fft_backward(const FftPlan& p, const SpecT* spec, long long sstride, T* phys, long long pstride, int nv,
                  int zoff, int nplanes, C2<T>* tmp, hipStream_t st, int scol, const int* gate) {
  fft_ypass_backward<T, SpecT>(p, spec, sstride, tmp, nv, zoff, nplanes, st, scol, gate);
  fft_xpass_backward<T>(p, tmp, phys, pstride, nv, nplanes, st, gate);
}

#define TBD_FFT_INST(T, S)                                                                              \
  template void fft_forward<T, S>(const FftPlan&, const T*, long long, S*, long long, int, int, int,  \
                                  C2<T>*, hipStream_t, int, const int*);                              \
  template void fft_backward<T, S>(const FftPlan&, const S*, long long, T*, long long, int, int, int, \
                                   C2<T>*, hipStream_t, int, const int*);                             \
  template void fft_ypass_forward<T, S>(const FftPlan&, const C2<T>*, S*, long long, int, int, int,   \
                                        hipStream_t, int, const int*);                                \
  template void fft_ypass_backward<T, S>(const FftPlan&, const S*, long long, C2<T>*, int, int, int,  \
                                         hipStream_t, int, const int*);

TBD_FFT_INST(double, double2)
TBD_FFT_INST(float, float2)
TBD_FFT_INST(float, double2)
template void fft_xpass_forward<double>(const FftPlan&, const double*, long long, double2*, int, int, hipStream_t,
                                        const int*);
template void fft_xpass_forward<float>(const FftPlan&, const float*, long long, float2*, int, int, hipStream_t,
                                       const int*);
template void fft_xpass_backward<double>(const FftPlan&, const double2*, double*, long long, int, int, hipStream_t,
                                         const int*);
template void fft_xpass_backward<float>(const FftPlan&, const float2*, float*, long long, int, int, hipStream_t,
                                        const int*);

}  // namespace tbd
