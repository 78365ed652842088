// Spin-up proposal network (see sno.h): BF16 MFMA GEMM, elementwise kernels,
// Adam, and the forward / backward sequences of the network.
#include "sno.h"

#include <cmath>
#include <cstring>

#include "pointwise.h"

namespace tbd {

// ---------------------------------------------------------------------------
// BF16 GEMM, FP32 accumulation: v_mfma_f32_16x16x16_bf16.  Lane l supplies
// A[l%16][4(l/16) + j] and B[4(l/16) + j][l%16], j = 0..3; accumulator
// register r holds row 4(l/16) + r, column l%16.  A workgroup (4 waves)
// computes a 64 x 32 tile, each wave 16 rows x 32 columns; K may be split
// over blockIdx.z (partial products added atomically).
// ---------------------------------------------------------------------------
typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);      // round to nearest even
  return (short)(u >> 16);
}

struct GemmArgs {
  int M, N, K, nb2, ksplit, kchunk;
  float alpha, beta;
  GemmOp A, B;
  GemmOut C;
  const float* bias;
};

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tn = (g.N + 31) / 32;
  const int tm = blockIdx.x / tn, tc = blockIdx.x - tm * tn;
  const int row0 = tm * 64 + wave * 16;
  const int col0 = tc * 32;
  const int i1 = blockIdx.y / g.nb2, i2 = blockIdx.y - i1 * g.nb2;
  const float* Ab = g.A.p + i1 * g.A.s1 + i2 * g.A.s2;
  const float* Bb = g.B.p + i1 * g.B.s1 + i2 * g.B.s2;
  float* Cb = g.C.p + i1 * g.C.s1 + i2 * g.C.s2;
  const int k0 = blockIdx.z * g.kchunk;
  const int k1 = min(g.K, k0 + g.kchunk);
  const int r = row0 + (lane & 15);
  const int kk = 4 * (lane >> 4);
  const int c0 = col0 + (lane & 15), c1 = c0 + 16;
  f4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
  if (row0 < g.M) {
    for (int kb = k0; kb < k1; kb += 16) {
      s4v a, b0, b1;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = kb + kk + j;
        const bool kin = k < k1;
        a[j] = (kin && r < g.M) ? to_bf16(Ab[(long long)r * g.A.rs + (long long)k * g.A.cs]) : (short)0;
        b0[j] = (kin && c0 < g.N) ? to_bf16(Bb[(long long)k * g.B.rs + (long long)c0 * g.B.cs]) : (short)0;
        b1[j] = (kin && c1 < g.N) ? to_bf16(Bb[(long long)k * g.B.rs + (long long)c1 * g.B.cs]) : (short)0;
      }
      acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b0, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1, acc1, 0, 0, 0);
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int row = row0 + 4 * (lane >> 4) + q;
    if (row >= g.M) continue;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int col = h == 0 ? c0 : c1;
      if (col >= g.N) continue;
      const float v = g.alpha * (h == 0 ? acc0[q] : acc1[q]);
      float* cp = Cb + (long long)row * g.C.rs + (long long)col * g.C.cs;
      const float bv = (g.bias && blockIdx.z == 0) ? g.bias[col] : 0.f;
      if (g.ksplit > 1) {
        atomicAdd(cp, v + bv);
      } else {
        *cp = v + bv + (g.beta != 0.f ? g.beta * *cp : 0.f);
      }
    }
  }
}

__global__ void k_scale_f32(float* p, long long n, float s) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = s == 0.f ? 0.f : s * p[i];
}

static void gemm_impl(int M, int N, int K, int nb1, int nb2, float alpha, const GemmOp& A, const GemmOp& B,
                      float beta, const GemmOut& C, const float* bias, hipStream_t st) {
  TBD_CHECK(M > 0 && N > 0 && K > 0 && nb1 > 0 && nb2 > 0 && (long long)nb1 * nb2 <= 65535,
            "gemm_bf16 > bad shape");
  GemmArgs g;
  g.M = M; g.N = N; g.K = K; g.nb2 = nb2;
  g.alpha = alpha; g.beta = beta; g.A = A; g.B = B; g.C = C; g.bias = bias;
  const long long tiles = (long long)((M + 63) / 64) * ((N + 31) / 32);
  const long long blocks = tiles * nb1 * nb2;
  int ks = 1;
  if (blocks < 512 && K >= 2048) {
    ks = (int)std::min<long long>(512 / blocks + 1, K / 1024);
    if (ks < 1) ks = 1;
  }
  g.kchunk = ((K + ks - 1) / ks + 15) / 16 * 16;
  ks = (K + g.kchunk - 1) / g.kchunk;
  g.ksplit = ks;
  if (ks > 1) {
    // C := beta C (beta is 0 or 1 here), then every split adds its part
    TBD_CHECK(beta == 0.f || beta == 1.f, "gemm_bf16 > split K needs beta 0 or 1");
    TBD_CHECK(nb1 == 1 && nb2 == 1 && C.cs == 1 && C.rs == N, "gemm_bf16 > split K needs one dense output");
    if (beta == 0.f) hipLaunchKernelGGL(k_scale_f32, dim3(64), dim3(256), 0, st, C.p, (long long)M * N, 0.f);
  }
  dim3 grid((unsigned)tiles, (unsigned)(nb1 * nb2), (unsigned)ks);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, g);
  TBD_HIP(hipGetLastError());
}

void gemm_bf16(int M, int N, int K, int nb1, int nb2, float alpha, const GemmOp& A, const GemmOp& B, float beta,
               const GemmOut& C, hipStream_t st) {
  gemm_impl(M, N, K, nb1, nb2, alpha, A, B, beta, C, nullptr, st);
}

// ---------------------------------------------------------------------------
// elementwise
// ---------------------------------------------------------------------------
#define SNO_GRID(i, n) \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

__device__ __forceinline__ float gelu(float s) { return 0.5f * s * (1.f + erff(s * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float s) {
  return 0.5f * (1.f + erff(s * 0.70710678118654752f)) + s * 0.39894228040143268f * expf(-0.5f * s * s);
}

// Hn = H + gelu(S)
__global__ void k_res_gelu(const float* __restrict__ H, const float* __restrict__ S, float* __restrict__ Hn,
                           long long n) {
  SNO_GRID(i, n) Hn[i] = H[i] + gelu(S[i]);
}
// dS = dH * gelu'(S)
__global__ void k_gelu_back(const float* __restrict__ dH, const float* __restrict__ S, float* __restrict__ dS,
                            long long n) {
  SNO_GRID(i, n) dS[i] = dH[i] * gelu_d(S[i]);
}
// dY = 2 w_o (Y - T) / count; partial sums of w_o (Y - T)^2
__global__ void k_loss(const float* __restrict__ Y, const float* __restrict__ T, const float* __restrict__ w,
                       float* __restrict__ dY, long long n, int cout, float inv, float* loss) {
  __shared__ float red[256];
  float acc = 0.f;
  SNO_GRID(i, n) {
    const float wv = w[i % cout];
    const float e = Y[i] - T[i];
    dY[i] = 2.f * wv * e * inv;
    acc += wv * e * e;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] * inv);
}
// complex mixing parameters (A_m, B_m; C x C each, per mode) -> real blocks
// [[A, B], [-B, A]] (2C x 2C, row-vector convention g = f R)
__global__ void k_expand(const float* __restrict__ p, float* __restrict__ Rb, int modes, int C) {
  const long long n = (long long)modes * 4 * C * C;
  SNO_GRID(i, n) {
    const int m = (int)(i / (4LL * C * C));
    const int rem = (int)(i - (long long)m * 4 * C * C);
    const int row = rem / (2 * C), col = rem - row * (2 * C);
    const float* A = p + (long long)m * 2 * C * C;
    const float* B = A + C * C;
    const int r = row % C, c = col % C;
    float v;
    if (row < C) v = col < C ? A[r * C + c] : B[r * C + c];
    else v = col < C ? -B[r * C + c] : A[r * C + c];
    Rb[i] = v;
  }
}
// gradient of the real blocks -> gradient of (A_m, B_m)
__global__ void k_fold(const float* __restrict__ dRb, float* __restrict__ dp, int modes, int C) {
  const long long n = (long long)modes * C * C;
  SNO_GRID(i, n) {
    const int m = (int)(i / ((long long)C * C));
    const int rc = (int)(i - (long long)m * C * C);
    const int r = rc / C, c = rc - r * C;
    const float* D = dRb + (long long)m * 4 * C * C;
    const int ld = 2 * C;
    const float dA = D[r * ld + c] + D[(C + r) * ld + C + c];
    const float dB = D[r * ld + C + c] - D[(C + r) * ld + c];
    dp[(long long)m * 2 * C * C + rc] = dA;
    dp[(long long)m * 2 * C * C + C * C + rc] = dB;
  }
}
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       long long n, float lr, float b1, float b2, float eps, float c1, float c2) {
  SNO_GRID(i, n) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

static dim3 egrid(long long n) {
  long long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  return dim3((unsigned)(b < 1 ? 1 : b));
}

// ---------------------------------------------------------------------------
// network
// ---------------------------------------------------------------------------
static int modes_of(int n) { return n / 2 + 1; }

SnoLayout SnoLayout::make(const SnoShape& s) {
  TBD_CHECK(s.L >= 1 && s.L <= 8, "Sno > 1..8 layers");
  SnoLayout y;
  long long o = 0;
  const long long C = s.C;
  y.P = o; o += (long long)s.cin * C;
  y.pb = o; o += C;
  for (int l = 0; l < s.L; ++l) {
    y.W[l] = o; o += C * C;
    y.b[l] = o; o += C;
    y.Rx[l] = o; o += (long long)modes_of(s.nx) * 2 * C * C;
    y.Ry[l] = o; o += (long long)modes_of(s.ny) * 2 * C * C;
    y.Rz[l] = o; o += (long long)s.nz * C * C;
  }
  y.Q = o; o += C * s.cout;
  y.qb = o; o += s.cout;
  y.total = o;
  return y;
}

// real DFT of n points: rows 2m, 2m+1 = Re, Im of (1/n) sum_x h_x e^{-2 pi i m x/n};
// inverse: h_x = sum_m c_m (F_re cos - F_im sin), c_m = 1 (m = 0, n/2), 2 otherwise
static void dft_mats(int n, std::vector<float>& D, std::vector<float>& Di) {
  const int mm = modes_of(n);
  D.assign((size_t)2 * mm * n, 0.f);
  Di.assign((size_t)n * 2 * mm, 0.f);
  for (int m = 0; m < mm; ++m)
    for (int x = 0; x < n; ++x) {
      const double th = 2.0 * M_PI * m * x / n;
      D[(size_t)(2 * m) * n + x] = (float)(std::cos(th) / n);
      D[(size_t)(2 * m + 1) * n + x] = (float)(-std::sin(th) / n);
      const double cm = (m == 0 || 2 * m == n) ? 1.0 : 2.0;
      Di[(size_t)x * 2 * mm + 2 * m] = (float)(cm * std::cos(th));
      Di[(size_t)x * 2 * mm + 2 * m + 1] = (float)(-cm * std::sin(th));
    }
}
// Chebyshev transform on n CGL points z_j = -cos(j pi/(n-1)) (bottom first):
// a_k = (2/(n-1)) sum_j'' h_j T_k(z_j) / c_k, h_j = sum_k a_k T_k(z_j)
static void cheb_mats(int n, std::vector<float>& D, std::vector<float>& Di) {
  D.assign((size_t)n * n, 0.f);
  Di.assign((size_t)n * n, 0.f);
  const int N = n - 1;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j) {
      const double z = -std::cos(M_PI * j / N);
      const double T = std::cos(k * std::acos(std::max(-1.0, std::min(1.0, z))));
      const double cj = (j == 0 || j == N) ? 2.0 : 1.0;
      const double ck = (k == 0 || k == N) ? 2.0 : 1.0;
      D[(size_t)k * n + j] = (float)(2.0 / N * T / (cj * ck));
      Di[(size_t)j * n + k] = (float)T;
    }
}

float* Sno::alloc(size_t n) {
  void* p = nullptr;
  TBD_HIP(hipMalloc(&p, std::max<size_t>(n * sizeof(float), 256)));
  allocs_.push_back(p);
  fill_zero(p, std::max<size_t>(n * sizeof(float), 256), st_);
  return (float*)p;
}

Sno::Sno(const SnoShape& s) : shape(s) {
  TBD_CHECK(s.nx >= 2 && s.ny >= 2 && s.nz >= 3 && s.C >= 1 && s.cin >= 1 && s.cout >= 1 && s.batch >= 1,
            "Sno > bad shape");
  lay = SnoLayout::make(s);
  TBD_HIP(hipStreamCreate(&st_));
  pts_ = s.nx * s.ny * s.nz;
  prm_ = alloc(lay.total);
  grd_ = alloc(lay.total);
  m_ = alloc(lay.total);
  v_ = alloc(lay.total);
  std::vector<float> a, b;
  auto up = [&](const std::vector<float>& h) {
    float* p = alloc(h.size());
    TBD_HIP(hipMemcpy(p, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    return p;
  };
  dft_mats(s.nx, a, b); Dx_ = up(a); Dxi_ = up(b);
  dft_mats(s.ny, a, b); Dy_ = up(a); Dyi_ = up(b);
  cheb_mats(s.nz, a, b); Dz_ = up(a); Dzi_ = up(b);
  const long long act = (long long)s.batch * pts_ * s.C;
  X_ = alloc((size_t)s.batch * pts_ * s.cin);
  T_ = alloc((size_t)s.batch * pts_ * s.cout);
  Y_ = alloc((size_t)s.batch * pts_ * s.cout);
  dY_ = alloc((size_t)s.batch * pts_ * s.cout);
  Wt_ = alloc(s.cout);
  for (int l = 0; l < s.L; ++l) {
    H_.push_back(alloc(act));
    S_.push_back(alloc(act));
  }
  HL_ = alloc(act);
  dH_ = alloc(act);
  dS_ = alloc(act);
  const long long lines = (long long)s.batch * std::max(s.nz * s.ny, std::max(s.nz * s.nx, s.ny * s.nx));
  const long long per = (long long)std::max(2 * modes_of(s.nx), std::max(2 * modes_of(s.ny), s.nz)) * s.C;
  F_ = alloc(lines * per);
  G_ = alloc(lines * per);
  dF_ = alloc(lines * per);
  dG_ = alloc(lines * per);
  loss_ = alloc(1);
  ones_ = alloc((size_t)s.batch * pts_);
  {
    std::vector<float> o((size_t)s.batch * pts_, 1.f);
    TBD_HIP(hipMemcpyAsync(ones_, o.data(), sizeof(float) * o.size(), hipMemcpyHostToDevice, st_));
    TBD_HIP(hipStreamSynchronize(st_));
  }
  const long long rb = (long long)std::max(modes_of(s.nx), modes_of(s.ny)) * 4 * s.C * s.C;
  Rb_ = alloc(rb);
  dRb_ = alloc(rb);
  sync();
}

Sno::~Sno() {
  if (st_) (void)hipStreamSynchronize(st_);
  for (void* p : allocs_) (void)hipFree(p);
  if (st_) (void)hipStreamDestroy(st_);
}

void Sno::sync() { TBD_HIP(hipStreamSynchronize(st_)); }

void Sno::set_params(const float* host) {
  TBD_HIP(hipMemcpyAsync(prm_, host, sizeof(float) * lay.total, hipMemcpyHostToDevice, st_));
  hipLaunchKernelGGL(k_scale_f32, egrid(lay.total), dim3(256), 0, st_, m_, lay.total, 0.f);
  hipLaunchKernelGGL(k_scale_f32, egrid(lay.total), dim3(256), 0, st_, v_, lay.total, 0.f);
  adam_t_ = 0;
  sync();
}
void Sno::get_params(float* host) {
  sync();
  TBD_HIP(hipMemcpy(host, prm_, sizeof(float) * lay.total, hipMemcpyDeviceToHost));
}
void Sno::get_grads(float* host) {
  sync();
  TBD_HIP(hipMemcpy(host, grd_, sizeof(float) * lay.total, hipMemcpyDeviceToHost));
}

// Line geometry of one axis: lines (i1, i2) with strides into the activation
// tensor, the element stride along the axis, the axis length and its modes.
struct AxisGeo {
  int n, modes, rows;        // points, modes, rows of the mode block (2*modes or modes)
  int nb1, nb2;
  long long s1, s2, es;      // line strides and element stride in H
  bool cplx;
};
static AxisGeo axis_geo(const SnoShape& s, int axis, int nb) {
  AxisGeo g;
  const long long C = s.C;
  if (axis == 0) {          // x: lines (b, z, y)
    g.n = s.nx; g.modes = modes_of(s.nx); g.rows = 2 * g.modes; g.cplx = true;
    g.nb1 = nb * s.nz * s.ny; g.nb2 = 1; g.s1 = (long long)s.nx * C; g.s2 = 0; g.es = C;
  } else if (axis == 1) {   // y: lines (b, z) x x
    g.n = s.ny; g.modes = modes_of(s.ny); g.rows = 2 * g.modes; g.cplx = true;
    g.nb1 = nb * s.nz; g.nb2 = s.nx; g.s1 = (long long)s.ny * s.nx * C; g.s2 = C; g.es = (long long)s.nx * C;
  } else {                  // z: lines b x (y, x)
    g.n = s.nz; g.modes = s.nz; g.rows = s.nz; g.cplx = false;
    g.nb1 = nb; g.nb2 = s.ny * s.nx; g.s1 = (long long)s.nz * s.ny * s.nx * C; g.s2 = C;
    g.es = (long long)s.ny * s.nx * C;
  }
  return g;
}

// S += Dinv Mix(D H) along one axis
void Sno::axis_forward(int l, int axis, const float* H, float* S, int nb) {
  const SnoShape& s = shape;
  const AxisGeo a = axis_geo(s, axis, nb);
  const int C = s.C;
  const float* D = axis == 0 ? Dx_ : (axis == 1 ? Dy_ : Dz_);
  const float* Di = axis == 0 ? Dxi_ : (axis == 1 ? Dyi_ : Dzi_);
  const long long fl = (long long)a.rows * C;        // F per line
  const long long nl = (long long)a.nb1 * a.nb2;
  // F[line] = D H[line]
  gemm_bf16(a.rows, C, a.n, a.nb1, a.nb2, 1.f, GemmOp{D, 0, 0, a.n, 1}, GemmOp{H, a.s1, a.s2, a.es, 1}, 0.f,
            GemmOut{F_, a.nb2 * fl, fl, C, 1}, st_);
  // G_m = F_m R_m for every mode
  if (a.cplx) {
    const float* par = prm_ + (axis == 0 ? lay.Rx[l] : lay.Ry[l]);
    hipLaunchKernelGGL(k_expand, egrid((long long)a.modes * 4 * C * C), dim3(256), 0, st_, par, Rb_, a.modes, C);
    gemm_bf16((int)nl, 2 * C, 2 * C, a.modes, 1, 1.f, GemmOp{F_, 2LL * C, 0, fl, 1},
              GemmOp{Rb_, 4LL * C * C, 0, 2LL * C, 1}, 0.f, GemmOut{G_, 2LL * C, 0, fl, 1}, st_);
  } else {
    const float* par = prm_ + lay.Rz[l];
    gemm_bf16((int)nl, C, C, a.modes, 1, 1.f, GemmOp{F_, (long long)C, 0, fl, 1},
              GemmOp{par, (long long)C * C, 0, C, 1}, 0.f, GemmOut{G_, (long long)C, 0, fl, 1}, st_);
  }
  // S[line] += Dinv G[line]
  gemm_bf16(a.n, C, a.rows, a.nb1, a.nb2, 1.f, GemmOp{Di, 0, 0, a.rows, 1}, GemmOp{G_, a.nb2 * fl, fl, C, 1}, 1.f,
            GemmOut{S, a.s1, a.s2, a.es, 1}, st_);
}

// given dS: dH += D^T Mix^T(Dinv^T dS), gradient of the mixing parameters
void Sno::axis_backward(int l, int axis, const float* H, const float* dS, float* dH, int nb) {
  const SnoShape& s = shape;
  const AxisGeo a = axis_geo(s, axis, nb);
  const int C = s.C;
  const float* D = axis == 0 ? Dx_ : (axis == 1 ? Dy_ : Dz_);
  const float* Di = axis == 0 ? Dxi_ : (axis == 1 ? Dyi_ : Dzi_);
  const long long fl = (long long)a.rows * C;
  const long long nl = (long long)a.nb1 * a.nb2;
  // forward modes again: F = D H
  gemm_bf16(a.rows, C, a.n, a.nb1, a.nb2, 1.f, GemmOp{D, 0, 0, a.n, 1}, GemmOp{H, a.s1, a.s2, a.es, 1}, 0.f,
            GemmOut{F_, a.nb2 * fl, fl, C, 1}, st_);
  // dG = Dinv^T dS
  gemm_bf16(a.rows, C, a.n, a.nb1, a.nb2, 1.f, GemmOp{Di, 0, 0, 1, a.rows}, GemmOp{dS, a.s1, a.s2, a.es, 1}, 0.f,
            GemmOut{dG_, a.nb2 * fl, fl, C, 1}, st_);
  if (a.cplx) {
    const float* par = prm_ + (axis == 0 ? lay.Rx[l] : lay.Ry[l]);
    float* gpar = grd_ + (axis == 0 ? lay.Rx[l] : lay.Ry[l]);
    hipLaunchKernelGGL(k_expand, egrid((long long)a.modes * 4 * C * C), dim3(256), 0, st_, par, Rb_, a.modes, C);
    // dF_m = dG_m R_m^T
    gemm_bf16((int)nl, 2 * C, 2 * C, a.modes, 1, 1.f, GemmOp{dG_, 2LL * C, 0, fl, 1},
              GemmOp{Rb_, 4LL * C * C, 0, 1, 2LL * C}, 0.f, GemmOut{dF_, 2LL * C, 0, fl, 1}, st_);
    // dR_m = F_m^T dG_m (per mode; K = lines)
    for (int m = 0; m < a.modes; ++m)
      gemm_bf16(2 * C, 2 * C, (int)nl, 1, 1, 1.f, GemmOp{F_ + 2LL * C * m, 0, 0, 1, fl},
                GemmOp{dG_ + 2LL * C * m, 0, 0, fl, 1}, 0.f, GemmOut{dRb_ + 4LL * C * C * m, 0, 0, 2LL * C, 1}, st_);
    hipLaunchKernelGGL(k_fold, egrid((long long)a.modes * C * C), dim3(256), 0, st_, dRb_, gpar, a.modes, C);
  } else {
    const float* par = prm_ + lay.Rz[l];
    float* gpar = grd_ + lay.Rz[l];
    gemm_bf16((int)nl, C, C, a.modes, 1, 1.f, GemmOp{dG_, (long long)C, 0, fl, 1},
              GemmOp{par, (long long)C * C, 0, 1, C}, 0.f, GemmOut{dF_, (long long)C, 0, fl, 1}, st_);
    for (int m = 0; m < a.modes; ++m)
      gemm_bf16(C, C, (int)nl, 1, 1, 1.f, GemmOp{F_ + (long long)C * m, 0, 0, 1, fl},
                GemmOp{dG_ + (long long)C * m, 0, 0, fl, 1}, 0.f, GemmOut{gpar + (long long)C * C * m, 0, 0, C, 1},
                st_);
  }
  // dH[line] += D^T dF[line]
  gemm_bf16(a.n, C, a.rows, a.nb1, a.nb2, 1.f, GemmOp{D, 0, 0, 1, a.n}, GemmOp{dF_, a.nb2 * fl, fl, C, 1}, 1.f,
            GemmOut{dH, a.s1, a.s2, a.es, 1}, st_);
}

void Sno::fwd_device(int nb) {
  const SnoShape& s = shape;
  const int C = s.C;
  const int np = nb * pts_;
  const long long act = (long long)np * C;
  // lift
  gemm_impl(np, C, s.cin, 1, 1, 1.f, GemmOp{X_, 0, 0, s.cin, 1}, GemmOp{prm_ + lay.P, 0, 0, C, 1}, 0.f,
            GemmOut{H_[0], 0, 0, C, 1}, prm_ + lay.pb, st_);
  for (int l = 0; l < s.L; ++l) {
    float* H = H_[l];
    float* S = S_[l];
    gemm_impl(np, C, C, 1, 1, 1.f, GemmOp{H, 0, 0, C, 1}, GemmOp{prm_ + lay.W[l], 0, 0, C, 1}, 0.f,
              GemmOut{S, 0, 0, C, 1}, prm_ + lay.b[l], st_);
    for (int ax = 0; ax < 3; ++ax) axis_forward(l, ax, H, S, nb);
    float* Hn = (l + 1 < s.L) ? H_[l + 1] : HL_;
    hipLaunchKernelGGL(k_res_gelu, egrid(act), dim3(256), 0, st_, H, S, Hn, act);
  }
  gemm_impl(np, s.cout, C, 1, 1, 1.f, GemmOp{HL_, 0, 0, C, 1}, GemmOp{prm_ + lay.Q, 0, 0, s.cout, 1}, 0.f,
            GemmOut{Y_, 0, 0, s.cout, 1}, prm_ + lay.qb, st_);
  TBD_HIP(hipGetLastError());
}

double Sno::bwd_device(int nb) {
  const SnoShape& s = shape;
  const int C = s.C;
  const int np = nb * pts_;
  const long long act = (long long)np * C;
  const long long ny = (long long)np * s.cout;
  hipLaunchKernelGGL(k_scale_f32, dim3(1), dim3(64), 0, st_, loss_, 1LL, 0.f);
  hipLaunchKernelGGL(k_loss, egrid(ny), dim3(256), 0, st_, Y_, T_, Wt_, dY_, ny, s.cout, 1.f / (float)ny, loss_);
  // projection
  gemm_bf16(C, s.cout, np, 1, 1, 1.f, GemmOp{HL_, 0, 0, 1, C}, GemmOp{dY_, 0, 0, s.cout, 1}, 0.f,
            GemmOut{grd_ + lay.Q, 0, 0, s.cout, 1}, st_);
  gemm_bf16(1, s.cout, np, 1, 1, 1.f, GemmOp{ones_, 0, 0, 0, 1}, GemmOp{dY_, 0, 0, s.cout, 1}, 0.f,
            GemmOut{grd_ + lay.qb, 0, 0, s.cout, 1}, st_);
  gemm_bf16(np, C, s.cout, 1, 1, 1.f, GemmOp{dY_, 0, 0, s.cout, 1}, GemmOp{prm_ + lay.Q, 0, 0, 1, s.cout}, 0.f,
            GemmOut{dH_, 0, 0, C, 1}, st_);
  for (int l = s.L - 1; l >= 0; --l) {
    float* H = H_[l];
    float* S = S_[l];
    // dS = dH_{l+1} gelu'(S); dH_l = dH_{l+1} + (contributions through S)
    hipLaunchKernelGGL(k_gelu_back, egrid(act), dim3(256), 0, st_, dH_, S, dS_, act);
    gemm_bf16(C, C, np, 1, 1, 1.f, GemmOp{H, 0, 0, 1, C}, GemmOp{dS_, 0, 0, C, 1}, 0.f,
              GemmOut{grd_ + lay.W[l], 0, 0, C, 1}, st_);
    gemm_bf16(1, C, np, 1, 1, 1.f, GemmOp{ones_, 0, 0, 0, 1}, GemmOp{dS_, 0, 0, C, 1}, 0.f,
              GemmOut{grd_ + lay.b[l], 0, 0, C, 1}, st_);
    gemm_bf16(np, C, C, 1, 1, 1.f, GemmOp{dS_, 0, 0, C, 1}, GemmOp{prm_ + lay.W[l], 0, 0, 1, C}, 1.f,
              GemmOut{dH_, 0, 0, C, 1}, st_);
    for (int ax = 0; ax < 3; ++ax) axis_backward(l, ax, H, dS_, dH_, nb);
  }
  // lift
  gemm_bf16(s.cin, C, np, 1, 1, 1.f, GemmOp{X_, 0, 0, 1, s.cin}, GemmOp{dH_, 0, 0, C, 1}, 0.f,
            GemmOut{grd_ + lay.P, 0, 0, C, 1}, st_);
  gemm_bf16(1, C, np, 1, 1, 1.f, GemmOp{ones_, 0, 0, 0, 1}, GemmOp{dH_, 0, 0, C, 1}, 0.f,
            GemmOut{grd_ + lay.pb, 0, 0, C, 1}, st_);
  float l = 0.f;
  TBD_HIP(hipMemcpyAsync(&l, loss_, sizeof(float), hipMemcpyDeviceToHost, st_));
  sync();
  return (double)l;
}

void Sno::forward(const float* x, float* y, int nb) {
  TBD_CHECK(nb >= 1 && nb <= shape.batch, "Sno::forward > batch size out of range");
  TBD_HIP(hipMemcpyAsync(X_, x, sizeof(float) * nb * pts_ * shape.cin, hipMemcpyHostToDevice, st_));
  fwd_device(nb);
  TBD_HIP(hipMemcpyAsync(y, Y_, sizeof(float) * nb * pts_ * shape.cout, hipMemcpyDeviceToHost, st_));
  sync();
}

double Sno::loss_and_grad(const float* x, const float* t, const float* w, int nb) {
  TBD_CHECK(nb >= 1 && nb <= shape.batch, "Sno::train > batch size out of range");
  TBD_HIP(hipMemcpyAsync(X_, x, sizeof(float) * nb * pts_ * shape.cin, hipMemcpyHostToDevice, st_));
  TBD_HIP(hipMemcpyAsync(T_, t, sizeof(float) * nb * pts_ * shape.cout, hipMemcpyHostToDevice, st_));
  TBD_HIP(hipMemcpyAsync(Wt_, w, sizeof(float) * shape.cout, hipMemcpyHostToDevice, st_));
  fwd_device(nb);
  return bwd_device(nb);
}

double Sno::train_step(const float* x, const float* t, const float* w, int nb, float lr) {
  const double loss = loss_and_grad(x, t, w, nb);
  ++adam_t_;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - std::pow(b1, (float)adam_t_), c2 = 1.f - std::pow(b2, (float)adam_t_);
  hipLaunchKernelGGL(k_adam, egrid(lay.total), dim3(256), 0, st_, prm_, grd_, m_, v_, lay.total, lr, b1, b2, eps, c1,
                     c2);
  TBD_HIP(hipGetLastError());
  return loss;
}

}  // namespace tbd
