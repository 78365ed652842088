// Learned spin-up proposal: a small factorized spectral neural operator on a
// reduced grid (Nx x Ny Fourier points x Nz Chebyshev points), trained and
// evaluated with hand-written BF16 MFMA kernels (FP32 accumulation).
//
// Tensors are FP32 in HBM, channels fastest: H[b][z][y][x][c].  GEMM
// operands are rounded to BF16 (round to nearest even) as they are loaded
// into the matrix cores; products accumulate in FP32.
//
// Network (per sample, grid points p, channels C):
//   H_0      = X P + p                                (lift, CIN -> C)
//   S_l      = H_l W_l + b_l + Kx_l(H_l) + Ky_l(H_l) + Kz_l(H_l)
//   H_{l+1}  = H_l + gelu(S_l)                        (l = 0 .. L-1)
//   Y        = H_L Q + q                              (projection, C -> COUT)
// with the axis operators K_a(H) = Dinv_a Mix_a(D_a H) along each line of
// axis a: D_a the real DFT (x, y; modes 0 .. n/2, real and imaginary rows
// interleaved) or the Chebyshev transform (z), Mix_a a C x C channel mixing
// per mode (complex for x and y, real for z).
#pragma once

#include <vector>

#include "common.h"

namespace tbd {

// C[i1,i2] = alpha * A[i1,i2] B[i1,i2] + beta * C[i1,i2]; each operand is
// addressed as base + i1*s1 + i2*s2 + row*rs + col*cs (elements).
struct GemmOp {
  const float* p = nullptr;
  long long s1 = 0, s2 = 0, rs = 0, cs = 0;
};
struct GemmOut {
  float* p = nullptr;
  long long s1 = 0, s2 = 0, rs = 0, cs = 0;
};
void gemm_bf16(int M, int N, int K, int nb1, int nb2, float alpha, const GemmOp& A, const GemmOp& B, float beta,
               const GemmOut& C, hipStream_t st);

struct SnoShape {
  int nx = 24, ny = 24, nz = 33;     // reduced grid
  int C = 32, L = 4, cin = 7, cout = 4;
  int batch = 8;                     // samples per launch
};

// Parameter block offsets (floats) of one network
struct SnoLayout {
  long long P, pb, Q, qb;                // lift, projection
  long long W[8], b[8];                  // pointwise per layer
  long long Rx[8], Ry[8], Rz[8];         // per-mode mixing per layer
  long long total;
  static SnoLayout make(const SnoShape& s);
  // real-block form of the complex mixing of one mode: 2C x 2C
};

class Sno {
 public:
  explicit Sno(const SnoShape& s);
  ~Sno();
  SnoShape shape;
  SnoLayout lay;
  void set_params(const float* host);            // lay.total floats
  void get_params(float* host);
  void get_grads(float* host);
  // forward for nb <= batch samples: x[nb][pts][cin] -> y[nb][pts][cout] (host)
  void forward(const float* x, float* y, int nb);
  // one training step on nb samples: loss = mean over samples, points and
  // outputs of w_o (y - t)^2; gradients, then Adam (lr, b1, b2, eps, step t)
  double train_step(const float* x, const float* t, const float* w, int nb, float lr);
  // gradient of the same loss without an update (for checks)
  double loss_and_grad(const float* x, const float* t, const float* w, int nb);
  void sync();

 private:
  void fwd_device(int nb);
  double bwd_device(int nb);
  void axis_forward(int l, int axis, const float* H, float* S, int nb);
  void axis_backward(int l, int axis, const float* H, const float* dS, float* dH, int nb);
  float* alloc(size_t n);

  hipStream_t st_ = nullptr;
  std::vector<void*> allocs_;
  int pts_ = 0;
  float *prm_ = nullptr, *grd_ = nullptr, *m_ = nullptr, *v_ = nullptr;
  float *Dx_ = nullptr, *Dxi_ = nullptr, *Dy_ = nullptr, *Dyi_ = nullptr, *Dz_ = nullptr, *Dzi_ = nullptr;
  float *X_ = nullptr, *T_ = nullptr, *Wt_ = nullptr, *Y_ = nullptr;
  std::vector<float*> H_, S_;          // per layer: input H_l, pre-activation S_l
  float *HL_ = nullptr;                // H_L
  float *F_ = nullptr, *G_ = nullptr;  // axis mode work
  float *dH_ = nullptr, *dS_ = nullptr, *dY_ = nullptr, *dF_ = nullptr, *dG_ = nullptr;
  float *loss_ = nullptr;
  float *ones_ = nullptr;              // batch*points ones (bias gradients)
  float *Rb_ = nullptr, *dRb_ = nullptr;   // real-block mixing matrices and their gradient
  long long adam_t_ = 0;
};

}  // namespace tbd
