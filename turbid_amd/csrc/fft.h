// Plane transforms of the spectral layout (reference RealCplxCube::DftForward /
// DftBackward, src/cube_real_cplx.cc:525-555, with P3DFFT's op "ffn"/"nff"
// truncation and ordering): x real-to-complex keeping kx < mx, y complex
// keeping nyc modes in FFT order, forward scaled by 1/(Nx*Ny).
//
// Each transform is two passes through an x-pencil intermediate
//   T[d][z][kx][y]   (complex, mx*Ny per (d,z) plane)
// so every global access is a contiguous run:
//   x pass : physical rows (x fastest)  <->  T (y fastest per kx)
//   y pass : T columns (y fastest)      <->  spectral (z fastest per (kx,ky))
#pragma once

#include "common.h"

namespace tbd {

struct FftLen {
  int n = 0;
  int nstage = 0;
  int radix[12] = {0};
};

FftLen make_fftlen(int n);

// Device-resident twiddle tables for one grid, in both precisions.
struct FftPlan {
  Dims d;
  FftLen lx, ly;
  double2* twx64 = nullptr;   // exp(-2 pi i t / Nx), t < Nx
  double2* twy64 = nullptr;
  float2* twx32 = nullptr;
  float2* twy32 = nullptr;
  // two-level register kernels where the length allows; false selects the
  // LDS Stockham kernels for every length (environment TBD_FFT_STOCKHAM=1)
  bool two_level = true;
  // two-level kernels of each pass (y backward, y forward, x backward, x
  // forward) with twice the default workgroup size (twice the sequences per
  // workgroup); chosen per grid and pass by the autotuner
  bool wide_yb = false, wide_yf = false, wide_xb = false, wide_xf = false;
  // the same passes with the two factors of the length in the other order
  // (N2 x N1: level A the smaller DFT), for lengths with two distinct factors
  // (fft_has_swap); chosen per grid and pass by the autotuner
  bool swap_yb = false, swap_yf = false, swap_xb = false, swap_xf = false;
  void create(const Dims& dims);
  void destroy();
};

// Whether a two-level length has the swapped factor order (FftPlan::swap_*).
// true for the lengths the two-level register kernels cover
bool fft_has_two_level(int n);
bool fft_has_swap(int n);

// Size of the x-pencil intermediate (complex elements) for nv components.
inline long long fft_tmp_elems(const Dims& d, int nv) { return (long long)nv * d.nz * d.mx * d.ny; }

// --- launchers ---------------------------------------------------------------
// T is the arithmetic type of the transform (double or float); spectral
// input/output may be double2 or float2 (converted on load/store).
// nplanes/zoff select a z-range [zoff, zoff+nplanes) of the spectral cube
// (the physical and intermediate arrays hold only those planes).

// physical (nv comps, each nplanes*Ny*Nx, stride phys_stride) -> spectral
// scol: distance between spectral columns (default nz; 1 for a single row
// stored compactly, e.g. the wall rows)
// gate (optional, device int): the forward passes do nothing when *gate == 0
template <typename T, typename SpecT>
void fft_forward(const FftPlan& p, const T* phys, long long phys_stride, SpecT* spec,
                 long long spec_stride, int nv, int zoff, int nplanes, C2<T>* tmp,
                 hipStream_t st, int scol = -1, const int* gate = nullptr);

// spectral -> physical (gate as for the forward direction)
template <typename T, typename SpecT>
void fft_backward(const FftPlan& p, const SpecT* spec, long long spec_stride, T* phys,
                  long long phys_stride, int nv, int zoff, int nplanes, C2<T>* tmp,
                  hipStream_t st, int scol = -1, const int* gate = nullptr);

// Individual passes (exposed for fused callers and tests).
template <typename T>
void fft_xpass_forward(const FftPlan& p, const T* phys, long long phys_stride, C2<T>* tmp,
                       int nv, int nplanes, hipStream_t st, const int* gate = nullptr);
template <typename T, typename SpecT>
void fft_ypass_forward(const FftPlan& p, const C2<T>* tmp, SpecT* spec, long long spec_stride,
                       int nv, int zoff, int nplanes, hipStream_t st, int scol = -1,
                       const int* gate = nullptr);
template <typename T, typename SpecT>
void fft_ypass_backward(const FftPlan& p, const SpecT* spec, long long spec_stride, C2<T>* tmp,
                        int nv, int zoff, int nplanes, hipStream_t st, int scol = -1,
                        const int* gate = nullptr);
template <typename T>
void fft_xpass_backward(const FftPlan& p, const C2<T>* tmp, T* phys, long long phys_stride,
                        int nv, int nplanes, hipStream_t st, const int* gate = nullptr);

}  // namespace tbd
