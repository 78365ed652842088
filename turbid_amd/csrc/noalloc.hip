// DftForward for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// tbd_dft_forward_noalloc takes the caller's DEVICE buffers for its input and
// output.  Everything else the transform needs -- the twiddle tables, the
// x-pencil intermediate T[d][z][kx][y] of the two passes and, for the FP32
// transform, the rounded input planes and the FP32 spectral planes -- lives in
// `scratch`, a fixed device array of kScratchBytes.  Each GPU holds its own
// copy of the array from the moment the library's code object is loaded on it
// until the process ends.  The z-planes are transformed in chunks of as many
// planes as fit beside the twiddles, so the only size limit is that one
// z-plane of nv components has to fit.
//
// The host reaches the array through the address hipGetSymbolAddress returns
// for the current device (looked up on the first call on each device and kept
// per device): in host code the array's own name is a host-side stand-in, not
// a device address.  Every call on one GPU shares its array and leaves the last
// call's contents in it; calls that overlap on one GPU race.  Nothing here
// reads a value an earlier call left: the twiddles are rewritten and every
// staged element is written before it is read.
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <string>

#include "common.h"
#include "fft.h"

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;   // 64 MiB per GPU
constexpr int kMaxDevices = 64;
constexpr size_t kAlign = 256;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's copy of `scratch`.
unsigned char* scratch_on_current_device() {
  int dev = -1;
  TBD_HIP(hipGetDevice(&dev));
  TBD_CHECK(dev >= 0 && dev < kMaxDevices, "tbd_dft_forward_noalloc > device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    TBD_CHECK(e == hipSuccess && p != nullptr,
              std::string("tbd_dft_forward_noalloc > hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  return g_scratch_ptr[dev];
}

size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// exp(-2 pi i t / n), t < n, in both precisions (the FP32 table is the FP64
// one rounded, as the library's plans build it)
__global__ void k_twiddles(double2* __restrict__ t64, float2* __restrict__ t32, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double ang = -2.0 * M_PI * (double)t / (double)n;
  const double c = cos(ang), s = sin(ang);
  t64[t] = make_double2(c, s);
  t32[t] = make_float2((float)c, (float)s);
}

// dst[v][e] = (float)src[v][e], e < n, for nv components
__global__ void k_round_planes(const double* __restrict__ src, long long sstride, float* __restrict__ dst,
                               long long dstride, long long n, int nv) {
  const long long total = n * nv;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int v = (int)(i / n);
    const long long e = i - (long long)v * n;
    dst[(long long)v * dstride + e] = (float)src[(long long)v * sstride + e];
  }
}

// spec[v][col][z0 + z] = fspec[v][col][z] for the n planes of a chunk
__global__ void k_widen_chunk(const float2* __restrict__ fspec, double2* __restrict__ spec, long long S, int nz,
                              int z0, int n, long long P, int nv) {
  const long long total = (long long)nv * P * n;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long vc = i / n;
    const int z = (int)(i - vc * n);
    const int v = (int)(vc / P);
    const long long col = vc - (long long)v * P;
    const float2 a = fspec[i];
    spec[(long long)v * S + col * nz + z0 + z] = make_double2((double)a.x, (double)a.y);
  }
}

int grid_for(long long n) {
  const long long b = (n + 255) / 256;
  return (int)(b < 4096 ? (b < 1 ? 1 : b) : 4096);
}

}  // namespace

namespace tbd {

// The body of tbd_dft_forward_noalloc (api.cpp turns its exceptions into
// return codes).
void dft_forward_noalloc(int nx, int ny, int nz, int nv, int fp32, const double* phys, double* spec) {
  TBD_CHECK(phys && spec, "tbd_dft_forward_noalloc > null buffer");
  TBD_CHECK(nv >= 1 && nx >= 2 && ny >= 2 && nz >= 1, "tbd_dft_forward_noalloc > bad size");
  const tbd::Dims d = tbd::Dims::make(nx, ny, nz);
  TBD_CHECK(d.nyc >= 2 && d.mx >= 1, "tbd_dft_forward_noalloc > grid too small for the 3/2 rule");

  // scratch layout: twiddles (x64, y64, x32, y32), then per chunk of n
  // planes the intermediate and, for FP32, the rounded input and the FP32
  // spectral planes
  const size_t tw = align_up(sizeof(double2) * nx) + align_up(sizeof(double2) * ny) +
                    align_up(sizeof(float2) * nx) + align_up(sizeof(float2) * ny);
  const size_t plane_xy = (size_t)nx * ny;
  const size_t tmp_per_plane = (size_t)nv * d.mx * ny * (fp32 ? sizeof(float2) : sizeof(double2));
  const size_t in_per_plane = fp32 ? (size_t)nv * plane_xy * sizeof(float) : 0;
  const size_t out_per_plane = fp32 ? (size_t)nv * d.P * sizeof(float2) : 0;
  auto chunk_bytes = [&](size_t n) {
    return align_up(n * tmp_per_plane) + (fp32 ? align_up(n * in_per_plane) + align_up(n * out_per_plane) : 0);
  };
  TBD_CHECK(tw < kScratchBytes && chunk_bytes(1) <= kScratchBytes - tw,
            "tbd_dft_forward_noalloc > one z-plane of " + std::to_string(nv) + " component(s) at " +
                std::to_string(nx) + "x" + std::to_string(ny) + " needs " + std::to_string(tw + chunk_bytes(1)) +
                " bytes of scratch; the limit is " + std::to_string(kScratchBytes));
  size_t n = (kScratchBytes - tw) / (tmp_per_plane + in_per_plane + out_per_plane);
  if (n > (size_t)nz) n = nz;
  while (n > 1 && chunk_bytes(n) > kScratchBytes - tw) --n;

  unsigned char* base = scratch_on_current_device();
  tbd::FftPlan p;
  p.d = d;
  const char* stockham = std::getenv("TBD_FFT_STOCKHAM");
  p.two_level = !(stockham && stockham[0] == '1');
  p.lx = tbd::make_fftlen(nx);
  p.ly = tbd::make_fftlen(ny);
  unsigned char* q = base;
  p.twx64 = (double2*)q; q += align_up(sizeof(double2) * nx);
  p.twy64 = (double2*)q; q += align_up(sizeof(double2) * ny);
  p.twx32 = (float2*)q;  q += align_up(sizeof(float2) * nx);
  p.twy32 = (float2*)q;  q += align_up(sizeof(float2) * ny);
  hipStream_t st = nullptr;
  hipLaunchKernelGGL(k_twiddles, dim3((nx + 255) / 256), dim3(256), 0, st, p.twx64, p.twx32, nx);
  hipLaunchKernelGGL(k_twiddles, dim3((ny + 255) / 256), dim3(256), 0, st, p.twy64, p.twy32, ny);
  TBD_HIP(hipGetLastError());

  unsigned char* tmp = q;
  for (int z0 = 0; z0 < nz; z0 += (int)n) {
    const int nc = (int)std::min<size_t>(n, (size_t)(nz - z0));
    const double* in = phys + (size_t)z0 * plane_xy;
    if (!fp32) {
      tbd::fft_forward<double, double2>(p, in, d.R, (double2*)spec, d.S, nv, z0, nc, (double2*)tmp, st);
    } else {
      float* fin = (float*)(tmp + align_up(n * tmp_per_plane));
      float2* fout = (float2*)((unsigned char*)fin + align_up(n * in_per_plane));
      const long long per = (long long)nc * (long long)plane_xy;
      hipLaunchKernelGGL(k_round_planes, dim3(grid_for(per * nv)), dim3(256), 0, st, in, d.R, fin, per, per, nv);
      TBD_HIP(hipGetLastError());
      // the FP32 planes of the chunk come out compact: column stride nc
      tbd::fft_forward<float, float2>(p, fin, per, fout, (long long)nc * d.P, nv, 0, nc, (float2*)tmp, st, nc);
      hipLaunchKernelGGL(k_widen_chunk, dim3(grid_for((long long)nv * d.P * nc)), dim3(256), 0, st, fout,
                         (double2*)spec, d.S, nz, z0, nc, d.P, nv);
      TBD_HIP(hipGetLastError());
    }
  }
  TBD_HIP(hipStreamSynchronize(st));
}

}  // namespace tbd
