// Shared helpers for the TURBID MI355X library: error handling, complex
// arithmetic on HIP vector types, layout constants.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace tbd {

// Errors are thrown inside the library and converted to return codes at the
// C ABI (the reference throws std::runtime_error("Class::Method > Message")).
struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

#define TBD_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw ::tbd::Error(std::string("HIP error ") + hipGetErrorString(e_) +      \
                         " at " + __FILE__ + ":" + std::to_string(__LINE__));      \
  } while (0)

#define TBD_CHECK(cond, msg)                 \
  do {                                       \
    if (!(cond)) throw ::tbd::Error(msg);    \
  } while (0)

// complex type per real type
template <typename T> struct cplx_t;
template <> struct cplx_t<double> { typedef double2 type; };
template <> struct cplx_t<float> { typedef float2 type; };
template <typename T> using C2 = typename cplx_t<T>::type;

template <typename T>
__host__ __device__ __forceinline__ C2<T> mk(T x, T y) { C2<T> r; r.x = x; r.y = y; return r; }
template <typename T>
__host__ __device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__host__ __device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__host__ __device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ C2<T> cscale(T s, C2<T> a) { return mk<T>(s * a.x, s * a.y); }
template <typename T>
__host__ __device__ __forceinline__ C2<T> cconj(C2<T> a) { return mk<T>(a.x, -a.y); }
// (0 + i k) * a  (the reference's complex multiplier product, stepper_space.cc:265-296)
template <typename T>
__host__ __device__ __forceinline__ C2<T> cmuli(T k, C2<T> a) {
  return mk<T>((T)0 * a.x - k * a.y, (T)0 * a.y + k * a.x);
}

template <typename To, typename Ti>
__host__ __device__ __forceinline__ C2<To> ccast(Ti a) { return mk<To>((To)a.x, (To)a.y); }

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Spectral dimensions (reference src/base.cc:233-258)
struct Dims {
  int nx = 0, ny = 0, nz = 0;
  int nxc = 0, nyc = 0, mx = 0;
  long long P = 0;   // spectral columns per component (nyc * mx)
  long long S = 0;   // complex entries per component (nz * P)
  long long R = 0;   // physical entries per component (nx * ny * nz)
  static Dims make(int nx, int ny, int nz) {
    Dims d;
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.nxc = nx / 3 * 2;
    d.nyc = ny / 3 * 2;
    d.mx = d.nxc / 2 + 1;
    d.P = (long long)d.nyc * d.mx;
    d.S = (long long)nz * d.P;
    d.R = (long long)nx * ny * nz;
    return d;
  }
};

}  // namespace tbd
