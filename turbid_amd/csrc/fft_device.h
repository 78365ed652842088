// Device-side building blocks of the plane transforms: the LDS Stockham
// transform shared by the transform passes (fft.hip) and the fused
// nonlinear kernels (nonlinear.hip).
#pragma once

#include "fft.h"

namespace tbd {

// ---------------------------------------------------------------------------
// Block-cooperative Stockham transform of nseq sequences (length L.n, leading
// dimension ldn) held in LDS.  tw: LDS copy of exp(-2 pi i t/n).  sign -1
// forward, +1 inverse (unnormalised).
// ---------------------------------------------------------------------------
template <typename T, int NT, int MAXB>
__device__ __forceinline__ void lds_fft(C2<T>* buf, int ldn, int nseq, const FftLen& L,
                                        const C2<T>* tw, int sign) {
  const int N = L.n;
  const T sg = (T)sign;
  int ns = 1;
  for (int s = 0; s < L.nstage; ++s) {
    const int R = L.radix[s];
    const int m = N / R;
    const int total = nseq * m;
    const int step = N / (ns * R);
    C2<T> v[MAXB][4];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const int q = threadIdx.x + b * NT;
      if (q < total) {
        const int sq = q / m;
        const int j = q - sq * m;
        const int k = j % ns;
        const C2<T>* src = buf + sq * ldn;
        C2<T> x0 = src[j];
        C2<T> x1 = src[j + m];
        C2<T> t1 = tw[k * step];
        if (sign > 0) t1.y = -t1.y;  // tw holds exp(-...): conjugate for the inverse
        x1 = cmul<T>(x1, t1);
        if (R == 2) {
          v[b][0] = cadd<T>(x0, x1);
          v[b][1] = csub<T>(x0, x1);
        } else if (R == 4) {
          C2<T> x2 = src[j + 2 * m];
          C2<T> x3 = src[j + 3 * m];
          C2<T> t2 = tw[2 * k * step];
          C2<T> t3 = tw[3 * k * step];
          if (sign > 0) { t2.y = -t2.y; t3.y = -t3.y; }
          x2 = cmul<T>(x2, t2);
          x3 = cmul<T>(x3, t3);
          const C2<T> a0 = cadd<T>(x0, x2), a1 = csub<T>(x0, x2);
          const C2<T> a2 = cadd<T>(x1, x3), a3 = csub<T>(x1, x3);
          // y1 = a1 + sign*i*a3 ; y3 = a1 - sign*i*a3
          const C2<T> ia3 = mk<T>(-sg * a3.y, sg * a3.x);
          v[b][0] = cadd<T>(a0, a2);
          v[b][1] = cadd<T>(a1, ia3);
          v[b][2] = csub<T>(a0, a2);
          v[b][3] = csub<T>(a1, ia3);
        } else {  // R == 3
          C2<T> x2 = src[j + 2 * m];
          C2<T> t2 = tw[2 * k * step];
          if (sign > 0) t2.y = -t2.y;
          x2 = cmul<T>(x2, t2);
          const T h = (T)0.86602540378443864676372317075294;  // sqrt(3)/2
          const C2<T> sm = cadd<T>(x1, x2), df = csub<T>(x1, x2);
          const C2<T> c0 = mk<T>(x0.x - (T)0.5 * sm.x, x0.y - (T)0.5 * sm.y);
          const C2<T> id = mk<T>(-sg * h * df.y, sg * h * df.x);
          v[b][0] = cadd<T>(x0, sm);
          v[b][1] = cadd<T>(c0, id);
          v[b][2] = csub<T>(c0, id);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const int q = threadIdx.x + b * NT;
      if (q < total) {
        const int sq = q / m;
        const int j = q - sq * m;
        const int k = j % ns;
        C2<T>* dst = buf + sq * ldn + (j / ns) * ns * R + k;
        dst[0] = v[b][0];
        dst[ns] = v[b][1];
        if (R >= 3) dst[2 * ns] = v[b][2];
        if (R == 4) dst[3 * ns] = v[b][3];
      }
    }
    __syncthreads();
    ns *= R;
  }
}

template <typename T, int NT>
__device__ __forceinline__ void load_tw(C2<T>* dst, const C2<T>* src, int n) {
  for (int i = threadIdx.x; i < n; i += NT) dst[i] = src[i];
}

// Number of LDS complex slots per sequence (odd padding spreads banks).
__host__ __device__ inline int lds_ld(int n) { return n + 1; }

}  // namespace tbd
