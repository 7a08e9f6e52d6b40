// Shared device/host helpers for the gfx950 ptychography kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>

#include "../../include/ptycho_hip.h"

namespace pty {

// ---------------------------------------------------------------- errors
int set_error(int code, const char* fmt, ...);
int check_launch(const char* what);

#define PTY_REQUIRE(cond, code, ...) \
  do { if (!(cond)) return ::pty::set_error((code), __VA_ARGS__); } while (0)

// ---------------------------------------------------------------- complex
struct c32 { float x, y; };

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

__device__ __forceinline__ void atomic_add_f32(float* p, float v) { unsafeAtomicAdd(p, v); }

// Compensated (Neumaier) running sum.
struct KSum {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float v) {
    float t = s + v;
    if (fabsf(s) >= fabsf(v)) c += (s - t) + v; else c += (v - t) + s;
    s = t;
  }
  __device__ __forceinline__ float value() const { return s + c; }
};

// ---------------------------------------------------------------- FFT in LDS
// Stockham autosort FFT (radix-4 passes, one radix-2 pass when log2 n is odd)
// over `lines` independent lines held in LDS at a + line*stride.  Ping-pongs
// between a and b and returns the buffer holding the result.  All threads of
// the workgroup must call it (it contains __syncthreads()).  tw holds
// exp(-2 pi i k / (n*tws)) for k < n*tws; DIR = -1 forward, +1 inverse
// (unnormalised).
template <int DIR>
__device__ __forceinline__ float2 twiddle(const float2* tw, int idx) {
  float2 w = tw[idx];
  if (DIR > 0) w.y = -w.y;
  return w;
}

template <int DIR>
__device__ float2* fft_lines(float2* a, float2* b, int n, int log2n, int lines, int stride,
                             const float2* tw, int tws, int tid, int nthreads) {
  int Ns = 1;
  int p = log2n;
  while (p > 0) {
    const int R = (p >= 2) ? 4 : 2;
    const int lgR = (R == 4) ? 2 : 1;
    const int lgnb = log2n - lgR;
    const int nb = 1 << lgnb;
    const int total = nb * lines;
    const int twscale = (n / (Ns * R)) * tws;
    for (int t = tid; t < total; t += nthreads) {
      const int line = t >> lgnb;
      const int j = t & (nb - 1);
      const float2* src = a + line * stride;
      float2* dst = b + line * stride;
      const int k = j & (Ns - 1);
      const int d = (j - k) * R + k;
      if (R == 4) {
        float2 v0 = src[j], v1 = src[j + nb], v2 = src[j + 2 * nb], v3 = src[j + 3 * nb];
        if (Ns > 1) {
          const int ti = k * twscale;
          v1 = cmul(v1, twiddle<DIR>(tw, ti));
          v2 = cmul(v2, twiddle<DIR>(tw, 2 * ti));
          v3 = cmul(v3, twiddle<DIR>(tw, 3 * ti));
        }
        float2 t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3), d13 = csub(v1, v3);
        // forward: t3 = -i*(v1-v3); inverse: t3 = +i*(v1-v3)
        float2 t3 = (DIR < 0) ? make_float2(d13.y, -d13.x) : make_float2(-d13.y, d13.x);
        dst[d] = cadd(t0, t2);
        dst[d + Ns] = cadd(t1, t3);
        dst[d + 2 * Ns] = csub(t0, t2);
        dst[d + 3 * Ns] = csub(t1, t3);
      } else {
        float2 v0 = src[j], v1 = src[j + nb];
        if (Ns > 1) v1 = cmul(v1, twiddle<DIR>(tw, k * twscale));
        dst[d] = cadd(v0, v1);
        dst[d + Ns] = csub(v0, v1);
      }
    }
    __syncthreads();
    float2* tmp = a; a = b; b = tmp;
    Ns *= R;
    p -= lgR;
  }
  return a;
}

// tw[k] = exp(-2 pi i k / n), k < n, accurate sincospi.
__device__ __forceinline__ void fill_twiddles(float2* tw, int n, int tid, int nthreads) {
  for (int k = tid; k < n; k += nthreads) {
    float s, c;
    sincospif(-2.0f * (float)k / (float)n, &s, &c);
    tw[k] = make_float2(c, s);
  }
}

__host__ __device__ __forceinline__ int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}
__host__ __device__ __forceinline__ bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

}  // namespace pty
