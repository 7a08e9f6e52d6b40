// Shared device/host helpers for the gfx950 ptychography kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>

#include "../../include/ptycho_hip.h"

namespace pty {

// ---------------------------------------------------------------- errors
int set_error(int code, const char* fmt, ...);
int check_launch(const char* what);
// compute units of the current device (256 on MI355X; 256 if the query fails)
int device_cu_count();

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

// ---------------------------------------------------------------- Adam
// torch.optim.Adam (amsgrad=False, maximize=False, weight_decay=0) on one
// real element: m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// p -= lr/bc1 * m / (sqrt(v)/sqrt(bc2) + eps), bc_i = 1 - b_i^step, with
// hyper = [lr, step] on the device.  Shared by the flat Adam kernel and the
// table-gradient kernels that apply the update in place.
struct AdamCoef {
  float b1, b2, eps, sbc2, step;
};
__device__ __forceinline__ AdamCoef adam_coef(const float* hyper, float b1, float b2, float eps) {
  const float lr = hyper[0], t = hyper[1];
  const float bc1 = 1.f - powf(b1, t);
  AdamCoef c;
  c.b1 = b1;
  c.b2 = b2;
  c.eps = eps;
  c.sbc2 = sqrtf(1.f - powf(b2, t));
  c.step = lr / bc1;
  return c;
}
// Contraction off: every caller rounds the update the same way whatever the
// surrounding code, so the in-place table update and the flat kernel agree
// to the bit.
__device__ __forceinline__ void adam_elem(float& p, float& m, float& v, float g, const AdamCoef& c) {
#pragma clang fp contract(off)
  m = c.b1 * m + (1.f - c.b1) * g;
  v = c.b2 * v + (1.f - c.b2) * g * g;
  const float denom = sqrtf(v) / c.sbc2 + c.eps;
  p = p - c.step * (m / denom);
}

// A table's Adam state for the table-gradient kernels (p == nullptr: store
// the gradient into dtable instead): p / m / v are the table's parameter and
// moment buffers, [n_entries, 2] like the table.
struct TableAdam {
  float2* p;
  float2* m;
  float2* v;
  const float* hyper;
  float b1, b2, eps;
};
// One table entry's Adam state, loaded when the entry's gradient work starts
// so that the loads overlap it.
struct EntryAdam {
  float2 p, m, v;
};
// The moments are touched once per step, so they move with nontemporal loads
// and stores (streamed past the caches, which keep the step's reused data);
// the parameter is stored normally, since the next forward gathers it.
typedef float nt_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 nt_load(const float2* q) {
  const nt_f2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f2*>(q));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void nt_store(float2* q, float2 v) {
  __builtin_nontemporal_store(nt_f2{v.x, v.y}, reinterpret_cast<nt_f2*>(q));
}
__device__ __forceinline__ EntryAdam table_adam_load(const TableAdam& a, uint32_t t) {
  return EntryAdam{nt_load(a.p + t), nt_load(a.m + t), nt_load(a.v + t)};
}
__device__ __forceinline__ void table_adam_store(const TableAdam& a, const AdamCoef& c, uint32_t t, EntryAdam e,
                                                 float2 g) {
  adam_elem(e.p.x, e.m.x, e.v.x, g.x, c);
  adam_elem(e.p.y, e.m.y, e.v.y, g.y, c);
  a.p[t] = e.p;
  nt_store(a.m + t, e.m);
  nt_store(a.v + t, e.v);
}

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

// 8-point DFT in registers: y_q = sum_r v_r W^{rq}, W = exp(DIR 2 pi i / 8).
// Split r = r' + 4s: y_{2q'} is the 4-point DFT of a_r' = v_r' + v_{r'+4};
// y_{2q'+1} the 4-point DFT of (v_r' - v_{r'+4}) W^{r'}.
template <int DIR>
__device__ __forceinline__ void dft4(const float2 x[4], float2 y[4]) {
  const float2 t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]), t2 = cadd(x[1], x[3]), d13 = csub(x[1], x[3]);
  // DIR * i * (x1 - x3)
  const float2 t3 = (DIR < 0) ? make_float2(d13.y, -d13.x) : make_float2(-d13.y, d13.x);
  y[0] = cadd(t0, t2);
  y[1] = cadd(t1, t3);
  y[2] = csub(t0, t2);
  y[3] = csub(t1, t3);
}

template <int DIR>
__device__ __forceinline__ void dft8(const float2 v[8], float2 y[8]) {
  constexpr float h = 0.70710678118654752f;
  float2 a[4], b[4], ya[4], yb[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    a[r] = cadd(v[r], v[r + 4]);
    b[r] = csub(v[r], v[r + 4]);
  }
  // b_r *= W^r: W = (1 + DIR i)/sqrt2, W^2 = DIR i, W^3 = (-1 + DIR i)/sqrt2
  const float s = (float)DIR;
  b[1] = make_float2(h * (b[1].x - s * b[1].y), h * (b[1].y + s * b[1].x));
  b[2] = make_float2(-s * b[2].y, s * b[2].x);
  b[3] = make_float2(h * (-b[3].x - s * b[3].y), h * (-b[3].y + s * b[3].x));
  dft4<DIR>(a, ya);
  dft4<DIR>(b, yb);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    y[2 * q] = ya[q];
    y[2 * q + 1] = yb[q];
  }
}

template <int DIR>
__device__ float2* fft_lines(float2* a, float2* b, int n, int log2n, int lines, int stride,
                             const float2* tw, int tws, int tid, int nthreads) {
  int Ns = 1;
  int p = log2n;
  while (p > 0) {
    if (p >= 3) {
      // radix-8 pass (register butterfly): fewer LDS round trips per line
      const int lgnb = log2n - 3;
      const int nb = 1 << lgnb;
      const int total = nb * lines;
      const int twscale = (n / (Ns * 8)) * tws;
      for (int t = tid; t < total; t += nthreads) {
        const int line = t >> lgnb;
        const int j = t & (nb - 1);
        const float2* src = a + line * stride;
        float2* dst = b + line * stride;
        const int k = j & (Ns - 1);
        const int d = (j - k) * 8 + k;
        float2 v[8], y[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = src[j + r * nb];
        if (Ns > 1) {
          const int ti = k * twscale;
#pragma unroll
          for (int r = 1; r < 8; ++r) v[r] = cmul(v[r], twiddle<DIR>(tw, r * ti));
        }
        dft8<DIR>(v, y);
#pragma unroll
        for (int q = 0; q < 8; ++q) dst[d + q * Ns] = y[q];
      }
      __syncthreads();
      float2* tmp = a; a = b; b = tmp;
      Ns *= 8;
      p -= 3;
      continue;
    }
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

// ---------------------------------------------------------------- loop indices
// (q, r) = (t / d, t % d) along t = t0, t0 + step, ...: one division per
// thread, then additions (GPU integer division is a long VALU sequence).
struct SplitIdx {
  int t, q, r, dq, dr;
  __device__ __forceinline__ SplitIdx(int t0, int step, int d)
      : t(t0), q(t0 / d), r(t0 % d), dq(step / d), dr(step % d) {}
  __device__ __forceinline__ void next(int step, int d) {
    t += step;
    q += dq;
    r += dr;
    if (r >= d) {
      r -= d;
      ++q;
    }
  }
};

// The indices of a Stockham pass loop, (line, j, k) = (t / nb, t % nb,
// (t % nb) % Ns), along t = t0, t0 + step, ...; nb is a multiple of Ns, so a
// wrap of j leaves k unchanged.
struct PassIdx {
  int t, line, j, k, dl, dj, dk;
  __device__ __forceinline__ PassIdx(int t0, int step, int nb, int Ns)
      : t(t0), line(t0 / nb), j(t0 % nb), k((t0 % nb) % Ns), dl(step / nb), dj(step % nb), dk((step % nb) % Ns) {}
  __device__ __forceinline__ void next(int step, int nb, int Ns) {
    t += step;
    line += dl;
    j += dj;
    k += dk;
    if (k >= Ns) k -= Ns;
    if (j >= nb) {
      j -= nb;
      ++line;
    }
  }
};

// ---------------------------------------------------------------- mixed radix
// Radix plan of a transform length: passes of radix 4 first, then 2, 3, 5 and
// any remaining prime factors (<= kMaxRadix).  Built on the host.
constexpr int kMaxPasses = 24;
constexpr int kMaxRadix = 64;

struct FftPlan {
  int n;
  int pow2;                 // 1 when n is a power of two (shift/mask fast path)
  int log2n;
  int n_passes;
  int radix[kMaxPasses];
};

// Host: the radix plan of n (PTY_EUNSUP when n has a prime factor above kMaxRadix).
int make_fft_plan(int n, FftPlan& P);

// One Stockham pass of prime radix R (5 <= R <= 31) over `lines` lines: each
// thread loads its R inputs into registers, applies the stage twiddles and
// forms the R-point DFT from the conjugate pairs a_r = v_r + v_{R-r},
// b_r = v_r - v_{R-r}:  y_q = v_0 + sum_r a_r cos(2 pi rq/R) + i DIR sum_r b_r sin(2 pi rq/R)
// and y_{R-q} the same with the sine term negated (q, r = 1..(R-1)/2).
template <int R, int DIR>
__device__ void fft_pass_prime(const float2* __restrict__ a, float2* __restrict__ b, int n, int Ns, int lines,
                               int stride, const float2* tw, int tws, int tid, int nthreads) {
  constexpr int Hh = (R - 1) / 2;
  const int nb = n / R;
  // cos / sin(2 pi e / R): half of them by sincospi, the rest by symmetry
  // (reading them from the twiddle table instead lets the compiler hoist R
  // loads, 235 VGPRs against 168, one wave per SIMD fewer and slower)
  float cs[R], sn[R];
  cs[0] = 1.f;
  sn[0] = 0.f;
#pragma unroll
  for (int e = 1; e <= Hh; ++e) {
    sincospif(2.0f * (float)e / (float)R, &sn[e], &cs[e]);
    cs[R - e] = cs[e];
    sn[R - e] = -sn[e];
  }
  const int total = nb * lines;
  const int twscale = (n / (Ns * R)) * tws;
  for (PassIdx x(tid, nthreads, nb, Ns); x.t < total; x.next(nthreads, nb, Ns)) {
    const int j = x.j, k = x.k;
    const float2* src = a + x.line * stride;
    float2* dst = b + x.line * stride;
    const int d = (j - k) * R + k;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[j + r * nb];
    if (k > 0) {
      const int ti = k * twscale;
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle<DIR>(tw, r * ti));
    }
    float2 pa[Hh], pb[Hh];
    float2 y0 = v[0];
#pragma unroll
    for (int r = 1; r <= Hh; ++r) {
      pa[r - 1] = cadd(v[r], v[R - r]);
      pb[r - 1] = csub(v[r], v[R - r]);
      y0 = cadd(y0, pa[r - 1]);
    }
    dst[d] = y0;
#pragma unroll
    for (int q = 1; q <= Hh; ++q) {
      float2 tq = v[0], uq = make_float2(0.f, 0.f);
#pragma unroll
      for (int r = 1; r <= Hh; ++r) {
        const int e = (r * q) % R;
        tq.x += pa[r - 1].x * cs[e];
        tq.y += pa[r - 1].y * cs[e];
        uq.x += pb[r - 1].x * sn[e];
        uq.y += pb[r - 1].y * sn[e];
      }
      // i * DIR * u
      const float2 iu = (DIR > 0) ? make_float2(-uq.y, uq.x) : make_float2(uq.y, -uq.x);
      dst[d + q * Ns] = cadd(tq, iu);
      dst[d + (R - q) * Ns] = csub(tq, iu);
    }
  }
}

// Mixed-radix Stockham autosort FFT over `lines` lines in LDS (same contract
// as fft_lines; tw may point to LDS or global memory and holds
// exp(-2 pi i k / (n*tws)), k < n*tws).  Radices 2, 3, 4, 8 have dedicated
// butterflies, primes 5..31 the register pass above; larger primes use a
// direct DFT of the R twiddled inputs.
template <int DIR>
__device__ float2* fft_lines_mixed(float2* a, float2* b, const FftPlan& P, int lines, int stride,
                                   const float2* tw, int tws, int tid, int nthreads) {
  const int n = P.n;
  int Ns = 1;
  for (int ps = 0; ps < P.n_passes; ++ps) {
    const int R = P.radix[ps];
    switch (R) {
      case 5: fft_pass_prime<5, DIR>(a, b, n, Ns, lines, stride, tw, tws, tid, nthreads); goto next;
      case 7: fft_pass_prime<7, DIR>(a, b, n, Ns, lines, stride, tw, tws, tid, nthreads); goto next;
      case 11: fft_pass_prime<11, DIR>(a, b, n, Ns, lines, stride, tw, tws, tid, nthreads); goto next;
      case 13: fft_pass_prime<13, DIR>(a, b, n, Ns, lines, stride, tw, tws, tid, nthreads); goto next;
      case 17: fft_pass_prime<17, DIR>(a, b, n, Ns, lines, stride, tw, tws, tid, nthreads); goto next;
      case 19: fft_pass_prime<19, DIR>(a, b, n, Ns, lines, stride, tw, tws, tid, nthreads); goto next;
      case 23: fft_pass_prime<23, DIR>(a, b, n, Ns, lines, stride, tw, tws, tid, nthreads); goto next;
      case 29: fft_pass_prime<29, DIR>(a, b, n, Ns, lines, stride, tw, tws, tid, nthreads); goto next;
      case 31: fft_pass_prime<31, DIR>(a, b, n, Ns, lines, stride, tw, tws, tid, nthreads); goto next;
      default: break;
    }
    {
      const int nb = n / R;
      const int total = nb * lines;
      const int twscale = (n / (Ns * R)) * tws;
      const int wR = (n / R) * tws;          // twiddle stride of the R-point DFT kernel
      for (PassIdx x(tid, nthreads, nb, Ns); x.t < total; x.next(nthreads, nb, Ns)) {
        const int j = x.j, k = x.k;
        const float2* src = a + x.line * stride;
        float2* dst = b + x.line * stride;
        const int d = (j - k) * R + k;
        if (R == 8) {
          float2 v[8], y[8];
#pragma unroll
          for (int r = 0; r < 8; ++r) v[r] = src[j + r * nb];
          if (k > 0) {
            const int ti = k * twscale;
#pragma unroll
            for (int r = 1; r < 8; ++r) v[r] = cmul(v[r], twiddle<DIR>(tw, r * ti));
          }
          dft8<DIR>(v, y);
#pragma unroll
          for (int q = 0; q < 8; ++q) dst[d + q * Ns] = y[q];
        } else if (R == 4) {
          float2 v0 = src[j], v1 = src[j + nb], v2 = src[j + 2 * nb], v3 = src[j + 3 * nb];
          if (k > 0) {
            const int ti = k * twscale;
            v1 = cmul(v1, twiddle<DIR>(tw, ti));
            v2 = cmul(v2, twiddle<DIR>(tw, 2 * ti));
            v3 = cmul(v3, twiddle<DIR>(tw, 3 * ti));
          }
          float2 t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3), d13 = csub(v1, v3);
          float2 t3 = (DIR < 0) ? make_float2(d13.y, -d13.x) : make_float2(-d13.y, d13.x);
          dst[d] = cadd(t0, t2);
          dst[d + Ns] = cadd(t1, t3);
          dst[d + 2 * Ns] = csub(t0, t2);
          dst[d + 3 * Ns] = csub(t1, t3);
        } else if (R == 2) {
          float2 v0 = src[j], v1 = src[j + nb];
          if (k > 0) v1 = cmul(v1, twiddle<DIR>(tw, k * twscale));
          dst[d] = cadd(v0, v1);
          dst[d + Ns] = csub(v0, v1);
        } else if (R == 3) {
          float2 v0 = src[j], v1 = src[j + nb], v2 = src[j + 2 * nb];
          if (k > 0) {
            const int ti = k * twscale;
            v1 = cmul(v1, twiddle<DIR>(tw, ti));
            v2 = cmul(v2, twiddle<DIR>(tw, 2 * ti));
          }
          // y0 = v0+v1+v2; y1,2 = v0 - (v1+v2)/2 -+ i*DIR*(sqrt3/2)(v1-v2)
          const float2 s12 = cadd(v1, v2), d12 = csub(v1, v2);
          const float2 m = make_float2(v0.x - 0.5f * s12.x, v0.y - 0.5f * s12.y);
          const float c = 0.86602540378443865f * (float)DIR;   // sin(2 pi / 3) * DIR
          const float2 r = make_float2(-c * d12.y, c * d12.x);  // i * c * d12
          dst[d] = cadd(v0, s12);
          dst[d + Ns] = cadd(m, r);
          dst[d + 2 * Ns] = csub(m, r);
        } else {
          for (int q = 0; q < R; ++q) {
            float2 acc = make_float2(0.f, 0.f);
            int e = 0;                         // (r * q) mod R, stepped
            for (int r = 0; r < R; ++r) {
              float2 v = src[j + r * nb];
              if (k > 0 && r > 0) v = cmul(v, twiddle<DIR>(tw, r * k * twscale));
              if (e) v = cmul(v, twiddle<DIR>(tw, e * wR));
              acc = cadd(acc, v);
              e += q;
              if (e >= R) e -= R;
            }
            dst[d + q * Ns] = acc;
          }
        }
      }
    }
  next:
    __syncthreads();
    float2* tmp = a; a = b; b = tmp;
    Ns *= R;
  }
  return a;
}

template <int DIR>
__device__ __forceinline__ float2* fft_any(float2* a, float2* b, const FftPlan& P, int lines, int stride,
                                           const float2* tw, int tws, int tid, int nthreads) {
  if (P.pow2) return fft_lines<DIR>(a, b, P.n, P.log2n, lines, stride, tw, tws, tid, nthreads);
  return fft_lines_mixed<DIR>(a, b, P, lines, stride, tw, tws, tid, nthreads);
}

// Frequency k of an n-point spectrum in the ds-times larger m-point spectrum
// of the reference upsample: fftshift (roll n//2), zero pad (m-n)//2 before,
// ifftshift (roll -(m//2)).  For even n this is k (k < n/2) or k - n + m.
__device__ __forceinline__ int embed_freq(int k, int n, int m) {
  int q = (k + n / 2) % n + (m - n) / 2 - m / 2;
  q %= m;
  return q < 0 ? q + m : q;
}

// Twiddles exp(-2 pi i k / n), k < n.  For n divisible by 8 every entry comes
// from one sincospi in the first octant, theta0 = 2 pi k0 / n with
// 0 <= k0 <= n/8, by exact symmetries (with C = cos theta0, S = sin theta0):
//   k = k0: (C, -S)      n/4 - k0: (S, -C)     n/4 + k0: (-S, -C)
//   n/2 - k0: (-C, -S)   n/2 + k0: (-C, S)     3n/4 - k0: (-S, C)
//   3n/4 + k0: (S, C)    n - k0: (C, S)
// so a table costs n/8 + 1 sincospi and twiddle_value(k, n) gives the same
// bits as fill_twiddles' entry k.
__device__ __forceinline__ float2 octant_twiddle(int oct, float C, float S) {
  switch (oct) {
    case 0: return make_float2(C, -S);
    case 1: return make_float2(S, -C);
    case 2: return make_float2(-S, -C);
    case 3: return make_float2(-C, -S);
    case 4: return make_float2(-C, S);
    case 5: return make_float2(-S, C);
    case 6: return make_float2(S, C);
    default: return make_float2(C, S);
  }
}

__device__ __forceinline__ float2 twiddle_value(int k, int n) {
  float s, c;
  if (n % 8) {
    sincospif(-2.0f * (float)k / (float)n, &s, &c);
    return make_float2(c, s);
  }
  const int q = n / 8, oct = k / q, r = k - oct * q;
  const int k0 = (oct & 1) ? q - r : r;
  sincospif(2.0f * (float)k0 / (float)n, &s, &c);
  return octant_twiddle(oct, c, s);
}

__device__ __forceinline__ void fill_twiddles(float2* tw, int n, int tid, int nthreads) {
  if (n % 8) {
    for (int k = tid; k < n; k += nthreads) tw[k] = twiddle_value(k, n);
    return;
  }
  const int q = n / 8;
  for (int k0 = tid; k0 <= q; k0 += nthreads) {
    float S, C;
    sincospif(2.0f * (float)k0 / (float)n, &S, &C);
    // octants with k = oct q + k0 (even oct, k0 < q) and (oct + 1) q - k0 (odd oct, k0 > 0)
#pragma unroll
    for (int oct = 0; oct < 8; oct += 2) {
      if (k0 < q) tw[oct * q + k0] = octant_twiddle(oct, C, S);
      if (k0 > 0) tw[(oct + 2) * q - k0] = octant_twiddle(oct + 1, C, S);
    }
  }
}

__host__ __device__ __forceinline__ int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}
__host__ __device__ __forceinline__ bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

}  // namespace pty
