// Register-resident DFTs of compile-time length, for one lane's own line of
// data: every loop is unrolled and every array index is a constant after
// unrolling, so the arrays stay in VGPRs, and the twiddle factors are
// compile-time constants (tables built in double precision, rounded once).
//   dft_reg<N, DIR>(x): x <- DFT_N(x) in natural order, DIR = -1 forward
//   (exp(-2 pi i n k / N)), +1 inverse (unnormalised).
// Lengths: powers of two up to 64 (radix 4 x (N/4) over dft4 / dft8), odd
// primes (3 by its closed form, 5 .. 31 by the conjugate-pair form), and
// products N = R * S of two of those (one twiddle stage between them).
#pragma once
#include "common.h"
#include "fft_wave.h"

namespace pty {

// exp(-2 pi i k / N), k < N
template <int N>
struct CTw {
  float re[N];
  float im[N];
};

template <int N>
constexpr CTw<N> make_ctw() {
  CTw<N> t{};
  for (int k = 0; k < N; ++k) {
    // angle 2 pi k / N = quad * pi/2 + (pi/2) r / N with 4k = quad N + r
    const int quad = (4 * k) / N, r = 4 * k - quad * N;
    double s = 0.0, c = 0.0;
    if (2 * r <= N) {
      cx_sincos_small(0.5 * cx_pi * r / N, s, c);
    } else {
      double s2 = 0.0, c2 = 0.0;
      cx_sincos_small(0.5 * cx_pi * (N - r) / N, s2, c2);
      s = c2;
      c = s2;
    }
    double cr = c, sr = s;
    for (int q = 0; q < quad; ++q) {
      const double tmp = cr;
      cr = -sr;
      sr = tmp;
    }
    t.re[k] = (float)cr;
    t.im[k] = (float)(-sr);
  }
  return t;
}

// the same table as (re, im) pairs, read as float2
template <int N>
struct CTwPairs {
  float v[2 * N];
};

template <int N>
constexpr CTwPairs<N> make_ctw_pairs() {
  const CTw<N> t = make_ctw<N>();
  CTwPairs<N> p{};
  for (int k = 0; k < N; ++k) {
    p.v[2 * k] = t.re[k];
    p.v[2 * k + 1] = t.im[k];
  }
  return p;
}

template <int N>
constexpr bool is_pow2_c() { return N > 0 && (N & (N - 1)) == 0; }

// x * exp(DIR 2 pi i e / N) for a compile-time table entry e
template <int N, int DIR>
__device__ __forceinline__ float2 rot_ct(float2 x, const CTw<N>& tw, int e) {
  if (e == 0) return x;
  if (4 * e == N) return DIR < 0 ? make_float2(x.y, -x.x) : make_float2(-x.y, x.x);
  if (2 * e == N) return make_float2(-x.x, -x.y);
  if (4 * e == 3 * N) return DIR < 0 ? make_float2(-x.y, x.x) : make_float2(x.y, -x.x);
  const float2 w = make_float2(tw.re[e], DIR < 0 ? tw.im[e] : -tw.im[e]);
  return cmul(x, w);
}

template <int N, int DIR>
__device__ __forceinline__ void dft_reg(float2 (&x)[N]);

// 3 points: y0 = x0 + x1 + x2, y1,2 = x0 - (x1 + x2)/2 -+ i DIR (sqrt3/2)(x1 - x2)
template <int DIR>
__device__ __forceinline__ void dft_reg3(float2 (&x)[3]) {
  const float2 s12 = cadd(x[1], x[2]), d12 = csub(x[1], x[2]);
  const float2 mid = make_float2(x[0].x - 0.5f * s12.x, x[0].y - 0.5f * s12.y);
  const float c = 0.86602540378443865f * (float)DIR;
  const float2 r = make_float2(-c * d12.y, c * d12.x);
  x[0] = cadd(x[0], s12);
  x[1] = cadd(mid, r);
  x[2] = csub(mid, r);
}

// odd prime R >= 5: a_r = x_r + x_{R-r}, b_r = x_r - x_{R-r} (r = 1 .. (R-1)/2),
//   y_q = x_0 + sum_r a_r cos(2 pi r q / R) + i DIR sum_r b_r sin(2 pi r q / R),
//   y_{R-q} = the same with the sine term negated.
template <int R, int DIR>
__device__ __forceinline__ void dft_reg_prime(float2 (&x)[R]) {
  constexpr int H = (R - 1) / 2;
  constexpr CTw<R> tw = make_ctw<R>();
  float2 a[H], b[H];
  float2 y0 = x[0];
#pragma unroll
  for (int r = 1; r <= H; ++r) {
    a[r - 1] = cadd(x[r], x[R - r]);
    b[r - 1] = csub(x[r], x[R - r]);
    y0 = cadd(y0, a[r - 1]);
  }
  const float2 x0 = x[0];
  x[0] = y0;
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    float2 t = x0, u = make_float2(0.f, 0.f);
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      const int e = (r * q) % R;
      const float cs = tw.re[e], sn = -tw.im[e];   // cos, sin of 2 pi e / R
      t.x += a[r - 1].x * cs;
      t.y += a[r - 1].y * cs;
      u.x += b[r - 1].x * sn;
      u.y += b[r - 1].y * sn;
    }
    const float2 iu = DIR > 0 ? make_float2(-u.y, u.x) : make_float2(u.y, -u.x);   // i DIR u
    x[q] = cadd(t, iu);
    x[R - q] = csub(t, iu);
  }
}

// N = R S (n = S n1 + n2, k = k1 + R k2):
//   X[k1 + R k2] = sum_n2 W_S^{n2 k2} W_N^{n2 k1} sum_n1 W_R^{n1 k1} x[S n1 + n2]
template <int R, int S, int DIR>
__device__ __forceinline__ void dft_reg_split(float2 (&x)[R * S]) {
  constexpr int N = R * S;
  constexpr CTw<N> tw = make_ctw<N>();
  float2 u[N];
#pragma unroll
  for (int n2 = 0; n2 < S; ++n2) {
    float2 v[R];
#pragma unroll
    for (int n1 = 0; n1 < R; ++n1) v[n1] = x[S * n1 + n2];
    dft_reg<R, DIR>(v);
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) u[k1 * S + n2] = rot_ct<N, DIR>(v[k1], tw, (n2 * k1) % N);
  }
#pragma unroll
  for (int k1 = 0; k1 < R; ++k1) {
    float2 v[S];
#pragma unroll
    for (int n2 = 0; n2 < S; ++n2) v[n2] = u[k1 * S + n2];
    dft_reg<S, DIR>(v);
#pragma unroll
    for (int k2 = 0; k2 < S; ++k2) x[k1 + R * k2] = v[k2];
  }
}

template <int N, int DIR>
__device__ __forceinline__ void dft_reg(float2 (&x)[N]) {
  if constexpr (N == 1) {
    return;
  } else if constexpr (N == 2) {
    const float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  } else if constexpr (N == 3) {
    dft_reg3<DIR>(x);
  } else if constexpr (N == 4) {
    float2 y[4];
    dft4<DIR>(x, y);
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = y[i];
  } else if constexpr (N == 8) {
    float2 y[8];
    dft8<DIR>(x, y);
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = y[i];
  } else if constexpr (is_pow2_c<N>()) {
    static_assert(N <= 64, "power-of-two register DFTs up to 64 points");
    dft_reg_split<4, N / 4, DIR>(x);
  } else if constexpr (N == 5 || N == 7 || N == 11 || N == 13 || N == 17 || N == 19 || N == 23 || N == 29 ||
                       N == 31) {
    dft_reg_prime<N, DIR>(x);
  } else if constexpr (N % 3 == 0) {
    dft_reg_split<3, N / 3, DIR>(x);
  } else if constexpr (N % 5 == 0) {
    dft_reg_split<5, N / 5, DIR>(x);
  } else {
    static_assert(N % 7 == 0, "register DFT length not supported");
    dft_reg_split<7, N / 7, DIR>(x);
  }
}

}  // namespace pty
