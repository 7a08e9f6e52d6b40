// Register-resident FFT of one wavefront (gfx950, 64 lanes) for lines of
// N = 64 E points, E in {1, 2, 4, 8}: a wavefront transforms L = 8 / E lines
// at once, 8 complex values per lane.
//
// With n = n1 + 64 n2 (lane n1 holds n2 = 0 .. E-1 of each of its lines) and
// k = k1 + E k2:
//   X[k1 + E k2] = sum_n1 W_64^{n1 k2} W_N^{n1 k1} sum_n2 W_E^{n2 k1} x[n1 + 64 n2]
// step 1: an E-point DFT over n2 in registers, times W_N^{n1 k1};
// step 2: for each of the 8 columns (line, k1), a 64-point DFT across the
//         lanes, as 8 x 8: n1 = a + 8 b, k2 = c + 8 d,
//         W_64^{n1 k2} = W_8^{b c} W_64^{a c} W_8^{a d},
//         i.e. two exchanges through the wavefront's own LDS slice (padded so
//         that every access is at most 2- or 3-way bank conflicted), each
//         followed by an 8-point DFT in registers.
// On return lane l holds X_line[(l & (8E - 1)) + 8 E d] for d = 0 .. 7 with
// line = l >> (3 + log2 E): stores of one d are contiguous across lanes.
// The exchanges are wavefront-synchronous: no workgroup barrier.
#pragma once
#include "common.h"

namespace pty {

// exp(-2 pi i k / 1024) for k < 1024, built at compile time in double
// precision (octant-reduced Taylor series) and rounded to float once.
struct TwTable {
  float re[1024];
  float im[1024];
};

constexpr double cx_pi = 3.14159265358979323846264338327950288;

// sin and cos of x in [0, pi/4] by Taylor series (|error| < 1e-17)
constexpr void cx_sincos_small(double x, double& s, double& c) {
  double term_s = x, term_c = 1.0, sum_s = 0.0, sum_c = 0.0;
  const double x2 = x * x;
  for (int i = 0; i < 14; ++i) {
    sum_s += term_s;
    sum_c += term_c;
    term_s *= -x2 / ((2.0 * i + 2.0) * (2.0 * i + 3.0));
    term_c *= -x2 / ((2.0 * i + 1.0) * (2.0 * i + 2.0));
  }
  s = sum_s;
  c = sum_c;
}

constexpr TwTable make_tw1024() {
  TwTable t{};
  for (int k = 0; k < 1024; ++k) {
    const int quad = k / 256, m = k % 256;   // angle = quad * pi/2 + 2 pi m / 1024
    double s = 0.0, c = 0.0;
    if (m <= 128) {
      cx_sincos_small(2.0 * cx_pi * m / 1024.0, s, c);
    } else {
      double s2 = 0.0, c2 = 0.0;
      cx_sincos_small(2.0 * cx_pi * (256 - m) / 1024.0, s2, c2);
      s = c2;
      c = s2;
    }
    // rotate (c, s) by quad quarter turns
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

__constant__ constexpr TwTable kTw1024 = make_tw1024();

// exp(DIR 2 pi i m / 1024)
template <int DIR>
__device__ __forceinline__ float2 tw1024(int m) {
  const float re = kTw1024.re[m], im = kTw1024.im[m];
  return make_float2(re, DIR < 0 ? im : -im);
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kWaveFftScratch = 576;   // float2 per wavefront (8 columns x 72)

// One set of forward twiddles per lane serves both directions: the inverse
// transform multiplies by their conjugates.
template <int E>
struct WaveFft {
  static constexpr int N = 64 * E;
  static constexpr int L = 8 / E;
  static constexpr int LG_E = E == 1 ? 0 : E == 2 ? 1 : E == 4 ? 2 : 3;
  float2 t1[E];   // W_N^{lane k1}, W = exp(-2 pi i / N)
  float2 t2[8];   // W_64^{(lane & 7) c}

  __device__ __forceinline__ void init(int lane) {
#pragma unroll
    for (int k1 = 0; k1 < E; ++k1) t1[k1] = tw1024<-1>(((lane * k1) & (N - 1)) * (1024 / N));
#pragma unroll
    for (int c = 0; c < 8; ++c) t2[c] = tw1024<-1>((((lane & 7) * c) & 63) * 16);
  }

  // x * w (DIR < 0) or x * conj(w) (DIR > 0)
  template <int DIR>
  __device__ __forceinline__ static float2 rot(float2 x, float2 w) {
    return DIR < 0 ? cmul(x, w) : cmulc(x, w);
  }

  // E-point DFT of x[0..E) in place
  template <int DIR>
  __device__ __forceinline__ static void dft_e(float2* x) {
    if (E == 8) {
      float2 y[8];
      dft8<DIR>(x, y);
#pragma unroll
      for (int i = 0; i < 8; ++i) x[i] = y[i];
    } else if (E == 4) {
      float2 y[4];
      dft4<DIR>(x, y);
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = y[i];
    } else if (E == 2) {
      const float2 u = x[0], w = x[1];
      x[0] = cadd(u, w);
      x[1] = csub(u, w);
    }
  }

  // DIR = -1 forward, +1 inverse (unnormalised).  v[line E + n2] =
  // x_line[lane + 64 n2] on entry; see the header comment for the output
  // layout.  s: this wavefront's kWaveFftScratch float2 of LDS.
  template <int DIR>
  __device__ __forceinline__ void run(float2 v[8], float2* s, int lane) const {
#pragma unroll
    for (int l = 0; l < L; ++l) dft_e<DIR>(&v[l * E]);
#pragma unroll
    for (int col = 0; col < 8; ++col) {
      const int k1 = col & (E - 1);
      if (k1) v[col] = rot<DIR>(v[col], t1[k1]);
    }
    const int a = lane & 7, hi = lane >> 3;
    // exchange 1: lane (a + 8 b) holds T[col][a + 8 b]; lane (a + 8 col) takes b = 0..7
#pragma unroll
    for (int col = 0; col < 8; ++col) s[col * 72 + a * 9 + hi] = v[col];
    wave_lds_sync();
    float2 u[8], w[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) u[b] = s[hi * 72 + a * 9 + b];
    wave_lds_sync();
    dft8<DIR>(u, w);
#pragma unroll
    for (int c = 1; c < 8; ++c) w[c] = rot<DIR>(w[c], t2[c]);
    // exchange 2: lane (a + 8 col) holds U[col][a][c]; the reader of (col, c) takes a = 0..7
#pragma unroll
    for (int c = 0; c < 8; ++c) s[hi * 72 + c * 9 + a] = w[c];
    wave_lds_sync();
    const int rem = lane & (8 * E - 1);
    const int col2 = (lane >> (3 + LG_E)) * E + (rem & (E - 1));
    const int c2 = rem >> LG_E;
#pragma unroll
    for (int aa = 0; aa < 8; ++aa) u[aa] = s[col2 * 72 + c2 * 9 + aa];
    wave_lds_sync();
    dft8<DIR>(u, v);
  }

  // line and first element of lane's outputs (element of output d = first + 8 E d)
  __device__ __forceinline__ static int out_line(int lane) { return lane >> (3 + LG_E); }
  __device__ __forceinline__ static int out_first(int lane) { return lane & (8 * E - 1); }
};

}  // namespace pty
