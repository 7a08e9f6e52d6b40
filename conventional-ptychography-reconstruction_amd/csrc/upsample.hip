// Fourier upsample (models/complex_inr.py:252-271) and its exact adjoint.
//
// Forward: big = ds^2 * ifft2_m(ifftshift(pad(fftshift(fft2_n(small))))), m = ds*n.
// The shift/pad/unshift chain maps small-grid frequency k to big-grid
// frequency q(k) = ((k + n//2) % n + (m-n)//2 - m//2) mod m (embed_freq; for
// even n: k below n/2, k - n + m above, Nyquist on the negative side; for odd
// n the DC bin lands on q = m-1 exactly as in the reference); every other big
// frequency is zero.  The map is separable: with the per-line operator
// U = IFFT_m o embed_q o FFT_n (n samples in, m out),
//   big = U_rows(U_cols(small)) / n^2,
// and its adjoint (what autograd of the reference computes, net scale 1) is
//   gsmall = V_rows(V_cols(gbig)) / n^2,  V = IFFT_n o crop_q o FFT_m.
// Each stage is one pass of batched LDS FFTs over contiguous rows (mixed
// radix: any n whose prime factors are <= 64), so columns are handled by
// LDS-tiled transposes around a row pass:
//   forward: T(small) -> U rows -> T -> U rows (scaled) -> big
//   adjoint: gbig -> V rows -> T -> V rows (scaled) -> T -> gsmall
// Twiddles (exp(-2 pi i k/m), the n-point transforms read every ds-th)
// live in LDS when the line buffers leave room, else in a global table in
// the workspace (filled at the start of every call).
//
// ds = 2 with n even (both workloads) has a cheaper exact form of U: the
// embedded spectrum is q = s(k), the signed frequency (s = k below n/2,
// k - n from n/2 on, Nyquist negative), so with m = 2n
//   U x [2j]     = sum_k X_k e^{2 pi i k j / n}                 = n x_j
//   U x [2j + 1] = sum_k X_k e^{2 pi i s(k) / m} e^{2 pi i k j / n}
// i.e. the even samples are the input and the odd ones the half-sample
// shift IFFT_n(phi X), phi_k = e^{2 pi i s(k)/m}; and V = U^H gives
//   V g [j] = n g_{2j} + IFFT_n(conj(phi) FFT_n(g_odd))_j.
// Two n-point transforms per line instead of an n- and an m-point one, and
// line buffers of n instead of m (k_up2_rows).
#include "common.h"
#include "fft_reg.h"
#include "fft_wave.h"

namespace pty {

constexpr size_t kLdsBudget = 65536;

// inverse of embed_freq on its image: small frequency of big frequency q, or -1
__device__ __forceinline__ int crop_freq(int q, int n, int m) {
  // q = (s + pad - m/2) mod m with s = (k + n/2) mod n in [0, n)
  int s = (q + m / 2 - (m - n) / 2) % m;
  if (s < 0) s += m;
  if (s >= n) return -1;
  int k = (s - n / 2) % n;
  return k < 0 ? k + n : k;
}

__global__ void k_twiddle_table(float2* tw, int n) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    tw[k] = twiddle_value(k, n);
}

// out [C, R] = in [R, C]^T (c64), 32 x 32 tiles through LDS; grid
// (ceil(C/32), ceil(R/32)), 256 threads.
__global__ __launch_bounds__(256) void k_transpose(const float2* __restrict__ in, float2* __restrict__ out, int R,
                                                   int C) {
  __shared__ float2 tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + tx;
    if (r < R && c < C) tile[k][tx] = in[(int64_t)r * C + c];
  }
  __syncthreads();
#pragma unroll
  for (int k = ty; k < 32; k += 8) {
    const int orow = c0 + k, ocol = r0 + tx;
    if (orow < C && ocol < R) out[(int64_t)orow * R + ocol] = tile[tx][k];
  }
}

// One upsample stage over `rows` contiguous lines, `lines` per workgroup:
//   MODE 0 (U): in rows of n -> out rows of m: IFFT_m(embed(FFT_n(row))) * scale
//   MODE 1 (V): in rows of m -> out rows of n: IFFT_n(crop(FFT_m(row))) * scale
// LDS: [twiddles m] + two ping-pong buffers of lines x m.
template <int MODE>
__global__ __launch_bounds__(256) void k_up_rows(const float2* __restrict__ in, float2* __restrict__ out, FftPlan Pn,
                                                 FftPlan Pm, int rows, int lines, float scale,
                                                 const float2* __restrict__ gtw) {
  extern __shared__ float2 smem[];
  const int n = Pn.n, m = Pm.n, ds = m / n;
  const int tid = threadIdx.x, nt = blockDim.x;
  const float2* tw;
  float2* a;
  if (gtw) {
    tw = gtw;
    a = smem;
  } else {
    fill_twiddles(smem, m, tid, nt);
    tw = smem;
    a = smem + m;
  }
  float2* b = a + lines * m;
  const int r0 = blockIdx.x * lines;
  const int nl = min(lines, rows - r0);
  const int Lin = (MODE == 0) ? n : m;
  for (SplitIdx x(tid, nt, Lin); x.t < lines * Lin; x.next(nt, Lin)) {
    const int li = x.q, c = x.r;
    a[li * m + c] = (li < nl) ? in[(int64_t)(r0 + li) * Lin + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  float2* res;
  if (MODE == 0) {
    float2* f = fft_any<-1>(a, b, Pn, lines, m, tw, ds, tid, nt);
    float2* g = (f == a) ? b : a;
    for (SplitIdx x(tid, nt, m); x.t < lines * m; x.next(nt, m)) {
      const int li = x.q, q = x.r;
      const int k = crop_freq(q, n, m);
      g[li * m + q] = (k >= 0) ? f[li * m + k] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    res = fft_any<+1>(g, f, Pm, lines, m, tw, 1, tid, nt);
  } else {
    float2* f = fft_any<-1>(a, b, Pm, lines, m, tw, 1, tid, nt);
    float2* g = (f == a) ? b : a;
    for (SplitIdx x(tid, nt, n); x.t < lines * n; x.next(nt, n)) {
      const int li = x.q, k = x.r;
      g[li * m + k] = f[li * m + embed_freq(k, n, m)];
    }
    __syncthreads();
    res = fft_any<+1>(g, f, Pn, lines, m, tw, ds, tid, nt);
  }
  const int Lout = (MODE == 0) ? m : n;
  for (SplitIdx x(tid, nt, Lout); x.t < lines * Lout; x.next(nt, Lout)) {
    const int li = x.q, c = x.r;
    if (li < nl) out[(int64_t)(r0 + li) * Lout + c] = cscale(res[li * m + c], scale);
  }
}

// ds = 2, n even: the half-sample-shift form above over `rows` lines, `lines`
// per workgroup, with the same scale convention as k_up_rows.
//   MODE 0: in rows of n -> out rows of m: out[2j] = n x_j, out[2j+1] = IFFT_n(phi FFT_n(x))_j
//   MODE 1: in rows of m -> out rows of n: out[j] = n g_{2j} + IFFT_n(conj(phi) FFT_n(g_odd))_j
// tw: exp(-2 pi i k/m), k < m (LDS copy when gtw is null).
template <int MODE>
__global__ __launch_bounds__(256) void k_up2_rows(const float2* __restrict__ in, float2* __restrict__ out, FftPlan Pn,
                                                  int rows, int lines, float scale, const float2* __restrict__ gtw) {
  extern __shared__ float2 smem[];
  const int n = Pn.n, m = 2 * n, h = n / 2;
  const int tid = threadIdx.x, nt = blockDim.x;
  const float2* tw;
  float2* a;
  if (gtw) {
    tw = gtw;
    a = smem;
  } else {
    fill_twiddles(smem, m, tid, nt);
    tw = smem;
    a = smem + m;
  }
  float2* b = a + lines * n;
  const int r0 = blockIdx.x * lines;
  const int nl = min(lines, rows - r0);
  for (SplitIdx x(tid, nt, n); x.t < lines * n; x.next(nt, n)) {
    const int li = x.q, j = x.r;
    float2 v = make_float2(0.f, 0.f);
    if (li < nl) v = (MODE == 0) ? in[(int64_t)(r0 + li) * n + j] : in[(int64_t)(r0 + li) * m + 2 * j + 1];
    a[x.t] = v;
  }
  __syncthreads();
  float2* f = fft_any<-1>(a, b, Pn, lines, n, tw, 2, tid, nt);
  for (SplitIdx x(tid, nt, n); x.t < lines * n; x.next(nt, n)) {
    const int k = x.r;
    const int sk = k < h ? k : k - n;          // signed frequency
    const float2 w = tw[sk < 0 ? sk + m : sk];  // exp(-2 pi i s/m) = conj(phi)
    f[x.t] = cmul(f[x.t], (MODE == 0) ? make_float2(w.x, -w.y) : w);
  }
  __syncthreads();
  float2* g = (f == a) ? b : a;
  float2* res = fft_any<+1>(f, g, Pn, lines, n, tw, 2, tid, nt);
  const float ns = (float)n * scale;
  for (SplitIdx x(tid, nt, n); x.t < lines * n; x.next(nt, n)) {
    const int li = x.q, j = x.r;
    if (li >= nl) continue;
    const float2 r = res[x.t];
    if (MODE == 0) {
      const float2 x = in[(int64_t)(r0 + li) * n + j];
      reinterpret_cast<float4*>(out + (int64_t)(r0 + li) * m)[j] =
          make_float4(x.x * ns, x.y * ns, r.x * scale, r.y * scale);
    } else {
      const float2 e = in[(int64_t)(r0 + li) * m + 2 * j];
      out[(int64_t)(r0 + li) * n + j] = make_float2(e.x * ns + r.x * scale, e.y * ns + r.y * scale);
    }
  }
}

// ---------------------------------------------------------------- host side
int make_fft_plan(int n, FftPlan& P) {
  P = FftPlan{};
  P.n = n;
  P.pow2 = is_pow2(n) ? 1 : 0;
  P.log2n = P.pow2 ? ilog2(n) : 0;
  int r = n, k = 0;
  auto push = [&](int R) { P.radix[k++] = R; r /= R; };
  while (r % 8 == 0 && k < kMaxPasses) push(8);
  while (r % 4 == 0 && k < kMaxPasses) push(4);
  while (r % 2 == 0 && k < kMaxPasses) push(2);
  for (int p = 3; p <= kMaxRadix && r > 1; p += 2)
    while (r % p == 0 && k < kMaxPasses) push(p);
  P.n_passes = k;
  return r == 1 ? PTY_OK : PTY_EUNSUP;
}

// k_up2_rows for n = 64 E (E in 1, 2, 4, 8; canvases 128 .. 1024 at ds = 2):
// each wavefront takes L = 8 / E rows per job in registers (fft_wave.h),
// jobs strided over the grid.  Forward transform, the half-sample phase of
// each frequency, inverse transform; for E < 8 the wavefront's LDS slice
// brings the transform outputs back to the input order (lane + 64 n2) before
// the phase-and-inverse and before the store (for E = 8 the two orders are
// the same).  Output and scales as k_up2_rows.
template <int E>
__device__ __forceinline__ void wave_reorder(float2 v[8], float2* s, int lane) {
  using F = WaveFft<E>;
  if (E == 8) return;
  const int line = F::out_line(lane), first = F::out_first(lane);
#pragma unroll
  for (int d = 0; d < 8; ++d) s[line * F::N + first + 8 * E * d] = v[d];
  wave_lds_sync();
#pragma unroll
  for (int l = 0; l < F::L; ++l)
#pragma unroll
    for (int n2 = 0; n2 < E; ++n2) v[l * E + n2] = s[l * F::N + lane + 64 * n2];
  wave_lds_sync();
}

template <int E, int MODE>
__global__ __launch_bounds__(256) void k_up2_rows_w(const float2* __restrict__ in, float2* __restrict__ out,
                                                    int rows, float scale) {
  using F = WaveFft<E>;
  constexpr int n = F::N, m = 2 * n, h = n / 2;
  __shared__ float2 scratch[4][kWaveFftScratch];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  float2* s = scratch[wv];
  F fft;
  fft.init(lane);
  const float ns = (float)n * scale;
  const int jobs = rows / F::L;
  for (int j = blockIdx.x * 4 + wv; j < jobs; j += gridDim.x * 4) {
    const int r0 = j * F::L;
    float2 v[8];
#pragma unroll
    for (int l = 0; l < F::L; ++l)
#pragma unroll
      for (int n2 = 0; n2 < E; ++n2) {
        const int k = lane + 64 * n2;
        v[l * E + n2] = (MODE == 0) ? in[(int64_t)(r0 + l) * n + k] : in[(int64_t)(r0 + l) * m + 2 * k + 1];
      }
    fft.template run<-1>(v, s, lane);
    wave_reorder<E>(v, s, lane);
#pragma unroll
    for (int l = 0; l < F::L; ++l)
#pragma unroll
      for (int n2 = 0; n2 < E; ++n2) {
        const int k = lane + 64 * n2;
        const int sk = k < h ? k : k - n;                         // signed frequency
        const float2 w = tw1024<-1>(((sk + m) & (m - 1)) * (1024 / m));   // exp(-2 pi i s / m)
        v[l * E + n2] = (MODE == 0) ? cmulc(v[l * E + n2], w) : cmul(v[l * E + n2], w);
      }
    fft.template run<+1>(v, s, lane);
    wave_reorder<E>(v, s, lane);
#pragma unroll
    for (int l = 0; l < F::L; ++l)
#pragma unroll
      for (int n2 = 0; n2 < E; ++n2) {
        const int k = lane + 64 * n2;
        const int64_t r = r0 + l;
        const float2 res = v[l * E + n2];
        if (MODE == 0) {
          const float2 x = in[r * n + k];
          reinterpret_cast<float4*>(out + r * m)[k] = make_float4(x.x * ns, x.y * ns, res.x * scale, res.y * scale);
        } else {
          const float2 e = in[r * m + 2 * k];
          out[r * n + k] = make_float2(e.x * ns + res.x * scale, e.y * ns + res.y * scale);
        }
      }
  }
}

// k_up2_rows for n = P Q, P a power of two (16 or 32), Q = 57 (n = 912,
// 1824: the INR grids of canvases 1824 and 3648), on register DFTs.  One
// wavefront per workgroup takes L lines (L P <= 64; one line at n = 1824, so
// that the 1824-line pass launches 1824 wavefronts rather than 912: the
// kernel's 183 VGPRs allow two per SIMD, and with two lines per wavefront it
// ran at 0.7 per SIMD) per job (n = Q n1 + n2 =
// m1 + Q m2, k = k1 + P k2):
//   pass A : lane n2 < Q of each line loads x[Q n1 + n2] (n1 < P; coalesced
//            across the lanes), P-point DFT over n1, times W_n^{n2 k1}, into
//            the line's LDS slot at Q k1 + n2;
//   pass B : lane (line, k1) < L P takes the Q values at Q k1 .. Q k1 + Q - 1:
//            Q-point DFT (X[k1 + P k2] for k2 < Q), the half-sample phase of
//            each frequency, inverse Q-point DFT, times W_n^{-m1 k1}, back to
//            the same Q slots;
//   pass A': lane m1 < Q reads Q k1 + m1 (k1 < P), inverse P-point DFT: the
//            shifted line at m1 + Q m2 (m2 < P), stored with the even samples
//            (coalesced across the lanes).
// The exchanges are wavefront-synchronous (no workgroup barrier); LDS is
// L n float2 (14.6 KB at n = 1824, L = 1).  Twiddles exp(-2 pi i k / m), k < m = 2n, come from a
// table built at compile time (double precision, rounded once).
// Output and scales as k_up2_rows.
__constant__ constexpr CTwPairs<3648> kTwUp3648 = make_ctw_pairs<3648>();
__constant__ constexpr CTwPairs<1824> kTwUp1824 = make_ctw_pairs<1824>();

template <int P, int Q, int L, int MODE>
__global__ __launch_bounds__(64) void k_up2_rows_r(const float2* __restrict__ in, float2* __restrict__ out,
                                                   int rows, float scale) {
  constexpr int n = P * Q, m = 2 * n, h = n / 2;
  static_assert(P <= 64 && L * P <= 64 && Q <= 64, "k_up2_rows_r: L P lanes at most 64, Q at most 64");
  static_assert(m == 3648 || m == 1824, "k_up2_rows_r: twiddle tables for n = 912, 1824");
  const float2* tw = reinterpret_cast<const float2*>(m == 3648 ? kTwUp3648.v : kTwUp1824.v);
  __shared__ float2 buf[L * n];
  const int lane = threadIdx.x;
  const int r0 = blockIdx.x * L;
  const float ns = (float)n * scale;
  // pass A
  for (int li = 0; li < L; ++li) {
    const int64_t row = r0 + li;
    if (lane < Q) {
      float2 v[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) {
        const int j = Q * n1 + lane;
        v[n1] = row < rows ? (MODE == 0 ? in[row * n + j] : in[row * m + 2 * j + 1]) : make_float2(0.f, 0.f);
      }
      dft_reg<P, -1>(v);
      float2* dst = buf + li * n + lane;
      int e = 0;   // lane k1 mod n
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) {
        dst[Q * k1] = k1 ? cmul(v[k1], tw[2 * e]) : v[k1];
        e += lane;
        if (e >= n) e -= n;
      }
    }
  }
  wave_lds_sync();
  // pass B
  if (lane < L * P) {
    const int k1 = lane % P;
    float2* slot = buf + (lane / P) * n + Q * k1;
    float2 x[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) x[i] = slot[i];
    dft_reg<Q, -1>(x);
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) {
      const int k = k1 + P * k2;
      const int sk = k < h ? k : k - n;                    // signed frequency
      const float2 w = tw[sk < 0 ? sk + m : sk];           // exp(-2 pi i s / m) = conj(phi)
      x[k2] = (MODE == 0) ? cmulc(x[k2], w) : cmul(x[k2], w);
    }
    dft_reg<Q, +1>(x);
    int e = 0;   // m1 k1 mod n
#pragma unroll
    for (int m1 = 0; m1 < Q; ++m1) {
      slot[m1] = m1 ? cmulc(x[m1], tw[2 * e]) : x[m1];
      e += k1;
      if (e >= n) e -= n;
    }
  }
  wave_lds_sync();
  // pass A'
  for (int li = 0; li < L; ++li) {
    const int64_t row = r0 + li;
    if (lane < Q && row < rows) {
      float2 v[P];
      const float2* src = buf + li * n + lane;
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) v[k1] = src[Q * k1];
      dft_reg<P, +1>(v);
#pragma unroll
      for (int m2 = 0; m2 < P; ++m2) {
        const int j = lane + Q * m2;
        const float2 r = v[m2];
        if (MODE == 0) {
          const float2 x = in[row * n + j];
          reinterpret_cast<float4*>(out + row * m)[j] = make_float4(x.x * ns, x.y * ns, r.x * scale, r.y * scale);
        } else {
          const float2 ev = in[row * m + 2 * j];
          out[row * n + j] = make_float2(ev.x * ns + r.x * scale, ev.y * ns + r.y * scale);
        }
      }
    }
  }
}

// The same line transform along the columns of a row-major array, for
// n = 64 E at ds = 2, so that the upsample needs no transposes: a workgroup of
// 8 / L wavefronts takes 8 columns (coalesced 64-byte row segments staged in
// LDS, column stride 584), each wavefront transforms its L columns in
// registers, and the workgroup writes the rows back.
//   MODE 0 (forward, columns of `in` [n x n]): out [2n x n], out[2j][c] =
//          n scale x_j, out[2j+1][c] = scale IFFT(phi FFT(x))_j
//   MODE 1 (adjoint, columns of `in` [2n x n]): out [n x n], out[j][c] =
//          n scale in[2j][c] + scale IFFT(conj(phi) FFT(in[2 . + 1][c]))_j
constexpr int kUpColStride = 584;

template <int E, int MODE>
__global__ __launch_bounds__(64 * E) void k_up2_cols_w(const float2* __restrict__ in, float2* __restrict__ out,
                                                       float scale) {
  using F = WaveFft<E>;
  constexpr int n = F::N, m = 2 * n, h = n / 2, S = kUpColStride, NT = 64 * E;
  __shared__ float2 xt[8 * S];   // the columns to transform
  __shared__ float2 rt[8 * S];   // MODE 0: results; MODE 1: the even rows
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int c0 = blockIdx.x * 8, c = tid & 7;
  const float ns = (float)n * scale;
  {
    // all of a thread's loads first, then its LDS stores (one round trip)
    constexpr int IT = 8 * n / NT;
    float2 va[IT], vb[IT];
#pragma unroll
    for (int u = 0; u < IT; ++u) {
      const int j = (tid + u * NT) >> 3;
      if (MODE == 0) {
        va[u] = in[(int64_t)j * n + c0 + c];
      } else {
        va[u] = in[(int64_t)(2 * j + 1) * n + c0 + c];
        vb[u] = in[(int64_t)(2 * j) * n + c0 + c];
      }
    }
#pragma unroll
    for (int u = 0; u < IT; ++u) {
      const int j = (tid + u * NT) >> 3;
      xt[c * S + j] = va[u];
      if (MODE == 1) rt[c * S + j] = vb[u];
    }
  }
  __syncthreads();
  F fft;
  fft.init(lane);
  float2 v[8];
#pragma unroll
  for (int l = 0; l < F::L; ++l)
#pragma unroll
    for (int n2 = 0; n2 < E; ++n2) v[l * E + n2] = xt[(wv * F::L + l) * S + lane + 64 * n2];
  float2* scr = (MODE == 0 ? rt : xt) + wv * F::L * S;   // a slot no other wavefront reads now
  fft.template run<-1>(v, scr, lane);
  wave_reorder<E>(v, scr, lane);
#pragma unroll
  for (int l = 0; l < F::L; ++l)
#pragma unroll
    for (int n2 = 0; n2 < E; ++n2) {
      const int k = lane + 64 * n2;
      const int sk = k < h ? k : k - n;
      const float2 w = tw1024<-1>(((sk + m) & (m - 1)) * (1024 / m));
      v[l * E + n2] = (MODE == 0) ? cmulc(v[l * E + n2], w) : cmul(v[l * E + n2], w);
    }
  fft.template run<+1>(v, scr, lane);
  wave_reorder<E>(v, scr, lane);
  float2* dst = (MODE == 0 ? rt : xt) + wv * F::L * S;
#pragma unroll
  for (int l = 0; l < F::L; ++l)
#pragma unroll
    for (int n2 = 0; n2 < E; ++n2) dst[l * S + lane + 64 * n2] = v[l * E + n2];
  __syncthreads();
  for (int t = tid; t < 8 * n; t += NT) {
    const int j = t >> 3;
    const float2 x = xt[c * S + j], r = rt[c * S + j];
    if (MODE == 0) {
      out[(int64_t)(2 * j) * n + c0 + c] = make_float2(x.x * ns, x.y * ns);
      out[(int64_t)(2 * j + 1) * n + c0 + c] = make_float2(r.x * scale, r.y * scale);
    } else {
      out[(int64_t)j * n + c0 + c] = make_float2(r.x * ns + x.x * scale, r.y * ns + x.y * scale);
    }
  }
}

template <int MODE>
static void launch_up2_cols(const float2* in, float2* out, int n, float scale, hipStream_t s) {
  const dim3 g(n / 8);
  switch (n) {
    case 64: hipLaunchKernelGGL((k_up2_cols_w<1, MODE>), g, dim3(64), 0, s, in, out, scale); break;
    case 128: hipLaunchKernelGGL((k_up2_cols_w<2, MODE>), g, dim3(128), 0, s, in, out, scale); break;
    case 256: hipLaunchKernelGGL((k_up2_cols_w<4, MODE>), g, dim3(256), 0, s, in, out, scale); break;
    default: hipLaunchKernelGGL((k_up2_cols_w<8, MODE>), g, dim3(512), 0, s, in, out, scale); break;
  }
}

// power-of-two n = 64 .. 512 at ds = 2: register-FFT row and column stages
static bool up_wave(int n, int ds) { return ds == 2 && n >= 64 && n <= 512 && is_pow2(n); }
// n = 912, 1824 at ds = 2: register-DFT row stages (k_up2_rows_r, compile-time twiddles)
static bool up_reg(int n) { return n == 912 || n == 1824; }

struct UpLaunch {
  int lines;     // lines per workgroup of the row stages
  bool lds_tw;   // twiddle table in LDS (else the global table)
};

static UpLaunch up_launch(int n, int m) {
  (void)n;
  UpLaunch u;
  int lines = 1;
  while (lines < 16 && lines * 2 * m <= 2048 && sizeof(float2) * (size_t)(m + 2 * (lines * 2) * m) <= kLdsBudget)
    lines *= 2;
  u.lines = lines;
  u.lds_tw = sizeof(float2) * (size_t)(m + 2 * lines * m) <= kLdsBudget;
  return u;
}

static int check_sizes(int32_t n, int32_t ds, FftPlan& Pn, FftPlan& Pm) {
  PTY_REQUIRE(n >= 2, PTY_EUNSUP, "upsample: n=%d must be >= 2", n);
  PTY_REQUIRE(ds >= 1, PTY_EUNSUP, "upsample: downsample_factor=%d must be >= 1", ds);
  PTY_REQUIRE((int64_t)n * ds <= 4096, PTY_EUNSUP, "upsample: canvas %d exceeds 4096", n * ds);
  PTY_REQUIRE(make_fft_plan(n, Pn) == PTY_OK && make_fft_plan(n * ds, Pm) == PTY_OK, PTY_EUNSUP,
              "upsample: sizes %d / %d need prime factors <= %d", n, n * ds, kMaxRadix);
  // one line of m (ping-pong) must fit the LDS budget
  PTY_REQUIRE(sizeof(float2) * 2 * (size_t)(n * ds) <= kLdsBudget, PTY_EUNSUP, "upsample: canvas too large");
  return PTY_OK;
}

template <int MODE>
static void launch_up_rows(const float2* in, float2* out, const FftPlan& Pn, const FftPlan& Pm, int rows,
                           float scale, const UpLaunch& u, const float2* gtw, hipStream_t s) {
  const int m = Pm.n;
  const size_t shm = sizeof(float2) * (size_t)((u.lds_tw ? m : 0) + 2 * u.lines * m);
  const int blocks = (rows + u.lines - 1) / u.lines;
  hipLaunchKernelGGL(k_up_rows<MODE>, dim3(blocks), dim3(256), shm, s, in, out, Pn, Pm, rows, u.lines, scale,
                     u.lds_tw ? nullptr : gtw);
}

// ds = 2 form: lines of n, one line per workgroup from n = 512 on (more
// independent workgroups beat fuller ones for these latency-bound passes:
// n = 512 with 1 line 33 us against 40 us with 2); the twiddle table goes to
// LDS only while it is small (a large LDS copy costs more in occupancy than
// the L1-resident global table costs in latency)
static UpLaunch up2_launch(int n) {
  const int m = 2 * n;
  UpLaunch u;
  int lines = 1;
  while (lines < 16 && lines * 2 * n <= 1024) lines *= 2;
  u.lines = lines;
  u.lds_tw = m <= 2048 && sizeof(float2) * (size_t)(m + 2 * lines * n) <= kLdsBudget;
  return u;
}

template <int MODE>
static void launch_up2_rows(const float2* in, float2* out, const FftPlan& Pn, int rows, float scale,
                            const UpLaunch& u, const float2* gtw, hipStream_t s) {
  const int n = Pn.n;
  if (n >= 64 && n <= 512 && Pn.pow2) {
    const int E = n / 64, jobs = rows / (8 / E);
    const int cap = device_cu_count() * 8, need = (jobs + 3) / 4;
    const dim3 g(need < cap ? need : cap), blk(256);
    switch (E) {
      case 1: hipLaunchKernelGGL((k_up2_rows_w<1, MODE>), g, blk, 0, s, in, out, rows, scale); break;
      case 2: hipLaunchKernelGGL((k_up2_rows_w<2, MODE>), g, blk, 0, s, in, out, rows, scale); break;
      case 4: hipLaunchKernelGGL((k_up2_rows_w<4, MODE>), g, blk, 0, s, in, out, rows, scale); break;
      default: hipLaunchKernelGGL((k_up2_rows_w<8, MODE>), g, blk, 0, s, in, out, rows, scale); break;
    }
    return;
  }
  if (up_reg(n)) {
    if (n == 1824)
      hipLaunchKernelGGL((k_up2_rows_r<32, 57, 1, MODE>), dim3(rows), dim3(64), 0, s, in, out, rows, scale);
    else
      hipLaunchKernelGGL((k_up2_rows_r<16, 57, 2, MODE>), dim3((rows + 1) / 2), dim3(64), 0, s, in, out, rows, scale);
    return;
  }
  const size_t shm = sizeof(float2) * (size_t)((u.lds_tw ? 2 * n : 0) + 2 * u.lines * n);
  const int blocks = (rows + u.lines - 1) / u.lines;
  // 128 threads for long lines (n = 1824: 243 us against 255 us with 256)
  const int threads = n > 1024 ? 128 : 256;
  hipLaunchKernelGGL(k_up2_rows<MODE>, dim3(blocks), dim3(threads), shm, s, in, out, Pn, rows, u.lines, scale,
                     u.lds_tw ? nullptr : gtw);
}

// one upsample row stage: the ds = 2 form when it applies
template <int MODE>
static void up_stage(const float2* in, float2* out, const FftPlan& Pn, const FftPlan& Pm, int rows, float scale,
                     const float2* gtw, hipStream_t s) {
  const int n = Pn.n, m = Pm.n;
  if (m == 2 * n && n % 2 == 0)
    launch_up2_rows<MODE>(in, out, Pn, rows, scale, up2_launch(n), gtw, s);
  else
    launch_up_rows<MODE>(in, out, Pn, Pm, rows, scale, up_launch(n, m), gtw, s);
}

// true when no stage of (n, m) reads the global twiddle table
static bool up_lds_twiddles(int n, int m) {
  if (m == 2 * n && up_reg(n)) return true;   // no global table needed
  return (m == 2 * n && n % 2 == 0) ? up2_launch(n).lds_tw : up_launch(n, m).lds_tw;
}

static void launch_transpose(const float2* in, float2* out, int R, int C, hipStream_t s) {
  hipLaunchKernelGGL(k_transpose, dim3((C + 31) / 32, (R + 31) / 32), dim3(256), 0, s, in, out, R, C);
}

// workspace: t_nn [n, n], y [m, n], yt [n, m], twiddles [m]
struct UpWork {
  size_t t_nn, y, yt, tw, total;
};

static UpWork up_work(int n, int ds) {
  const size_t m = (size_t)n * ds;
  UpWork w;
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
  w.t_nn = take(sizeof(float2) * (size_t)n * n);
  w.y = take(sizeof(float2) * m * n);
  w.yt = take(sizeof(float2) * m * n);
  w.tw = take(sizeof(float2) * m);
  w.total = off;
  return w;
}

}  // namespace pty

using namespace pty;

extern "C" size_t pty_upsample_workspace(int32_t n, int32_t ds) { return up_work(n, ds).total; }

extern "C" int pty_upsample_fwd(const float* small, int32_t n, int32_t ds, float* big, void* work,
                                size_t work_bytes, hipStream_t stream) {
  FftPlan Pn, Pm;
  int rc = check_sizes(n, ds, Pn, Pm);
  if (rc) return rc;
  PTY_REQUIRE(small && big, PTY_EINVAL, "upsample_fwd: null pointer");
  const UpWork W = up_work(n, ds);
  PTY_REQUIRE(work && work_bytes >= W.total, PTY_EWORK, "upsample_fwd: workspace too small");
  const int m = n * ds;
  char* w = (char*)work;
  if (up_wave(n, ds)) {
    float2* Y = (float2*)(w + W.y);     // U along columns   [m, n]
    launch_up2_cols<0>((const float2*)small, Y, n, 1.f, stream);
    up_stage<0>(Y, (float2*)big, Pn, Pm, m, 1.f / ((float)n * (float)n), nullptr, stream);
    return check_launch("upsample_fwd");
  }
  float2* sT = (float2*)(w + W.t_nn);   // small^T        [n, n]
  float2* YT = (float2*)(w + W.yt);     // U along columns, transposed [n, m]
  float2* Y = (float2*)(w + W.y);       // [m, n]
  float2* tw = (float2*)(w + W.tw);
  if (!up_lds_twiddles(n, m)) hipLaunchKernelGGL(k_twiddle_table, dim3((m + 255) / 256), dim3(256), 0, stream, tw, m);
  launch_transpose((const float2*)small, sT, n, n, stream);
  up_stage<0>(sT, YT, Pn, Pm, n, 1.f, tw, stream);
  launch_transpose(YT, Y, n, m, stream);
  up_stage<0>(Y, (float2*)big, Pn, Pm, m, 1.f / ((float)n * (float)n), tw, stream);
  return check_launch("upsample_fwd");
}

extern "C" int pty_upsample_adj(const float* gbig, int32_t n, int32_t ds, float* gsmall, void* work,
                                size_t work_bytes, hipStream_t stream) {
  FftPlan Pn, Pm;
  int rc = check_sizes(n, ds, Pn, Pm);
  if (rc) return rc;
  PTY_REQUIRE(gbig && gsmall, PTY_EINVAL, "upsample_adj: null pointer");
  const UpWork W = up_work(n, ds);
  PTY_REQUIRE(work && work_bytes >= W.total, PTY_EWORK, "upsample_adj: workspace too small");
  const int m = n * ds;
  char* w = (char*)work;
  if (up_wave(n, ds)) {
    float2* Wr = (float2*)(w + W.y);    // V along rows      [m, n]
    up_stage<1>((const float2*)gbig, Wr, Pn, Pm, m, 1.f, nullptr, stream);
    launch_up2_cols<1>(Wr, (float2*)gsmall, n, 1.f / ((float)n * (float)n), stream);
    return check_launch("upsample_adj");
  }
  float2* Wr = (float2*)(w + W.y);      // V along rows            [m, n]
  float2* WT = (float2*)(w + W.yt);     // its transpose           [n, m]
  float2* gT = (float2*)(w + W.t_nn);   // gsmall^T                [n, n]
  float2* tw = (float2*)(w + W.tw);
  if (!up_lds_twiddles(n, m)) hipLaunchKernelGGL(k_twiddle_table, dim3((m + 255) / 256), dim3(256), 0, stream, tw, m);
  up_stage<1>((const float2*)gbig, Wr, Pn, Pm, m, 1.f, tw, stream);
  launch_transpose(Wr, WT, m, n, stream);
  up_stage<1>(WT, gT, Pn, Pm, n, 1.f / ((float)n * (float)n), tw, stream);
  launch_transpose(gT, (float2*)gsmall, n, n, stream);
  return check_launch("upsample_adj");
}
