// Fourier upsample (models/complex_inr.py:252-271) and its exact adjoint.
//
// Forward: big = ds^2 * ifft2_m(ifftshift(pad(fftshift(fft2_n(small))))), m = ds*n.
// The shift/pad/unshift chain maps small-grid frequency k to big-grid
// frequency q(k) = ((k + n//2) % n + (m-n)//2 - m//2) mod m (embed_freq; for
// even n: k below n/2, k - n + m above, Nyquist on the negative side; for odd
// n the DC bin lands on q = m-1 exactly as in the reference); every other big
// frequency is zero.  The 2-D transform is three passes of batched LDS FFTs
// (mixed radix: any n whose prime factors are <= 64):
//   U1 rows   : X1 = FFT_n along rows of small                      [n, n]
//   U2 columns: Z  = IFFT_m(embed_q(FFT_n(X1 column)))              [m, n]
//   U3 rows   : big = IFFT_m(embed_q(Z row)) / n^2                  [m, m]
// Adjoint (what autograd of the reference computes), net scale 1:
//   A1 rows   : W1 = crop_q(FFT_m(gbig row))                        [m, n]
//   A2 columns: W2 = IFFT_n(crop_q(FFT_m(W1 column)))               [n, n]
//   A3 rows   : gsmall = IFFT_n(W2 row) / n^2                       [n, n]
// Twiddles live in LDS when the line buffers leave room, else in a global
// table in the workspace (filled at the start of every call).
#include "common.h"

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
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    float s, c;
    sincospif(-2.0f * (float)k / (float)n, &s, &c);
    tw[k] = make_float2(c, s);
  }
}

// ---------------------------------------------------------------- row kernels
//   MODE 0: out = FFT_n(in)                         (U1)
//   MODE 1: out = IFFT_m(embed(in)) * scale         (U3; in rows of n, out rows of m)
//   MODE 2: out = crop(FFT_m(in))                   (A1; in rows of m, out rows of n)
//   MODE 3: out = IFFT_n(in) * scale                (A3)
template <int MODE>
__global__ __launch_bounds__(256) void k_rows(const float2* __restrict__ in, float2* __restrict__ out, FftPlan Pn,
                                              FftPlan Pm, int rows, int lines, float scale,
                                              const float2* __restrict__ gtw) {
  extern __shared__ float2 smem[];
  const int n = Pn.n, m = Pm.n;
  const FftPlan& P = (MODE == 0 || MODE == 3) ? Pn : Pm;
  const int L = P.n;  // transform length
  const int tid = threadIdx.x, nt = blockDim.x;
  const float2* tw;
  float2* a;
  if (gtw) {
    tw = gtw;
    a = smem;
  } else {
    fill_twiddles(smem, L, tid, nt);
    tw = smem;
    a = smem + L;
  }
  float2* b = a + lines * L;
  const int r0 = blockIdx.x * lines;
  const int nl = min(lines, rows - r0);
  for (int t = tid; t < lines * L; t += nt) {
    const int li = t / L, c = t - li * L;
    float2 v = make_float2(0.f, 0.f);
    if (li < nl) {
      const int r = r0 + li;
      if (MODE == 1) {
        const int k = crop_freq(c, n, m);   // small frequency placed at big frequency c
        if (k >= 0) v = in[(int64_t)r * n + k];
      } else {
        v = in[(int64_t)r * L + c];
      }
    }
    a[t] = v;
  }
  __syncthreads();
  float2* res = (MODE == 0 || MODE == 2) ? fft_any<-1>(a, b, P, lines, L, tw, 1, tid, nt)
                                         : fft_any<+1>(a, b, P, lines, L, tw, 1, tid, nt);
  const int W = (MODE == 2) ? n : L;  // output row length
  for (int t = tid; t < lines * W; t += nt) {
    const int li = t / W, c = t - li * W;
    if (li >= nl) continue;
    const int r = r0 + li;
    float2 v = (MODE == 2) ? res[li * L + embed_freq(c, n, m)] : res[li * L + c];
    if (MODE == 1 || MODE == 3) v = cscale(v, scale);
    out[(int64_t)r * W + c] = v;
  }
}

// ------------------------------------------------------------- column kernels
//   MODE 0 (U2): in [n, n] -> out [m, n]: IFFT_m(embed(FFT_n(col)))
//   MODE 1 (A2): in [m, n] -> out [n, n]: IFFT_n(crop(FFT_m(col)))
template <int MODE>
__global__ __launch_bounds__(256) void k_cols(const float2* __restrict__ in, float2* __restrict__ out, FftPlan Pn,
                                              FftPlan Pm, int C, const float2* __restrict__ gtw) {
  extern __shared__ float2 smem[];
  const int n = Pn.n, m = Pm.n;
  const int S = m + 1;  // padded line stride in LDS
  const int tid = threadIdx.x, nt = blockDim.x;
  const float2* tw;     // exp(-2 pi i k/m), k < m; the n-point FFT uses stride m/n
  float2* a;
  if (gtw) {
    tw = gtw;
    a = smem;
  } else {
    fill_twiddles(smem, m, tid, nt);
    tw = smem;
    a = smem + m;
  }
  float2* b = a + C * S;
  const int c0 = blockIdx.x * C;
  const int rows_in = (MODE == 0) ? n : m;
  for (int t = tid; t < rows_in * C; t += nt) {
    const int r = t / C, c = t - r * C;
    a[c * S + r] = in[(int64_t)r * n + c0 + c];
  }
  __syncthreads();
  if (MODE == 0) {
    float2* f = fft_any<-1>(a, b, Pn, C, S, tw, m / n, tid, nt);
    float2* g = (f == a) ? b : a;
    for (int t = tid; t < C * m; t += nt) {
      const int c = t / m, q = t - c * m;
      const int k = crop_freq(q, n, m);
      g[c * S + q] = (k >= 0) ? f[c * S + k] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2* h = fft_any<+1>(g, f, Pm, C, S, tw, 1, tid, nt);
    for (int t = tid; t < m * C; t += nt) {
      const int r = t / C, c = t - r * C;
      out[(int64_t)r * n + c0 + c] = h[c * S + r];
    }
  } else {
    float2* f = fft_any<-1>(a, b, Pm, C, S, tw, 1, tid, nt);
    float2* g = (f == a) ? b : a;
    for (int t = tid; t < C * n; t += nt) {
      const int c = t / n, k = t - c * n;
      g[c * S + k] = f[c * S + embed_freq(k, n, m)];
    }
    __syncthreads();
    float2* h = fft_any<+1>(g, f, Pn, C, S, tw, m / n, tid, nt);
    for (int t = tid; t < n * C; t += nt) {
      const int r = t / C, c = t - r * C;
      out[(int64_t)r * n + c0 + c] = h[c * S + r];
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
  while (r % 4 == 0 && k < kMaxPasses) push(4);
  while (r % 2 == 0 && k < kMaxPasses) push(2);
  for (int p = 3; p <= kMaxRadix && r > 1; p += 2)
    while (r % p == 0 && k < kMaxPasses) push(p);
  P.n_passes = k;
  return r == 1 ? PTY_OK : PTY_EUNSUP;
}

struct UpLaunch {
  int lines_n, lines_m;   // rows per workgroup for n- and m-length row kernels
  bool lds_tw_n, lds_tw_m, lds_tw_c;
  int C;                  // columns per workgroup
};

static int rows_per_wg(int L) { return L >= 1024 ? 1 : 1024 / L; }

static UpLaunch up_launch(int n, int m) {
  UpLaunch u;
  u.lines_n = rows_per_wg(n);
  u.lines_m = rows_per_wg(m);
  u.lds_tw_n = sizeof(float2) * (size_t)(n + 2 * u.lines_n * n) <= kLdsBudget;
  u.lds_tw_m = sizeof(float2) * (size_t)(m + 2 * u.lines_m * m) <= kLdsBudget;
  int C = 16;
  while (C > 1 && (sizeof(float2) * (size_t)(m + 2 * C * (m + 1)) > kLdsBudget || n % C)) C >>= 1;
  u.C = C;
  u.lds_tw_c = sizeof(float2) * (size_t)(m + 2 * C * (m + 1)) <= kLdsBudget;
  return u;
}

static int check_sizes(int32_t n, int32_t ds, FftPlan& Pn, FftPlan& Pm) {
  PTY_REQUIRE(n >= 2, PTY_EUNSUP, "upsample: n=%d must be >= 2", n);
  PTY_REQUIRE(ds >= 1, PTY_EUNSUP, "upsample: downsample_factor=%d must be >= 1", ds);
  PTY_REQUIRE((int64_t)n * ds <= 4096, PTY_EUNSUP, "upsample: canvas %d exceeds 4096", n * ds);
  PTY_REQUIRE(make_fft_plan(n, Pn) == PTY_OK && make_fft_plan(n * ds, Pm) == PTY_OK, PTY_EUNSUP,
              "upsample: sizes %d / %d need prime factors <= %d", n, n * ds, kMaxRadix);
  // one line of m (ping-pong) must fit the LDS budget
  PTY_REQUIRE(sizeof(float2) * 2 * (size_t)(n * ds + 1) <= kLdsBudget, PTY_EUNSUP, "upsample: canvas too large");
  return PTY_OK;
}

template <int MODE>
static void launch_rows(const float2* in, float2* out, const FftPlan& Pn, const FftPlan& Pm, int rows, float scale,
                        const UpLaunch& u, const float2* gtw_n, const float2* gtw_m, hipStream_t s) {
  const bool on_n = (MODE == 0 || MODE == 3);
  const int L = on_n ? Pn.n : Pm.n;
  const int lines = on_n ? u.lines_n : u.lines_m;
  const bool lds_tw = on_n ? u.lds_tw_n : u.lds_tw_m;
  const size_t shm = sizeof(float2) * (size_t)((lds_tw ? L : 0) + 2 * lines * L);
  const int blocks = (rows + lines - 1) / lines;
  hipLaunchKernelGGL(k_rows<MODE>, dim3(blocks), dim3(256), shm, s, in, out, Pn, Pm, rows, lines, scale,
                     lds_tw ? nullptr : (on_n ? gtw_n : gtw_m));
}

template <int MODE>
static void launch_cols(const float2* in, float2* out, const FftPlan& Pn, const FftPlan& Pm, const UpLaunch& u,
                        const float2* gtw_m, hipStream_t s) {
  const int m = Pm.n;
  const size_t shm = sizeof(float2) * (size_t)((u.lds_tw_c ? m : 0) + 2 * u.C * (m + 1));
  hipLaunchKernelGGL(k_cols<MODE>, dim3(Pn.n / u.C), dim3(256), shm, s, in, out, Pn, Pm, u.C,
                     u.lds_tw_c ? nullptr : gtw_m);
}

struct UpWork {
  size_t x1, z, tw_n, tw_m, total;
};

static UpWork up_work(int n, int ds) {
  const size_t m = (size_t)n * ds;
  UpWork w;
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
  w.x1 = take(sizeof(float2) * (size_t)n * n);
  w.z = take(sizeof(float2) * m * n);
  w.tw_n = take(sizeof(float2) * (size_t)n);
  w.tw_m = take(sizeof(float2) * m);
  w.total = off;
  return w;
}

static void fill_global_twiddles(float2* tw_n, int n, float2* tw_m, int m, const UpLaunch& u, hipStream_t s) {
  if (!u.lds_tw_n) hipLaunchKernelGGL(k_twiddle_table, dim3((n + 255) / 256), dim3(256), 0, s, tw_n, n);
  if (!u.lds_tw_m || !u.lds_tw_c) hipLaunchKernelGGL(k_twiddle_table, dim3((m + 255) / 256), dim3(256), 0, s, tw_m, m);
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
  const UpLaunch u = up_launch(n, m);
  char* w = (char*)work;
  float2* X1 = (float2*)(w + W.x1);
  float2* Z = (float2*)(w + W.z);
  float2* twn = (float2*)(w + W.tw_n);
  float2* twm = (float2*)(w + W.tw_m);
  fill_global_twiddles(twn, n, twm, m, u, stream);
  launch_rows<0>((const float2*)small, X1, Pn, Pm, n, 1.f, u, twn, twm, stream);
  launch_cols<0>(X1, Z, Pn, Pm, u, twm, stream);
  launch_rows<1>(Z, (float2*)big, Pn, Pm, m, 1.f / ((float)n * (float)n), u, twn, twm, stream);
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
  const UpLaunch u = up_launch(n, m);
  char* w = (char*)work;
  float2* W1 = (float2*)(w + W.z);     // [m, n]
  float2* W2 = (float2*)(w + W.x1);    // [n, n]
  float2* twn = (float2*)(w + W.tw_n);
  float2* twm = (float2*)(w + W.tw_m);
  fill_global_twiddles(twn, n, twm, m, u, stream);
  launch_rows<2>((const float2*)gbig, W1, Pn, Pm, m, 1.f, u, twn, twm, stream);
  launch_cols<1>(W1, W2, Pn, Pm, u, twm, stream);
  launch_rows<3>(W2, (float2*)gsmall, Pn, Pm, n, 1.f / ((float)n * (float)n), u, twn, twm, stream);
  return check_launch("upsample_adj");
}
