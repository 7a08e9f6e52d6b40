// Fourier upsample (models/complex_inr.py:252-271) and its exact adjoint.
//
// Forward: big = ds^2 * ifft2_m(ifftshift(pad(fftshift(fft2_n(small))))), m = ds*n.
// With even n the shift/pad/unshift chain maps small-grid frequency k to big-
// grid frequency q(k) = k for k < n/2 and k - n + m otherwise (the Nyquist
// bin n/2 lands on the negative side), every other big frequency is zero.
// The 2-D transform is done as three passes of batched LDS FFTs:
//   U1 rows   : X1 = FFT_n along rows of small                      [n, n]
//   U2 columns: Z  = IFFT_m(embed_q(FFT_n(X1 column)))              [m, n]
//   U3 rows   : big = IFFT_m(embed_q(Z row)) / n^2                  [m, m]
// Adjoint (what autograd of the reference computes), net scale 1:
//   A1 rows   : W1 = crop_q(FFT_m(gbig row))                        [m, n]
//   A2 columns: W2 = IFFT_n(crop_q(FFT_m(W1 column)))               [n, n]
//   A3 rows   : gsmall = IFFT_n(W2 row) / n^2                       [n, n]
#include "common.h"

namespace pty {

__device__ __forceinline__ int embed_q(int k, int n, int m) { return k < n / 2 ? k : k - n + m; }

// ---------------------------------------------------------------- row kernels
// Generic row transform: each workgroup takes `lines` rows of length n_in from
// `in` (row stride ld_in), optionally (mode) pre-transforms, and writes rows of
// length n_out.
//   MODE 0: out = FFT_n(in)                         (U1; n_in = n_out = n)
//   MODE 1: out = IFFT_m(embed(in)) * scale         (U3; n_in = n, n_out = m)
//   MODE 2: out = crop(FFT_m(in))                   (A1; n_in = m, n_out = n)
//   MODE 3: out = IFFT_n(in) * scale                (A3)
template <int MODE>
__global__ __launch_bounds__(256) void k_rows(const float2* __restrict__ in, float2* __restrict__ out, int n,
                                              int m, int rows, int lines, float scale) {
  extern __shared__ float2 smem[];
  const int L = (MODE == 0 || MODE == 3) ? n : m;  // transform length
  const int lg = ilog2(L);
  float2* tw = smem;
  float2* a = tw + L;
  float2* b = a + lines * L;
  const int tid = threadIdx.x, nt = blockDim.x;
  fill_twiddles(tw, L, tid, nt);
  const int r0 = blockIdx.x * lines;
  const int nl = min(lines, rows - r0);
  // load
  for (int t = tid; t < lines * L; t += nt) {
    const int li = t / L, c = t - li * L;
    float2 v = make_float2(0.f, 0.f);
    if (li < nl) {
      const int r = r0 + li;
      if (MODE == 1) {
        // c is a big-grid frequency; find the small one mapping to it
        int k = -1;
        if (c < n / 2) k = c;
        else if (c >= m - n / 2) k = c - m + n;
        if (k >= 0) v = in[(int64_t)r * n + k];
      } else {
        v = in[(int64_t)r * L + c];
      }
    }
    a[t] = v;
  }
  __syncthreads();
  float2* res = (MODE == 0 || MODE == 2) ? fft_lines<-1>(a, b, L, lg, lines, L, tw, 1, tid, nt)
                                         : fft_lines<+1>(a, b, L, lg, lines, L, tw, 1, tid, nt);
  const int W = (MODE == 2) ? n : L;  // output row length
  for (int t = tid; t < lines * W; t += nt) {
    const int li = t / W, c = t - li * W;
    if (li >= nl) continue;
    const int r = r0 + li;
    float2 v;
    if (MODE == 2) v = res[li * L + embed_q(c, n, m)];
    else v = res[li * L + c];
    if (MODE == 1 || MODE == 3) v = cscale(v, scale);
    out[(int64_t)r * W + c] = v;
  }
}

// ------------------------------------------------------------- column kernels
// Each workgroup takes C adjacent columns of a row-major array.
//   MODE 0 (U2): in [n, n]  -> out [m, n]: IFFT_m(embed(FFT_n(col)))
//   MODE 1 (A2): in [m, n]  -> out [n, n]: IFFT_n(crop(FFT_m(col)))
template <int MODE>
__global__ __launch_bounds__(256) void k_cols(const float2* __restrict__ in, float2* __restrict__ out, int n,
                                              int m, int C) {
  extern __shared__ float2 smem[];
  const int lgn = ilog2(n), lgm = ilog2(m);
  const int S = m + 1;  // padded line stride in LDS
  float2* tw = smem;    // exp(-2 pi i k/m), k < m; the n-point FFT uses stride m/n
  float2* a = tw + m;
  float2* b = a + C * S;
  const int tid = threadIdx.x, nt = blockDim.x;
  fill_twiddles(tw, m, tid, nt);
  const int c0 = blockIdx.x * C;
  const int rows_in = (MODE == 0) ? n : m;
  for (int t = tid; t < rows_in * C; t += nt) {
    const int r = t / C, c = t - r * C;
    a[c * S + r] = in[(int64_t)r * n + c0 + c];
  }
  __syncthreads();
  if (MODE == 0) {
    float2* f = fft_lines<-1>(a, b, n, lgn, C, S, tw, m / n, tid, nt);
    float2* g = (f == a) ? b : a;
    for (int t = tid; t < C * m; t += nt) {
      const int c = t / m, q = t - c * m;
      int k = -1;
      if (q < n / 2) k = q;
      else if (q >= m - n / 2) k = q - m + n;
      g[c * S + q] = (k >= 0) ? f[c * S + k] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2* h = fft_lines<+1>(g, f, m, lgm, C, S, tw, 1, tid, nt);
    for (int t = tid; t < m * C; t += nt) {
      const int r = t / C, c = t - r * C;
      out[(int64_t)r * n + c0 + c] = h[c * S + r];
    }
  } else {
    float2* f = fft_lines<-1>(a, b, m, lgm, C, S, tw, 1, tid, nt);
    float2* g = (f == a) ? b : a;
    for (int t = tid; t < C * n; t += nt) {
      const int c = t / n, k = t - c * n;
      g[c * S + k] = f[c * S + embed_q(k, n, m)];
    }
    __syncthreads();
    float2* h = fft_lines<+1>(g, f, n, lgn, C, S, tw, m / n, tid, nt);
    for (int t = tid; t < n * C; t += nt) {
      const int r = t / C, c = t - r * C;
      out[(int64_t)r * n + c0 + c] = h[c * S + r];
    }
  }
}

static int row_lines(int L) { return L >= 1024 ? 1 : 1024 / L; }
// Columns per workgroup: the largest power of two keeping LDS (twiddles + two
// padded line buffers) within 64 KiB.
static int col_tile(int m, int n) {
  int C = 16;
  while (C > 1 && (size_t)(m + 2 * C * (m + 1)) * sizeof(float2) > 65536) C >>= 1;
  while (C > 1 && n % C) C >>= 1;
  return C;
}

template <int MODE>
static void launch_rows(const float2* in, float2* out, int n, int m, int rows, float scale, hipStream_t s) {
  const int L = (MODE == 0 || MODE == 3) ? n : m;
  const int lines = row_lines(L);
  const size_t shm = sizeof(float2) * (size_t)(L + 2 * lines * L);
  const int blocks = (rows + lines - 1) / lines;
  hipLaunchKernelGGL(k_rows<MODE>, dim3(blocks), dim3(256), shm, s, in, out, n, m, rows, lines, scale);
}

template <int MODE>
static void launch_cols(const float2* in, float2* out, int n, int m, hipStream_t s) {
  const int C = col_tile(m, n);
  const size_t shm = sizeof(float2) * (size_t)(m + 2 * C * (m + 1));
  hipLaunchKernelGGL(k_cols<MODE>, dim3(n / C), dim3(256), shm, s, in, out, n, m, C);
}

static int check_sizes(int32_t n, int32_t ds) {
  PTY_REQUIRE(n >= 2 && is_pow2(n), PTY_EUNSUP, "upsample: n=%d must be a power of two >= 2", n);
  PTY_REQUIRE(ds >= 1 && is_pow2(ds), PTY_EUNSUP, "upsample: downsample_factor=%d must be a power of two", ds);
  PTY_REQUIRE((int64_t)n * ds <= 2048, PTY_EUNSUP, "upsample: canvas %d exceeds 2048", n * ds);
  return PTY_OK;
}

}  // namespace pty

using namespace pty;

extern "C" size_t pty_upsample_workspace(int32_t n, int32_t ds) {
  const size_t m = (size_t)n * ds;
  return sizeof(float2) * ((size_t)n * n + m * n);
}

extern "C" int pty_upsample_fwd(const float* small, int32_t n, int32_t ds, float* big, void* work,
                                size_t work_bytes, hipStream_t stream) {
  int rc = check_sizes(n, ds);
  if (rc) return rc;
  PTY_REQUIRE(small && big, PTY_EINVAL, "upsample_fwd: null pointer");
  PTY_REQUIRE(work && work_bytes >= pty_upsample_workspace(n, ds), PTY_EWORK, "upsample_fwd: workspace too small");
  const int m = n * ds;
  float2* X1 = (float2*)work;
  float2* Z = X1 + (size_t)n * n;
  launch_rows<0>((const float2*)small, X1, n, m, n, 1.f, stream);
  launch_cols<0>(X1, Z, n, m, stream);
  launch_rows<1>(Z, (float2*)big, n, m, m, 1.f / ((float)n * (float)n), stream);
  return check_launch("upsample_fwd");
}

extern "C" int pty_upsample_adj(const float* gbig, int32_t n, int32_t ds, float* gsmall, void* work,
                                size_t work_bytes, hipStream_t stream) {
  int rc = check_sizes(n, ds);
  if (rc) return rc;
  PTY_REQUIRE(gbig && gsmall, PTY_EINVAL, "upsample_adj: null pointer");
  PTY_REQUIRE(work && work_bytes >= pty_upsample_workspace(n, ds), PTY_EWORK, "upsample_adj: workspace too small");
  const int m = n * ds;
  float2* W1 = (float2*)work;                 // [m, n]
  float2* W2 = W1 + (size_t)m * n;            // [n, n]
  launch_rows<2>((const float2*)gbig, W1, n, m, m, 1.f, stream);
  launch_cols<1>(W1, W2, n, m, stream);
  launch_rows<3>(W2, (float2*)gsmall, n, m, n, 1.f / ((float)n * (float)n), stream);
  return check_launch("upsample_adj");
}
