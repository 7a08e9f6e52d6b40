// Learned warm start: the conv encoder-decoder that maps pooled diffraction
// patterns to object patches (no counterpart in the reference, whose object
// starts from the INR initialisation, Recovery.py:201-222; its output feeds
// ComplexINRModel2D(object_initial=..., use_residual=True),
// models/complex_inr.py:209-250).  Forward and backward on gfx950:
//
//   * every convolution is an im2col (activations NHWC f32 -> bf16 rows of
//     K = taps x C_in, padded to 32) followed by a BF16 GEMM on
//     v_mfma_f32_32x32x16_bf16 with FP32 accumulation (k_gemm_bf16);
//     transposed convolutions (2x2, stride 2) are one GEMM plus a
//     depth-to-space;
//   * bias + exact-erf GELU, 2x2 average pooling, the skip concatenations,
//     the MSE loss and every backward step are small elementwise kernels;
//   * weight gradients are GEMMs over the pixels with split-K partials
//     summed in a fixed order: the whole step is deterministic.
//
// Network (base width c): 64^2 x 3 -> [conv3 c, conv3 c] -> pool ->
// [2c, 2c] -> pool -> [4c, 4c] -> pool -> [8c, 8c] -> up 4c, cat -> [4c, 4c]
// -> up 2c, cat -> [2c, 2c] -> conv1 2 (32^2 x 2 output).  Parameters are one
// f32 vector in torch's parameter order with conv weights as [Co][kh][kw][Ci]
// and transposed-conv weights as [kh][kw][Co][Ci].
#include "common.h"

#include <algorithm>
#include <vector>

namespace pty {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int kGemmTile = 128;   // M and N tile of a workgroup
constexpr int kGemmK = 32;       // K step
constexpr int kLdsStride = 40;   // bf16 per LDS row (80 B: 16-B aligned, spreads the banks)

__device__ __forceinline__ int mfma_row(int v, int h) { return 8 * (v >> 2) + 4 * h + (v & 3); }

// C[m][n] (+ z * strideC) = sum_{k in chunk z} A(m, k) B(n, k), bf16 operands,
// f32 result.  A(m, k) = TA ? A[k*lda + m] : A[m*lda + k]; B likewise.
// M, N, K and every leading dimension are multiples of 32 (callers pad).
template <int TA, int TB>
__global__ __launch_bounds__(256) void k_gemm_bf16(const __bf16* __restrict__ A, const __bf16* __restrict__ B,
                                                   float* __restrict__ C, int M, int N, int K, int lda, int ldb,
                                                   int ldc, int kchunk, int64_t strideC) {
  __shared__ __attribute__((aligned(16))) __bf16 As[kGemmTile * kLdsStride];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[kGemmTile * kLdsStride];
  const int m0 = blockIdx.y * kGemmTile, n0 = blockIdx.x * kGemmTile;
  const int k_lo = blockIdx.z * kchunk;
  const int k_hi = min(K, k_lo + kchunk);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  const int j = lane & 31, h = lane >> 5;
  f16v acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;

  for (int k0 = k_lo; k0 < k_hi; k0 += kGemmK) {
    // stage A and B tiles (128 rows x 32 k) as [row][k] in LDS
#pragma unroll
    for (int rep = 0; rep < 2; ++rep) {
      const int c = t + rep * 256;   // 512 chunks of 8 bf16
      if (!TA) {
        const int r = c >> 2, kk = (c & 3) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (m0 + r < M) v = *(const uint4*)(A + (int64_t)(m0 + r) * lda + k0 + kk);
        *(uint4*)(As + r * kLdsStride + kk) = v;
      } else {
        const int kk = c >> 4, r = (c & 15) * 8;   // 32 k rows of 128 m
        uint4 v = make_uint4(0, 0, 0, 0);
        if (m0 + r < M) v = *(const uint4*)(A + (int64_t)(k0 + kk) * lda + m0 + r);
        const __bf16* e = (const __bf16*)&v;
#pragma unroll
        for (int q = 0; q < 8; ++q) As[(r + q) * kLdsStride + kk] = e[q];
      }
      if (!TB) {
        const int r = c >> 2, kk = (c & 3) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (n0 + r < N) v = *(const uint4*)(B + (int64_t)(n0 + r) * ldb + k0 + kk);
        *(uint4*)(Bs + r * kLdsStride + kk) = v;
      } else {
        const int kk = c >> 4, r = (c & 15) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (n0 + r < N) v = *(const uint4*)(B + (int64_t)(k0 + kk) * ldb + n0 + r);
        const __bf16* e = (const __bf16*)&v;
#pragma unroll
        for (int q = 0; q < 8; ++q) Bs[(r + q) * kLdsStride + kk] = e[q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int kh = 0; kh < 2; ++kh) {
      const int kk = kh * 16 + 8 * h;
      bf16x8_t a[2], b[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        a[x] = *(const bf16x8_t*)(As + (wm + x * 32 + j) * kLdsStride + kk);
        b[x] = *(const bf16x8_t*)(Bs + (wn + x * 32 + j) * kLdsStride + kk);
      }
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[x], b[y], acc[x][y], 0, 0, 0);
    }
    __syncthreads();
  }
  float* Cz = C + (int64_t)blockIdx.z * strideC;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int n = n0 + wn + y * 32 + j;
      if (n >= N) continue;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int m = m0 + wm + x * 32 + mfma_row(v, h);
        if (m < M) Cz[(int64_t)m * ldc + n] = acc[x][y][v];
      }
    }
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

#define GRID_LOOP(i, n) for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
                             i += (int64_t)gridDim.x * blockDim.x)

// col[p][k] (bf16, Kp columns) of a k x k (k = 1 or 3, pad k/2) convolution
// of in [B, H, W, C] (f32, NHWC); k = tap * C + ci, zero beyond taps * C.
__global__ void k_im2col(const float* __restrict__ in, int B, int H, int W, int C, int ks, int Kp,
                         __bf16* __restrict__ col) {
  const int64_t chunks = (int64_t)B * H * W * (Kp / 8);
  const int r = ks / 2, taps = ks * ks;
  GRID_LOOP(t, chunks) {
    const int64_t p = t / (Kp / 8);
    const int k0 = (int)(t - p * (Kp / 8)) * 8;
    const int x = (int)(p % W), y = (int)((p / W) % H);
    const int64_t b = p / ((int64_t)H * W);
    bf16x8_t o;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = k0 + q;
      float v = 0.f;
      if (k < taps * C) {
        const int tap = k / C, ci = k - tap * C;
        const int yy = y + tap / ks - r, xx = x + tap % ks - r;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = in[((b * H + yy) * W + xx) * C + ci];
      }
      o[q] = (__bf16)v;
    }
    *(bf16x8_t*)(col + p * Kp + k0) = o;
  }
}

// in-gradient of the im2col: din[b, y, x, ci] += sum over taps of dcol at
// the output pixel that read it (one thread per input element).
__global__ void k_col2im(const float* __restrict__ dcol, int B, int H, int W, int C, int ks, int Kp,
                         float* __restrict__ din) {
  const int r = ks / 2;
  GRID_LOOP(t, (int64_t)B * H * W * C) {
    const int ci = (int)(t % C);
    const int64_t p = t / C;
    const int x = (int)(p % W), y = (int)((p / W) % H);
    const int64_t b = p / ((int64_t)H * W);
    float s = 0.f;
    for (int tap = 0; tap < ks * ks; ++tap) {
      const int yy = y - (tap / ks - r), xx = x - (tap % ks - r);
      if (yy >= 0 && yy < H && xx >= 0 && xx < W) s += dcol[((b * H + yy) * W + xx) * Kp + tap * C + ci];
    }
    din[t] += s;
  }
}

// y = acc[:, :Co] + bias; pre = y; post = gelu(y) (act) or post = y.  acc has
// ldc columns; pre / post have Co (post may be a slice of a wider tensor:
// post_ld columns starting at post_off).
__global__ void k_bias_act(const float* __restrict__ acc, int64_t P, int Co, int ldc, const float* __restrict__ bias,
                           int act, float* __restrict__ pre, float* __restrict__ post, int post_ld, int post_off) {
  GRID_LOOP(t, P * Co) {
    const int64_t p = t / Co;
    const int c = (int)(t - p * Co);
    const float y = acc[p * ldc + c] + bias[c];
    if (pre) pre[t] = y;
    post[p * post_ld + post_off + c] = act ? gelu(y) : y;
  }
}

// dpre = dpost * gelu'(pre) (or dpost without the activation), as bf16 [P][ldd]
// (zero in the padding columns Co..ldd) for the GEMMs, plus per-block
// partial column sums for the bias gradient.  A workgroup takes a contiguous
// range of pixels; thread (r, x) = (t >> 5, t & 31) walks rows r, r + 8, ...
// of column 32 g + x, and the eight rows' sums are added in LDS.
__global__ __launch_bounds__(256) void k_act_bwd(const float* __restrict__ dpost, int dpost_ld, int dpost_off,
                                                 const float* __restrict__ pre, int64_t P, int Co, int ldd, int act,
                                                 __bf16* __restrict__ dpre, float* __restrict__ partial) {
  __shared__ float red[8][33];
  const int64_t chunk = (P + gridDim.x - 1) / gridDim.x;
  const int64_t p0 = blockIdx.x * chunk, p1 = min(P, p0 + chunk);
  const int x = threadIdx.x & 31, r = threadIdx.x >> 5;
  for (int g = 0; g < ldd / 32; ++g) {
    const int c = 32 * g + x;
    float s = 0.f;
    for (int64_t p = p0 + r; p < p1; p += 8) {
      float v = 0.f;
      if (c < Co) {
        v = dpost[p * dpost_ld + dpost_off + c];
        if (act) v *= gelu_grad(pre[p * Co + c]);
      }
      s += v;
      dpre[p * ldd + c] = (__bf16)v;
    }
    red[r][x] = s;
    __syncthreads();
    if (r == 0) {
      float t = 0.f;
      for (int q = 0; q < 8; ++q) t += red[q][x];
      partial[(int64_t)blockIdx.x * ldd + c] = t;
    }
    __syncthreads();
  }
}

__global__ void k_ws_zero(float* __restrict__ p, int64_t n) {
  GRID_LOOP(t, n) p[t] = 0.f;
}

__global__ void k_colsum_finish(const float* __restrict__ partial, int blocks, int ld, int Co, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Co) return;
  float s = 0.f;
  for (int b = 0; b < blocks; ++b) s += partial[(int64_t)b * ld + c];
  out[c] = s;
}

// Weight gradient: the split-K partials [S][Mr][Nr] summed in order and
// written in the parameter layout.  mode 0 (conv): rows co < Co, columns
// k < taps * Ci -> g[co][k]; mode 1 (transposed conv): rows (tap, co) over
// 4 * Co, columns ci -> g[tap][co][ci] (the same row-major order).
__global__ void k_wgrad_finish(const float* __restrict__ part, int S, int Mr, int Nr, int rows, int cols,
                               float* __restrict__ g) {
  GRID_LOOP(t, (int64_t)rows * cols) {
    const int r = (int)(t / cols), c = (int)(t - (int64_t)r * cols);
    float s = 0.f;
    for (int z = 0; z < S; ++z) s += part[((int64_t)z * Mr + r) * Nr + c];
    g[t] = s;
  }
}

// bf16 copy of conv weights [Co][taps*Ci] f32 into [rows][Kp] (zero padded).
__global__ void k_weights_bf16(const float* __restrict__ w, int Co, int K, int rows, int Kp, __bf16* __restrict__ out) {
  GRID_LOOP(t, (int64_t)rows * Kp) {
    const int r = (int)(t / Kp), k = (int)(t - (int64_t)r * Kp);
    out[t] = (__bf16)((r < Co && k < K) ? w[(int64_t)r * K + k] : 0.f);
  }
}

__global__ void k_to_bf16(const float* __restrict__ in, int64_t n, __bf16* __restrict__ out) {
  GRID_LOOP(t, n) out[t] = (__bf16)in[t];
}

// 2x2 average pooling of [B, H, W, C] -> [B, H/2, W/2, C]
__global__ void k_pool(const float* __restrict__ in, int B, int H, int W, int C, float* __restrict__ out) {
  const int h = H / 2, w = W / 2;
  GRID_LOOP(t, (int64_t)B * h * w * C) {
    const int c = (int)(t % C);
    const int64_t p = t / C;
    const int x = (int)(p % w), y = (int)((p / w) % h);
    const int64_t b = p / ((int64_t)h * w);
    const float* r0 = in + ((b * H + 2 * y) * W + 2 * x) * C + c;
    out[t] = 0.25f * (r0[0] + r0[C] + r0[(int64_t)W * C] + r0[(int64_t)W * C + C]);
  }
}

__global__ void k_pool_bwd(const float* __restrict__ dout, int B, int H, int W, int C, float* __restrict__ din) {
  const int h = H / 2, w = W / 2;
  GRID_LOOP(t, (int64_t)B * H * W * C) {
    const int c = (int)(t % C);
    const int64_t p = t / C;
    const int x = (int)(p % W), y = (int)((p / W) % H);
    const int64_t b = p / ((int64_t)H * W);
    din[t] += 0.25f * dout[((b * h + y / 2) * w + x / 2) * C + c];
  }
}

// depth-to-space of the transposed convolution: acc [P_in][4*Co] (rows
// (tap, co)) + bias -> out [B, 2h, 2w, out_ld] at channel offset 0
__global__ void k_up_scatter(const float* __restrict__ acc, int B, int h, int w, int Co, const float* __restrict__ bias,
                             float* __restrict__ out, int out_ld) {
  GRID_LOOP(t, (int64_t)B * 2 * h * 2 * w * Co) {
    const int c = (int)(t % Co);
    const int64_t q = t / Co;
    const int X = (int)(q % (2 * w)), Y = (int)((q / (2 * w)) % (2 * h));
    const int64_t b = q / ((int64_t)4 * h * w);
    const int tap = (Y & 1) * 2 + (X & 1);
    const int64_t pin = (b * h + Y / 2) * w + X / 2;
    out[q * out_ld + c] = acc[pin * 4 * Co + tap * Co + c] + bias[c];
  }
}

// space-to-depth of the output gradient: dc[p_in][(tap, co)] (bf16) and the
// per-block bias partials (sum over pixels, per (tap, co) column); the
// workgroup layout of k_act_bwd
__global__ __launch_bounds__(256) void k_up_gather(const float* __restrict__ dout, int out_ld, int B, int h, int w,
                                                   int Co, __bf16* __restrict__ dc, float* __restrict__ partial) {
  __shared__ float red[8][33];
  const int64_t P = (int64_t)B * h * w;
  const int64_t chunk = (P + gridDim.x - 1) / gridDim.x;
  const int64_t p0 = blockIdx.x * chunk, p1 = min(P, p0 + chunk);
  const int x = threadIdx.x & 31, r = threadIdx.x >> 5;
  for (int g = 0; g < 4 * Co / 32; ++g) {
    const int col = 32 * g + x;
    const int tap = col / Co, c = col - tap * Co;
    float s = 0.f;
    for (int64_t p = p0 + r; p < p1; p += 8) {
      const int px = (int)(p % w), py = (int)((p / w) % h);
      const int64_t b = p / ((int64_t)h * w);
      const int64_t q = (b * 2 * h + 2 * py + (tap >> 1)) * 2 * w + 2 * px + (tap & 1);
      const float v = dout[q * out_ld + c];
      s += v;
      dc[p * 4 * Co + col] = (__bf16)v;
    }
    red[r][x] = s;
    __syncthreads();
    if (r == 0) {
      float t = 0.f;
      for (int qq = 0; qq < 8; ++qq) t += red[qq][x];
      partial[(int64_t)blockIdx.x * 4 * Co + col] = t;
    }
    __syncthreads();
  }
}

// per-channel bias gradient of the transposed conv from the (tap, co) partials
__global__ void k_up_bias_finish(const float* __restrict__ partial, int blocks, int Co, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Co) return;
  float s = 0.f;
  for (int b = 0; b < blocks; ++b)
    for (int tap = 0; tap < 4; ++tap) s += partial[(int64_t)b * 4 * Co + tap * Co + c];
  out[c] = s;
}

// copy a [P][C] tensor into channels [off, off + C) of a [P][ld] tensor, or
// (bwd) add that slice back into a [P][C] gradient
__global__ void k_slice_copy(const float* __restrict__ src, int64_t P, int C, float* __restrict__ dst, int ld, int off) {
  GRID_LOOP(t, P * C) {
    const int64_t p = t / C;
    dst[p * ld + off + (t - p * C)] = src[t];
  }
}
__global__ void k_slice_add(const float* __restrict__ src, int ld, int off, int64_t P, int C, float* __restrict__ dst) {
  GRID_LOOP(t, P * C) {
    const int64_t p = t / C;
    dst[t] += src[p * ld + off + (t - p * C)];
  }
}

// out [B, 2, 32, 32] (NCHW) from the output layer's f32 accumulator [P][ldc]
__global__ void k_out_nchw(const float* __restrict__ acc, int ldc, int B, int HW, const float* __restrict__ bias,
                           float* __restrict__ out) {
  GRID_LOOP(t, (int64_t)B * 2 * HW) {
    const int q = (int)(t % HW);
    const int c = (int)((t / HW) % 2);
    const int64_t b = t / (2 * HW);
    out[t] = acc[(b * HW + q) * ldc + c] + bias[c];
  }
}

// d out (NCHW [B, 2, HW]) -> NHWC [P][2] f32
__global__ void k_nchw_to_nhwc2(const float* __restrict__ d, int B, int HW, float* __restrict__ o) {
  GRID_LOOP(t, (int64_t)B * HW * 2) {
    const int c = (int)(t % 2);
    const int64_t p = t / 2;
    const int64_t b = p / HW;
    o[t] = d[(b * 2 + c) * HW + (p - b * HW)];
  }
}

// loss = mean (out - y)^2, dout = 2 (out - y) / n; fixed-order block sums
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ out, const float* __restrict__ y, int64_t n,
                                             float* __restrict__ dout, float* __restrict__ partial) {
  __shared__ float red[256];
  float s = 0.f;
  const float inv = 1.f / (float)n;
  GRID_LOOP(t, n) {
    const float d = out[t] - y[t];
    s += d * d;
    if (dout) dout[t] = 2.f * d * inv;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_mse_finish(const float* __restrict__ partial, int blocks, int64_t n, float* __restrict__ loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float s = 0.f;
  for (int b = 0; b < blocks; ++b) s += partial[b];
  loss[0] = s / (float)n;
}

// Net input: log(1 + I) of a fftshifted pattern [M, M] average-pooled to
// 64 x 64, and the window centre's offset from the canvas centre in units of
// the canvas (two constant planes); out NHWC [P, 64, 64, 3].
__global__ void k_ws_inputs(const float* __restrict__ intens, int P, int M, const int32_t* __restrict__ tl,
                            int canvas, float* __restrict__ out) {
  const int f = M / 64;
  GRID_LOOP(t, (int64_t)P * 64 * 64) {
    const int x = (int)(t % 64), y = (int)((t / 64) % 64);
    const int64_t p = t / 4096;
    const float* src = intens + p * M * M;
    float s = 0.f;
    for (int a = 0; a < f; ++a)
      for (int b = 0; b < f; ++b) s += log1pf(src[(int64_t)(y * f + a) * M + x * f + b]);
    out[t * 3] = s / (float)(f * f);
    out[t * 3 + 1] = ((float)tl[2 * p] + 0.5f * M) / (float)canvas - 0.5f;
    out[t * 3 + 2] = ((float)tl[2 * p + 1] + 0.5f * M) / (float)canvas - 0.5f;
  }
}

// ------------------------------------------------------------------ host plan
static unsigned grid1(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 16384) b = 16384;
  return (unsigned)(b > 0 ? b : 1);
}
static int pad32(int v) { return (v + 31) / 32 * 32; }

enum LayerKind { CONV3, CONV1, UP };

struct NetLayer {
  LayerKind kind;
  int ci, co, res;    // input channels, output channels, input resolution
  size_t w, b;        // parameter offsets (floats)
};

struct Net {
  int c;
  std::vector<NetLayer> L;   // e1a e1b e2a e2b e3a e3b e4a e4b u3 d3a d3b u2 d2a d2b out
  size_t n_params;
};

static Net make_net(int c) {
  Net n;
  n.c = c;
  size_t off = 0;
  auto add = [&](LayerKind k, int ci, int co, int res) {
    NetLayer l;
    l.kind = k; l.ci = ci; l.co = co; l.res = res;
    const int taps = k == CONV3 ? 9 : (k == UP ? 4 : 1);
    l.w = off; off += (size_t)taps * ci * co;
    l.b = off; off += co;
    n.L.push_back(l);
  };
  add(CONV3, 3, c, 64);          add(CONV3, c, c, 64);
  add(CONV3, c, 2 * c, 32);      add(CONV3, 2 * c, 2 * c, 32);
  add(CONV3, 2 * c, 4 * c, 16);  add(CONV3, 4 * c, 4 * c, 16);
  add(CONV3, 4 * c, 8 * c, 8);   add(CONV3, 8 * c, 8 * c, 8);
  add(UP, 8 * c, 4 * c, 8);
  add(CONV3, 8 * c, 4 * c, 16);  add(CONV3, 4 * c, 4 * c, 16);
  add(UP, 4 * c, 2 * c, 16);
  add(CONV3, 4 * c, 2 * c, 32);  add(CONV3, 2 * c, 2 * c, 32);
  add(CONV1, 2 * c, 2, 32);
  n.n_params = off;
  return n;
}

// Workspace layout for a batch B (all offsets 256-B aligned).
struct NetWork {
  // forward activations (f32 NHWC); index by layer: pre[i], post[i]
  size_t x, pre[15], post[15], pool[3], cat3, cat2, out_acc;
  // gradients of the post activations / pool outputs / cat buffers
  size_t dpost[15], dpool[3], dcat3, dcat2, dx;
  // scratch: im2col (bf16), GEMM accumulator / dcol (f32), bf16 dpre,
  // bf16 weights, split-K partials, column-sum partials, mse partials
  size_t col, acc, dpre, wb, part, csum, mse;
  size_t total;
};

constexpr int kSplitMax = 64;
constexpr int kColBlocks = 1024;
constexpr int kMseBlocks = 1024;

static NetWork plan_work(const Net& n, int B) {
  NetWork w;
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
  auto act = [&](int res, int C) { return take((size_t)B * res * res * C * sizeof(float)); };
  const int c = n.c;
  w.x = act(64, 3);
  size_t col_max = 0, acc_max = 0, dpre_max = 0, wb_max = 0, part_max = 0;
  for (int i = 0; i < 15; ++i) {
    const NetLayer& l = n.L[i];
    const int ores = l.kind == UP ? 2 * l.res : l.res;
    const int64_t P = (int64_t)B * l.res * l.res;   // GEMM rows (input pixels)
    const int taps = l.kind == CONV3 ? 9 : (l.kind == UP ? 4 : 1);
    const int Kp = l.kind == UP ? pad32(l.ci) : pad32(taps * l.ci);
    const int Np = l.kind == UP ? 4 * l.co : pad32(l.co);
    w.pre[i] = l.kind == UP ? 0 : act(ores, l.co);
    w.post[i] = (l.kind == UP || i == 14) ? 0 : act(ores, l.co);
    col_max = std::max(col_max, (size_t)P * Kp * 2);
    // f32 accumulator: forward [P][Np], dcol [P][Kp]
    acc_max = std::max(acc_max, (size_t)P * std::max(Np, Kp) * 4);
    dpre_max = std::max(dpre_max, (size_t)P * Np * 2);
    wb_max = std::max(wb_max, (size_t)Np * Kp * 2);
    part_max = std::max(part_max, (size_t)kSplitMax * Np * Kp * 4);
  }
  w.pool[0] = act(32, c); w.pool[1] = act(16, 2 * c); w.pool[2] = act(8, 4 * c);
  w.cat3 = act(16, 8 * c); w.cat2 = act(32, 4 * c);
  w.out_acc = 0;
  for (int i = 0; i < 15; ++i) {
    const NetLayer& l = n.L[i];
    const int ores = l.kind == UP ? 2 * l.res : l.res;
    w.dpost[i] = (l.kind == UP || i == 14) ? 0 : act(ores, l.co);
  }
  w.dpool[0] = act(32, c); w.dpool[1] = act(16, 2 * c); w.dpool[2] = act(8, 4 * c);
  w.dcat3 = act(16, 8 * c); w.dcat2 = act(32, 4 * c);
  w.dx = act(64, 3);
  w.col = take(col_max);
  w.acc = take(acc_max);
  w.dpre = take(dpre_max);
  w.wb = take(wb_max);
  w.part = take(part_max);
  w.csum = take((size_t)kColBlocks * 1024 * 4);   // up to 1024 columns
  w.mse = take((size_t)kMseBlocks * 4);
  w.total = off;
  return w;
}

struct Ctx {
  const Net& n;
  const NetWork& w;
  char* base;
  hipStream_t s;
  int B;
  float* f(size_t off) const { return (float*)(base + off); }
  __bf16* h(size_t off) const { return (__bf16*)(base + off); }
};

template <int TA, int TB>
static void gemm(const Ctx& x, const __bf16* A, const __bf16* Bm, float* C, int M, int N, int K, int lda, int ldb,
                 int ldc, int splits, int64_t strideC) {
  const int kchunk = ((K + splits - 1) / splits + kGemmK - 1) / kGemmK * kGemmK;
  const int S = (K + kchunk - 1) / kchunk;
  dim3 grid((N + kGemmTile - 1) / kGemmTile, (M + kGemmTile - 1) / kGemmTile, S);
  hipLaunchKernelGGL((k_gemm_bf16<TA, TB>), grid, dim3(256), 0, x.s, A, Bm, C, M, N, K, lda, ldb, ldc, kchunk,
                     strideC);
}

// splits of a weight-gradient GEMM over P pixels: enough workgroups to fill
// the GPU, chunks of at least 512 pixels
static int wgrad_splits(int64_t P, int M, int N) {
  const int tiles = ((M + kGemmTile - 1) / kGemmTile) * ((N + kGemmTile - 1) / kGemmTile);
  int s = (1024 + tiles - 1) / tiles;
  const int64_t maxs = P / 512;
  if (s > maxs) s = (int)(maxs > 0 ? maxs : 1);
  if (s > kSplitMax) s = kSplitMax;
  return s > 0 ? s : 1;
}

// forward of layer i from input act `in` [B, res, res, ci]; output written to
// post (and pre) or, for UP, to channels [0, co) of `out_cat`
static void layer_fwd(const Ctx& x, int i, const float* params, const float* in, float* out_cat, int cat_ld) {
  const NetLayer& l = x.n.L[i];
  const int64_t P = (int64_t)x.B * l.res * l.res;
  if (l.kind == UP) {
    const int Kp = pad32(l.ci), Np = 4 * l.co;
    hipLaunchKernelGGL(k_to_bf16, dim3(grid1(P * l.ci)), dim3(256), 0, x.s, in, P * l.ci, x.h(x.w.col));
    hipLaunchKernelGGL(k_weights_bf16, dim3(grid1((int64_t)Np * Kp)), dim3(256), 0, x.s, params + l.w, Np, l.ci, Np,
                       Kp, x.h(x.w.wb));
    gemm<0, 0>(x, x.h(x.w.col), x.h(x.w.wb), x.f(x.w.acc), (int)P, Np, Kp, Kp, Kp, Np, 1, 0);
    hipLaunchKernelGGL(k_up_scatter, dim3(grid1(P * 4 * l.co)), dim3(256), 0, x.s, x.f(x.w.acc), x.B, l.res, l.res,
                       l.co, params + l.b, out_cat, cat_ld);
    return;
  }
  const int taps = l.kind == CONV3 ? 9 : 1;
  const int Kp = pad32(taps * l.ci), Np = pad32(l.co);
  hipLaunchKernelGGL(k_im2col, dim3(grid1(P * (Kp / 8))), dim3(256), 0, x.s, in, x.B, l.res, l.res, l.ci,
                     l.kind == CONV3 ? 3 : 1, Kp, x.h(x.w.col));
  hipLaunchKernelGGL(k_weights_bf16, dim3(grid1((int64_t)Np * Kp)), dim3(256), 0, x.s, params + l.w, l.co,
                     taps * l.ci, Np, Kp, x.h(x.w.wb));
  gemm<0, 0>(x, x.h(x.w.col), x.h(x.w.wb), x.f(x.w.acc), (int)P, Np, Kp, Kp, Kp, Np, 1, 0);
}

// backward of a conv layer: dpost -> dpre (bf16), bias / weight gradients,
// input gradient added into din.  `in` is the layer's forward input.
static void conv_bwd(const Ctx& x, int i, const float* params, float* grads, const float* in, const float* dpost,
                     int dpost_ld, int dpost_off, int act, float* din) {
  const NetLayer& l = x.n.L[i];
  const int64_t P = (int64_t)x.B * l.res * l.res;
  const int taps = l.kind == CONV3 ? 9 : 1, ks = l.kind == CONV3 ? 3 : 1;
  const int Kp = pad32(taps * l.ci), Np = pad32(l.co);
  hipLaunchKernelGGL(k_act_bwd, dim3(kColBlocks), dim3(256), 0, x.s, dpost, dpost_ld, dpost_off,
                     act ? x.f(x.w.pre[i]) : nullptr, P, l.co, Np, act, x.h(x.w.dpre), x.f(x.w.csum));
  hipLaunchKernelGGL(k_colsum_finish, dim3((l.co + 255) / 256), dim3(256), 0, x.s, x.f(x.w.csum), kColBlocks, Np, l.co,
                     grads + l.b);
  // weight gradient: dW[co][k] = sum_p dpre[p][co] col[p][k]
  hipLaunchKernelGGL(k_im2col, dim3(grid1(P * (Kp / 8))), dim3(256), 0, x.s, in, x.B, l.res, l.res, l.ci, ks, Kp,
                     x.h(x.w.col));
  const int S = wgrad_splits(P, Np, Kp);
  gemm<1, 1>(x, x.h(x.w.dpre), x.h(x.w.col), x.f(x.w.part), Np, Kp, (int)P, Np, Kp, Kp, S, (int64_t)Np * Kp);
  const int kchunk = (((int)P + S - 1) / S + kGemmK - 1) / kGemmK * kGemmK;
  const int Sr = ((int)P + kchunk - 1) / kchunk;
  hipLaunchKernelGGL(k_wgrad_finish, dim3(grid1((int64_t)l.co * taps * l.ci)), dim3(256), 0, x.s, x.f(x.w.part), Sr,
                     Np, Kp, l.co, taps * l.ci, grads + l.w);
  if (!din) return;
  // input gradient: dcol[p][k] = sum_co dpre[p][co] W[co][k]; col2im into din
  hipLaunchKernelGGL(k_weights_bf16, dim3(grid1((int64_t)Np * Kp)), dim3(256), 0, x.s, params + l.w, l.co,
                     taps * l.ci, Np, Kp, x.h(x.w.wb));
  gemm<0, 1>(x, x.h(x.w.dpre), x.h(x.w.wb), x.f(x.w.acc), (int)P, Kp, Np, Np, Kp, Kp, 1, 0);
  hipLaunchKernelGGL(k_col2im, dim3(grid1(P * l.ci)), dim3(256), 0, x.s, x.f(x.w.acc), x.B, l.res, l.res, l.ci, ks, Kp,
                     din);
}

// backward of a transposed conv whose output is channels [0, co) of dcat
static void up_bwd(const Ctx& x, int i, const float* params, float* grads, const float* in, const float* dcat,
                   int cat_ld, float* din) {
  const NetLayer& l = x.n.L[i];
  const int64_t P = (int64_t)x.B * l.res * l.res;
  const int Kp = pad32(l.ci), Np = 4 * l.co;
  hipLaunchKernelGGL(k_up_gather, dim3(kColBlocks), dim3(256), 0, x.s, dcat, cat_ld, x.B, l.res, l.res, l.co,
                     x.h(x.w.dpre), x.f(x.w.csum));
  hipLaunchKernelGGL(k_up_bias_finish, dim3((l.co + 255) / 256), dim3(256), 0, x.s, x.f(x.w.csum), kColBlocks, l.co,
                     grads + l.b);
  hipLaunchKernelGGL(k_to_bf16, dim3(grid1(P * l.ci)), dim3(256), 0, x.s, in, P * l.ci, x.h(x.w.col));
  const int S = wgrad_splits(P, Np, Kp);
  gemm<1, 1>(x, x.h(x.w.dpre), x.h(x.w.col), x.f(x.w.part), Np, Kp, (int)P, Np, Kp, Kp, S, (int64_t)Np * Kp);
  const int kchunk = (((int)P + S - 1) / S + kGemmK - 1) / kGemmK * kGemmK;
  const int Sr = ((int)P + kchunk - 1) / kchunk;
  hipLaunchKernelGGL(k_wgrad_finish, dim3(grid1((int64_t)Np * l.ci)), dim3(256), 0, x.s, x.f(x.w.part), Sr, Np, Kp,
                     Np, l.ci, grads + l.w);
  hipLaunchKernelGGL(k_weights_bf16, dim3(grid1((int64_t)Np * Kp)), dim3(256), 0, x.s, params + l.w, Np, l.ci, Np, Kp,
                     x.h(x.w.wb));
  gemm<0, 1>(x, x.h(x.w.dpre), x.h(x.w.wb), x.f(x.w.acc), (int)P, Kp, Np, Np, Kp, Kp, 1, 0);
  // din[p][ci] += acc[p][ci] (Kp = ci here: channel counts are multiples of 32)
  hipLaunchKernelGGL(k_slice_add, dim3(grid1(P * l.ci)), dim3(256), 0, x.s, x.f(x.w.acc), Kp, 0, P, l.ci, din);
}

static int check_net(int base, int batch, const char* what) {
  PTY_REQUIRE(base >= 16 && base <= 64 && base % 16 == 0, PTY_EUNSUP, "%s: base width %d (16, 32, 48 or 64)", what,
              base);
  // the GEMM grid's y extent is (batch * 64^2) / 128 workgroups: <= 32768
  PTY_REQUIRE(batch >= 1 && batch <= 1024, PTY_EUNSUP, "%s: batch %d (1..1024)", what, batch);
  return PTY_OK;
}

}  // namespace pty

using namespace pty;

extern "C" size_t pty_wsnet_param_count(int32_t base) {
  if (base < 16 || base > 64 || base % 16) return 0;
  return make_net(base).n_params;
}

extern "C" size_t pty_wsnet_workspace(int32_t base, int32_t batch) {
  if (check_net(base, batch, "wsnet_workspace")) return 0;
  return plan_work(make_net(base), batch).total;
}

extern "C" int pty_wsnet_inputs(const float* intens, int32_t P, int32_t M, const int32_t* tl, int32_t canvas,
                                float* x, hipStream_t stream) {
  PTY_REQUIRE(intens && tl && x, PTY_EINVAL, "wsnet_inputs: null pointer");
  PTY_REQUIRE(P >= 1 && M >= 64 && M % 64 == 0 && canvas >= M, PTY_EUNSUP,
              "wsnet_inputs: %d patterns of %d (a multiple of 64) on canvas %d", P, M, canvas);
  hipLaunchKernelGGL(k_ws_inputs, dim3(grid1((int64_t)P * 4096)), dim3(256), 0, stream, intens, P, M, tl, canvas, x);
  return check_launch("wsnet_inputs");
}

extern "C" int pty_wsnet_forward(int32_t base, const float* params, const float* x, int32_t batch, float* out,
                                 void* work, size_t work_bytes, hipStream_t stream) {
  int rc = check_net(base, batch, "wsnet_forward");
  if (rc) return rc;
  PTY_REQUIRE(params && x && out, PTY_EINVAL, "wsnet_forward: null pointer");
  const Net n = make_net(base);
  const NetWork w = plan_work(n, batch);
  PTY_REQUIRE(work && work_bytes >= w.total, PTY_EWORK, "wsnet_forward: workspace %zu < %zu", work_bytes, w.total);
  Ctx c{n, w, (char*)work, stream, batch};
  const int B = batch;
  // the input is kept with the activations: the backward needs it
  const hipError_t e = hipMemcpyAsync(c.f(w.x), x, (size_t)B * 64 * 64 * 3 * sizeof(float), hipMemcpyDeviceToDevice,
                                      stream);
  if (e != hipSuccess) return set_error(PTY_ELAUNCH, "wsnet_forward: input copy: %s", hipGetErrorString(e));
  const float* cur = c.f(w.x);
  int pi = 0;
  for (int i = 0; i < 15; ++i) {
    const NetLayer& l = n.L[i];
    const int64_t P = (int64_t)B * l.res * l.res;
    if (i == 8 || i == 11) {
      float* cat = i == 8 ? c.f(w.cat3) : c.f(w.cat2);
      const int ld = 2 * l.co;
      layer_fwd(c, i, params, cur, cat, ld);
      // skip connection: e3b (layer 5) or e2b (layer 3) into channels [co, 2co)
      const int skip = i == 8 ? 5 : 3;
      const int64_t Po = (int64_t)B * 4 * l.res * l.res;
      hipLaunchKernelGGL(k_slice_copy, dim3(grid1(Po * l.co)), dim3(256), 0, stream, c.f(w.post[skip]), Po, l.co, cat,
                         ld, l.co);
      cur = cat;
      continue;
    }
    layer_fwd(c, i, params, cur, nullptr, 0);
    const int Np = pad32(l.co);
    if (i == 14) {
      hipLaunchKernelGGL(k_out_nchw, dim3(grid1(P * 2)), dim3(256), 0, stream, c.f(w.acc), Np, B, l.res * l.res,
                         params + l.b, out);
      break;
    }
    hipLaunchKernelGGL(k_bias_act, dim3(grid1(P * l.co)), dim3(256), 0, stream, c.f(w.acc), P, l.co, Np,
                       params + l.b, 1, c.f(w.pre[i]), c.f(w.post[i]), l.co, 0);
    cur = c.f(w.post[i]);
    if (i == 1 || i == 3 || i == 5) {   // pool after each encoder level but the last
      hipLaunchKernelGGL(k_pool, dim3(grid1(P / 4 * l.co)), dim3(256), 0, stream, cur, B, l.res, l.res, l.co,
                         c.f(w.pool[pi]));
      cur = c.f(w.pool[pi++]);
    }
  }
  return check_launch("wsnet_forward");
}

extern "C" int pty_wsnet_mse(const float* out, const float* target, int64_t n, float* loss, float* dout, void* work,
                             size_t work_bytes, hipStream_t stream) {
  PTY_REQUIRE(out && target && loss && work && n > 0, PTY_EINVAL, "wsnet_mse: null pointer or n = 0");
  PTY_REQUIRE(work_bytes >= (size_t)kMseBlocks * 4, PTY_EWORK, "wsnet_mse: workspace %zu < %zu", work_bytes,
              (size_t)kMseBlocks * 4);
  hipLaunchKernelGGL(k_mse, dim3(kMseBlocks), dim3(256), 0, stream, out, target, n, dout, (float*)work);
  hipLaunchKernelGGL(k_mse_finish, dim3(1), dim3(64), 0, stream, (const float*)work, kMseBlocks, n, loss);
  return check_launch("wsnet_mse");
}

extern "C" int pty_wsnet_backward(int32_t base, const float* params, int32_t batch, const float* dout, float* grads,
                                  void* work, size_t work_bytes, hipStream_t stream) {
  int rc = check_net(base, batch, "wsnet_backward");
  if (rc) return rc;
  PTY_REQUIRE(params && dout && grads, PTY_EINVAL, "wsnet_backward: null pointer");
  const Net n = make_net(base);
  const NetWork w = plan_work(n, batch);
  PTY_REQUIRE(work && work_bytes >= w.total, PTY_EWORK, "wsnet_backward: workspace %zu < %zu", work_bytes, w.total);
  Ctx c{n, w, (char*)work, stream, batch};
  const int B = batch, ch = base;
  // gradient buffers start at zero (every backward step adds into them)
  auto zero = [&](size_t off, int64_t count) {
    hipLaunchKernelGGL(k_ws_zero, dim3(grid1(count)), dim3(256), 0, stream, c.f(off), count);
  };
  for (int i = 0; i < 14; ++i) {
    const NetLayer& l = n.L[i];
    if (l.kind == UP) continue;
    zero(w.dpost[i], (int64_t)B * l.res * l.res * l.co);
  }
  zero(w.dpool[0], (int64_t)B * 32 * 32 * ch);
  zero(w.dpool[1], (int64_t)B * 16 * 16 * 2 * ch);
  zero(w.dpool[2], (int64_t)B * 8 * 8 * 4 * ch);
  zero(w.dcat3, (int64_t)B * 16 * 16 * 8 * ch);
  zero(w.dcat2, (int64_t)B * 32 * 32 * 4 * ch);
  // output layer (1x1, no activation): d out NCHW -> NHWC [P][2]
  const int64_t P32 = (int64_t)B * 32 * 32;
  // the output gradient NCHW -> NHWC [P][2], staged in the (otherwise
  // unused) input-gradient buffer
  float* d_out = c.f(w.dx);
  hipLaunchKernelGGL(k_nchw_to_nhwc2, dim3(grid1(P32 * 2)), dim3(256), 0, stream, dout, B, 32 * 32, d_out);
  conv_bwd(c, 14, params, grads, c.f(w.post[13]), d_out, 2, 0, 0, c.f(w.dpost[13]));
  conv_bwd(c, 13, params, grads, c.f(w.post[12]), c.f(w.dpost[13]), n.L[13].co, 0, 1, c.f(w.dpost[12]));
  conv_bwd(c, 12, params, grads, c.f(w.cat2), c.f(w.dpost[12]), n.L[12].co, 0, 1, c.f(w.dcat2));
  // cat2 = [u2 (2c), e2b (2c)]
  hipLaunchKernelGGL(k_slice_add, dim3(grid1(P32 * 2 * ch)), dim3(256), 0, stream, c.f(w.dcat2), 4 * ch, 2 * ch, P32,
                     2 * ch, c.f(w.dpost[3]));
  up_bwd(c, 11, params, grads, c.f(w.post[10]), c.f(w.dcat2), 4 * ch, c.f(w.dpost[10]));
  conv_bwd(c, 10, params, grads, c.f(w.post[9]), c.f(w.dpost[10]), n.L[10].co, 0, 1, c.f(w.dpost[9]));
  conv_bwd(c, 9, params, grads, c.f(w.cat3), c.f(w.dpost[9]), n.L[9].co, 0, 1, c.f(w.dcat3));
  const int64_t P16 = (int64_t)B * 16 * 16;
  hipLaunchKernelGGL(k_slice_add, dim3(grid1(P16 * 4 * ch)), dim3(256), 0, stream, c.f(w.dcat3), 8 * ch, 4 * ch, P16,
                     4 * ch, c.f(w.dpost[5]));
  up_bwd(c, 8, params, grads, c.f(w.post[7]), c.f(w.dcat3), 8 * ch, c.f(w.dpost[7]));
  // encoder, deepest first
  conv_bwd(c, 7, params, grads, c.f(w.post[6]), c.f(w.dpost[7]), n.L[7].co, 0, 1, c.f(w.dpost[6]));
  conv_bwd(c, 6, params, grads, c.f(w.pool[2]), c.f(w.dpost[6]), n.L[6].co, 0, 1, c.f(w.dpool[2]));
  hipLaunchKernelGGL(k_pool_bwd, dim3(grid1(P16 * 4 * ch)), dim3(256), 0, stream, c.f(w.dpool[2]), B, 16, 16, 4 * ch,
                     c.f(w.dpost[5]));
  conv_bwd(c, 5, params, grads, c.f(w.post[4]), c.f(w.dpost[5]), n.L[5].co, 0, 1, c.f(w.dpost[4]));
  conv_bwd(c, 4, params, grads, c.f(w.pool[1]), c.f(w.dpost[4]), n.L[4].co, 0, 1, c.f(w.dpool[1]));
  hipLaunchKernelGGL(k_pool_bwd, dim3(grid1(P32 * 2 * ch)), dim3(256), 0, stream, c.f(w.dpool[1]), B, 32, 32, 2 * ch,
                     c.f(w.dpost[3]));
  conv_bwd(c, 3, params, grads, c.f(w.post[2]), c.f(w.dpost[3]), n.L[3].co, 0, 1, c.f(w.dpost[2]));
  conv_bwd(c, 2, params, grads, c.f(w.pool[0]), c.f(w.dpost[2]), n.L[2].co, 0, 1, c.f(w.dpool[0]));
  const int64_t P64 = (int64_t)B * 64 * 64;
  hipLaunchKernelGGL(k_pool_bwd, dim3(grid1(P64 * ch)), dim3(256), 0, stream, c.f(w.dpool[0]), B, 64, 64, ch,
                     c.f(w.dpost[1]));
  conv_bwd(c, 1, params, grads, c.f(w.post[0]), c.f(w.dpost[1]), n.L[1].co, 0, 1, c.f(w.dpost[0]));
  conv_bwd(c, 0, params, grads, c.f(w.x), c.f(w.dpost[0]), n.L[0].co, 0, 1, nullptr);
  return check_launch("wsnet_backward");
}
