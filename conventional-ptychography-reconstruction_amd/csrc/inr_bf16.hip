// BF16 variant of the fused INR kernels: v_mfma_f32_32x32x16_bf16 (gfx950)
// with FP32 accumulation.  Same tile layout and chain rule as the FP32
// kernels in inr.hip; per MFMA step a lane supplies 8 consecutive K values,
// lane (x, h) holding K = 8h + t of the step (for A and for B alike), so the
// kappa maps are applied to value index 8q + t.
//
// * Weight operands are converted once per workgroup into LDS in operand
//   order (16 B per lane per step): one ds_read_b128 per MFMA operand.
// * Activations are packed to bf16 in registers for the B operand; the
//   elementwise Gabor activation and its adjoint stay FP32.
// * The weight-gradient GEMMs take their point-major operands from bf16 LDS
//   staging rows of 32 points (80 B stride: 16-B aligned, conflict-free).
// * Bias gradients are FP32 sums of the bf16 staged values.
#include "inr_common.h"

namespace pty {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ v16f mfma16(uint4 a, uint4 b, v16f c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                  0, 0);
}

__device__ __forceinline__ uint4 pack8(const float* x) {
  bf16x8 r;
#pragma unroll
  for (int t = 0; t < 8; ++t) r[t] = (__bf16)x[t];
  return __builtin_bit_cast(uint4, r);
}

__device__ __forceinline__ uint4 neg8(uint4 a) {
  a.x ^= 0x80008000u;
  a.y ^= 0x80008000u;
  a.z ^= 0x80008000u;
  a.w ^= 0x80008000u;
  return a;
}

__device__ __forceinline__ float sum8(uint4 a) {
  const uint32_t w[4] = {a.x, a.y, a.z, a.w};
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    s += __uint_as_float(w[k] << 16);
    s += __uint_as_float(w[k] & 0xFFFF0000u);
  }
  return s;
}

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f2bf(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }

// ---------------------------------------------------------------- LDS
template <int H>
struct Bf16Fwd {
  uint4 w0[2][64];         // layer 0: lane (j,h), step q: W0[j][kappa0(8q+t, h)]
  uint4 a[H][2][4][64];    // hidden k, M block (Re rows, Im rows), step q
};

template <int H>
struct Bf16Bwd {
  uint4 w0t[2][64];        // layer-0 transpose: W0[crow(8q+t, h)][phi0(j)]
  uint4 at[H][2][4][64];   // hidden transpose, block = (Re inputs, Im inputs)
};

template <int H>
struct SmallLds {
  float b0[32];
  float cr[H][32];
  float ci[H][32];
  float om[H + 1];
  float wor[32], woi[32];
  float bo[2];
};

template <int H>
__device__ void load_small(SmallLds<H>& s, const PtyMlp& m) {
  for (int t = threadIdx.x; t < 32; t += blockDim.x) {
    s.b0[t] = m.b0[t];
    for (int k = 0; k < H; ++k) {
      s.cr[k][t] = m.bh[k][2 * t];
      s.ci[k][t] = m.bh[k][2 * t + 1];
    }
    s.wor[t] = m.wo[2 * t];
    s.woi[t] = m.wo[2 * t + 1];
  }
  if (threadIdx.x <= H) s.om[threadIdx.x] = *m.omega[threadIdx.x];
  if (threadIdx.x == 0) { s.bo[0] = m.bo[0]; s.bo[1] = m.bo[1]; }
}

template <int H>
__device__ void build_fwd_ops(Bf16Fwd<H>& o, const PtyMlp& m) {
  for (int idx = threadIdx.x; idx < 2 * 64; idx += blockDim.x) {
    const int q = idx >> 6, lane = idx & 63, j = lane & 31, h = lane >> 5;
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int f = kappa0(8 * q + t, h);
      v[t] = f < m.n_in ? m.w0[j * m.n_in + f] : 0.f;
    }
    o.w0[q][lane] = pack8(v);
  }
  for (int idx = threadIdx.x; idx < H * 2 * 4 * 64; idx += blockDim.x) {
    const int lane = idx & 63, q = (idx >> 6) & 3, mb = (idx >> 8) & 1, k = idx >> 9;
    const int j = lane & 31, h = lane >> 5;
    const float2* A = (const float2*)m.wh[k];
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int s = 8 * q + t;
      const float2 w = A[j * 32 + crow(s & 15, h)];   // A[out j][in u]
      v[t] = mb == 0 ? (s < 16 ? w.x : -w.y) : (s < 16 ? w.y : w.x);
    }
    o.a[k][mb][q][lane] = pack8(v);
  }
}

template <int H>
__device__ void build_bwd_ops(Bf16Bwd<H>& o, const PtyMlp& m) {
  for (int idx = threadIdx.x; idx < 2 * 64; idx += blockDim.x) {
    const int q = idx >> 6, lane = idx & 63, j = lane & 31, h = lane >> 5;
    const int f = phi0(j);
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = f < m.n_in ? m.w0[crow(8 * q + t, h) * m.n_in + f] : 0.f;
    o.w0t[q][lane] = pack8(v);
  }
  for (int idx = threadIdx.x; idx < H * 2 * 4 * 64; idx += blockDim.x) {
    const int lane = idx & 63, q = (idx >> 6) & 3, mb = (idx >> 8) & 1, k = idx >> 9;
    const int j = lane & 31, h = lane >> 5;
    const float2* A = (const float2*)m.wh[k];
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int s = 8 * q + t;
      const float2 w = A[crow(s & 15, h) * 32 + j];   // A[out u][in j]
      v[t] = mb == 0 ? (s < 16 ? w.x : w.y) : (s < 16 ? -w.y : w.x);
    }
    o.at[k][mb][q][lane] = pack8(v);
  }
}

template <int H>
__device__ __forceinline__ v16f layer0_bf16(const Bf16Fwd<H>& o, const SmallLds<H>& s, const float feat[16],
                                            int lane, int h) {
  v16f acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = s.b0[crow(v, h)];
  acc = mfma16(o.w0[0][lane], pack8(feat), acc);
  acc = mfma16(o.w0[1][lane], pack8(feat + 8), acc);
  return acc;
}

template <int H>
__device__ __forceinline__ void hidden_bf16(const Bf16Fwd<H>& o, const SmallLds<H>& s, int k, const float zr[16],
                                            const float zi[16], int lane, int h, v16f& vr, v16f& vi) {
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    vr[v] = s.cr[k][crow(v, h)];
    vi[v] = s.ci[k][crow(v, h)];
  }
  const uint4 b[4] = {pack8(zr), pack8(zr + 8), pack8(zi), pack8(zi + 8)};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    vr = mfma16(o.a[k][0][q][lane], b[q], vr);
    vi = mfma16(o.a[k][1][q][lane], b[q], vi);
  }
}

// Hidden layer k from the packed bf16 input (4 operands in kappa1 order).
template <int H>
__device__ __forceinline__ void hidden_bf16p(const Bf16Fwd<H>& o, const SmallLds<H>& s, int k, const uint4 b[4],
                                             int lane, int h, v16f& vr, v16f& vi) {
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    vr[v] = s.cr[k][crow(v, h)];
    vi[v] = s.ci[k][crow(v, h)];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    vr = mfma16(o.a[k][0][q][lane], b[q], vr);
    vi = mfma16(o.a[k][1][q][lane], b[q], vi);
  }
}

// unpack value v (0..15) of the Re (part 0) or Im (part 1) half of a packed operand set
__device__ __forceinline__ float unpack_val(const uint4 b[4], int part, int v) {
  const uint4 w = b[2 * part + (v >> 3)];
  const int t = v & 7;
  const uint32_t word = (t < 2) ? w.x : (t < 4) ? w.y : (t < 6) ? w.z : w.w;
  return (t & 1) ? __uint_as_float(word & 0xFFFF0000u) : __uint_as_float(word << 16);
}

// ============================================================== forward
template <int H>
__global__ __launch_bounds__(kThreads) void k_inr_fwd_bf16(PtyGrid g, PtyMlp m, const float2* __restrict__ table,
                                                           PtyPoints pts, float2* __restrict__ out) {
  __shared__ Bf16Fwd<H> o;
  __shared__ SmallLds<H> s;
  build_fwd_ops<H>(o, m);
  load_small<H>(s, m);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  const int64_t ntiles = (pts.n + 31) / 32;
  for (int64_t tile = (int64_t)blockIdx.x * kWaves + wave; tile < ntiles; tile += (int64_t)gridDim.x * kWaves) {
    asm volatile("" ::: "memory");
    const int64_t p = tile * 32 + j;
    const bool valid = p < pts.n;
    float x0 = 0.5f, x1 = 0.5f;
    if (valid) point_coords(pts, p, x0, x1);
    float feat[16];
    encode_lane(g, x0, x1, h, table, feat);
    const v16f u0 = layer0_bf16<H>(o, s, feat, lane, h);
    float zr[16], zi[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) cis(s.om[0] * u0[v], zr[v], zi[v]);
#pragma unroll
    for (int k = 0; k < H; ++k) {
      v16f vr, vi;
      hidden_bf16<H>(o, s, k, zr, zi, lane, h, vr, vi);
      gabor_act(s.om[k + 1], vr, vi, zr, zi);
    }
    float orr = 0.f, oi = 0.f;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int u = crow(v, h);
      orr += s.wor[u] * zr[v] - s.woi[u] * zi[v];
      oi += s.wor[u] * zi[v] + s.woi[u] * zr[v];
    }
    orr += __shfl_xor(orr, 32);
    oi += __shfl_xor(oi, 32);
    if (valid && h == 0) out[p] = make_float2(orr + s.bo[0], oi + s.bo[1]);
  }
}

// ============================================================== backward
constexpr int BS = 40;  // bf16 row stride of the staging arrays (80 B)

template <int H>
struct alignas(16) Bf16Stage {
  uint16_t X[H][64 * BS];  // input z of hidden layer k, o-major, column = point
  uint16_t F[32 * BS];     // features, feature-major
  uint16_t G[64 * BS];     // gradient staging
};

__device__ __forceinline__ uint4 row8(const uint16_t* base, int row, int col0) {
  return *(const uint4*)(base + row * BS + col0);
}

template <int H>
__global__ __launch_bounds__(kThreads) void k_inr_bwd_bf16(PtyGrid g, PtyMlp m, const float2* __restrict__ table,
                                                           float2* __restrict__ gfeat, PtyPoints pts,
                                                           const float2* __restrict__ dout,
                                                           float* __restrict__ partial) {
  __shared__ Bf16Fwd<H> o;
  __shared__ Bf16Bwd<H> ot;
  __shared__ SmallLds<H> s;
  __shared__ Bf16Stage<H> st[kWaves];
  build_fwd_ops<H>(o, m);
  build_bwd_ops<H>(ot, m);
  load_small<H>(s, m);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  Bf16Stage<H>& S = st[wave];

  v16f dWr[H], dWi[H];
  v16f dW0;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    dW0[v] = 0.f;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      dWr[k][v] = 0.f;
      dWi[k][v] = 0.f;
    }
  }
  float acc_c[H][2];
  float acc_om[H + 1];
#pragma unroll
  for (int k = 0; k < H; ++k) acc_c[k][0] = acc_c[k][1] = 0.f;
#pragma unroll
  for (int k = 0; k <= H; ++k) acc_om[k] = 0.f;
  float acc_b0 = 0.f, acc_wo = 0.f, acc_bor = 0.f, acc_boi = 0.f;

  const int64_t ntiles = (pts.n + 31) / 32;
  for (int64_t tile = (int64_t)blockIdx.x * kWaves + wave; tile < ntiles; tile += (int64_t)gridDim.x * kWaves) {
    asm volatile("" ::: "memory");
    const int64_t p = tile * 32 + j;
    const bool valid = p < pts.n;
    float x0 = 0.5f, x1 = 0.5f;
    if (valid) point_coords(pts, p, x0, x1);
    // ---------------- recompute forward; keep only the packed bf16 inputs of
    // every layer (the pre-activations are recomputed where needed)
    uint4 pf[2];
    {
      float feat[16];
      encode_lane_batched(g, x0, x1, h, table, feat);
#pragma unroll
      for (int v = 0; v < 16; ++v) S.F[kappa0(v, h) * BS + j] = f2bf(feat[v]);
      pf[0] = pack8(feat);
      pf[1] = pack8(feat + 8);
    }
    float zr[16], zi[16];
    {
      v16f u0;
#pragma unroll
      for (int v = 0; v < 16; ++v) u0[v] = s.b0[crow(v, h)];
      u0 = mfma16(o.w0[0][lane], pf[0], u0);
      u0 = mfma16(o.w0[1][lane], pf[1], u0);
#pragma unroll
      for (int v = 0; v < 16; ++v) cis(s.om[0] * u0[v], zr[v], zi[v]);
    }
    uint4 px[H][4];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      px[k][0] = pack8(zr);
      px[k][1] = pack8(zr + 8);
      px[k][2] = pack8(zi);
      px[k][3] = pack8(zi + 8);
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        S.X[k][crow(v, h) * BS + j] = f2bf(zr[v]);
        S.X[k][(32 + crow(v, h)) * BS + j] = f2bf(zi[v]);
      }
      v16f vr, vi;
      hidden_bf16p<H>(o, s, k, px[k], lane, h, vr, vi);
      gabor_act(s.om[k + 1], vr, vi, zr, zi);
    }
    // ---------------- output layer
    float2 go = make_float2(0.f, 0.f);
    if (valid) go = dout[p];
    if (h == 0) { acc_bor += go.x; acc_boi += go.y; }
    float gzr[16], gzi[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int u = crow(v, h);
      S.G[u * BS + j] = f2bf(go.x * zr[v] + go.y * zi[v]);
      S.G[(32 + u) * BS + j] = f2bf(go.y * zr[v] - go.x * zi[v]);
      gzr[v] = s.wor[u] * go.x + s.woi[u] * go.y;
      gzi[v] = s.wor[u] * go.y - s.woi[u] * go.x;
    }
    wave_sync();
    {
      float t = 0.f;
#pragma unroll
      for (int c = 0; c < 32; c += 8) t += sum8(row8(S.G, lane, c));
      acc_wo += t;
    }
    wave_sync();
    // ---------------- hidden layers, last to first
#pragma unroll
    for (int k = H - 1; k >= 0; --k) {
      const float w = s.om[k + 1];
      // pre-activation of layer k, recomputed from its packed input
      v16f vr, vi;
      hidden_bf16p<H>(o, s, k, px[k], lane, h, vr, vi);
      float gvr[16], gvi[16];
      float om_acc = 0.f;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const float gwr = gzr[v] * zr[v] + gzi[v] * zi[v];
        const float gwi = gzi[v] * zr[v] - gzr[v] * zi[v];
        om_acc += gwi * vr[v] - gwr * vi[v];
        gvr[v] = w * gwi;
        gvi[v] = -w * gwr;
        const int u = crow(v, h);
        S.G[u * BS + j] = f2bf(gvr[v]);
        S.G[(32 + u) * BS + j] = f2bf(gvi[v]);
      }
      acc_om[k + 1] += om_acc;
      wave_sync();
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int c0 = 16 * q + 8 * h;
        const uint4 a0 = row8(S.G, j, c0), a1 = row8(S.G, 32 + j, c0);
        const uint4 b0 = row8(S.X[k], j, c0), b1 = row8(S.X[k], 32 + j, c0);
        dWr[k] = mfma16(a0, b0, dWr[k]);
        dWr[k] = mfma16(a1, b1, dWr[k]);
        dWi[k] = mfma16(a1, b0, dWi[k]);
        dWi[k] = mfma16(neg8(a0), b1, dWi[k]);
        acc_c[k][0] += sum8(a0);
        acc_c[k][1] += sum8(a1);
      }
      v16f gxr, gxi;
#pragma unroll
      for (int v = 0; v < 16; ++v) { gxr[v] = 0.f; gxi[v] = 0.f; }
      const uint4 bq[4] = {pack8(gvr), pack8(gvr + 8), pack8(gvi), pack8(gvi + 8)};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        gxr = mfma16(ot.at[k][0][q][lane], bq[q], gxr);
        gxi = mfma16(ot.at[k][1][q][lane], bq[q], gxi);
      }
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        gzr[v] = gxr[v];
        gzi[v] = gxi[v];
        zr[v] = unpack_val(px[k], 0, v);
        zi[v] = unpack_val(px[k], 1, v);
      }
      wave_sync();
    }
    // ---------------- layer 0 (real)
    {
      const float w = s.om[0];
      v16f u0;
#pragma unroll
      for (int v = 0; v < 16; ++v) u0[v] = s.b0[crow(v, h)];
      u0 = mfma16(o.w0[0][lane], pf[0], u0);
      u0 = mfma16(o.w0[1][lane], pf[1], u0);
      float gu[16];
      float om_acc = 0.f;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const float gwi = gzi[v] * zr[v] - gzr[v] * zi[v];
        om_acc += u0[v] * gwi;
        gu[v] = w * gwi;
        S.G[crow(v, h) * BS + j] = f2bf(gu[v]);
      }
      acc_om[0] += om_acc;
      wave_sync();
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int c0 = 16 * q + 8 * h;
        const uint4 a = row8(S.G, j, c0);
        dW0 = mfma16(a, row8(S.F, j, c0), dW0);
        acc_b0 += sum8(a);
      }
      v16f gh;
#pragma unroll
      for (int v = 0; v < 16; ++v) gh[v] = 0.f;
      gh = mfma16(ot.w0t[0][lane], pack8(gu), gh);
      gh = mfma16(ot.w0t[1][lane], pack8(gu + 8), gh);
      wave_sync();
      if (valid) {
#pragma unroll
        for (int mm = 0; mm < 8; ++mm) {
          const int l = 2 * mm + h;
          if (l < g.n_levels) gfeat[(int64_t)l * pts.n + p] = make_float2(gh[2 * mm], gh[2 * mm + 1]);
        }
      }
    }
  }

  // ---------------- workgroup reduction into partial[blockIdx.x][GradLayout::total]
  using GL = GradLayout<H>;
  static_assert(sizeof(st) / (sizeof(float)) >= (size_t)GL::total, "reduction buffer does not fit");
  __syncthreads();
  float* red = (float*)&st[0];
  for (int t = threadIdx.x; t < GL::total; t += blockDim.x) red[t] = 0.f;
  __syncthreads();
  for (int w = 0; w < kWaves; ++w) {
    if (wave == w) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int oo = crow(v, h);
        red[GL::w0 + oo * 32 + j] += dW0[v];
#pragma unroll
        for (int k = 0; k < H; ++k) {
          red[GL::A(k) + 2 * (oo * 32 + j)] += dWr[k][v];
          red[GL::A(k) + 2 * (oo * 32 + j) + 1] += dWi[k][v];
        }
      }
#pragma unroll
      for (int k = 0; k < H; ++k) {
        const float cr = acc_c[k][0] + __shfl_xor(acc_c[k][0], 32);
        const float ci = acc_c[k][1] + __shfl_xor(acc_c[k][1], 32);
        if (h == 0) {
          red[GL::c(k) + 2 * j] += cr;
          red[GL::c(k) + 2 * j + 1] += ci;
        }
      }
      const float b0s = acc_b0 + __shfl_xor(acc_b0, 32);
      if (h == 0) red[GL::b0 + j] += b0s;
      red[GL::wo + 2 * (lane & 31) + (lane >> 5)] += acc_wo;
#pragma unroll
      for (int k = 0; k <= H; ++k) {
        const float t = wave_sum(acc_om[k]);
        if (lane == 0) red[GL::om + k] += t;
      }
      const float br = wave_sum(acc_bor), bi = wave_sum(acc_boi);
      if (lane == 0) { red[GL::bo] += br; red[GL::bo + 1] += bi; }
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < GL::total; t += blockDim.x) partial[(int64_t)blockIdx.x * GL::total + t] = red[t];
}

void launch_inr_fwd_bf16(const PtyGrid& g, const PtyMlp& m, const float2* table, const PtyPoints& pts, float2* out,
                         int blocks, hipStream_t stream) {
  if (m.n_hidden == 1)
    hipLaunchKernelGGL(k_inr_fwd_bf16<1>, dim3(blocks), dim3(kThreads), 0, stream, g, m, table, pts, out);
  else
    hipLaunchKernelGGL(k_inr_fwd_bf16<2>, dim3(blocks), dim3(kThreads), 0, stream, g, m, table, pts, out);
}

void launch_inr_bwd_bf16(const PtyGrid& g, const PtyMlp& m, const float2* table, float2* gfeat, const PtyPoints& pts,
                         const float2* dout, float* partial, int blocks, hipStream_t stream) {
  if (m.n_hidden == 1)
    hipLaunchKernelGGL(k_inr_bwd_bf16<1>, dim3(blocks), dim3(kThreads), 0, stream, g, m, table, gfeat, pts, dout,
                       partial);
  else
    hipLaunchKernelGGL(k_inr_bwd_bf16<2>, dim3(blocks), dim3(kThreads), 0, stream, g, m, table, gfeat, pts, dout,
                       partial);
}

}  // namespace pty
