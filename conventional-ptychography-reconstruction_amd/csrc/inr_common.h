// Shared pieces of the fused INR kernels (FP32 and BF16 MFMA variants).
//
// Layout.  A wavefront owns a tile of 32 points; lane l = (j = l&31, h = l>>5)
// holds point j.  Every layer is computed transposed, Y^T = W X^T, with the
// points as the MFMA N dimension.  The C layout of the 32x32 MFMAs gives lane
// (j, h) rows crow(v, h) = 8(v>>2) + 4h + (v&3), v < 16, of column j; with the
// K index of the next weight matrix permuted to match (kappa maps), that
// register set is the next layer's B operand, so activations stay in
// registers through the forward.
#pragma once
#include "hashgrid.h"

namespace pty {


typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int WS = 33;  // padded LDS row stride (floats)
constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;

// C-layout row of value v for lane half h
__device__ __forceinline__ int crow(int v, int h) { return 8 * (v >> 2) + 4 * h + (v & 3); }
// feature held by lane half h as value s (layer-0 B operand): level 2(s>>1)+h, feature s&1
__device__ __forceinline__ int kappa0(int s, int h) { return 4 * (s >> 1) + 2 * h + (s & 1); }
// inverse of crow, composed with kappa0: feature index of M row rho in g_h = W0^T g_u0
__device__ __forceinline__ int phi0(int rho) { return kappa0(4 * (rho >> 3) + (rho & 3), (rho >> 2) & 1); }

__device__ __forceinline__ v16f mfma(float a, float b, v16f c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

// exp(i*w*x) for real x: (cos wx, sin wx).  Range-reduced to revolutions in
// [-0.5, 0.5] before the hardware sin/cos.
__device__ __forceinline__ void cis(float theta, float& c, float& s) {
  float r = theta * 0.15915494309189535f;
  r = r - rintf(r);
  s = __builtin_amdgcn_sinf(r);
  c = __builtin_amdgcn_cosf(r);
}

__device__ __forceinline__ float exp_f(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

template <int H>
struct MlpLds {
  float w0[32 * WS];   // W0[o][f], f zero-padded to 32
  float b0[32];
  float ar[H][32 * WS];
  float ai[H][32 * WS];
  float nai[H][32 * WS];
  float cr[H][32];
  float ci[H][32];
  float om[H + 1];
  float wor[32], woi[32];
  float bo[2];
};

template <int H>
__device__ void load_weights(MlpLds<H>& s, const PtyMlp& m) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int t = tid; t < 32 * 32; t += nt) {
    const int o = t >> 5, f = t & 31;
    s.w0[o * WS + f] = f < m.n_in ? m.w0[o * m.n_in + f] : 0.f;
    for (int k = 0; k < H; ++k) {
      const float2 a = ((const float2*)m.wh[k])[t];
      s.ar[k][o * WS + f] = a.x;
      s.ai[k][o * WS + f] = a.y;
      s.nai[k][o * WS + f] = -a.y;
    }
  }
  for (int t = tid; t < 32; t += nt) {
    s.b0[t] = m.b0[t];
    for (int k = 0; k < H; ++k) {
      s.cr[k][t] = m.bh[k][2 * t];
      s.ci[k][t] = m.bh[k][2 * t + 1];
    }
    s.wor[t] = m.wo[2 * t];
    s.woi[t] = m.wo[2 * t + 1];
  }
  if (tid <= H) s.om[tid] = *m.omega[tid];
  if (tid == 0) { s.bo[0] = m.bo[0]; s.bo[1] = m.bo[1]; }
}

// Features of the lane's 8 levels: feat[s] = feature kappa0(s, h), one level
// at a time (few registers; the forward kernels hide the gather latency with
// other waves).
__device__ __forceinline__ void encode_lane(const PtyGrid& g, float x0, float x1, int h,
                                            const float2* __restrict__ table, float feat[16]) {
  const bool f16 = g.table_precision == PTY_TABLE_F16;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int l = 2 * m + h;
    float2 acc = make_float2(0.f, 0.f);
    if (l < g.n_levels) {
      const LevelC c = level_c(g, 2 * m, 2 * m + 1, h != 0);
      const LevelPos p = level_pos_c(c, x0, x1);
      uint32_t idx[4];
      float w[4];
      corner_indices_c(c, p, idx);
      corner_weights(p, w);
      float2 t[4];
      if (f16) {
        const __half2* t16 = reinterpret_cast<const __half2*>(table);
#pragma unroll
        for (int k = 0; k < 4; ++k) t[k] = __half22float2(t16[idx[k]]);
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) t[k] = table[idx[k]];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        acc.x += w[k] * t[k].x;
        acc.y += w[k] * t[k].y;
      }
    }
    feat[2 * m] = acc.x;
    feat[2 * m + 1] = acc.y;
  }
}

// The same features with all 32 corner indices formed first and the 32
// gathers issued back to back, so that their latencies overlap (the backward
// kernels run one wave per SIMD and have no other wave to hide them); the
// interpolation order per level is that of level_feature.
__device__ __forceinline__ void encode_lane_batched(const PtyGrid& g, float x0, float x1, int h,
                                                    const float2* __restrict__ table, float feat[16]) {
  uint32_t idx[8][4];
  float fx[8], fy[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int l = 2 * m + h;
    if (l < g.n_levels) {
      const LevelC c = level_c(g, 2 * m, 2 * m + 1, h != 0);
      const LevelPos p = level_pos_c(c, x0, x1);
      corner_indices_c(c, p, idx[m]);
      fx[m] = p.fx;
      fy[m] = p.fy;
    } else {
      idx[m][0] = idx[m][1] = idx[m][2] = idx[m][3] = 0u;   // in bounds; the result is discarded
      fx[m] = fy[m] = 0.f;
    }
  }
  float2 t[8][4];
  if (g.table_precision == PTY_TABLE_F16) {
    const __half2* t16 = reinterpret_cast<const __half2*>(table);
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int c = 0; c < 4; ++c) t[m][c] = __half22float2(t16[idx[m][c]]);
  } else {
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int c = 0; c < 4; ++c) t[m][c] = table[idx[m][c]];
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int l = 2 * m + h;
    float2 acc = make_float2(0.f, 0.f);
    if (l < g.n_levels) {
      LevelPos p;
      p.fx = fx[m];
      p.fy = fy[m];
      float w[4];
      corner_weights(p, w);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        acc.x += w[c] * t[m][c].x;
        acc.y += w[c] * t[m][c].y;
      }
    }
    feat[2 * m] = acc.x;
    feat[2 * m + 1] = acc.y;
  }
}

// Feature cache of the 32-point tile layout: the 16 features of lane `lane`
// of tile `tile` at features[(tile * 64 + lane) * 16] (64 coalesced bytes).
__device__ __forceinline__ void store_features32(float* f, int64_t tile, int lane, const float feat[16]) {
  float4* q = reinterpret_cast<float4*>(f + (tile * 64 + lane) * 16);
#pragma unroll
  for (int i = 0; i < 4; ++i) q[i] = make_float4(feat[4 * i], feat[4 * i + 1], feat[4 * i + 2], feat[4 * i + 3]);
}
__device__ __forceinline__ void load_features32(const float* f, int64_t tile, int lane, float feat[16]) {
  const float4* q = reinterpret_cast<const float4*>(f + (tile * 64 + lane) * 16);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float4 a = q[i];
    feat[4 * i] = a.x; feat[4 * i + 1] = a.y; feat[4 * i + 2] = a.z; feat[4 * i + 3] = a.w;
  }
}

// Layer 0: u0[v] (unit crow(v,h)) = b0 + W0 feat.
template <int H>
__device__ __forceinline__ v16f layer0(const MlpLds<H>& s, const float feat[16], int j, int h) {
  v16f acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = s.b0[crow(v, h)];
#pragma unroll
  for (int st = 0; st < 16; ++st) acc = mfma(s.w0[j * WS + kappa0(st, h)], feat[st], acc);
  return acc;
}

// Hidden layer k: (vr, vi)[v] for unit crow(v,h), input x[s] in kappa1 order
// (s < 16: Re z of unit crow(s,h); s >= 16: Im z of unit crow(s-16,h)).
template <int H>
__device__ __forceinline__ void hidden_fwd(const MlpLds<H>& s, int k, const float zr[16], const float zi[16], int j,
                                           int h, v16f& vr, v16f& vi) {
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    vr[v] = s.cr[k][crow(v, h)];
    vi[v] = s.ci[k][crow(v, h)];
  }
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    const int u = crow(st, h);
    vr = mfma(s.ar[k][j * WS + u], zr[st], vr);
    vi = mfma(s.ai[k][j * WS + u], zr[st], vi);
  }
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    const int u = crow(st, h);
    vr = mfma(s.nai[k][j * WS + u], zi[st], vr);
    vi = mfma(s.ar[k][j * WS + u], zi[st], vi);
  }
}

__device__ __forceinline__ void gabor_act(float w, const v16f& vr, const v16f& vi, float zr[16], float zi[16]) {
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    float c, sn;
    cis(w * vr[v], c, sn);
    const float mag = exp_f(-w * vi[v]);
    zr[v] = mag * c;
    zi[v] = mag * sn;
  }
}

// Flat gradient layout of one workgroup partial (floats):
//   w0 [32*32] | b0 [32] | per hidden k: A [32*32*2] (re,im) | c [32*2] | omega [H+1] | wo [32*2] | bo [2]
template <int H>
struct GradLayout {
  static constexpr int w0 = 0;
  static constexpr int b0 = 1024;
  static constexpr int A(int k) { return 1056 + k * 2112; }
  static constexpr int c(int k) { return 1056 + k * 2112 + 2048; }
  static constexpr int om = 1056 + H * 2112;
  static constexpr int wo = om + H + 1;
  static constexpr int bo = wo + 64;
  static constexpr int total = bo + 2;
};

__device__ __forceinline__ float wave_sum(float v);

// Frequency-gradient terms of one workgroup's complete partial `red`
// (GradLayout<H>, in LDS): layer 0 sum W0 P0 + sum b0 q0; hidden layer k the
// sum over its (re, im) slots of -Im(W) Re(.) and Re(W) Im(.), W = A_k or c_k
// (the homogeneity identity applied to this workgroup's share of the sums).
// Stored into the partial's frequency slots red[om + k], which the gradient
// finish then totals over the workgroups like every other slot.  All threads
// of the workgroup call it (at most 16 waves); it ends with a barrier.
template <int H>
__device__ void omega_terms(float* red, const PtyMlp& m) {
  using GL = GradLayout<H>;
  __shared__ float part[H + 1][16];
  float acc[H + 1];
#pragma unroll
  for (int k = 0; k <= H; ++k) acc[k] = 0.f;
  for (int e = threadIdx.x; e < GL::om; e += blockDim.x) {
    const float v = red[e];
    if (e < GL::b0) {
      const int o = e >> 5, f = e & 31;
      if (f < m.n_in) acc[0] += m.w0[o * m.n_in + f] * v;
    } else if (e < GL::A(0)) {
      acc[0] += m.b0[e - GL::b0] * v;
    } else {
      const int k = (e - GL::A(0)) / 2112, r = (e - GL::A(0)) % 2112;
#pragma unroll
      for (int kk = 0; kk < H; ++kk) {
        if (kk != k) continue;
        const float* W = r < 2048 ? m.wh[kk] + (r & ~1) : m.bh[kk] + ((r - 2048) & ~1);
        acc[kk + 1] += (r & 1) ? W[0] * v : -W[1] * v;
      }
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k <= H; ++k) {
    const float sw = wave_sum(acc[k]);
    if (lane == 0) part[k][wave] = sw;
  }
  __syncthreads();
  if ((int)threadIdx.x <= H) {
    float sum = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) sum += part[threadIdx.x][w];
    red[GL::om + threadIdx.x] = sum;
  }
  __syncthreads();
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// FP32 launchers (inr_f32.hip, 16-point wave tiles)
void launch_inr_fwd_f32(const PtyGrid& g, const PtyMlp& m, const float2* table, const PtyPoints& pts, float2* out,
                        hipStream_t stream);
void launch_inr_bwd_f32(const PtyGrid& g, const PtyMlp& m, const float2* table, float2* gfeat, const PtyPoints& pts,
                        const float2* dout, float* partial, int blocks, hipStream_t stream);

// BF16 launchers (inr_bf16.hip, 16-point wave tiles)
void launch_inr_fwd_b16(const PtyGrid& g, const PtyMlp& m, const float2* table, const PtyPoints& pts, float2* out,
                        hipStream_t stream);
void launch_inr_bwd_b16(const PtyGrid& g, const PtyMlp& m, const float2* table, float2* gfeat, const PtyPoints& pts,
                        const float2* dout, float* partial, int blocks, hipStream_t stream);

// FP8 (and the 32-point BF16) launchers (inr_lowp.hip)
void launch_inr_fwd_lowp(const PtyGrid& g, const PtyMlp& m, const float2* table, const PtyPoints& pts, float2* out,
                         int blocks, hipStream_t stream);
void launch_inr_bwd_lowp(const PtyGrid& g, const PtyMlp& m, const float2* table, float2* gfeat, const PtyPoints& pts,
                         const float2* dout, float* partial, int blocks, hipStream_t stream);

}  // namespace pty
