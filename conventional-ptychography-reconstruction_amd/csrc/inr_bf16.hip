// BF16 fused hash-grid encode + complex-exponential MLP, forward and backward,
// on gfx950's v_mfma_f32_16x16x32_bf16 (K = 32 units in one instruction) and
// v_mfma_f32_16x16x16_bf16 (K = 16 points for the weight gradients), FP32
// accumulation, 16-point wave tiles (inr16.h).
//
// Reference: ComplexINRModel2D.forward -> HashGrid (models/complex_inr.py:279,
// 294), ComplexMLP.net (models/complex_inr.py:57-78, 313); the chain rule is
// the FP32 kernels' (inr_f32.hip).
//
// The activations are transcendental-bound (sin, cos, exp2 per unit at a
// quarter of the VALU rate), not MFMA-bound, so the design removes VALU work:
//   * the layer's frequency and the transcendentals' scales are folded into
//     the bf16 weight operands and the f32 bias (built once per workgroup):
//     the Re rows come out in revolutions (w / 2pi) for the hardware sin /
//     cos, the Im rows as -w log2(e) Im v for exp2, so an activation is
//     three transcendentals and two products;
//   * complex products use the four-product real form ([[Ar, -Ai], [Ai, Ar]],
//     bias as the accumulator's initial value), which needs no VALU to
//     combine results (the three-product form trades two MFMAs for ten VALU
//     operations per unit block, the wrong trade here);
//   * activations stay in registers between layers: lane group g's K values
//     are its own C rows (unit16), so the weight operands are permuted to
//     match and the next layer's B operand is a register pack;
//   * the backward's input-gradient operands carry -i w and w0 (likewise
//     folded); bias sums are taken from the staged bf16 operands the weight
//     gradients read anyway, output-layer sums from f32 staging.
// Weight gradients need the points as K: the layer inputs, the features and
// the staged gradients go to LDS as bf16 rows of 16 points (40-byte stride)
// and are read back as 8-byte MFMA operands.
#include "inr16.h"

namespace pty {

typedef __bf16 b16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));

constexpr int kB16Waves = 8;      // backward workgroup: two waves per SIMD
constexpr int kB16FwdWaves = 4;   // forward workgroup
constexpr int SR = 20;            // staging row stride in bf16 (16 points + pad)
constexpr float kInv2Pi = 0.15915494309189535f;
constexpr float kLog2e = 1.4426950408889634f;

__device__ __forceinline__ v4f mfma32(const uint4& a, const b16x8& b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b16x8, a), b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma16(const s16x4& a, const s16x4& b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ b16x8 pack8(const float x[8]) {
  b16x8 r;
#pragma unroll
  for (int t = 0; t < 8; ++t) r[t] = (__bf16)x[t];
  return r;
}
__device__ __forceinline__ void unpack8(const b16x8& b, float x[8]) {
#pragma unroll
  for (int t = 0; t < 8; ++t) x[t] = (float)b[t];
}
__device__ __forceinline__ s16x4 neg4bf(s16x4 a) {
  uint2 u = __builtin_bit_cast(uint2, a);
  u.x ^= 0x80008000u;
  u.y ^= 0x80008000u;
  return __builtin_bit_cast(s16x4, u);
}
__device__ __forceinline__ s16x4 lds_s4(const __bf16* p) { return *reinterpret_cast<const s16x4*>(p); }
__device__ __forceinline__ v4f lds_v4(const float* p) { return *reinterpret_cast<const v4f*>(p); }

// Forward operands (per workgroup, bf16 in MFMA operand order: entry [lane]
// holds the 8 K values of lane (i, g) = (lane & 15, lane >> 4) for M row
// 16 mb + i).
template <int H>
struct B16Fwd {
  uint4 w0[2][64];        // W0 w0/(2 pi): [mb][lane] = W0[16 mb + i][feat16(t, g)]
  uint4 a[H][4][2][64];   // k: 0 Ar s1, 1 -Ai s1, 2 Ai s2, 3 Ar s2  (s1 = w/(2 pi), s2 = -w log2 e)
  float b0[32];           // b0 w0/(2 pi)
  float cr[H][32];        // c_r s1
  float ci[H][32];        // c_i s2
  float wor[32], woi[32];
  float bo[2];
};

// Backward operands: input gradients, M = inputs, K = outputs (unit16).
template <int H>
struct B16Bwd {
  uint4 w0t[2][64];       // w0 W0^T: [fb][lane] = w0 W0[unit16(t, g)][16 fb + i]
  uint4 at[H][3][2][64];  // k: 0 w Ar^T, 1 -w Ai^T, 2 -w Ar^T: [mb][lane] = . [unit16(t, g)][16 mb + i]
};

__device__ __forceinline__ uint4 pack_op(const float v[8]) {
  const b16x8 b = pack8(v);
  return __builtin_bit_cast(uint4, b);
}

template <int H>
__device__ void build_fwd16(B16Fwd<H>& o, const PtyMlp& m, int tid, int nt) {
  const float w0s = *m.omega[0] * kInv2Pi;
  for (int e = tid; e < 2 * 64; e += nt) {
    const int mb = e >> 6, lane = e & 63, i = lane & 15, g = lane >> 4;
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int f = feat16(t, g);
      v[t] = f < m.n_in ? m.w0[(16 * mb + i) * m.n_in + f] * w0s : 0.f;
    }
    o.w0[mb][lane] = pack_op(v);
  }
  for (int e = tid; e < H * 4 * 2 * 64; e += nt) {
    const int lane = e & 63, mb = (e >> 6) & 1, part = (e >> 7) & 3, k = e >> 9;
    const int i = lane & 15, g = lane >> 4;
    const float w = *m.omega[k + 1];
    const float s = part < 2 ? w * kInv2Pi : -w * kLog2e;
    const int comp = (part == 0 || part == 3) ? 0 : 1;      // Ar or Ai
    const float sign = part == 1 ? -1.f : 1.f;
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = sign * s * m.wh[k][2 * ((16 * mb + i) * 32 + unit16(t, g)) + comp];
    o.a[k][part][mb][lane] = pack_op(v);
  }
  for (int t = tid; t < 32; t += nt) {
    o.b0[t] = m.b0[t] * w0s;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const float w = *m.omega[k + 1];
      o.cr[k][t] = m.bh[k][2 * t] * w * kInv2Pi;
      o.ci[k][t] = m.bh[k][2 * t + 1] * (-w * kLog2e);
    }
    o.wor[t] = m.wo[2 * t];
    o.woi[t] = m.wo[2 * t + 1];
  }
  if (tid == 0) { o.bo[0] = m.bo[0]; o.bo[1] = m.bo[1]; }
}

template <int H>
__device__ void build_bwd16(B16Bwd<H>& o, const PtyMlp& m, int tid, int nt) {
  const float w0 = *m.omega[0];
  for (int e = tid; e < 2 * 64; e += nt) {
    const int fb = e >> 6, lane = e & 63, i = lane & 15, g = lane >> 4;
    const int f = 16 * fb + i;
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = f < m.n_in ? w0 * m.w0[unit16(t, g) * m.n_in + f] : 0.f;
    o.w0t[fb][lane] = pack_op(v);
  }
  for (int e = tid; e < H * 3 * 2 * 64; e += nt) {
    const int lane = e & 63, mb = (e >> 6) & 1, part = (e >> 7) % 3, k = (e >> 7) / 3;
    const int i = lane & 15, g = lane >> 4;
    const float w = *m.omega[k + 1];
    const float s = part == 0 ? w : -w;
    const int comp = part == 1 ? 1 : 0;
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = s * m.wh[k][2 * (unit16(t, g) * 32 + 16 * mb + i) + comp];
    o.at[k][part][mb][lane] = pack_op(v);
  }
}

// exp(i w v) from the folded rows: r in revolutions, e = -w log2(e) Im v.
// The hardware sin / cos reduce their argument (in revolutions) themselves
// over |r| <= 256; no separate range reduction (the layers' phases stay far
// inside that, and the reduction's f32 rounding, r * 2^-23, is small against
// the bf16 rounding of the products that formed r).
__device__ __forceinline__ void act_rev(float r, float e, float& zr, float& zi) {
  const float mag = __builtin_amdgcn_exp2f(e);
  zr = mag * __builtin_amdgcn_cosf(r);
  zi = mag * __builtin_amdgcn_sinf(r);
}
__device__ __forceinline__ void act_rev0(float r, float& zr, float& zi) {
  zr = __builtin_amdgcn_cosf(r);
  zi = __builtin_amdgcn_sinf(r);
}

// Stage the 8 values of a pack at column j of rows row(t): the 16-bit halves
// of the packed dwords are stored as they are (no second conversion).
template <class Row>
__device__ __forceinline__ void stage8(__bf16* base, int j, const b16x8& v, Row row) {
  const uint4 u = __builtin_bit_cast(uint4, v);
  const uint32_t w[4] = {u.x, u.y, u.z, u.w};
  unsigned short* b = reinterpret_cast<unsigned short*>(base);
#pragma unroll
  for (int t = 0; t < 8; ++t) b[row(t) * SR + j] = (t & 1) ? (unsigned short)(w[t >> 1] >> 16)
                                                           : (unsigned short)(w[t >> 1] & 0xFFFFu);
}

// Forward chain of one tile: features -> z_H (f32), with the bf16 packs of
// every layer input kept in zb[k] (k = 0 .. H-1: inputs of the hidden layers,
// zb[H]: z_H).  STAGE: also write the hidden-layer inputs to X (bf16 rows).
template <int H, bool STAGE>
__device__ __forceinline__ void chain16(const B16Fwd<H>& o, const float feat[8], int lane, int j, int g,
                                        b16x8 (&zbr)[H + 1], b16x8 (&zbi)[H + 1], float zr[8], float zi[8],
                                        __bf16* X, __bf16* F) {
  const b16x8 hb = pack8(feat);
  if (STAGE) stage8(F, j, hb, [g](int t) { return feat16(t, g); });
  {
    v4f u[2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) u[mb] = mfma32(o.w0[mb][lane], hb, lds_v4(&o.b0[16 * mb + 4 * g]));
#pragma unroll
    for (int t = 0; t < 8; ++t) act_rev0(u[t >> 2][t & 3], zr[t], zi[t]);
  }
#pragma unroll
  for (int k = 0; k < H; ++k) {
    zbr[k] = pack8(zr);
    zbi[k] = pack8(zi);
    if (STAGE) {
      __bf16* Xk = X + k * 64 * SR;
      stage8(Xk, j, zbr[k], [g](int t) { return unit16(t, g); });
      stage8(Xk, j, zbi[k], [g](int t) { return 32 + unit16(t, g); });
    }
    // the layer's eight weight operands are read back to back, so that the
    // MFMAs wait for one LDS round trip rather than one each
    uint4 A[2][4];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int q = 0; q < 4; ++q) A[mb][q] = o.a[k][q][mb][lane];
    v4f vr[2], vi[2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      vr[mb] = mfma32(A[mb][0], zbr[k], lds_v4(&o.cr[k][16 * mb + 4 * g]));
      vr[mb] = mfma32(A[mb][1], zbi[k], vr[mb]);
      vi[mb] = mfma32(A[mb][2], zbr[k], lds_v4(&o.ci[k][16 * mb + 4 * g]));
      vi[mb] = mfma32(A[mb][3], zbi[k], vi[mb]);
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) act_rev(vr[t >> 2][t & 3], vi[t >> 2][t & 3], zr[t], zi[t]);
  }
  zbr[H] = pack8(zr);
  zbi[H] = pack8(zi);
}

// ============================================================== forward
template <int H>
__global__ __launch_bounds__(64 * kB16FwdWaves) void k_inr_fwd_b16(PtyGrid gr, PtyMlp m,
                                                                  const float2* __restrict__ table, PtyPoints pts,
                                                                  float2* __restrict__ out) {
  __shared__ B16Fwd<H> o;
  __shared__ uint4 s_lvl[kLevelTab];
  build_fwd16<H>(o, m, threadIdx.x, blockDim.x);
  load_level_tab(gr, s_lvl);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int64_t ntiles = (pts.n + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * kB16FwdWaves + wave; tile < ntiles;
       tile += (int64_t)gridDim.x * kB16FwdWaves) {
    asm volatile("" ::: "memory");
    const int64_t p = tile * 16 + j;
    const bool valid = p < pts.n;
    float x0 = 0.5f, x1 = 0.5f;
    if (valid) point_coords(pts, p, x0, x1);
    float feat[8];
    encode16(gr, s_lvl, x0, x1, g, table, feat);
    if (pts.features) store_features16(pts.features, tile, lane, feat);
    b16x8 zbr[H + 1], zbi[H + 1];
    float zr[8], zi[8];
    chain16<H, false>(o, feat, lane, j, g, zbr, zbi, zr, zi, nullptr, nullptr);
    float orr = 0.f, oi = 0.f;
    const v4f wr[2] = {lds_v4(&o.wor[4 * g]), lds_v4(&o.wor[16 + 4 * g])};
    const v4f wi[2] = {lds_v4(&o.woi[4 * g]), lds_v4(&o.woi[16 + 4 * g])};
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const float a = wr[t >> 2][t & 3], b = wi[t >> 2][t & 3];
      orr += a * zr[t] - b * zi[t];
      oi += a * zi[t] + b * zr[t];
    }
    orr += __shfl_xor(orr, 16);
    oi += __shfl_xor(oi, 16);
    orr += __shfl_xor(orr, 32);
    oi += __shfl_xor(oi, 32);
    if (valid && g == 0) out[p] = make_float2(orr + o.bo[0], oi + o.bo[1]);
  }
}

// ============================================================== backward
template <int H>
struct B16Stage {
  __bf16 F[32 * SR];        // features, feature-major rows of 16 points
  __bf16 X[H][64 * SR];     // hidden-layer inputs: rows 0..31 Re, 32..63 Im
  __bf16 G[64 * SR];        // staged gradient
  float W[64 * 16];         // output-layer weight gradient terms (f32), rows as G
};

// sum of the four bf16 values of an MFMA operand (f32)
__device__ __forceinline__ float sum4bf(const s16x4& a) {
  const uint2 u = __builtin_bit_cast(uint2, a);
  return (__uint_as_float(u.x << 16) + __uint_as_float(u.x & 0xFFFF0000u)) +
         (__uint_as_float(u.y << 16) + __uint_as_float(u.y & 0xFFFF0000u));
}

template <int H, bool ADD>
__device__ __forceinline__ void put_b16(float* R, const v4f (&P0)[2][2], const v4f (&Pr)[H][2][2],
                                        const v4f (&Pi)[H][2][2], const float (&q0)[2], const float (&qr)[H][2],
                                        const float (&qi)[H][2], float acc_wo, float br, float bi, int lane, int j,
                                        int g) {
  using GL = GradLayout<H>;
  auto put = [](float& r, float v) { if (ADD) r += v; else r = v; };
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * mb + 4 * g + v, col = 16 * nb + j;
        put(R[GL::w0 + row * 32 + col], P0[mb][nb][v]);
#pragma unroll
        for (int k = 0; k < H; ++k) {
          put(R[GL::A(k) + 2 * (row * 32 + col)], Pr[k][mb][nb][v]);
          put(R[GL::A(k) + 2 * (row * 32 + col) + 1], Pi[k][mb][nb][v]);
        }
      }
  if (g == 0) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      const int u = 16 * mb + j;
      put(R[GL::b0 + u], q0[mb]);
#pragma unroll
      for (int k = 0; k < H; ++k) {
        put(R[GL::c(k) + 2 * u], qr[k][mb]);
        put(R[GL::c(k) + 2 * u + 1], qi[k][mb]);
      }
    }
  }
  // lane l holds row l of the staged output-layer terms: Re of unit l (l < 32)
  // or Im of unit l - 32
  put(R[GL::wo + 2 * (lane & 31) + (lane >> 5)], acc_wo);
  if (lane == 0) {
    put(R[GL::bo], br);
    put(R[GL::bo + 1], bi);
  }
}

template <int H>
__global__ __launch_bounds__(64 * kB16Waves) void k_inr_bwd_b16(PtyGrid gr, PtyMlp m, const float2* __restrict__ table,
                                                               float2* __restrict__ gfeat, PtyPoints pts,
                                                               const float2* __restrict__ dout,
                                                               float* __restrict__ partial) {
  __shared__ B16Fwd<H> o;
  __shared__ B16Bwd<H> ob;
  __shared__ B16Stage<H> st[kB16Waves];
  __shared__ uint4 s_lvl[kLevelTab];
  build_fwd16<H>(o, m, threadIdx.x, blockDim.x);
  build_bwd16<H>(ob, m, threadIdx.x, blockDim.x);
  load_level_tab(gr, s_lvl);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  B16Stage<H>& S = st[wave];

  v4f Pr[H][2][2], Pi[H][2][2], P0[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      P0[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < H; ++k) Pr[k][a][b] = Pi[k][a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    }
  // bias sums: lane (j, g) holds unit 16 mb + j over points 4 g .. 4 g + 3
  float qr[H][2], qi[H][2], q0[2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    q0[a] = 0.f;
#pragma unroll
    for (int k = 0; k < H; ++k) qr[k][a] = qi[k][a] = 0.f;
  }
  float acc_wo = 0.f, acc_bor = 0.f, acc_boi = 0.f;

  const int64_t ntiles = (pts.n + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * kB16Waves + wave; tile < ntiles; tile += (int64_t)gridDim.x * kB16Waves) {
    asm volatile("" ::: "memory");
    const int64_t p = tile * 16 + j;
    const bool valid = p < pts.n;
    float feat[8];
    if (pts.features) {
      load_features16(pts.features, tile, lane, feat);
    } else {
      float x0 = 0.5f, x1 = 0.5f;
      if (valid) point_coords(pts, p, x0, x1);
      encode16(gr, s_lvl, x0, x1, g, table, feat);
    }
    b16x8 zbr[H + 1], zbi[H + 1];
    float zr[8], zi[8];
    __builtin_amdgcn_s_setprio(1);
    chain16<H, true>(o, feat, lane, j, g, zbr, zbi, zr, zi, &S.X[0][0], S.F);
    __builtin_amdgcn_s_setprio(0);
    // ---------------- output layer (f32): g_wo += g_o conj(z_H), g_z = conj(wo) g_o
    float2 go = make_float2(0.f, 0.f);
    if (valid) go = dout[p];
    if (g == 0) { acc_bor += go.x; acc_boi += go.y; }
    float gzr[8], gzi[8];
    const v4f wr[2] = {lds_v4(&o.wor[4 * g]), lds_v4(&o.wor[16 + 4 * g])};
    const v4f wi[2] = {lds_v4(&o.woi[4 * g]), lds_v4(&o.woi[16 + 4 * g])};
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int u = unit16(t, g);
      S.W[u * 16 + j] = go.x * zr[t] + go.y * zi[t];
      S.W[(32 + u) * 16 + j] = go.y * zr[t] - go.x * zi[t];
      const float a = wr[t >> 2][t & 3], b = wi[t >> 2][t & 3];
      gzr[t] = a * go.x + b * go.y;
      gzi[t] = a * go.y - b * go.x;
    }
    wave_sync();
    {
      const float* row = &S.W[lane * 16];
      const v4f a = lds_v4(row), b = lds_v4(row + 4), c = lds_v4(row + 8), d = lds_v4(row + 12);
      acc_wo += ((a[0] + a[1]) + (a[2] + a[3])) + ((b[0] + b[1]) + (b[2] + b[3])) +
                (((c[0] + c[1]) + (c[2] + c[3])) + ((d[0] + d[1]) + (d[2] + d[3])));
    }
    // ---------------- hidden layers, last to first: g_w = g_z conj(z_{k+1})
#pragma unroll
    for (int k = H - 1; k >= 0; --k) {
      if (k < H - 1) {
        unpack8(zbr[k + 1], zr);
        unpack8(zbi[k + 1], zi);
      }
      float gwr[8], gwi[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        gwr[t] = gzr[t] * zr[t] + gzi[t] * zi[t];
        gwi[t] = gzi[t] * zr[t] - gzr[t] * zi[t];
      }
      const b16x8 gbr = pack8(gwr), gbi = pack8(gwi);
      stage8(S.G, j, gbr, [g](int t) { return unit16(t, g); });
      stage8(S.G, j, gbi, [g](int t) { return 32 + unit16(t, g); });
      wave_sync();
      __builtin_amdgcn_s_setprio(1);
      // P += G X^H over the tile's 16 points: Pr = Gr Xr^T + Gi Xi^T, Pi = Gi Xr^T - Gr Xi^T
      const __bf16* Xk = &S.X[k][0];
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        const s16x4 ar = lds_s4(&S.G[(16 * mb + j) * SR + 4 * g]);
        const s16x4 ai = lds_s4(&S.G[(32 + 16 * mb + j) * SR + 4 * g]);
        qr[k][mb] += sum4bf(ar);
        qi[k][mb] += sum4bf(ai);
        const s16x4 nar = neg4bf(ar);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
          const s16x4 xr = lds_s4(&Xk[(16 * nb + j) * SR + 4 * g]);
          const s16x4 xi = lds_s4(&Xk[(32 + 16 * nb + j) * SR + 4 * g]);
          Pr[k][mb][nb] = mfma16(ar, xr, Pr[k][mb][nb]);
          Pr[k][mb][nb] = mfma16(ai, xi, Pr[k][mb][nb]);
          Pi[k][mb][nb] = mfma16(ai, xr, Pi[k][mb][nb]);
          Pi[k][mb][nb] = mfma16(nar, xi, Pi[k][mb][nb]);
        }
      }
      // g_z of the layer input = -i w A^H g_w (w folded into the operands)
      uint4 T[2][3];
#pragma unroll
      for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int q = 0; q < 3; ++q) T[mb][q] = ob.at[k][q][mb][lane];
      v4f tr[2], ti[2];
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        tr[mb] = mfma32(T[mb][0], gbi, v4f{0.f, 0.f, 0.f, 0.f});
        tr[mb] = mfma32(T[mb][1], gbr, tr[mb]);
        ti[mb] = mfma32(T[mb][2], gbr, v4f{0.f, 0.f, 0.f, 0.f});
        ti[mb] = mfma32(T[mb][1], gbi, ti[mb]);
      }
      __builtin_amdgcn_s_setprio(0);
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        gzr[t] = tr[t >> 2][t & 3];
        gzi[t] = ti[t >> 2][t & 3];
      }
      wave_sync();
    }
    // ---------------- layer 0 (real): staged Im(g_z conj(z0)) = g_u0 / w0
    {
      unpack8(zbr[0], zr);
      unpack8(zbi[0], zi);
      float gu[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) gu[t] = gzi[t] * zr[t] - gzr[t] * zi[t];
      const b16x8 gub = pack8(gu);
      stage8(S.G, j, gub, [g](int t) { return unit16(t, g); });
      wave_sync();
      __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        const s16x4 a = lds_s4(&S.G[(16 * mb + j) * SR + 4 * g]);
        q0[mb] += sum4bf(a);
#pragma unroll
        for (int fb = 0; fb < 2; ++fb) P0[mb][fb] = mfma16(a, lds_s4(&S.F[(16 * fb + j) * SR + 4 * g]), P0[mb][fb]);
      }
      // g_h (features 16 fb + 4 g + v) = w0 W0^T Im(.)
      v4f gh[2];
#pragma unroll
      for (int fb = 0; fb < 2; ++fb) gh[fb] = mfma32(ob.w0t[fb][lane], gub, v4f{0.f, 0.f, 0.f, 0.f});
      __builtin_amdgcn_s_setprio(0);
      wave_sync();
      if (valid) {
#pragma unroll
        for (int fb = 0; fb < 2; ++fb)
#pragma unroll
          for (int hh = 0; hh < 2; ++hh) {
            const int l = 8 * fb + 2 * g + hh;
            if (l < gr.n_levels) gfeat[(int64_t)l * pts.n + p] = make_float2(gh[fb][2 * hh], gh[fb][2 * hh + 1]);
          }
      }
    }
  }

  // ---------------- workgroup reduction into partial[blockIdx.x][GradLayout::total]
  // Per-lane sums over the wave's points first, then the waves in pairs into
  // two LDS regions (four phases), then the two regions in a fixed order.
  // bias sums over the four lane groups (same unit, different points)
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    q0[a] += __shfl_xor(q0[a], 16);
    q0[a] += __shfl_xor(q0[a], 32);
#pragma unroll
    for (int k = 0; k < H; ++k) {
      qr[k][a] += __shfl_xor(qr[k][a], 16);
      qr[k][a] += __shfl_xor(qr[k][a], 32);
      qi[k][a] += __shfl_xor(qi[k][a], 16);
      qi[k][a] += __shfl_xor(qi[k][a], 32);
    }
  }
  const float br = wave_sum(acc_bor), bi = wave_sum(acc_boi);
  using GL = GradLayout<H>;
  constexpr int kRegions = 2;
  static_assert(sizeof(st) / sizeof(float) >= (size_t)kRegions * GL::total, "reduction buffer does not fit");
  __syncthreads();
  float* red = reinterpret_cast<float*>(&st[0]);
  float* R = red + (wave % kRegions) * GL::total;
  for (int phase = 0; phase < kB16Waves / kRegions; ++phase) {
    if (wave / kRegions == phase) {
      if (phase == 0) put_b16<H, false>(R, P0, Pr, Pi, q0, qr, qi, acc_wo, br, bi, lane, j, g);
      else put_b16<H, true>(R, P0, Pr, Pi, q0, qr, qi, acc_wo, br, bi, lane, j, g);
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < GL::total; t += blockDim.x)
    if (t < GL::om || t >= GL::wo) red[t] = red[t] + red[GL::total + t];
  __syncthreads();
  omega_terms<H>(red, m);
  for (int t = threadIdx.x; t < GL::total; t += blockDim.x) partial[(int64_t)blockIdx.x * GL::total + t] = red[t];
}

// ============================================================== launchers
void launch_inr_fwd_b16(const PtyGrid& g, const PtyMlp& m, const float2* table, const PtyPoints& pts, float2* out,
                        hipStream_t stream) {
  const int64_t tiles = (pts.n + 15) / 16;
  int64_t blocks = (tiles + kB16FwdWaves - 1) / kB16FwdWaves;
  const int64_t cap = 4 * (int64_t)device_cu_count();
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  if (m.n_hidden == 1)
    hipLaunchKernelGGL(k_inr_fwd_b16<1>, dim3((unsigned)blocks), dim3(64 * kB16FwdWaves), 0, stream, g, m, table,
                       pts, out);
  else
    hipLaunchKernelGGL(k_inr_fwd_b16<2>, dim3((unsigned)blocks), dim3(64 * kB16FwdWaves), 0, stream, g, m, table,
                       pts, out);
}

void launch_inr_bwd_b16(const PtyGrid& g, const PtyMlp& m, const float2* table, float2* gfeat, const PtyPoints& pts,
                        const float2* dout, float* partial, int blocks, hipStream_t stream) {
  if (m.n_hidden == 1)
    hipLaunchKernelGGL(k_inr_bwd_b16<1>, dim3(blocks), dim3(64 * kB16Waves), 0, stream, g, m, table, gfeat, pts, dout,
                       partial);
  else
    hipLaunchKernelGGL(k_inr_bwd_b16<2>, dim3(blocks), dim3(64 * kB16Waves), 0, stream, g, m, table, gfeat, pts, dout,
                       partial);
}

}  // namespace pty
