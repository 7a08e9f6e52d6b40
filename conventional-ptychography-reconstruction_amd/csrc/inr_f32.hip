// FP32 fused hash-grid encode + complex-exponential MLP, forward and backward,
// on v_mfma_f32_16x16x4f32 with 16-point wave tiles (gfx950).
//
// Reference: ComplexINRModel2D.forward -> HashGrid (models/complex_inr.py:279,
// 294), ComplexMLP.net (models/complex_inr.py:57-78, 313):
//   u0 = W0 h + b0            z0 = exp(i w0 u0)                 (real first layer)
//   v_k = A_k z_{k-1} + c_k   z_k = exp(i w_k v_k) = e^{-w Im v}(cos w Re v, sin w Re v)
//   o = wo . z_H + bo
//
// Tile layout.  A wavefront owns 16 points; lane l = (j, g) = (l & 15, l >> 4)
// holds point j.  Every layer is computed transposed (Y^T = W X^T) with the
// points as the MFMA N dimension.  The 16x16 C layout gives lane (j, g) rows
// 4g + v (v < 4) of column j, so with two 16-row M blocks a lane holds the 8
// units unit16(s, g) = 16 (s >> 2) + 4 g + (s & 3), s < 8.  Taking K chunk s
// of the next layer to be {unit16(s, g) : g < 4} makes that register set the
// next layer's B operand: activations stay in registers through the forward.
// Complex layers use the real form [[Ar, -Ai], [Ai, Ar]].  The first layer's
// K chunk c holds features feat16(c, g) = 2 (4 (c >> 1) + g) + (c & 1): each
// lane interpolates 4 levels (4 m + g, m < 4) of its point.
//
// Against 32-point tiles on the 32x32x2 instruction the MFMA work per point
// is the same, but a wave holds half the activations and half the LDS
// staging, so the backward fits two waves per SIMD (eight per workgroup) and
// one wave's MFMA chains overlap the other's transcendental and memory work.
//
// Backward (PyTorch's complex-gradient convention g = dL/dRe + i dL/dIm),
// per tile after recomputing the forward:
//   g_z_H = conj(wo) g_o,  g_wo += g_o conj(z_H),  g_bo += g_o
//   g_w = g_z conj(z);  P += g_w x^H (points as K), q += sum g_w;
//   g_x = -i w A^H g_w
//   layer 0: Im(g_z conj(z0)) = g_u0 / w0;  P0 += Im(.) h^T, q0 += Im(.);
//   g_h = w0 W0^T Im(.), written level-major for the table-gradient kernels.
// Each workgroup forms its share of the frequency gradients from its partial
// sums (omega_terms); the finish kernel of inr.hip applies -i w (w0) to the
// totals of P, q (P0, q0).
#include "inr16.h"

namespace pty {

constexpr int kW16 = 8;            // waves per backward workgroup (two per SIMD)
constexpr int kFwd16Waves = 4;     // waves per forward workgroup
constexpr int WS36 = 36;           // LDS row stride of the weight matrices (floats)
constexpr int PS = 16;             // LDS row stride of the per-point staging (16 points)

// W0 is staged with each lane group's layer-0 K values contiguous: row o,
// position 8 g + c holds W0[o][feat16(c, g)] (zero past n_in), so layer 0
// reads them as two 16-byte loads per M block.
__device__ __forceinline__ int w0pos(int f) { return 8 * ((f >> 1) & 3) + 2 * (f >> 3) + (f & 1); }

template <int H>
struct __align__(16) Lds16 {   // read with 16-byte loads
  float w0[32 * WS36];     // W0[o][feat16(c, g)] at column 8 g + c
  float ar[H][32 * WS36];  // Re A_k[o][i]
  float ai[H][32 * WS36];  // Im A_k[o][i]
  float b0[32];
  float cr[H][32], ci[H][32];
  float cs[H][32];         // c_r + c_i (the three-product form's third accumulator start)
  float om[H + 1];
  float wor[32], woi[32];
  float bo[2];
};

template <int H>
__device__ void load16(Lds16<H>& s, const PtyMlp& m, int tid, int nt) {
  for (int t = tid; t < 32 * 32; t += nt) {
    const int o = t >> 5, f = t & 31;
    const int fw = feat16(f & 7, f >> 3);
    s.w0[o * WS36 + f] = fw < m.n_in ? m.w0[o * m.n_in + fw] : 0.f;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const float2 a = ((const float2*)m.wh[k])[t];
      s.ar[k][o * WS36 + f] = a.x;
      s.ai[k][o * WS36 + f] = a.y;
    }
  }
  for (int t = tid; t < 32; t += nt) {
    s.b0[t] = m.b0[t];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      s.cr[k][t] = m.bh[k][2 * t];
      s.ci[k][t] = m.bh[k][2 * t + 1];
      s.cs[k][t] = m.bh[k][2 * t] + m.bh[k][2 * t + 1];
    }
    s.wor[t] = m.wo[2 * t];
    s.woi[t] = m.wo[2 * t + 1];
  }
  if (tid <= H) s.om[tid] = *m.omega[tid];
  if (tid == 0) { s.bo[0] = m.bo[0]; s.bo[1] = m.bo[1]; }
}

__device__ __forceinline__ float4 lds4(const float* p) { return *reinterpret_cast<const float4*>(p); }

// u[mb] (units 16 mb + 4 g + v) = b0 + W0 h
template <int H>
__device__ __forceinline__ void layer0_16(const Lds16<H>& s, const float feat[8], int i, int g, v4f u[2]) {
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int v = 0; v < 4; ++v) u[mb][v] = s.b0[16 * mb + 4 * g + v];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      const float4 w = lds4(&s.w0[(16 * mb + i) * WS36 + 8 * g + 4 * h]);
      u[mb] = mfma4(w.x, feat[4 * h], u[mb]);
      u[mb] = mfma4(w.y, feat[4 * h + 1], u[mb]);
      u[mb] = mfma4(w.z, feat[4 * h + 2], u[mb]);
      u[mb] = mfma4(w.w, feat[4 * h + 3], u[mb]);
    }
}

// (vr, vi)[mb] = A_k z + c_k, three real products per K step (Gauss):
// T1 = Ar zr, T2 = Ai zi, T3 = (Ar + Ai)(zr + zi); vr = c_r + T1 - T2,
// vi = c_i + T3 - T1 - T2.  The biases start the accumulators (t1 = c_r + T1,
// t3 = c_r + c_i + T3), so vr = t1 - t2 and vi = t3 - t1 - t2.  Ar + Ai is
// formed from the two LDS values the four-product form reads anyway.
template <int H>
__device__ __forceinline__ void hidden16(const Lds16<H>& s, int k, const float zr[8], const float zi[8], int i, int g,
                                         v4f vr[2], v4f vi[2]) {
  v4f t1[2], t2[2], t3[2];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) {
    t1[mb] = *reinterpret_cast<const v4f*>(&s.cr[k][16 * mb + 4 * g]);
    t3[mb] = *reinterpret_cast<const v4f*>(&s.cs[k][16 * mb + 4 * g]);
    t2[mb] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  // K chunk st = 4 qd + e reads unit16(st, g) = 16 qd + 4 g + e: the four
  // e of one row are contiguous (one 16-byte load, conflict-free at row
  // stride 36)
#pragma unroll
  for (int qd = 0; qd < 2; ++qd) {
    float zs[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) zs[e] = zr[4 * qd + e] + zi[4 * qd + e];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      const float4 r4 = lds4(&s.ar[k][(16 * mb + i) * WS36 + 16 * qd + 4 * g]);
      const float4 i4 = lds4(&s.ai[k][(16 * mb + i) * WS36 + 16 * qd + 4 * g]);
      const float a_r[4] = {r4.x, r4.y, r4.z, r4.w}, a_i[4] = {i4.x, i4.y, i4.z, i4.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        t1[mb] = mfma4(a_r[e], zr[4 * qd + e], t1[mb]);
        t2[mb] = mfma4(a_i[e], zi[4 * qd + e], t2[mb]);
        t3[mb] = mfma4(a_r[e] + a_i[e], zs[e], t3[mb]);
      }
    }
  }
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      vr[mb][v] = t1[mb][v] - t2[mb][v];
      vi[mb][v] = t3[mb][v] - t1[mb][v] - t2[mb][v];
    }
}

// exp(i 2 pi r) for a phase r in revolutions: r - rint(r) is exact in f32, so
// the only roundings are those of the scaled phase (as with the radian form).
__device__ __forceinline__ void cis_rev(float r, float& c, float& s) {
  r = r - rintf(r);
  s = __builtin_amdgcn_sinf(r);
  c = __builtin_amdgcn_cosf(r);
}

__device__ __forceinline__ void act16(float w, const v4f vr[2], const v4f vi[2], float zr[8], float zi[8]) {
  const float wl = -w * 1.4426950408889634f;   // exp(-w v) = 2^(-w log2(e) v)
  const float wr = w * 0.15915494309189535f;   // w v in revolutions
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    float c, sn;
    cis_rev(wr * vr[s >> 2][s & 3], c, sn);
    const float mag = __builtin_amdgcn_exp2f(wl * vi[s >> 2][s & 3]);
    zr[s] = mag * c;
    zi[s] = mag * sn;
  }
}

// ============================================================== forward
template <int H>
__global__ __launch_bounds__(64 * kFwd16Waves, 5) void k_inr_fwd16(PtyGrid gr, PtyMlp m,
                                                                const float2* __restrict__ table, PtyPoints pts,
                                                                float2* __restrict__ out) {
  __shared__ Lds16<H> s;
  __shared__ uint4 s_lvl[kLevelTab];
  load16<H>(s, m, threadIdx.x, blockDim.x);
  load_level_tab(gr, s_lvl);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int64_t ntiles = (pts.n + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * kFwd16Waves + wave; tile < ntiles;
       tile += (int64_t)gridDim.x * kFwd16Waves) {
    asm volatile("" ::: "memory");
    const int64_t p = tile * 16 + j;
    const bool valid = p < pts.n;
    float x0 = 0.5f, x1 = 0.5f;
    if (valid) point_coords(pts, p, x0, x1);
    float feat[8];
    encode16(gr, s_lvl, x0, x1, g, table, feat);
    if (pts.features) store_features16(pts.features, tile, lane, feat);
    float zr[8], zi[8];
    // a wave raises its issue priority while it feeds the matrix core, so
    // that the other waves' encoder and activation work fills in around it
    {
      v4f u[2];
      const float w0r = s.om[0] * 0.15915494309189535f;
      __builtin_amdgcn_s_setprio(2);
      layer0_16<H>(s, feat, j, g, u);
      __builtin_amdgcn_s_setprio(0);
#pragma unroll
      for (int q = 0; q < 8; ++q) cis_rev(w0r * u[q >> 2][q & 3], zr[q], zi[q]);
    }
#pragma unroll
    for (int k = 0; k < H; ++k) {
      v4f vr[2], vi[2];
      __builtin_amdgcn_s_setprio(2);
      hidden16<H>(s, k, zr, zi, j, g, vr, vi);
      __builtin_amdgcn_s_setprio(0);
      act16(s.om[k + 1], vr, vi, zr, zi);
    }
    float orr = 0.f, oi = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int u = unit16(q, g);
      orr += s.wor[u] * zr[q] - s.woi[u] * zi[q];
      oi += s.wor[u] * zi[q] + s.woi[u] * zr[q];
    }
    orr += __shfl_xor(orr, 16);
    oi += __shfl_xor(oi, 16);
    orr += __shfl_xor(orr, 32);
    oi += __shfl_xor(oi, 32);
    if (valid && g == 0) out[p] = make_float2(orr + s.bo[0], oi + s.bo[1]);
  }
}

// ============================================================== backward
template <int H>
struct __align__(16) Stage16 {
  // rows of 16 points, element (row, point) at st_elem(row, point)
  float X[H][64 * PS];   // input of hidden layer k: rows 0..31 Re, 32..63 Im
  float F[32 * PS];      // features, one row per feature
  float G[64 * PS];      // gradient staging
};

// Per-wave staging rows of 16 points.  Element (row r, point p) lives at
//   16 P(r) + 4 ((p >> 2) ^ hq(P(r))) + (p & 3),  P(r) = r ^ ((r >> 2) & 3),
//   hq(x) = ((x >> 1) & 3) ^ ((x >> 3) & 1)
// (a permutation inside each aligned 4-row group and of the four point quads
// of a row).  The 16-byte reads of rows 16 b + j (lane j) at quad g are then
// conflict-free (as are those of row = lane), and the C-layout writes of rows
// 4 g + c (lane group g) fall into distinct bank groups; unpermuted, both had
// 4-way bank conflicts.  The quad a read returns holds points 4 g .. 4 g + 3
// in order, so the sums are unchanged.
static_assert(PS == 16, "the staging permutation assumes 16-point rows");
__device__ __forceinline__ int st_hq(int x) { return ((x >> 1) & 3) ^ ((x >> 3) & 1); }
__device__ __forceinline__ int st_elem(int r, int p) {
  const int P = r ^ ((r >> 2) & 3);
  return 16 * P + 4 * ((p >> 2) ^ st_hq(P)) + (p & 3);
}
__device__ __forceinline__ int st_quad(int r, int q) {
  const int P = r ^ ((r >> 2) & 3);
  return 16 * P + 4 * (q ^ st_hq(P));
}

__device__ __forceinline__ float sum4(const float4& a) { return (a.x + a.y) + (a.z + a.w); }

// acc += A B over the 16 points: A rows (ra), B rows (rb) staged point-major;
// MFMA kc contracts points {4 g' + kc : g' < 4}.
__device__ __forceinline__ v4f mfma_pts(const float4& a, const float4& b, v4f acc) {
  acc = mfma4(a.x, b.x, acc);
  acc = mfma4(a.y, b.y, acc);
  acc = mfma4(a.z, b.z, acc);
  acc = mfma4(a.w, b.w, acc);
  return acc;
}
__device__ __forceinline__ float4 neg4(const float4& a) { return make_float4(-a.x, -a.y, -a.z, -a.w); }

// One wave's share of the weight-gradient sums into the GradLayout slots at R
// (every slot but the frequency terms): stored (ADD false) or added.
template <bool ADD>
__device__ __forceinline__ void put_slot(float& r, float v) {
  if (ADD) r += v;
  else r = v;
}

template <int H, bool ADD>
__device__ __forceinline__ void put_wave_sums(float* R, const v4f (&P0)[2][2], const v4f (&Pr)[H][2][2],
                                              const v4f (&Pi)[H][2][2], const float (&q0s)[2],
                                              const float (&qrs)[H][2], const float (&qis)[H][2], float acc_wo,
                                              float br, float bi, int lane, int g, int j) {
  using GL = GradLayout<H>;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int o = 16 * mb + 4 * g + v, in = 16 * nb + j;
        put_slot<ADD>(R[GL::w0 + o * 32 + in], P0[mb][nb][v]);
#pragma unroll
        for (int k = 0; k < H; ++k) {
          put_slot<ADD>(R[GL::A(k) + 2 * (o * 32 + in)], Pr[k][mb][nb][v]);
          put_slot<ADD>(R[GL::A(k) + 2 * (o * 32 + in) + 1], Pi[k][mb][nb][v]);
        }
      }
  if (g == 0) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      put_slot<ADD>(R[GL::b0 + 16 * mb + j], q0s[mb]);
#pragma unroll
      for (int k = 0; k < H; ++k) {
        put_slot<ADD>(R[GL::c(k) + 2 * (16 * mb + j)], qrs[k][mb]);
        put_slot<ADD>(R[GL::c(k) + 2 * (16 * mb + j) + 1], qis[k][mb]);
      }
    }
  }
  put_slot<ADD>(R[GL::wo + 2 * (lane & 31) + (lane >> 5)], acc_wo);
  if (lane == 0) {
    put_slot<ADD>(R[GL::bo], br);
    put_slot<ADD>(R[GL::bo + 1], bi);
  }
}

template <int H>
__global__ __launch_bounds__(64 * kW16) void k_inr_bwd16(PtyGrid gr, PtyMlp m, const float2* __restrict__ table,
                                                         float2* __restrict__ gfeat, PtyPoints pts,
                                                         const float2* __restrict__ dout,
                                                         float* __restrict__ partial) {
  __shared__ Lds16<H> s;
  __shared__ Stage16<H> st[kW16];
  __shared__ uint4 s_lvl[kLevelTab];
  load16<H>(s, m, threadIdx.x, blockDim.x);
  load_level_tab(gr, s_lvl);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  Stage16<H>& S = st[wave];
  // Loop-invariant staging offsets of this lane (rows 16 apart differ by 256):
  // wrow[e] = (row 4 g + e, point j), rrow = (row j, quad g), frow[e] =
  // (row 8 (e >> 1) + 2 g + (e & 1), point j).
  int wrow[4], frow[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    wrow[e] = st_elem(4 * g + e, j);
    frow[e] = st_elem(8 * (e >> 1) + 2 * g + (e & 1), j);
  }
  const int rrow = st_quad(j, g);

  // P_k[mb][nb] (re, im): rows 16 mb + 4 g + v (out units), column 16 nb + j (in units)
  v4f Pr[H][2][2], Pi[H][2][2];
  v4f P0[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      P0[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < H; ++k) Pr[k][a][b] = Pi[k][a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    }
  // bias-like sums: lane (i, g) holds unit 16 mb + i over points 4 g .. 4 g + 3
  float qr[H][2], qi[H][2], q0[2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    q0[a] = 0.f;
#pragma unroll
    for (int k = 0; k < H; ++k) qr[k][a] = qi[k][a] = 0.f;
  }
  float acc_wo = 0.f, acc_bor = 0.f, acc_boi = 0.f;

  const int64_t ntiles = (pts.n + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * kW16 + wave; tile < ntiles; tile += (int64_t)gridDim.x * kW16) {
    asm volatile("" ::: "memory");
    const int64_t p = tile * 16 + j;
    const bool valid = p < pts.n;
    // ---------------- recompute forward (inputs of every layer staged); the
    // features come from the forward's cache when there is one
    float feat[8];
    if (pts.features) {
      load_features16(pts.features, tile, lane, feat);
    } else {
      float x0 = 0.5f, x1 = 0.5f;
      if (valid) point_coords(pts, p, x0, x1);
      encode16(gr, s_lvl, x0, x1, g, table, feat);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) S.F[256 * (c >> 2) + frow[c & 3]] = feat[c];
    float zr[8], zi[8];
    // The two waves of a SIMD alternate MFMA blocks with transcendental
    // activation work; a wave raises its issue priority while it feeds the
    // matrix core, so that the other wave's VALU work fills in around it
    // rather than the reverse.
    {
      v4f u[2];
      const float w0r = s.om[0] * 0.15915494309189535f;
      __builtin_amdgcn_s_setprio(2);
      layer0_16<H>(s, feat, j, g, u);
      __builtin_amdgcn_s_setprio(0);
#pragma unroll
      for (int q = 0; q < 8; ++q) cis_rev(w0r * u[q >> 2][q & 3], zr[q], zi[q]);
    }
#pragma unroll
    for (int k = 0; k < H; ++k) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        S.X[k][256 * (q >> 2) + wrow[q & 3]] = zr[q];
        S.X[k][512 + 256 * (q >> 2) + wrow[q & 3]] = zi[q];
      }
      v4f vr[2], vi[2];
      __builtin_amdgcn_s_setprio(2);
      hidden16<H>(s, k, zr, zi, j, g, vr, vi);
      __builtin_amdgcn_s_setprio(0);
      act16(s.om[k + 1], vr, vi, zr, zi);
    }
    // ---------------- output layer
    float2 go = make_float2(0.f, 0.f);
    if (valid) go = dout[p];
    if (g == 0) { acc_bor += go.x; acc_boi += go.y; }
    float gzr[8], gzi[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int u = unit16(q, g);
      // g_wo[u] += g_o conj(z_u); g_z = conj(wo_u) g_o
      S.G[256 * (q >> 2) + wrow[q & 3]] = go.x * zr[q] + go.y * zi[q];
      S.G[512 + 256 * (q >> 2) + wrow[q & 3]] = go.y * zr[q] - go.x * zi[q];
      gzr[q] = s.wor[u] * go.x + s.woi[u] * go.y;
      gzi[q] = s.wor[u] * go.y - s.woi[u] * go.x;
    }
    wave_sync();
    // lane l sums staged row l (Re of unit l for l < 32, else Im of unit l - 32)
    acc_wo += (sum4(lds4(&S.G[st_quad(lane, 0)])) + sum4(lds4(&S.G[st_quad(lane, 1)]))) +
              (sum4(lds4(&S.G[st_quad(lane, 2)])) + sum4(lds4(&S.G[st_quad(lane, 3)])));
    wave_sync();
    // ---------------- hidden layers, last to first
#pragma unroll
    for (int k = H - 1; k >= 0; --k) {
      const float w = s.om[k + 1];
      float gwr[8], gwi[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        gwr[q] = gzr[q] * zr[q] + gzi[q] * zi[q];
        gwi[q] = gzi[q] * zr[q] - gzr[q] * zi[q];
        const int u = unit16(q, g);
        S.G[256 * (q >> 2) + wrow[q & 3]] = gwr[q];
        S.G[512 + 256 * (q >> 2) + wrow[q & 3]] = gwi[q];
      }
      wave_sync();
      // P += G X^H: Pr = Gr Xr^T + Gi Xi^T, Pi = Gi Xr^T - Gr Xi^T
      __builtin_amdgcn_s_setprio(2);
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        const float4 ar = lds4(&S.G[256 * mb + rrow]);
        const float4 ai = lds4(&S.G[512 + 256 * mb + rrow]);
        qr[k][mb] += sum4(ar);
        qi[k][mb] += sum4(ai);
        const float4 nar = neg4(ar);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
          const float4 xr = lds4(&S.X[k][256 * nb + rrow]);
          const float4 xi = lds4(&S.X[k][512 + 256 * nb + rrow]);
          Pr[k][mb][nb] = mfma_pts(ar, xr, Pr[k][mb][nb]);
          Pr[k][mb][nb] = mfma_pts(ai, xi, Pr[k][mb][nb]);
          Pi[k][mb][nb] = mfma_pts(ai, xr, Pi[k][mb][nb]);
          Pi[k][mb][nb] = mfma_pts(nar, xi, Pi[k][mb][nb]);
        }
      }
      // t = A^H g_w (in units 16 mb + 4 g + v), g_x = -i w t; three real
      // products (Gauss): T1 = Ar^T gwr, T2 = Ai^T gwi,
      // T3 = (Ar - Ai)^T (gwr + gwi); Re t = T1 + T2, Im t = T3 - T1 + T2
      v4f t1[2], t2[2], t3[2];
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) t1[mb] = t2[mb] = t3[mb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int u = unit16(q, g);
        const float gs = gwr[q] + gwi[q];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
          const float a_r = s.ar[k][u * WS36 + 16 * mb + j];
          const float a_i = s.ai[k][u * WS36 + 16 * mb + j];
          t1[mb] = mfma4(a_r, gwr[q], t1[mb]);
          t2[mb] = mfma4(a_i, gwi[q], t2[mb]);
          t3[mb] = mfma4(a_r - a_i, gs, t3[mb]);
        }
      }
      __builtin_amdgcn_s_setprio(0);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float a1 = t1[q >> 2][q & 3], a2 = t2[q >> 2][q & 3], a3 = t3[q >> 2][q & 3];
        gzr[q] = w * ((a3 - a1) + a2);
        gzi[q] = -w * (a1 + a2);
        zr[q] = S.X[k][256 * (q >> 2) + wrow[q & 3]];
        zi[q] = S.X[k][512 + 256 * (q >> 2) + wrow[q & 3]];
      }
      wave_sync();
    }
    // ---------------- layer 0 (real): staged Im(g_z conj(z0)) = g_u0 / w0
    {
      const float w = s.om[0];
      float gu[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        gu[q] = gzi[q] * zr[q] - gzr[q] * zi[q];
        S.G[256 * (q >> 2) + wrow[q & 3]] = gu[q];
      }
      wave_sync();
      __builtin_amdgcn_s_setprio(2);
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        const float4 a = lds4(&S.G[256 * mb + rrow]);
        q0[mb] += sum4(a);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) P0[mb][nb] = mfma_pts(a, lds4(&S.F[256 * nb + rrow]), P0[mb][nb]);
      }
      // g_h (features 16 fb + 4 g + v) = W0^T Im(.)
      v4f gh[2];
#pragma unroll
      for (int fb = 0; fb < 2; ++fb) gh[fb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int u = unit16(q, g);
#pragma unroll
        for (int fb = 0; fb < 2; ++fb) gh[fb] = mfma4(s.w0[u * WS36 + w0pos(16 * fb + j)], gu[q], gh[fb]);
      }
      __builtin_amdgcn_s_setprio(0);
      wave_sync();
      // feature 16 fb + 4 g + 2 h + e is feature e of level 8 fb + 2 g + h
      if (valid) {
#pragma unroll
        for (int fb = 0; fb < 2; ++fb)
#pragma unroll
          for (int hh = 0; hh < 2; ++hh) {
            const int l = 8 * fb + 2 * g + hh;
            if (l < gr.n_levels)
              gfeat[(int64_t)l * pts.n + p] = make_float2(w * gh[fb][2 * hh], w * gh[fb][2 * hh + 1]);
          }
      }
    }
  }

  // ---------------- workgroup reduction into partial[blockIdx.x][GradLayout::total]
  // Waves 0-3 store their sums into regions 0-3 of the staging area, waves 4-7
  // add theirs into region wave - 4, then every thread totals its slots over
  // the four regions in a fixed order: three barriers instead of one per wave.
  using GL = GradLayout<H>;
  constexpr int kRegions = kW16 / 2;
  static_assert(sizeof(st) / sizeof(float) >= (size_t)kRegions * GL::total, "reduction buffer does not fit");
  __syncthreads();
  float* red = &st[0].X[0][0];
  // bias sums over the four lane groups (same unit, different points)
  float q0s[2], qrs[H][2], qis[H][2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    q0s[a] = q0[a] + __shfl_xor(q0[a], 16);
    q0s[a] += __shfl_xor(q0s[a], 32);
#pragma unroll
    for (int k = 0; k < H; ++k) {
      qrs[k][a] = qr[k][a] + __shfl_xor(qr[k][a], 16);
      qrs[k][a] += __shfl_xor(qrs[k][a], 32);
      qis[k][a] = qi[k][a] + __shfl_xor(qi[k][a], 16);
      qis[k][a] += __shfl_xor(qis[k][a], 32);
    }
  }
  const float br = wave_sum(acc_bor), bi = wave_sum(acc_boi);
  float* R = red + (wave % kRegions) * GL::total;
  if (wave < kRegions)
    put_wave_sums<H, false>(R, P0, Pr, Pi, q0s, qrs, qis, acc_wo, br, bi, lane, g, j);
  __syncthreads();
  if (wave >= kRegions)
    put_wave_sums<H, true>(R, P0, Pr, Pi, q0s, qrs, qis, acc_wo, br, bi, lane, g, j);
  __syncthreads();
  static_assert(kRegions == 4, "the fixed summation order below assumes four regions");
  for (int t = threadIdx.x; t < GL::total; t += blockDim.x)
    if (t < GL::om || t >= GL::wo)
      red[t] = (red[t] + red[GL::total + t]) + (red[2 * GL::total + t] + red[3 * GL::total + t]);
  __syncthreads();
  omega_terms<H>(red, m);
  for (int t = threadIdx.x; t < GL::total; t += blockDim.x) partial[(int64_t)blockIdx.x * GL::total + t] = red[t];
}

// ============================================================== launchers
void launch_inr_fwd_f32(const PtyGrid& g, const PtyMlp& m, const float2* table, const PtyPoints& pts, float2* out,
                        hipStream_t stream) {
  // five workgroups per CU (the kernel fits five waves per SIMD): each wave
  // walks several tiles, so the weights are staged in LDS once per
  // workgroup rather than once per 64 points
  const int64_t tiles = (pts.n + 15) / 16;
  int64_t blocks = (tiles + kFwd16Waves - 1) / kFwd16Waves;
  const int64_t cap = 5 * (int64_t)device_cu_count();
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  if (m.n_hidden == 1)
    hipLaunchKernelGGL(k_inr_fwd16<1>, dim3((unsigned)blocks), dim3(64 * kFwd16Waves), 0, stream, g, m, table, pts,
                       out);
  else
    hipLaunchKernelGGL(k_inr_fwd16<2>, dim3((unsigned)blocks), dim3(64 * kFwd16Waves), 0, stream, g, m, table, pts,
                       out);
}

// `blocks` workgroups of 8 waves; each walks 16-point tiles
void launch_inr_bwd_f32(const PtyGrid& g, const PtyMlp& m, const float2* table, float2* gfeat, const PtyPoints& pts,
                        const float2* dout, float* partial, int blocks, hipStream_t stream) {
  if (m.n_hidden == 1)
    hipLaunchKernelGGL(k_inr_bwd16<1>, dim3(blocks), dim3(64 * kW16), 0, stream, g, m, table, gfeat, pts, dout,
                       partial);
  else
    hipLaunchKernelGGL(k_inr_bwd16<2>, dim3(blocks), dim3(64 * kW16), 0, stream, g, m, table, gfeat, pts, dout,
                       partial);
}

}  // namespace pty
