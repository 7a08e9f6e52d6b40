// 16-point wave tiles shared by the FP32 and BF16 fused INR kernels: a
// wavefront owns 16 points, lane l = (j, g) = (l & 15, l >> 4) holds point j;
// the 16x16 MFMA C layout gives lane (j, g) rows 4 g + v (v < 4) of column j.
//   unit16(s, g): the 8 MLP units lane group g holds (two 16-row M blocks);
//   feat16(c, g): the 8 hash features lane group g interpolates (levels
//                 4 m + g, m < 4), in the order they enter layer 0;
// encode16 forms the lane's features, and the feature cache keeps them
// between a forward and the backward of the same step (8 floats per lane).
#pragma once
#include "inr_common.h"

namespace pty {

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int unit16(int s, int g) { return 16 * (s >> 2) + 4 * g + (s & 3); }
__device__ __forceinline__ int feat16(int c, int g) { return 2 * (4 * (c >> 1) + g) + (c & 1); }

// The encoder's per-level constants (scale bits, res, size, offset) staged in
// LDS once per workgroup: lane group g reads level 4 m + g's entry (four
// distinct addresses per read, a broadcast).  Kept in LDS rather than as
// uniform kernel-argument values, which the compiler had to hold in 64 SGPRs
// (spilling some into VGPR lanes inside the tile loop) and select per lane.
constexpr int kLevelTab = 16;

__device__ __forceinline__ void load_level_tab(const PtyGrid& gr, uint4* tab) {
  const int l = threadIdx.x;
  if (l < kLevelTab)
    tab[l] = l < gr.n_levels ? make_uint4(__float_as_uint(gr.scale[l]), gr.res[l], gr.size[l], gr.offset[l])
                             : make_uint4(0u, 0u, 1u, 0u);   // scale 0, res 0, size 1: every corner is entry 0
}

__device__ __forceinline__ LevelC level_c_tab(const uint4* tab, uint32_t dense_mask, int l) {
  const uint4 q = tab[l];
  LevelC c;
  c.scale = __uint_as_float(q.x);
  c.res = q.y;
  c.size = q.z;
  c.offset = q.w;
  c.dense = (dense_mask >> l) & 1u;
  return c;
}

// Corner indices and fractions of the lane's levels 4 m + g (m < 4).
template <bool UNIT>
__device__ __forceinline__ void corners16(const PtyGrid& gr, const uint4* lvl, uint32_t dm, float x0, float x1, int g,
                                          uint32_t (&idx)[4][4], float (&fx)[4], float (&fy)[4]) {
  // Levels past n_levels read the padding entries of the table (scale 0,
  // res 0, size 1, offset 0), whose corners are all entry 0: in bounds, and
  // their features are discarded.  No per-level branch, so the four LDS reads
  // issue together.
  LevelC c[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) c[m] = level_c_tab(lvl, dm, 4 * m + g);
  // Lane group g holds level 4 m + g: when the four levels of m are all dense
  // or all hashed (a wave-uniform test of the dense mask; padding levels count
  // as hashed) only that index form is computed, else both and a select.
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const LevelPos p = level_pos_c(c[m], x0, x1);
    const uint32_t kinds = (dm >> (4 * m)) & 0xFu;
    if (kinds == 0u)
      corner_indices_k<UNIT, false>(c[m], p, idx[m]);
    else
      corner_indices_u<UNIT>(c[m], p, idx[m]);
    fx[m] = p.fx;
    fy[m] = p.fy;
  }
}

// The lane's 8 features feat[c] = feature feat16(c, g): the 16 corner indices
// are formed first and the 16 gathers issued back to back.  When every
// coordinate of the wavefront lies in [0, 1] (the INR grids always do) the
// dense-level modulo is one conditional subtraction; one uniform test per
// tile picks that form.
__device__ __forceinline__ void encode16(const PtyGrid& gr, const uint4* lvl, float x0, float x1, int g,
                                         const float2* __restrict__ table, float feat[8]) {
  uint32_t idx[4][4];
  float fx[4], fy[4];
  const uint32_t dm = __builtin_amdgcn_readfirstlane(gr.dense_mask);
  const bool in_unit = x0 >= 0.f && x0 <= 1.f && x1 >= 0.f && x1 <= 1.f;
  if (__builtin_amdgcn_ballot_w64(!in_unit) == 0)
    corners16<true>(gr, lvl, dm, x0, x1, g, idx, fx, fy);
  else
    corners16<false>(gr, lvl, dm, x0, x1, g, idx, fx, fy);
  // Gathers: below 2^29 entries the byte offset of an entry fits 32 bits, so
  // the load takes the table base as a scalar and a 32-bit lane offset (no
  // 64-bit address arithmetic per gather).
  float2 t[4][4];
  const char* tb = reinterpret_cast<const char*>(table);
  const bool off32 = gr.n_entries <= (1u << 29);
  if (gr.table_precision == PTY_TABLE_F16) {
    const __half2* t16 = reinterpret_cast<const __half2*>(table);
    if (off32) {
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int c = 0; c < 4; ++c)
          t[m][c] = __half22float2(*reinterpret_cast<const __half2*>(tb + (idx[m][c] << 2)));
    } else {
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int c = 0; c < 4; ++c) t[m][c] = __half22float2(t16[idx[m][c]]);
    }
  } else if (off32) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int c = 0; c < 4; ++c) t[m][c] = *reinterpret_cast<const float2*>(tb + (idx[m][c] << 3));
  } else {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int c = 0; c < 4; ++c) t[m][c] = table[idx[m][c]];
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    float2 acc = make_float2(0.f, 0.f);
    if (4 * m + g < gr.n_levels) {
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

// Feature cache: the 8 features of lane `lane` of tile `tile` at
// features[(tile * 64 + lane) * 8] (32 coalesced bytes per lane).
__device__ __forceinline__ void store_features16(float* f, int64_t tile, int lane, const float feat[8]) {
  float4* p = reinterpret_cast<float4*>(f + (tile * 64 + lane) * 8);
  p[0] = make_float4(feat[0], feat[1], feat[2], feat[3]);
  p[1] = make_float4(feat[4], feat[5], feat[6], feat[7]);
}
__device__ __forceinline__ void load_features16(const float* f, int64_t tile, int lane, float feat[8]) {
  const float4* p = reinterpret_cast<const float4*>(f + (tile * 64 + lane) * 8);
  const float4 a = p[0], b = p[1];
  feat[0] = a.x; feat[1] = a.y; feat[2] = a.z; feat[3] = a.w;
  feat[4] = b.x; feat[5] = b.y; feat[6] = b.z; feat[7] = b.w;
}


}  // namespace pty
