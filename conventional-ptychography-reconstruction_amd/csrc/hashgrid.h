// Hash-grid index and interpolation math shared by the standalone encoder
// kernels and the fused INR kernels.
//
// Restates the multiresolution hash encoding that the reference obtains from
// tiny-cuda-nn's HashGrid (models/complex_inr.py:130-145): per level l,
// pos = fma(x, scale_l, 0.5) (one rounding, as tiny-cuda-nn's pos_fract),
// g = floor(pos),
// frac = pos - g; corner c = (c&1, c>>1); dense index (cx + cy*res) reduced
// modulo the level size (the edge corner cx = res wraps into the next row),
// hashed index (cx ^ cy*2654435761) mod 2^32 mod 2^log2T; bilinear weights
// w_c = (c&1 ? fx : 1-fx) * (c&2 ? fy : 1-fy).
#pragma once
#include "common.h"

namespace pty {

int validate_grid(const PtyGrid* g);
// table gradient from level-major feature gradients gfeat[l][p] (hashbwd.hip)
size_t points_plan_bytes(const PtyGrid& g, const PtyPoints& pts);
size_t table_bwd_workspace(const PtyGrid& g, const PtyPoints& pts);
// With ad.p set, each entry's gradient goes straight into the table's Adam
// update (ad) instead of being added to dtable.
int launch_table_bwd(const PtyGrid& g, const PtyPoints& pts, const float2* gfeat, float* dtable, char* work,
                     hipStream_t stream, const TableAdam& ad);

struct LevelPos {
  uint32_t gx, gy;
  float fx, fy;
};

__device__ __forceinline__ LevelPos level_pos(const PtyGrid& g, int l, float x0, float x1) {
  // x*scale + 0.5 with a single rounding (fused multiply-add), the form of
  // tiny-cuda-nn's pos_fract
  const float s = g.scale[l];
  float p0 = __builtin_fmaf(x0, s, 0.5f);
  float p1 = __builtin_fmaf(x1, s, 0.5f);
  float f0 = floorf(p0), f1 = floorf(p1);
  LevelPos r;
  r.gx = (uint32_t)(int)f0;
  r.gy = (uint32_t)(int)f1;
  r.fx = p0 - f0;
  r.fy = p1 - f1;
  return r;
}

// Entry index (level offset included) of corner (cx, cy) at level l.
__device__ __forceinline__ uint32_t level_index(const PtyGrid& g, int l, uint32_t cx, uint32_t cy) {
  const uint32_t size = g.size[l];
  uint32_t local;
  if ((g.dense_mask >> l) & 1u) {
    // For inputs in [0, 1], cx, cy <= res, so cx + cy*res < 2*size and one
    // subtraction implements the modulo; other inputs take the full modulo.
    local = cx + cy * g.res[l];
    if (local >= size) local = (local < 2u * size) ? local - size : local % size;
  } else {
    // hashed levels have size == 2^log2T
    local = (cx ^ (cy * 2654435761u)) & (size - 1u);
  }
  return g.offset[l] + local;
}

// Constants of one level for lanes that may sit on different levels: la and
// lb are compile-time level indices (both candidates are then uniform kernel
// arguments, read through the scalar path) and the lane picks one with a
// select instead of indexing the descriptor per lane.
struct LevelC {
  float scale;
  uint32_t res, size, offset, dense;
};

// readfirstlane makes each candidate a value of its own: without it the
// compiler turns select(load a, load b) into a per-lane load of a selected
// address.
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ LevelC level_c(const PtyGrid& g, int la, int lb, bool pick_b) {
  LevelC c;
  const uint32_t sa = uniform_u32(__float_as_uint(g.scale[la])), sb = uniform_u32(__float_as_uint(g.scale[lb]));
  const uint32_t ra = uniform_u32(g.res[la]), rb = uniform_u32(g.res[lb]);
  const uint32_t za = uniform_u32(g.size[la]), zb = uniform_u32(g.size[lb]);
  const uint32_t oa = uniform_u32(g.offset[la]), ob = uniform_u32(g.offset[lb]);
  const uint32_t dm = uniform_u32(g.dense_mask);
  c.scale = __uint_as_float(pick_b ? sb : sa);
  c.res = pick_b ? rb : ra;
  c.size = pick_b ? zb : za;
  c.offset = pick_b ? ob : oa;
  c.dense = (dm >> (pick_b ? lb : la)) & 1u;
  return c;
}

__device__ __forceinline__ LevelPos level_pos_c(const LevelC& c, float x0, float x1) {
  float p0 = __builtin_fmaf(x0, c.scale, 0.5f);
  float p1 = __builtin_fmaf(x1, c.scale, 0.5f);
  float f0 = floorf(p0), f1 = floorf(p1);
  LevelPos r;
  r.gx = (uint32_t)(int)f0;
  r.gy = (uint32_t)(int)f1;
  r.fx = p0 - f0;
  r.fy = p1 - f1;
  return r;
}

// level_index with the level's constants in registers; both index forms are
// computed and selected.  For inputs in [0, 1] the dense index is below
// 2 size, so one conditional subtraction (no branch) is the modulo; the full
// modulo runs only when some lane of the wavefront has an input outside
// [0, 1] (a wave-uniform branch: no per-lane exec-mask nesting in the common
// path).  (dl - size) mod size = dl mod size, so the two steps compose.
__device__ __forceinline__ uint32_t level_index_c(const LevelC& c, uint32_t cx, uint32_t cy) {
  uint32_t dl = cx + cy * c.res;
  dl = dl >= c.size ? dl - c.size : dl;
  if (__builtin_amdgcn_ballot_w64(dl >= c.size)) {
    if (dl >= c.size) dl %= c.size;
  }
  const uint32_t hl = (cx ^ (cy * 2654435761u)) & (c.size - 1u);
  return c.offset + (c.dense ? dl : hl);
}

// level_index_c split by a wave-uniform property of the inputs.  UNIT: every
// input of the wavefront lies in [0, 1], so (validate_grid requires scale >= 0,
// floor(scale + 0.5) + 1 <= res and res (res + 1) < 2 size on dense levels)
// cx, cy <= res, cx + cy res < 2 size and one conditional subtraction is the
// modulo.  Otherwise the full modulo of the 32-bit wrapped index.
template <bool UNIT>
__device__ __forceinline__ uint32_t level_index_u(const LevelC& c, uint32_t cx, uint32_t cy) {
  uint32_t dl = cx + cy * c.res;
  if constexpr (UNIT) dl = dl >= c.size ? dl - c.size : dl;
  else dl %= c.size;
  const uint32_t hl = (cx ^ (cy * 2654435761u)) & (c.size - 1u);
  return c.offset + (c.dense ? dl : hl);
}

template <bool UNIT>
__device__ __forceinline__ void corner_indices_u(const LevelC& c, const LevelPos& p, uint32_t idx[4]) {
  idx[0] = level_index_u<UNIT>(c, p.gx, p.gy);
  idx[1] = level_index_u<UNIT>(c, p.gx + 1u, p.gy);
  idx[2] = level_index_u<UNIT>(c, p.gx, p.gy + 1u);
  idx[3] = level_index_u<UNIT>(c, p.gx + 1u, p.gy + 1u);
}

// The same corners when every lane's level is known to be dense (DENSE) or
// hashed (!DENSE): only that index form is computed.
template <bool UNIT, bool DENSE>
__device__ __forceinline__ uint32_t level_index_k(const LevelC& c, uint32_t cx, uint32_t cy) {
  if constexpr (DENSE) {
    uint32_t dl = cx + cy * c.res;
    if constexpr (UNIT) dl = dl >= c.size ? dl - c.size : dl;
    else dl %= c.size;
    return c.offset + dl;
  } else {
    return c.offset + ((cx ^ (cy * 2654435761u)) & (c.size - 1u));
  }
}

template <bool UNIT, bool DENSE>
__device__ __forceinline__ void corner_indices_k(const LevelC& c, const LevelPos& p, uint32_t idx[4]) {
  idx[0] = level_index_k<UNIT, DENSE>(c, p.gx, p.gy);
  idx[1] = level_index_k<UNIT, DENSE>(c, p.gx + 1u, p.gy);
  idx[2] = level_index_k<UNIT, DENSE>(c, p.gx, p.gy + 1u);
  idx[3] = level_index_k<UNIT, DENSE>(c, p.gx + 1u, p.gy + 1u);
}

__device__ __forceinline__ void corner_indices_c(const LevelC& c, const LevelPos& p, uint32_t idx[4]) {
  idx[0] = level_index_c(c, p.gx, p.gy);
  idx[1] = level_index_c(c, p.gx + 1u, p.gy);
  idx[2] = level_index_c(c, p.gx, p.gy + 1u);
  idx[3] = level_index_c(c, p.gx + 1u, p.gy + 1u);
}

__device__ __forceinline__ void corner_weights(const LevelPos& p, float w[4]) {
  const float ax = 1.f - p.fx, ay = 1.f - p.fy;
  w[0] = ax * ay;
  w[1] = p.fx * ay;
  w[2] = ax * p.fy;
  w[3] = p.fx * p.fy;
}

__device__ __forceinline__ void corner_indices(const PtyGrid& g, int l, const LevelPos& p, uint32_t idx[4]) {
  idx[0] = level_index(g, l, p.gx, p.gy);
  idx[1] = level_index(g, l, p.gx + 1u, p.gy);
  idx[2] = level_index(g, l, p.gx, p.gy + 1u);
  idx[3] = level_index(g, l, p.gx + 1u, p.gy + 1u);
}

__device__ __forceinline__ void level_scatter(const PtyGrid& g, int l, float x0, float x1, float2 dy,
                                              float* __restrict__ dtable) {
  LevelPos p = level_pos(g, l, x0, x1);
  uint32_t idx[4];
  float w[4];
  corner_indices(g, l, p, idx);
  corner_weights(p, w);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    atomic_add_f32(dtable + 2 * idx[c], w[c] * dy.x);
    atomic_add_f32(dtable + 2 * idx[c] + 1, w[c] * dy.y);
  }
}

// Input coordinates of point n (explicit list or the regular grid).
__device__ __forceinline__ void point_coords(const PtyPoints& pts, int64_t n, float& x0, float& x1) {
  if (pts.coords) {
    x0 = pts.coords[2 * n];
    x1 = pts.coords[2 * n + 1];
  } else {
    const int i = (int)(n / pts.n_axis0);
    const int j = (int)(n - (int64_t)i * pts.n_axis0);
    x0 = pts.axis0[j];
    x1 = pts.axis1[i];
  }
}

}  // namespace pty
