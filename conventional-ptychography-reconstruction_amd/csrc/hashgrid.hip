// Standalone hash-grid encoder kernels (drop-in for tcnn.Encoding "HashGrid",
// reference call sites models/complex_inr.py:138-145, 279, 294).
#include "hashgrid.h"

namespace pty {

// Level constants of the descriptor staged in LDS, so that lanes on
// different levels read their own without indexing the kernel arguments.
struct LevelTab {
  float scale[PTY_MAX_LEVELS];
  uint32_t res[PTY_MAX_LEVELS], size[PTY_MAX_LEVELS], offset[PTY_MAX_LEVELS], dense[PTY_MAX_LEVELS];
};

__device__ __forceinline__ void stage_levels(const PtyGrid& g, LevelTab& t) {
  const int l = threadIdx.x;
  if (l < g.n_levels) {
    t.scale[l] = g.scale[l];
    t.res[l] = g.res[l];
    t.size[l] = g.size[l];
    t.offset[l] = g.offset[l];
    t.dense[l] = (g.dense_mask >> l) & 1u;
  }
}

__device__ __forceinline__ LevelC level_from(const LevelTab& t, int l) {
  LevelC c;
  c.scale = t.scale[l];
  c.res = t.res[l];
  c.size = t.size[l];
  c.offset = t.offset[l];
  c.dense = t.dense[l];
  return c;
}

// One lane per (point, level): lane t of a workgroup handles level t % 16 of
// point t / 16, so a point's output row [L x 2] is written by consecutive
// lanes (coalesced) and each lane issues the 4 gathers of its level.
constexpr int kPtsPerBlock = 256 / PTY_MAX_LEVELS;

__global__ __launch_bounds__(256) void k_hashgrid_fwd(PtyGrid g, const float* __restrict__ coords,
                                                       int64_t n, const float2* __restrict__ table,
                                                       float2* __restrict__ out) {
  __shared__ LevelTab lt;
  stage_levels(g, lt);
  __syncthreads();
  const int l = threadIdx.x % PTY_MAX_LEVELS;
  const int64_t i = (int64_t)blockIdx.x * kPtsPerBlock + threadIdx.x / PTY_MAX_LEVELS;
  if (i >= n || l >= g.n_levels) return;
  const float x0 = coords[2 * i], x1 = coords[2 * i + 1];
  const LevelC c = level_from(lt, l);
  const LevelPos p = level_pos_c(c, x0, x1);
  uint32_t idx[4];
  float w[4];
  corner_indices_c(c, p, idx);
  corner_weights(p, w);
  float2 acc = make_float2(0.f, 0.f);
  if (g.table_precision == PTY_TABLE_F16) {
    const __half2* t16 = reinterpret_cast<const __half2*>(table);
    float2 t[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) t[k] = __half22float2(t16[idx[k]]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc.x += w[k] * t[k].x;
      acc.y += w[k] * t[k].y;
    }
  } else {
    float2 t[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) t[k] = table[idx[k]];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc.x += w[k] * t[k].x;
      acc.y += w[k] * t[k].y;
    }
  }
  out[i * g.n_levels + l] = acc;
}

__global__ __launch_bounds__(256) void k_hashgrid_bwd(PtyGrid g, const float* __restrict__ coords,
                                                       int64_t n, const float2* __restrict__ dout,
                                                       float* __restrict__ dtable) {
  __shared__ LevelTab lt;
  stage_levels(g, lt);
  __syncthreads();
  const int l = threadIdx.x % PTY_MAX_LEVELS;
  const int64_t i = (int64_t)blockIdx.x * kPtsPerBlock + threadIdx.x / PTY_MAX_LEVELS;
  if (i >= n || l >= g.n_levels) return;
  const float2 dy = dout[i * g.n_levels + l];
  if (dy.x == 0.f && dy.y == 0.f) return;
  const float x0 = coords[2 * i], x1 = coords[2 * i + 1];
  const LevelC c = level_from(lt, l);
  const LevelPos p = level_pos_c(c, x0, x1);
  uint32_t idx[4];
  float w[4];
  corner_indices_c(c, p, idx);
  corner_weights(p, w);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    atomic_add_f32(dtable + 2 * idx[k], w[k] * dy.x);
    atomic_add_f32(dtable + 2 * idx[k] + 1, w[k] * dy.y);
  }
}

__global__ __launch_bounds__(256) void k_hashgrid_indices(PtyGrid g, const float* __restrict__ coords,
                                                           int64_t n, uint4* __restrict__ idx_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x0 = coords[2 * i], x1 = coords[2 * i + 1];
  for (int l = 0; l < g.n_levels; ++l) {
    LevelPos p = level_pos(g, l, x0, x1);
    uint32_t idx[4];
    corner_indices(g, l, p, idx);
    idx_out[i * g.n_levels + l] = make_uint4(idx[0], idx[1], idx[2], idx[3]);
  }
}

int validate_grid(const PtyGrid* g) {
  PTY_REQUIRE(g != nullptr, PTY_EINVAL, "grid descriptor is NULL");
  PTY_REQUIRE(g->abi_version == PTY_ABI_VERSION, PTY_EINVAL, "grid abi_version %d != %d",
              g->abi_version, PTY_ABI_VERSION);
  PTY_REQUIRE(g->n_levels >= 1 && g->n_levels <= PTY_MAX_LEVELS, PTY_EUNSUP,
              "n_levels %d outside 1..%d", g->n_levels, PTY_MAX_LEVELS);
  PTY_REQUIRE(g->n_features == 2, PTY_EUNSUP, "n_features_per_level must be 2 (got %d)", g->n_features);
  PTY_REQUIRE(g->table_precision == PTY_TABLE_F32 || g->table_precision == PTY_TABLE_F16, PTY_EUNSUP,
              "table_precision %u is neither PTY_TABLE_F32 nor PTY_TABLE_F16", g->table_precision);
  PTY_REQUIRE(g->log2_hashmap_size >= 3 && g->log2_hashmap_size <= 28, PTY_EUNSUP,
              "log2_hashmap_size %d outside 3..28", g->log2_hashmap_size);
  for (int l = 0; l < g->n_levels; ++l) {
    const bool dense = (g->dense_mask >> l) & 1u;
    // inputs in [0, 1] must reach at most corner res on each axis (the
    // encoders' dense-index modulo relies on it)
    PTY_REQUIRE(g->scale[l] >= 0.f && floorf(g->scale[l] + 0.5f) + 1.f <= (float)g->res[l], PTY_EINVAL,
                "level %d: scale %g needs res >= floor(scale + 0.5) + 1 (res %u)", l, (double)g->scale[l],
                g->res[l]);
    if (!dense)
      PTY_REQUIRE(g->size[l] == (1u << g->log2_hashmap_size), PTY_EINVAL,
                  "hashed level %d must have size 2^log2T", l);
    else
      PTY_REQUIRE((uint64_t)g->res[l] * g->res[l] <= g->size[l] &&
                      (uint64_t)g->res[l] * (g->res[l] + 1u) < 2ull * g->size[l] &&
                      (uint64_t)g->res[l] * (g->res[l] + 1u) < (1ull << 32),
                  PTY_EINVAL, "dense level %d: res^2 > size, res (res + 1) >= 2 size or >= 2^32", l);
  }
  return PTY_OK;
}

}  // namespace pty

using namespace pty;

extern "C" int pty_hashgrid_fwd(const PtyGrid* g, const float* coords, int64_t n, const float* table,
                                float* out, hipStream_t stream) {
  int rc = validate_grid(g);
  if (rc) return rc;
  PTY_REQUIRE(n >= 0 && (n == 0 || (coords && table && out)), PTY_EINVAL, "null pointer");
  if (n == 0) return PTY_OK;
  const int64_t blocks = (n + kPtsPerBlock - 1) / kPtsPerBlock;
  PTY_REQUIRE(blocks <= 0x7fffffff, PTY_EUNSUP, "hashgrid_fwd: %lld points exceed one launch", (long long)n);
  hipLaunchKernelGGL(k_hashgrid_fwd, dim3((unsigned)blocks), dim3(256), 0, stream, *g, coords, n,
                     (const float2*)table, (float2*)out);
  return check_launch("hashgrid_fwd");
}

extern "C" int pty_hashgrid_bwd(const PtyGrid* g, const float* coords, int64_t n, const float* dout,
                                float* dtable, hipStream_t stream) {
  int rc = validate_grid(g);
  if (rc) return rc;
  PTY_REQUIRE(n >= 0 && (n == 0 || (coords && dout && dtable)), PTY_EINVAL, "null pointer");
  if (n == 0) return PTY_OK;
  const int64_t blocks = (n + kPtsPerBlock - 1) / kPtsPerBlock;
  PTY_REQUIRE(blocks <= 0x7fffffff, PTY_EUNSUP, "hashgrid_bwd: %lld points exceed one launch", (long long)n);
  hipLaunchKernelGGL(k_hashgrid_bwd, dim3((unsigned)blocks), dim3(256), 0, stream, *g, coords, n,
                     (const float2*)dout, dtable);
  return check_launch("hashgrid_bwd");
}

extern "C" int pty_hashgrid_indices(const PtyGrid* g, const float* coords, int64_t n, uint32_t* idx,
                                    hipStream_t stream) {
  int rc = validate_grid(g);
  if (rc) return rc;
  PTY_REQUIRE(n >= 0 && (n == 0 || (coords && idx)), PTY_EINVAL, "null pointer");
  if (n == 0) return PTY_OK;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_hashgrid_indices, dim3((unsigned)blocks), dim3(256), 0, stream, *g, coords, n,
                     (uint4*)idx);
  return check_launch("hashgrid_indices");
}
