// Standalone hash-grid encoder kernels (drop-in for tcnn.Encoding "HashGrid",
// reference call sites models/complex_inr.py:138-145, 279, 294).
#include "hashgrid.h"

namespace pty {

__global__ __launch_bounds__(256) void k_hashgrid_fwd(PtyGrid g, const float* __restrict__ coords,
                                                       int64_t n, const float2* __restrict__ table,
                                                       float2* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x0 = coords[2 * i], x1 = coords[2 * i + 1];
  float2* o = out + i * g.n_levels;
  for (int l = 0; l < g.n_levels; ++l) o[l] = level_feature(g, l, x0, x1, table);
}

__global__ __launch_bounds__(256) void k_hashgrid_bwd(PtyGrid g, const float* __restrict__ coords,
                                                       int64_t n, const float2* __restrict__ dout,
                                                       float* __restrict__ dtable) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x0 = coords[2 * i], x1 = coords[2 * i + 1];
  const float2* d = dout + i * g.n_levels;
  for (int l = 0; l < g.n_levels; ++l) level_scatter(g, l, x0, x1, d[l], dtable);
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
    if (!dense)
      PTY_REQUIRE(g->size[l] == (1u << g->log2_hashmap_size), PTY_EINVAL,
                  "hashed level %d must have size 2^log2T", l);
    else
      PTY_REQUIRE((uint64_t)g->res[l] * g->res[l] <= g->size[l], PTY_EINVAL,
                  "dense level %d: res^2 > size", l);
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
  const int64_t blocks = (n + 255) / 256;
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
  const int64_t blocks = (n + 255) / 256;
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
