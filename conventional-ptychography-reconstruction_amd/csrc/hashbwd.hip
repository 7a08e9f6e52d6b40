// Hash-table gradient from per-point feature gradients gfeat[l][p] (float2,
// level-major), the adjoint of the encoding in the fused INR backward.
//
// Hashed levels (and every level for explicit coordinates) scatter
// w_c * gfeat into the 4 corner entries with atomics; those entries are
// spread over the table so the atomics hardly ever collide.
//
// Dense levels on the regular grid of init_scale_grids use the separable
// structure of the bilinear weights instead of atomics: with
//   gx(j), fx(j) = cell / fraction of axis0[j],  gy(i), fy(i) of axis1[i],
// the gradient of virtual corner (cx, cy) (cx, cy in [0, res]) is
//   sum_{dy} sum_{i: gy(i)=cy-dy} wy(i,dy) * A[i][cx],
//   A[i][cx] = sum_{dx} sum_{j: gx(j)=cx-dx} wx(j,dx) * gfeat[l][i*n0 + j],
// w(.,0) = 1 - f, w(.,1) = f.  gx is non-decreasing in j (the axes are
// increasing and f32 x*s+0.5 is monotone), so each cell's points form a
// contiguous range found by binary search.  Each virtual corner is summed by
// one thread in a fixed order and added once to its entry (the only
// collisions are the edge corners cx = res that wrap onto the next row).
#include "hashgrid.h"

namespace pty {

struct TableBwdLayout {
  int n_dense;                  // dense levels handled by the gather path
  int lev[PTY_MAX_LEVELS];      // level of each dense slot
  int off_s[PTY_MAX_LEVELS];    // offset of the level's start table (res + 2 ints)
  int64_t off_a[PTY_MAX_LEVELS + 1];  // offset of the level's A block (n1 * (res + 1) float2)
  int64_t off_c[PTY_MAX_LEVELS + 1];  // prefix of (res + 1)^2 over dense slots
  int n_starts;                 // total ints per axis start table
};

__device__ __forceinline__ uint32_t axis_cell(const PtyGrid& g, int l, float x, float& frac) {
#pragma clang fp contract(off)
  const float p = x * g.scale[l] + 0.5f;
  const float f = floorf(p);
  frac = p - f;
  return (uint32_t)(int)f;
}

// starts[off_s[d] + c] = first index j with cell(axis[j]) >= c, c in [0, res + 2)
__global__ void k_axis_starts(PtyGrid g, TableBwdLayout L, const float* __restrict__ axis, int n,
                              int* __restrict__ starts) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.n_starts) return;
  int d = 0;
  while (d + 1 < L.n_dense && t >= L.off_s[d + 1]) ++d;
  const int l = L.lev[d];
  const int c = t - L.off_s[d];
  int lo = 0, hi = n;  // first j in [0, n) with cell >= c, else n
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    float fr;
    if ((int)axis_cell(g, l, axis[mid], fr) >= c) hi = mid; else lo = mid + 1;
  }
  starts[t] = lo;
}

// A[i][cx] for every dense slot
__global__ __launch_bounds__(256) void k_dense_pass1(PtyGrid g, TableBwdLayout L, PtyPoints pts,
                                                     const float2* __restrict__ gfeat, const int* __restrict__ s0,
                                                     float2* __restrict__ A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.off_a[L.n_dense]) return;
  int d = 0;
  while (d + 1 < L.n_dense && t >= L.off_a[d + 1]) ++d;
  const int l = L.lev[d];
  const int res1 = (int)g.res[l] + 1;
  const int64_t r = t - L.off_a[d];
  const int i = (int)(r / res1), cx = (int)(r - (int64_t)i * res1);
  const int* st = s0 + L.off_s[d];
  const float2* gl = gfeat + (int64_t)l * pts.n + (int64_t)i * pts.n_axis0;
  float ax = 0.f, ay = 0.f;
#pragma unroll
  for (int dx = 0; dx < 2; ++dx) {
    const int cell = cx - dx;
    if (cell < 0) continue;
    const int j0 = st[cell], j1 = st[cell + 1];
    for (int j = j0; j < j1; ++j) {
      float fr;
      axis_cell(g, l, pts.axis0[j], fr);
      const float w = dx ? fr : 1.f - fr;
      const float2 v = gl[j];
      ax += w * v.x;
      ay += w * v.y;
    }
  }
  A[t] = make_float2(ax, ay);
}

// corner sums and one add per virtual corner
__global__ __launch_bounds__(256) void k_dense_pass2(PtyGrid g, TableBwdLayout L, PtyPoints pts,
                                                     const int* __restrict__ s1, const float2* __restrict__ A,
                                                     float* __restrict__ dtable) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.off_c[L.n_dense]) return;
  int d = 0;
  while (d + 1 < L.n_dense && t >= L.off_c[d + 1]) ++d;
  const int l = L.lev[d];
  const int res1 = (int)g.res[l] + 1;
  const int64_t r = t - L.off_c[d];
  const int cy = (int)(r / res1), cx = (int)(r - (int64_t)cy * res1);
  const int* st = s1 + L.off_s[d];
  const float2* Al = A + L.off_a[d];
  float ax = 0.f, ay = 0.f;
#pragma unroll
  for (int dy = 0; dy < 2; ++dy) {
    const int cell = cy - dy;
    if (cell < 0) continue;
    const int i0 = st[cell], i1 = st[cell + 1];
    for (int i = i0; i < i1; ++i) {
      float fr;
      axis_cell(g, l, pts.axis1[i], fr);
      const float w = dy ? fr : 1.f - fr;
      const float2 v = Al[(int64_t)i * res1 + cx];
      ax += w * v.x;
      ay += w * v.y;
    }
  }
  if (ax != 0.f || ay != 0.f) {
    const uint32_t e = level_index(g, l, (uint32_t)cx, (uint32_t)cy);
    atomic_add_f32(dtable + 2 * e, ax);
    atomic_add_f32(dtable + 2 * e + 1, ay);
  }
}

// scatter for the levels in `mask`
__global__ __launch_bounds__(256) void k_level_scatter(PtyGrid g, uint32_t mask, PtyPoints pts,
                                                       const float2* __restrict__ gfeat,
                                                       float* __restrict__ dtable) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int l = blockIdx.y;
  if (p >= pts.n || !((mask >> l) & 1u)) return;
  float x0, x1;
  point_coords(pts, p, x0, x1);
  const float2 dy = gfeat[(int64_t)l * pts.n + p];
  if (dy.x == 0.f && dy.y == 0.f) return;
  level_scatter(g, l, x0, x1, dy, dtable);
}

static TableBwdLayout table_layout(const PtyGrid& g, const PtyPoints& pts) {
  TableBwdLayout L{};
  L.n_dense = 0;
  if (!pts.coords) {
    int s = 0;
    int64_t a = 0, c = 0;
    for (int l = 0; l < g.n_levels; ++l) {
      if (!((g.dense_mask >> l) & 1u)) continue;
      const int d = L.n_dense++;
      L.lev[d] = l;
      L.off_s[d] = s;
      L.off_a[d] = a;
      L.off_c[d] = c;
      s += (int)g.res[l] + 2;
      a += (int64_t)pts.n_axis1 * (g.res[l] + 1);
      c += (int64_t)(g.res[l] + 1) * (g.res[l] + 1);
    }
    L.n_starts = s;
    L.off_a[L.n_dense] = a;
    L.off_c[L.n_dense] = c;
  }
  return L;
}

size_t table_bwd_workspace(const PtyGrid& g, const PtyPoints& pts) {
  const TableBwdLayout L = table_layout(g, pts);
  if (L.n_dense == 0) return 0;
  return sizeof(int) * 2 * (size_t)L.n_starts + 256 + sizeof(float2) * (size_t)L.off_a[L.n_dense];
}

int launch_table_bwd(const PtyGrid& g, const PtyPoints& pts, const float2* gfeat, float* dtable, char* work,
                     hipStream_t stream) {
  const TableBwdLayout L = table_layout(g, pts);
  uint32_t scatter_mask = (g.n_levels >= 32) ? 0xFFFFFFFFu : ((1u << g.n_levels) - 1u);
  if (L.n_dense > 0) {
    int* s0 = (int*)work;
    int* s1 = s0 + L.n_starts;
    float2* A = (float2*)(work + (((sizeof(int) * 2 * (size_t)L.n_starts) + 255) & ~size_t(255)));
    hipLaunchKernelGGL(k_axis_starts, dim3((L.n_starts + 255) / 256), dim3(256), 0, stream, g, L, pts.axis0,
                       pts.n_axis0, s0);
    hipLaunchKernelGGL(k_axis_starts, dim3((L.n_starts + 255) / 256), dim3(256), 0, stream, g, L, pts.axis1,
                       pts.n_axis1, s1);
    const int64_t na = L.off_a[L.n_dense], nc = L.off_c[L.n_dense];
    hipLaunchKernelGGL(k_dense_pass1, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, stream, g, L, pts, gfeat,
                       (const int*)s0, A);
    hipLaunchKernelGGL(k_dense_pass2, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, stream, g, L, pts,
                       (const int*)s1, (const float2*)A, dtable);
    for (int d = 0; d < L.n_dense; ++d) scatter_mask &= ~(1u << L.lev[d]);
  }
  if (scatter_mask) {
    hipLaunchKernelGGL(k_level_scatter, dim3((unsigned)((pts.n + 255) / 256), g.n_levels), dim3(256), 0, stream, g,
                       scatter_mask, pts, gfeat, dtable);
  }
  return check_launch("table_bwd");
}

}  // namespace pty
