// Hash-table gradient from per-point feature gradients gfeat[l][p] (float2,
// level-major): the adjoint of the encoding in the fused INR backward.
//
// Explicit coordinates: every level scatters w_c * gfeat into its 4 corner
// entries with atomics.
//
// Regular grid (init_scale_grids): owner-computes, no atomics, deterministic.
// With gx(j), fx(j) the cell / fraction of axis0[j] and gy(i), fy(i) those of
// axis1[i] at level l, the gradient of virtual corner (cx, cy), cx, cy in
// [0, res], is
//   G(cx,cy) = sum_{dx,dy} sum_{j: gx(j)=cx-dx} sum_{i: gy(i)=cy-dy} w(fx,dx) w(fy,dy) g[l][i*n0+j]
// with w(f,0) = 1-f, w(f,1) = f (the bilinear weights are separable).  The
// cells are monotone in the axis index, so each cell's points are a
// contiguous index range (plan: starts[c] = first index with cell >= c).
// One thread owns each table entry e and sums the virtual corners mapping to
// it:
//   dense level : local = cx + cy*res (minus size once past the end), so the
//                 preimages are cx + cy*res in {e, e+size}: at most 4 corners;
//                 the levels first contract along axis0 (A[i][cx], pass 1)
//                 so that the coarse levels (many points per cell) stay cheap;
//   hashed level: local = (cx ^ h(cy)) & (T-1), h(cy) = cy*2654435761 mod 2^32,
//                 so for each cy the corner is cx = (e ^ h(cy)) & (T-1) (+ k T).
//                 Only rows cy touched by a grid point can contribute; the
//                 plan buckets them by the bits of h(cy) above log2(res+1),
//                 which must equal those of e for cx <= res.  Since the grid is
//                 static, the plan build runs this enumeration once and stores
//                 every entry's (point, corner) terms as a CSR list (exactly
//                 4 terms per point and level); each step then gathers them.
#include "hashgrid.h"

#include <algorithm>

namespace pty {

constexpr uint32_t kPrime = 2654435761u;

// Plan layout, a pure function of (grid, axes), shared by build and use.
struct PlanLayout {
  int n0, n1;
  int off_s[PTY_MAX_LEVELS];    // starts tables: res + 2 ints per level (axis0 and axis1 alike)
  int n_starts;
  int bbits[PTY_MAX_LEVELS];    // hashed: bits of cx (2^bbits > res), dense: 0
  int off_b[PTY_MAX_LEVELS];    // hashed: bucket starts (nb + 1 ints)
  int off_cy[PTY_MAX_LEVELS];   // hashed: bucketed rows (res + 1 ints capacity)
  int off_m[PTY_MAX_LEVELS];    // hashed: touched-column bitmask, (res + 32) / 32 words
  int64_t off_co[PTY_MAX_LEVELS];   // hashed: CSR offsets (T + 1 ints)
  int64_t off_cp[PTY_MAX_LEVELS];   // hashed: CSR terms (4 n ints)
  int n_bucket_ints;
  int n_cy_ints;
  int n_mask_words;
  int64_t n_co, n_cp;
  // byte offsets
  size_t s0, s1, fx, fy, bstart, bcy, mask, csr_o, csr_p, total;
  // dense pass-1 buffer (workspace, not plan): A[l] of n1 * (res + 1) float2
  int64_t off_a[PTY_MAX_LEVELS + 1];
};

static size_t a256(size_t b) { return (b + 255) & ~size_t(255); }

static int log2_ceil(uint32_t v) {
  int b = 0;
  while ((1u << b) < v && b < 31) ++b;
  return b;
}

static PlanLayout plan_layout(const PtyGrid& g, const PtyPoints& pts) {
  PlanLayout L{};
  L.n0 = pts.n_axis0;
  L.n1 = pts.n_axis1;
  int s = 0, bo = 0, co = 0, mo = 0;
  int64_t a = 0, ro = 0, rp = 0;
  const int64_t npts = (int64_t)L.n0 * L.n1;
  for (int l = 0; l < g.n_levels; ++l) {
    L.off_s[l] = s;
    s += (int)g.res[l] + 2;
    const bool dense = (g.dense_mask >> l) & 1u;
    L.off_a[l] = a;
    if (dense) {
      L.bbits[l] = 0;
      L.off_b[l] = bo;
      L.off_cy[l] = co;
      L.off_m[l] = mo;
      L.off_co[l] = ro;
      L.off_cp[l] = rp;
      a += (int64_t)L.n1 * (g.res[l] + 1);
    } else {
      const int log2T = g.log2_hashmap_size;
      int b = log2_ceil(g.res[l] + 1);
      if (b > log2T) b = log2T;
      L.bbits[l] = b;
      const int nb = 1 << (log2T - b);
      L.off_b[l] = bo;
      bo += nb + 1;
      L.off_cy[l] = co;
      co += (int)g.res[l] + 1;
      L.off_m[l] = mo;
      mo += ((int)g.res[l] + 32) / 32;
      L.off_co[l] = ro;
      ro += (int64_t)g.size[l] + 1;
      L.off_cp[l] = rp;
      rp += 4 * npts;
    }
  }
  L.n_co = ro;
  L.n_cp = rp;
  L.off_a[g.n_levels] = a;
  L.n_starts = s;
  L.n_bucket_ints = bo;
  L.n_cy_ints = co;
  L.n_mask_words = mo;
  size_t off = 0;
  L.s0 = off; off += a256(sizeof(int) * (size_t)s);
  L.s1 = off; off += a256(sizeof(int) * (size_t)s);
  L.fx = off; off += a256(sizeof(float) * (size_t)g.n_levels * L.n0);
  L.fy = off; off += a256(sizeof(float) * (size_t)g.n_levels * L.n1);
  L.bstart = off; off += a256(sizeof(int) * (size_t)(bo > 0 ? bo : 1));
  L.bcy = off; off += a256(sizeof(int) * (size_t)(co > 0 ? co : 1));
  L.mask = off; off += a256(sizeof(uint32_t) * (size_t)(mo > 0 ? mo : 1));
  L.csr_o = off; off += a256(sizeof(int) * (size_t)(ro > 0 ? ro : 1));
  L.csr_p = off; off += a256(sizeof(uint32_t) * (size_t)(rp > 0 ? rp : 1));
  L.total = off;
  return L;
}

struct PlanPtrs {
  const int* s0;
  const int* s1;
  const float* fx;
  const float* fy;
  const int* bstart;
  const int* bcy;
  const uint32_t* mask;
  const int* csr_o;
  const uint32_t* csr_p;
};

static PlanPtrs plan_ptrs(const PlanLayout& L, const void* plan) {
  const char* p = (const char*)plan;
  PlanPtrs P;
  P.s0 = (const int*)(p + L.s0);
  P.s1 = (const int*)(p + L.s1);
  P.fx = (const float*)(p + L.fx);
  P.fy = (const float*)(p + L.fy);
  P.bstart = (const int*)(p + L.bstart);
  P.bcy = (const int*)(p + L.bcy);
  P.mask = (const uint32_t*)(p + L.mask);
  P.csr_o = (const int*)(p + L.csr_o);
  P.csr_p = (const uint32_t*)(p + L.csr_p);
  return P;
}

__device__ __forceinline__ uint32_t axis_cell(const PtyGrid& g, int l, float x, float& frac) {
  const float p = __builtin_fmaf(x, g.scale[l], 0.5f);   // one rounding, as level_pos
  const float f = floorf(p);
  frac = p - f;
  return (uint32_t)(int)f;
}

__device__ __forceinline__ int level_of_starts(const PlanLayout& L, int n_levels, int t) {
  int l = 0;
  while (l + 1 < n_levels && t >= L.off_s[l + 1]) ++l;
  return l;
}

// ------------------------------------------------------------------ plan build
// starts (both axes) and fractions (both axes)
__global__ void k_plan_axes(PtyGrid g, PlanLayout L, const float* __restrict__ a0, const float* __restrict__ a1,
                            int* __restrict__ s0, int* __restrict__ s1, float* __restrict__ fx,
                            float* __restrict__ fy) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < L.n_starts) {
    const int l = level_of_starts(L, g.n_levels, t);
    const int c = t - L.off_s[l];
    for (int ax = 0; ax < 2; ++ax) {
      const float* a = ax ? a1 : a0;
      const int n = ax ? L.n1 : L.n0;
      int lo = 0, hi = n;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        float fr;
        if ((int)axis_cell(g, l, a[mid], fr) >= c) hi = mid; else lo = mid + 1;
      }
      (ax ? s1 : s0)[t] = lo;
    }
  }
  const int u = t;
  if (u < g.n_levels * L.n0) {
    const int l = u / L.n0, j = u - l * L.n0;
    float fr;
    axis_cell(g, l, a0[j], fr);
    fx[u] = fr;
  }
  if (u < g.n_levels * L.n1) {
    const int l = u / L.n1, i = u - l * L.n1;
    float fr;
    axis_cell(g, l, a1[i], fr);
    fy[u] = fr;
  }
}

// bucketed rows of each hashed level: rows cy in [0, res] with a point in
// cell cy-1 or cy, grouped by ((h(cy) & (T-1)) >> bbits), ascending cy.
// One workgroup per level; counting in LDS-free sequential passes by lane 0
// (a one-time build).
__global__ void k_plan_buckets(PtyGrid g, PlanLayout L, const int* __restrict__ s0, const int* __restrict__ s1,
                               int* __restrict__ bstart, int* __restrict__ bcy, uint32_t* __restrict__ mask) {
  const int l = blockIdx.x;
  if (((g.dense_mask >> l) & 1u) || threadIdx.x != 0) return;
  const int res = (int)g.res[l];
  const uint32_t T = 1u << g.log2_hashmap_size;
  const int b = L.bbits[l];
  const int nb = 1 << (g.log2_hashmap_size - b);
  const int* st = s1 + L.off_s[l];
  int* bs = bstart + L.off_b[l];
  int* out = bcy + L.off_cy[l];
  for (int k = 0; k <= nb; ++k) bs[k] = 0;
  for (int cy = 0; cy <= res; ++cy) {
    const bool touched = (st[cy + 1] > st[cy]) || (cy > 0 && st[cy] > st[cy - 1]);
    if (!touched) continue;
    const uint32_t key = ((cy * kPrime) & (T - 1u)) >> b;
    bs[key + 1] += 1;
  }
  for (int k = 0; k < nb; ++k) bs[k + 1] += bs[k];
  // fill (bs[k] used as a cursor, then restored)
  for (int cy = 0; cy <= res; ++cy) {
    const bool touched = (st[cy + 1] > st[cy]) || (cy > 0 && st[cy] > st[cy - 1]);
    if (!touched) continue;
    const uint32_t key = ((cy * kPrime) & (T - 1u)) >> b;
    out[bs[key]++] = cy;
  }
  for (int k = nb; k > 0; --k) bs[k] = bs[k - 1];
  bs[0] = 0;
  // column cx is touched when axis0 cell cx-1 or cx holds a point
  const int* s0l = s0 + L.off_s[l];
  uint32_t* m = mask + L.off_m[l];
  const int words = (res + 32) / 32;
  for (int w = 0; w < words; ++w) {
    uint32_t bits = 0;
    for (int k = 0; k < 32; ++k) {
      const int cx = w * 32 + k;
      if (cx > res) break;
      const bool touched = (s0l[cx + 1] > s0l[cx]) || (cx > 0 && s0l[cx] > s0l[cx - 1]);
      if (touched) bits |= 1u << k;
    }
    m[w] = bits;
  }
}

// ------------------------------------------------------------------ gradient
// Dense pass 1: A[l][i][cx] = S0(cx) + S1(cx - 1), with the cell sums
// S0(c) = sum_{j in cell c} (1 - fx_j) g[l][i n0 + j] and S1(c) = sum fx_j g:
// each thread reads its own cell's points once (the cells of a row are
// consecutive index ranges, so a workgroup's reads are contiguous) and takes
// the left neighbour's S1 from LDS; the first thread of a workgroup forms it.
__device__ __forceinline__ void cell_sums(const int* st, const float* fxl, const float2* gl, int cell, float2& s0,
                                          float2& s1) {
  s0 = s1 = make_float2(0.f, 0.f);
  if (cell < 0) return;
  const int j1 = st[cell + 1];
  // unrolled so that the loads of coarse cells (tens of points) overlap
#pragma unroll 8
  for (int j = st[cell]; j < j1; ++j) {
    const float f = fxl[j];
    const float2 v = gl[j];
    s0.x += (1.f - f) * v.x;
    s0.y += (1.f - f) * v.y;
    s1.x += f * v.x;
    s1.y += f * v.y;
  }
}

__global__ __launch_bounds__(256) void k_dense_pass1(PtyGrid g, PlanLayout L, PtyPoints pts, PlanPtrs P,
                                                     const float2* __restrict__ gfeat, float2* __restrict__ A) {
  __shared__ float2 s1_lds[256];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = t < L.off_a[g.n_levels];
  int l = 0, i = 0, cx = 0;
  float2 s0 = make_float2(0.f, 0.f), s1 = make_float2(0.f, 0.f);
  const int* st = nullptr;
  const float* fxl = nullptr;
  const float2* gl = nullptr;
  if (active) {
    while (l + 1 < g.n_levels && t >= L.off_a[l + 1]) ++l;
    const int res1 = (int)g.res[l] + 1;
    const int64_t r = t - L.off_a[l];
    i = (int)(r / res1);
    cx = (int)(r - (int64_t)i * res1);
    st = P.s0 + L.off_s[l];
    fxl = P.fx + (int64_t)l * L.n0;
    gl = gfeat + (int64_t)l * pts.n + (int64_t)i * L.n0;
    cell_sums(st, fxl, gl, cx, s0, s1);
  }
  s1_lds[threadIdx.x] = s1;
  __syncthreads();
  if (!active) return;
  float2 left = make_float2(0.f, 0.f);
  if (cx > 0) {
    if (threadIdx.x > 0) {
      left = s1_lds[threadIdx.x - 1];   // thread t-1 holds cell cx-1 of the same row
    } else {
      float2 d0;
      cell_sums(st, fxl, gl, cx - 1, d0, left);
    }
  }
  A[t] = make_float2(s0.x + left.x, s0.y + left.y);
}

// Row form of pass 1: one workgroup per (dense level l, row i).  The row's
// feature gradients and x fractions and the level's cell bounds are staged in
// LDS (each thread issues all of its loads of a batch before its first LDS
// store, so the staging is one round trip to memory, not one per element),
// then each thread sums its cells' S0 / S1 from LDS in point order (the same
// sums as cell_sums; cells c = tid + 256 k, k < KC) and A[l][i][cx] =
// S0(cx) + S1(cx - 1) is written, S1 passed through LDS.  LDS:
// max(n0, res + 1) * 8 B + n0 * 4 B + (res + 2) * 4 B.
template <int KC>
__global__ __launch_bounds__(256) void k_dense_pass1_rows(PtyGrid g, PlanLayout L, PtyPoints pts, PlanPtrs P,
                                                          const float2* __restrict__ gfeat, float2* __restrict__ A) {
  extern __shared__ float2 srow[];
  const int n0 = L.n0, tid = threadIdx.x;
  int b = blockIdx.x, l = -1;
  for (int q = 0; q < g.n_levels; ++q) {
    if (!((g.dense_mask >> q) & 1u)) continue;
    if (b < L.n1) { l = q; break; }
    b -= L.n1;
  }
  if (l < 0) return;   // uniform across the workgroup
  const int i = b;
  const int res1 = (int)g.res[l] + 1;
  float2* sg = srow;                                          // [n0] row gradients, later [res1] S1
  float* sf = reinterpret_cast<float*>(srow + max(n0, res1));  // [n0] x fractions
  int* sst = reinterpret_cast<int*>(sf + n0);                  // [res1 + 1] first point of each cell
  const float2* gl = gfeat + (int64_t)l * pts.n + (int64_t)i * n0;
  const float* fxl = P.fx + (int64_t)l * n0;
  const int* st = P.s0 + L.off_s[l];
  constexpr int kB = 8;   // row elements per thread and batch
  // the first batch and the cell bounds in one round trip (clamped loads,
  // guarded stores), then any further batches
  {
    float2 vg[kB];
    float vf[kB];
    int vs[KC + 1];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int t = min(u * 256 + tid, n0 - 1);
      vg[u] = gl[t];
      vf[u] = fxl[t];
    }
#pragma unroll
    for (int u = 0; u <= KC; ++u) vs[u] = st[min(u * 256 + tid, res1)];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int t = u * 256 + tid;
      if (t < n0) {
        sg[t] = vg[u];
        sf[t] = vf[u];
      }
    }
#pragma unroll
    for (int u = 0; u <= KC; ++u) {
      const int t = u * 256 + tid;
      if (t <= res1) sst[t] = vs[u];
    }
  }
  for (int t0 = kB * 256; t0 < n0; t0 += kB * 256) {
    float2 vg[kB];
    float vf[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int t = min(t0 + u * 256 + tid, n0 - 1);
      vg[u] = gl[t];
      vf[u] = fxl[t];
    }
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int t = t0 + u * 256 + tid;
      if (t < n0) {
        sg[t] = vg[u];
        sf[t] = vf[u];
      }
    }
  }
  __syncthreads();
  float2 s0r[KC], s1r[KC];
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const int c = tid + 256 * k;
    float2 s0 = make_float2(0.f, 0.f), s1 = make_float2(0.f, 0.f);
    if (c < res1) {
      const int j1 = sst[c + 1];
#pragma unroll 4
      for (int j = sst[c]; j < j1; ++j) {
        const float f = sf[j];
        const float2 v = sg[j];
        s0.x += (1.f - f) * v.x;
        s0.y += (1.f - f) * v.y;
        s1.x += f * v.x;
        s1.y += f * v.y;
      }
    }
    s0r[k] = s0;
    s1r[k] = s1;
  }
  __syncthreads();   // every thread is done reading the row
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const int c = tid + 256 * k;
    if (c < res1) sg[c] = s1r[k];
  }
  __syncthreads();
  float2* Ar = A + L.off_a[l] + (int64_t)i * res1;
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const int c = tid + 256 * k;
    if (c < res1) {
      const float2 left = c > 0 ? sg[c - 1] : make_float2(0.f, 0.f);
      Ar[c] = make_float2(s0r[k].x + left.x, s0r[k].y + left.y);
    }
  }
}

// G(cx, cy) from A (dense levels), or the part of it over the rows i with
// i - st[cell] = s mod S (lane s of an entry split over S lanes): the rows of
// cell cy (weight 1 - fy) and of cell cy - 1 (weight fy), each summed in row
// order and the two sums added.  The three cell bounds are loaded together,
// then the rows of both cells in predicated batches of kRowBatch, so a
// typical entry (a few rows per cell) waits on memory twice, not once per row.
constexpr int kRowBatch = 4;

__device__ __forceinline__ float2 corner_from_A(const PlanLayout& L, const PlanPtrs& P, int l, int res,
                                                const float2* __restrict__ Al, int cx, int cy, int s, int S) {
  const int* st = P.s1 + L.off_s[l];
  const float* fyl = P.fy + (int64_t)l * L.n1;
  const int64_t stride = res + 1;
  // cy in [0, res]: cell cy is rows [b0, b1), cell cy - 1 rows [a0, b0) (empty at cy = 0)
  const int b0 = st[cy], b1 = st[cy + 1];
  const int sa = st[max(cy - 1, 0)];
  const int a0 = cy > 0 ? sa : b0;
  float2 sum0 = make_float2(0.f, 0.f), sum1 = make_float2(0.f, 0.f);
  for (int i = b0 + s, j = a0 + s; i < b1 || j < b0; i += kRowBatch * S, j += kRowBatch * S) {
    float2 v0[kRowBatch], v1[kRowBatch];
    float f0[kRowBatch], f1[kRowBatch];
#pragma unroll
    for (int u = 0; u < kRowBatch; ++u) {
      const int ii = i + u * S, jj = j + u * S;
      const int ic = ii < b1 ? ii : 0, jc = jj < b0 ? jj : 0;   // row 0 stands in for a missing row
      v0[u] = Al[ic * stride + cx];
      f0[u] = fyl[ic];
      v1[u] = Al[jc * stride + cx];
      f1[u] = fyl[jc];
    }
#pragma unroll
    for (int u = 0; u < kRowBatch; ++u) {
      if (i + u * S < b1) {
        const float w = 1.f - f0[u];
        sum0.x += w * v0[u].x;
        sum0.y += w * v0[u].y;
      }
      if (j + u * S < b0) {
        sum1.x += f1[u] * v1[u].x;
        sum1.y += f1[u] * v1[u].y;
      }
    }
  }
  return make_float2(sum0.x + sum1.x, sum0.y + sum1.y);
}

// Dense levels: one thread per table entry (entries of hashed levels are
// skipped here and handled by k_table_hashed).  The coarse levels, whose cells
// hold many rows of points, come first and get kDenseSplit lanes per entry:
// lane s sums the rows s, s + kDenseSplit, ... of every corner and the lanes
// of an entry add up by a fixed butterfly, so a coarse entry's chain of
// dependent loads is kDenseSplit times shorter and the result deterministic.
constexpr int kDenseSplit = 8;

__global__ __launch_bounds__(256) void k_table_dense(PtyGrid g, PlanLayout L, PlanPtrs P,
                                                     const float2* __restrict__ A, uint32_t n_dense_entries,
                                                     uint32_t n_split_entries, float2* __restrict__ dtable,
                                                     TableAdam ad) {
  const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n_split_lanes = n_split_entries * kDenseSplit;
  const bool split = gt < n_split_lanes;
  const uint32_t t = split ? gt / kDenseSplit : n_split_entries + (gt - n_split_lanes);
  const int s = split ? (int)(gt % kDenseSplit) : 0;
  const int S = split ? kDenseSplit : 1;
  const bool active = t < n_dense_entries;
  const bool owner = active && s == 0;
  // Adam's step coefficients are uniform: formed once per workgroup
  __shared__ AdamCoef s_coef;
  if (ad.p && threadIdx.x == 0) s_coef = adam_coef(ad.hyper, ad.b1, ad.b2, ad.eps);
  EntryAdam st{};
  if (ad.p && owner) st = table_adam_load(ad, t);
  // level: the wave's first entry's by a uniform search, then one step for
  // the lanes past a level boundary (a wave spans at most 64 entries, every
  // dense level holds more)
  const uint32_t t_first = __builtin_amdgcn_readfirstlane(min(t, n_dense_entries - 1));
  int l0 = 0;
  while (l0 + 1 < g.n_levels && t_first >= g.offset[l0 + 1]) ++l0;
  const uint32_t next_off = l0 + 1 < g.n_levels ? g.offset[l0 + 1] : 0xFFFFFFFFu;
  float2 acc = make_float2(0.f, 0.f);
  if (active) {
    const int l = t >= next_off ? l0 + 1 : l0;
    const uint32_t e = t - g.offset[l];
    const int res = (int)g.res[l];
    const uint32_t size = g.size[l];
    const float2* Al = A + L.off_a[l];
    // 32-bit index math: dense levels have res^2 <= T <= 2^28, so every
    // target below is under 2^29 (64-bit divisions cost several times more)
    const uint32_t ures = (uint32_t)res;
    const uint32_t top = ures * ures + ures;  // largest cx + cy*res
    for (int k = 0; k < 2; ++k) {
      const uint32_t target = e + (uint32_t)k * size;
      if (target > top) break;
      // cy with cx = target - cy*res in [0, res]
      const uint32_t cy_lo = target > ures ? (target - 1u) / ures : 0u;
      const uint32_t cy_hi = target / ures;
      for (uint32_t cy = cy_lo; cy <= cy_hi && cy <= ures; ++cy) {
        const float2 v = corner_from_A(L, P, l, res, Al, (int)(target - cy * ures), (int)cy, s, S);
        acc.x += v.x;
        acc.y += v.y;
      }
    }
  }
  // the butterfly runs with every lane of the wave converged; the groups of
  // kDenseSplit lanes are aligned, so split and unsplit lanes never mix
  float2 sum = acc;
#pragma unroll
  for (int d = 1; d < kDenseSplit; d <<= 1) {
    sum.x += __shfl_xor(sum.x, d);
    sum.y += __shfl_xor(sum.y, d);
  }
  __syncthreads();   // s_coef
  if (!owner) return;
  if (split) acc = sum;
  if (ad.p) {
    table_adam_store(ad, s_coef, t, st, acc);
  } else if (acc.x != 0.f || acc.y != 0.f) {
    float2 d = dtable[t];
    d.x += acc.x;
    d.y += acc.y;
    dtable[t] = d;
  }
}

// ------------------------------------------------------------------ hashed levels
// blockIdx.x -> (level, chunk of `per` entries) over the hashed levels in
// order; false when past the last one.
__device__ __forceinline__ bool hashed_chunk(const PtyGrid& g, uint32_t bid, int& l, uint32_t& chunk,
                                             uint32_t per = 256) {
  chunk = bid;
  for (l = 0; l < g.n_levels; ++l) {
    if ((g.dense_mask >> l) & 1u) continue;
    const uint32_t nch = (g.size[l] + per - 1) / per;
    if (chunk < nch) return true;
    chunk -= nch;
  }
  return false;
}

// Term of one grid point at one corner: (i, j) point, (dx, dy) corner offset.
__device__ __forceinline__ uint32_t pack_term(int i, int j, int dx, int dy) {
  return ((uint32_t)i << 17) | ((uint32_t)j << 2) | ((uint32_t)dy << 1) | (uint32_t)dx;
}

// Every (point, corner) term of hashed-level entry e, in a fixed order:
// bucket rows ascending, candidate columns ascending, dx = 0 then 1, points j
// ascending, then the points i of cell cy-1 (dy = 1) and of cell cy (dy = 0).
template <class Visit>
__device__ void hashed_terms(const PtyGrid& g, const PlanLayout& L, const PlanPtrs& P, int l, uint32_t e,
                             Visit&& visit) {
  const int res = (int)g.res[l];
  const uint32_t T = 1u << g.log2_hashmap_size;
  const int* s0 = P.s0 + L.off_s[l];
  const int* s1 = P.s1 + L.off_s[l];
  const int* bs = P.bstart + L.off_b[l];
  const int* rows = P.bcy + L.off_cy[l];
  const uint32_t* mk = P.mask + L.off_m[l];
  const uint32_t key = e >> L.bbits[l];
  const int r1 = bs[key + 1];
  for (int r = bs[key]; r < r1; ++r) {
    const int cy = rows[r];
    const int i0a = cy > 0 ? s1[cy - 1] : 0, i1a = cy > 0 ? s1[cy] : 0;
    const int i0b = s1[cy], i1b = s1[cy + 1];
    for (uint32_t cx = (e ^ ((uint32_t)cy * kPrime)) & (T - 1u); cx <= (uint32_t)res; cx += T) {
      if (!((mk[cx >> 5] >> (cx & 31)) & 1u)) continue;
      for (int dx = 0; dx < 2; ++dx) {
        const int cxx = (int)cx - dx;
        if (cxx < 0 || cxx > res) continue;
        const int j1 = s0[cxx + 1];
        for (int j = s0[cxx]; j < j1; ++j) visit(j, dx, i0a, i1a, i0b, i1b);
      }
    }
  }
}

// plan build, pass 1: csr_o[l][e + 1] = number of terms of entry e
__global__ __launch_bounds__(256) void k_csr_count(PtyGrid g, PlanLayout L, PlanPtrs P, int* __restrict__ csr_o) {
  int l;
  uint32_t chunk;
  if (!hashed_chunk(g, blockIdx.x, l, chunk)) return;
  const uint32_t e = chunk * 256 + threadIdx.x;
  int* off = csr_o + L.off_co[l];
  if (e == 0) off[0] = 0;
  if (e >= g.size[l]) return;
  int cnt = 0;
  hashed_terms(g, L, P, l, e, [&](int, int, int i0a, int i1a, int i0b, int i1b) {
    cnt += (i1a - i0a) + (i1b - i0b);
  });
  off[e + 1] = cnt;
}

// plan build, pass 2: inclusive scan of csr_o[l][1..T] (one workgroup per
// hashed level, 1024 threads, tiles of 1024 with a carried total)
__global__ __launch_bounds__(1024) void k_csr_scan(PtyGrid g, PlanLayout L, int* __restrict__ csr_o) {
  __shared__ int s_tmp[1024];
  __shared__ int s_carry;
  int l = -1;
  for (int k = 0, h = 0; k < g.n_levels; ++k) {
    if ((g.dense_mask >> k) & 1u) continue;
    if (h++ == (int)blockIdx.x) { l = k; break; }
  }
  if (l < 0) return;
  int* off = csr_o + L.off_co[l] + 1;
  const int n = (int)g.size[l];
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    const int i = base + (int)threadIdx.x;
    int v = i < n ? off[i] : 0;
    s_tmp[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      const int add = threadIdx.x >= (unsigned)d ? s_tmp[threadIdx.x - d] : 0;
      __syncthreads();
      s_tmp[threadIdx.x] += add;
      __syncthreads();
    }
    const int carry = s_carry;
    if (i < n) off[i] = s_tmp[threadIdx.x] + carry;
    __syncthreads();
    if (threadIdx.x == 1023) s_carry = carry + s_tmp[1023];
    __syncthreads();
  }
}

// plan build, pass 3: the terms themselves, in the order of hashed_terms
__global__ __launch_bounds__(256) void k_csr_fill(PtyGrid g, PlanLayout L, PlanPtrs P, uint32_t* __restrict__ csr_p) {
  int l;
  uint32_t chunk;
  if (!hashed_chunk(g, blockIdx.x, l, chunk)) return;
  const uint32_t e = chunk * 256 + threadIdx.x;
  if (e >= g.size[l]) return;
  // every point has exactly 4 terms per level; the guard keeps a miscount
  // inside this level's region
  const int64_t cap = 4 * (int64_t)L.n0 * L.n1;
  uint32_t* base = csr_p + L.off_cp[l];
  int64_t pos = P.csr_o[L.off_co[l] + e];
  hashed_terms(g, L, P, l, e, [&](int j, int dx, int i0a, int i1a, int i0b, int i1b) {
    for (int i = i0a; i < i1a; ++i, ++pos)
      if (pos < cap) base[pos] = pack_term(i, j, dx, 1);
    for (int i = i0b; i < i1b; ++i, ++pos)
      if (pos < cap) base[pos] = pack_term(i, j, dx, 0);
  });
}

// Hashed-level table gradient over a chunk of 256 EPT consecutive entries of
// one level, EPT adjacent entries per thread.  The kernel is a chain of
// dependent memory round trips per workgroup (offsets, terms, gathers, Adam
// state), so with few terms per entry the number of workgroup rounds sets its
// time and two entries per thread (half the workgroups; the pair's Adam state
// as 16-byte loads and stores) is faster; with many terms per entry the
// entry-serial sums dominate and one entry per thread is (the launcher picks
// EPT from the level's mean terms per entry).  The chunk's CSR terms are one
// contiguous range; the workgroup walks it in windows of kHashWin terms:
// every thread forms the product w(fx_j, dx) * w(fy_i, dy) * g[l][i n0 + j]
// of one term (coalesced term loads, neighbouring entries read neighbouring
// cells of the same rows) into LDS, then each entry's terms are added from
// LDS in plan order by the entry's thread.  The sum order per entry is fixed,
// so the result is deterministic and the same for either EPT.
constexpr int kHashWin = 2048;
constexpr int kHashPer = kHashWin / 256;   // terms per thread and window

typedef float nt_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 nt_load4(const float2* q) {
  const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(q));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void nt_store4(float2* q, float4 v) {
  __builtin_nontemporal_store(nt_f4{v.x, v.y, v.z, v.w}, reinterpret_cast<nt_f4*>(q));
}

template <int EPT>
__global__ __launch_bounds__(256) void k_table_hashed(PtyGrid g, PlanLayout L, PtyPoints pts, PlanPtrs P,
                                                      const float2* __restrict__ gfeat,
                                                      float2* __restrict__ dtable, TableAdam ad) {
  static_assert(EPT == 1 || EPT == 2, "one or two entries per thread");
  constexpr int CH = 256 * EPT;
  __shared__ int s_off[CH + 1];
  __shared__ float2 s_val[kHashWin];
  __shared__ AdamCoef s_coef;   // Adam's step coefficients, formed once per workgroup
  int l;
  uint32_t chunk;
  if (!hashed_chunk(g, blockIdx.x, l, chunk, CH)) return;   // uniform across the workgroup
  if (ad.p && threadIdx.x == 0) s_coef = adam_coef(ad.hyper, ad.b1, ad.b2, ad.eps);
  const int tid = threadIdx.x;
  const uint32_t e0 = chunk * CH;
  // a hashed level's size is a power of two >= 8, so ne is a multiple of EPT
  const int ne = (int)min((uint32_t)CH, g.size[l] - e0);
  const bool mine = EPT * tid < ne;
  // even for EPT 2: level offsets and sizes are multiples of 8
  const uint32_t t0 = g.offset[l] + e0 + EPT * tid;
  const int* off = P.csr_o + L.off_co[l] + e0;
  const uint32_t* terms = P.csr_p + L.off_cp[l];
  const float* fxl = P.fx + (int64_t)l * L.n0;
  const float* fyl = P.fy + (int64_t)l * L.n1;
  const float2* gl = gfeat + (int64_t)l * pts.n;
  const int cap = 4 * L.n0 * L.n1;
  float2 sp[EPT], sm[EPT], sv[EPT];
  if (ad.p && mine) {
    if constexpr (EPT == 2) {
      const float4 p4 = nt_load4(ad.p + t0), m4 = nt_load4(ad.m + t0), v4 = nt_load4(ad.v + t0);
      sp[0] = make_float2(p4.x, p4.y); sp[1] = make_float2(p4.z, p4.w);
      sm[0] = make_float2(m4.x, m4.y); sm[1] = make_float2(m4.z, m4.w);
      sv[0] = make_float2(v4.x, v4.y); sv[1] = make_float2(v4.z, v4.w);
    } else {
      sp[0] = nt_load(ad.p + t0);
      sm[0] = nt_load(ad.m + t0);
      sv[0] = nt_load(ad.v + t0);
    }
  }
  for (int t = tid; t <= ne; t += 256) s_off[t] = min(off[t], cap);
  __syncthreads();
  const int k0 = s_off[0], k1 = s_off[ne];
  int r[EPT + 1];   // the thread's entries' term ranges [r[q], r[q + 1])
#pragma unroll
  for (int q = 0; q <= EPT; ++q) r[q] = mine ? s_off[EPT * tid + q] : 0;
  float2 acc[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) acc[q] = make_float2(0.f, 0.f);
  for (int w0 = k0; w0 < k1; w0 += kHashWin) {
    const int wn = min(kHashWin, k1 - w0);
    const int nu = (wn + 255) / 256;   // term slots in use, uniform across the workgroup
    // all of a thread's term loads first, then all of their gathers, so that
    // each phase is one round trip to memory instead of one per term; a slot
    // past the window's end reads its last term again and stores nothing
    uint32_t c[kHashPer];
#pragma unroll
    for (int u = 0; u < kHashPer; ++u)
      if (u < nu) c[u] = terms[w0 + min(tid + 256 * u, wn - 1)];
    float wt[kHashPer];
    float2 v[kHashPer];
#pragma unroll
    for (int u = 0; u < kHashPer; ++u)
      if (u < nu) {
        const int i = (int)(c[u] >> 17), j = (int)((c[u] >> 2) & 0x7FFFu);
        const float fx = fxl[j], fy = fyl[i];
        const float wx = (c[u] & 1u) ? fx : 1.f - fx;
        wt[u] = wx * ((c[u] & 2u) ? fy : 1.f - fy);
        v[u] = gl[(int64_t)i * L.n0 + j];
      }
#pragma unroll
    for (int u = 0; u < kHashPer; ++u) {
      const int t = tid + 256 * u;
      if (u < nu && t < wn) s_val[t] = make_float2(wt[u] * v[u].x, wt[u] * v[u].y);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < EPT; ++q)
      for (int k = max(r[q], w0), kend = min(r[q + 1], w0 + wn); k < kend; ++k) {
        const float2 x = s_val[k - w0];
        acc[q].x += x.x;
        acc[q].y += x.y;
      }
    __syncthreads();
  }
  if (!mine) return;
  if (ad.p) {
    const AdamCoef co = s_coef;
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      adam_elem(sp[q].x, sm[q].x, sv[q].x, acc[q].x, co);
      adam_elem(sp[q].y, sm[q].y, sv[q].y, acc[q].y, co);
    }
    if constexpr (EPT == 2) {
      *reinterpret_cast<float4*>(ad.p + t0) = make_float4(sp[0].x, sp[0].y, sp[1].x, sp[1].y);
      nt_store4(ad.m + t0, make_float4(sm[0].x, sm[0].y, sm[1].x, sm[1].y));
      nt_store4(ad.v + t0, make_float4(sv[0].x, sv[0].y, sv[1].x, sv[1].y));
    } else {
      ad.p[t0] = sp[0];
      nt_store(ad.m + t0, sm[0]);
      nt_store(ad.v + t0, sv[0]);
    }
  } else {
#pragma unroll
    for (int q = 0; q < EPT; ++q)
      if (acc[q].x != 0.f || acc[q].y != 0.f) {
        float2 d = dtable[t0 + q];
        d.x += acc[q].x;
        d.y += acc[q].y;
        dtable[t0 + q] = d;
      }
  }
}

// explicit coordinates: scatter with atomics
__global__ __launch_bounds__(256) void k_level_scatter(PtyGrid g, PtyPoints pts, const float2* __restrict__ gfeat,
                                                       float* __restrict__ dtable) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int l = blockIdx.y;
  if (p >= pts.n) return;
  float x0, x1;
  point_coords(pts, p, x0, x1);
  const float2 dy = gfeat[(int64_t)l * pts.n + p];
  if (dy.x == 0.f && dy.y == 0.f) return;
  level_scatter(g, l, x0, x1, dy, dtable);
}

size_t points_plan_bytes(const PtyGrid& g, const PtyPoints& pts) {
  if (pts.coords) return 0;
  return plan_layout(g, pts).total;
}

int build_points_plan(const PtyGrid& g, const PtyPoints& pts, void* plan, hipStream_t stream) {
  // CSR terms pack the point as 15-bit (i, j)
  PTY_REQUIRE(pts.n_axis0 <= 32768 && pts.n_axis1 <= 32768, PTY_EUNSUP,
              "regular grid %d x %d exceeds 32768 per axis", pts.n_axis0, pts.n_axis1);
  const PlanLayout L = plan_layout(g, pts);
  char* p = (char*)plan;
  const int nax = L.n0 > L.n1 ? L.n0 : L.n1;
  const int nt = L.n_starts > g.n_levels * nax ? L.n_starts : g.n_levels * nax;
  hipLaunchKernelGGL(k_plan_axes, dim3((nt + 255) / 256), dim3(256), 0, stream, g, L, pts.axis0, pts.axis1,
                     (int*)(p + L.s0), (int*)(p + L.s1), (float*)(p + L.fx), (float*)(p + L.fy));
  hipLaunchKernelGGL(k_plan_buckets, dim3(g.n_levels), dim3(64), 0, stream, g, L, (const int*)(p + L.s0),
                     (const int*)(p + L.s1), (int*)(p + L.bstart), (int*)(p + L.bcy), (uint32_t*)(p + L.mask));
  uint32_t hashed_chunks = 0;
  int n_hashed = 0;
  for (int l = 0; l < g.n_levels; ++l)
    if (!((g.dense_mask >> l) & 1u)) {
      hashed_chunks += (g.size[l] + 255) / 256;
      ++n_hashed;
    }
  if (n_hashed > 0) {
    const PlanPtrs P = plan_ptrs(L, plan);
    hipLaunchKernelGGL(k_csr_count, dim3(hashed_chunks), dim3(256), 0, stream, g, L, P, (int*)(p + L.csr_o));
    hipLaunchKernelGGL(k_csr_scan, dim3(n_hashed), dim3(1024), 0, stream, g, L, (int*)(p + L.csr_o));
    hipLaunchKernelGGL(k_csr_fill, dim3(hashed_chunks), dim3(256), 0, stream, g, L, P, (uint32_t*)(p + L.csr_p));
  }
  return check_launch("points_plan_build");
}

size_t table_bwd_workspace(const PtyGrid& g, const PtyPoints& pts) {
  if (pts.coords) return 0;
  const PlanLayout L = plan_layout(g, pts);
  size_t bytes = a256(sizeof(float2) * (size_t)L.off_a[g.n_levels]);
  if (!pts.plan) bytes += L.total;
  return bytes;
}

int launch_table_bwd(const PtyGrid& g, const PtyPoints& pts, const float2* gfeat, float* dtable, char* work,
                     hipStream_t stream, const TableAdam& ad) {
  if (pts.coords) {
    PTY_REQUIRE(!ad.p, PTY_EUNSUP, "table gradient with Adam in place needs grid points (a plan), not coordinates");
    hipLaunchKernelGGL(k_level_scatter, dim3((unsigned)((pts.n + 255) / 256), g.n_levels), dim3(256), 0, stream, g,
                       pts, gfeat, dtable);
    return check_launch("table_bwd_scatter");
  }
  const PlanLayout L = plan_layout(g, pts);
  float2* A = (float2*)work;
  const void* plan = pts.plan;
  if (!plan) {
    char* pw = work + a256(sizeof(float2) * (size_t)L.off_a[g.n_levels]);
    int rc = build_points_plan(g, pts, pw, stream);
    if (rc) return rc;
    plan = pw;
  }
  const PlanPtrs P = plan_ptrs(L, plan);
  const int64_t na = L.off_a[g.n_levels];
  if (na > 0) {
    int n_dense = 0, res1_max = 0;
    for (int l = 0; l < g.n_levels; ++l)
      if ((g.dense_mask >> l) & 1u) {
        ++n_dense;
        if ((int)g.res[l] + 1 > res1_max) res1_max = (int)g.res[l] + 1;
      }
    const size_t shm = sizeof(float2) * (size_t)std::max(L.n0, res1_max) + sizeof(float) * (size_t)L.n0 +
                       sizeof(int) * ((size_t)res1_max + 1);
    const int kc = (res1_max + 255) / 256;
    const dim3 grid((unsigned)(n_dense * L.n1)), blk(256);
    if (shm <= 65536 && kc <= 4) {
      switch (kc) {
        case 1: hipLaunchKernelGGL(k_dense_pass1_rows<1>, grid, blk, shm, stream, g, L, pts, P, gfeat, A); break;
        case 2: hipLaunchKernelGGL(k_dense_pass1_rows<2>, grid, blk, shm, stream, g, L, pts, P, gfeat, A); break;
        case 3: hipLaunchKernelGGL(k_dense_pass1_rows<3>, grid, blk, shm, stream, g, L, pts, P, gfeat, A); break;
        default: hipLaunchKernelGGL(k_dense_pass1_rows<4>, grid, blk, shm, stream, g, L, pts, P, gfeat, A); break;
      }
    } else {
      hipLaunchKernelGGL(k_dense_pass1, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, stream, g, L, pts, P,
                         gfeat, A);
    }
  }
  // entries per thread of the hashed kernel: two while the hashed levels hold
  // at most 8 terms per entry on average (4 per point and level), else one
  int64_t hashed_entries = 0, n_hashed = 0;
  for (int l = 0; l < g.n_levels; ++l)
    if (!((g.dense_mask >> l) & 1u)) {
      hashed_entries += g.size[l];
      ++n_hashed;
    }
  const int64_t n_pts = (int64_t)L.n0 * L.n1;
  const int ept = 4 * n_pts * n_hashed <= 8 * hashed_entries ? 2 : 1;
  uint32_t n_dense_entries = 0, hashed_chunks = 0, n_split_entries = 0;
  bool coarse = true;
  for (int l = 0; l < g.n_levels; ++l) {
    if ((g.dense_mask >> l) & 1u) n_dense_entries = g.offset[l] + g.size[l];
    else hashed_chunks += (g.size[l] + 256 * ept - 1) / (256 * ept);
    // leading dense levels with at least kDenseSplit rows of points per cell
    coarse = coarse && ((g.dense_mask >> l) & 1u) && L.n1 >= kDenseSplit * ((int)g.res[l] + 1);
    if (coarse) n_split_entries = g.offset[l] + g.size[l];
  }
  if (n_dense_entries > 0) {
    const uint32_t lanes = n_split_entries * kDenseSplit + (n_dense_entries - n_split_entries);
    hipLaunchKernelGGL(k_table_dense, dim3((lanes + 255) / 256), dim3(256), 0, stream, g, L, P, (const float2*)A,
                       n_dense_entries, n_split_entries, (float2*)dtable, ad);
  }
  if (hashed_chunks > 0) {
    if (ept == 2)
      hipLaunchKernelGGL(k_table_hashed<2>, dim3(hashed_chunks), dim3(256), 0, stream, g, L, pts, P, gfeat,
                         (float2*)dtable, ad);
    else
      hipLaunchKernelGGL(k_table_hashed<1>, dim3(hashed_chunks), dim3(256), 0, stream, g, L, pts, P, gfeat,
                         (float2*)dtable, ad);
  }
  return check_launch("table_bwd");
}

}  // namespace pty

using namespace pty;

extern "C" size_t pty_points_plan_bytes(const PtyGrid* g, const PtyPoints* pts) {
  if (!g || !pts) return 0;
  return points_plan_bytes(*g, *pts);
}

extern "C" int pty_points_plan_build(const PtyGrid* g, const PtyPoints* pts, void* plan, size_t plan_bytes,
                                     hipStream_t stream) {
  int rc = validate_grid(g);
  if (rc) return rc;
  PTY_REQUIRE(pts && !pts->coords && pts->axis0 && pts->axis1 && pts->n_axis0 > 0 && pts->n_axis1 > 0, PTY_EINVAL,
              "points_plan_build: needs a regular grid (axis0/axis1)");
  PTY_REQUIRE(plan && plan_bytes >= points_plan_bytes(*g, *pts), PTY_EWORK, "points_plan_build: buffer too small");
  return build_points_plan(*g, *pts, plan, stream);
}
