// 2-D Gaussian-field rasterizer for gfx950: the reference's gscp CUDA
// rasterizer (model_type gaussian_field), re-designed for 64-wide wavefronts
// and a deterministic, atomic-free backward.
//
// Stages (reference: gscp/csrc/gscp_rasterizer/*):
//   preprocess   one lane per Gaussian: clamped mean, conic, radius, tile rect,
//                tile count; per-tile pair counts from a 2-D difference array
//                of the rects (four integer atomics per Gaussian) and its
//                prefix sums (forward.cu:18-170, auxiliary.h:17-30)
//   scans        exclusive scans of the per-Gaussian and per-tile counts
//                (rasterizer_impl.cu:186-192)
//   binning      each Gaussian writes its id into the lists of the tiles its
//                rect covers (slots from per-chunk LDS counters and a column
//                scan of the chunk counts; global atomics only on canvases
//                beyond 16384 tiles), then every tile sorts its list: ids are
//                unique, so the order is the ascending index the reference's
//                stable radix sort gives (rasterizer_impl.cu:45-106, 212-240)
//   render       one 256-lane workgroup (4 waves) per 16 x 16 tile, one pixel
//                per lane; the tile's Gaussians staged in LDS in batches of 256
//                (forward.cu:176-258)
//   backward     per tile, each (Gaussian, tile) entry is owned by a group of
//                P lanes that walk the tile's pixels with stride P (consecutive
//                lanes on consecutive pixels), sum the entry's seven gradient
//                terms and reduce them with cross-lane shuffles; the result is
//                stored at the entry's slot in its Gaussian's rect order, and a
//                finish kernel sums each Gaussian's slots in that fixed order
//                and applies the preprocess adjoint (backward.cu:21-266).  No
//                atomics: every gradient is deterministic.
#include "common.h"

namespace pty {
namespace {

constexpr int kTile = 16;              // 16 x 16 pixel tiles (gscp config.h:13-14)
constexpr int kTilePix = kTile * kTile;
constexpr int kScanBlock = 1024;       // elements per workgroup in the scans
constexpr int kSortCap = 8192;         // tile lists up to this length sort in LDS
constexpr int kLdsTiles = 16384;       // canvases up to this many tiles bin through LDS counters
constexpr int kMaxChunks = 512;        // Gaussian chunks of the LDS binning
constexpr float kGMin = 1e-7f;         // forward.cu:243
constexpr float kLog2e = 1.4426950408889634f;

inline size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

struct GeomLayout {
  size_t mean, conic, rect, goff, touched, gsums, diff, count, start, cursor, tsums, total, chunk, bytes;
  int tiles_x, tiles_y, tiles, nb_g, nb_t;
  int chunk_size, n_chunks;   // LDS binning: Gaussians per chunk and chunks (0: global-atomic binning)
};

GeomLayout geom_layout(const PtyGs* d) {
  GeomLayout L{};
  const size_t n = (size_t)d->n;
  L.tiles_x = (d->width + kTile - 1) / kTile;
  L.tiles_y = (d->height + kTile - 1) / kTile;
  L.tiles = L.tiles_x * L.tiles_y;
  L.nb_g = (int)((n + kScanBlock - 1) / kScanBlock);
  L.nb_t = (L.tiles + kScanBlock - 1) / kScanBlock;
  size_t o = 0;
  L.mean = o;    o = align_up(o + n * sizeof(float2));
  L.conic = o;   o = align_up(o + n * sizeof(float4));
  L.rect = o;    o = align_up(o + n * sizeof(int4));
  L.goff = o;    o = align_up(o + n * sizeof(uint32_t));
  L.touched = o; o = align_up(o + n * sizeof(uint32_t));
  L.gsums = o;   o = align_up(o + (size_t)(L.nb_g + 1) * sizeof(uint32_t));
  L.diff = o;    o = align_up(o + (size_t)(L.tiles_x + 1) * (L.tiles_y + 1) * sizeof(int));
  L.count = o;   o = align_up(o + (size_t)L.tiles * sizeof(uint32_t));
  L.start = o;   o = align_up(o + (size_t)L.tiles * sizeof(uint32_t));
  L.cursor = o;  o = align_up(o + (size_t)L.tiles * sizeof(uint32_t));
  L.tsums = o;   o = align_up(o + (size_t)(L.nb_t + 1) * sizeof(uint32_t));
  L.total = o;   o = align_up(o + 2 * sizeof(unsigned long long));
  L.chunk_size = 0;
  L.n_chunks = 0;
  if (L.tiles <= kLdsTiles && n > 0) {
    L.chunk_size = 256;
    while ((n + L.chunk_size - 1) / L.chunk_size > (size_t)kMaxChunks) L.chunk_size *= 2;
    L.n_chunks = (int)((n + L.chunk_size - 1) / L.chunk_size);
  }
  L.chunk = o;   o = align_up(o + (size_t)L.n_chunks * L.tiles * sizeof(uint32_t));
  L.bytes = o;
  return L;
}

struct BinLayout {
  size_t list_u, list, partial, bytes;
};

BinLayout bin_layout(int64_t R, int with_grad) {
  BinLayout B{};
  size_t o = 0;
  B.list_u = o;  o = align_up(o + (size_t)R * sizeof(uint32_t));
  B.list = o;    o = align_up(o + (size_t)R * sizeof(uint32_t));
  B.partial = o; if (with_grad) o = align_up(o + (size_t)R * 2 * sizeof(float4));
  B.bytes = o;
  return B;
}

template <typename T>
T* at(void* base, size_t off) { return reinterpret_cast<T*>(static_cast<char*>(base) + off); }
template <typename T>
const T* at(const void* base, size_t off) { return reinterpret_cast<const T*>(static_cast<const char*>(base) + off); }

__global__ void k_zero_u32(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// ------------------------------------------------------------------ preprocess
// getRect (auxiliary.h:17-30) with the reference's float steps: the lower
// corner from (p - r) / 16, the upper from ((p + r) + 16 - 1) / 16, each
// truncated toward zero and clamped to [0, grid].
__device__ __forceinline__ int4 tile_rect(float px, float py, int r, int gx, int gy) {
#pragma clang fp contract(off)
  const float rf = (float)r;
  const int x0 = min(gx, max(0, (int)((px - rf) / 16.f)));
  const int y0 = min(gy, max(0, (int)((py - rf) / 16.f)));
  const int x1 = min(gx, max(0, (int)((((px + rf) + 16.f) - 1.f) / 16.f)));
  const int y1 = min(gy, max(0, (int)((((py + rf) + 16.f) - 1.f) / 16.f)));
  return make_int4(x0, y0, x1, y1);
}

// Covariance [[a, b], [b, c]] of Gaussian i (forward.cu:49-64 and :127-140);
// unfused float operations in the reference's order, so that the integer
// radius derived from them matches a float32 restatement.
template <int PARAM>
__device__ __forceinline__ void covariance(const float* __restrict__ p0, const float* __restrict__ p1, int i,
                                           float smax, float smin, float& a, float& b, float& c, float& det) {
#pragma clang fp contract(off)
  if (PARAM == PTY_GS_RS) {
    const float sx = fminf(smax, fmaxf(smin, expf(p0[2 * i + 0])));
    const float sy = fminf(smax, fmaxf(smin, expf(p0[2 * i + 1])));
    const float sx2 = sx * sx, sy2 = sy * sy;
    const float th = p1[i];
    const float ct = cosf(th), st = sinf(th);
    a = ct * ct * sx2 + st * st * sy2;
    b = ct * st * (sx2 - sy2);
    c = st * st * sx2 + ct * ct * sy2;
    det = a * c - b * b;
  } else {
    const float l11 = fminf(smax, fmaxf(smin, expf(p0[2 * i + 0])));
    const float l22 = fminf(smax, fmaxf(smin, expf(p0[2 * i + 1])));
    const float l21 = p1[i];
    a = l11 * l11;
    b = l11 * l21;
    c = l21 * l21 + l22 * l22;
    det = a * l22 * l22;
  }
}

template <int PARAM>
__global__ __launch_bounds__(256) void k_gs_preprocess(
    int n, int W, int H, const float* __restrict__ xy, const float* __restrict__ p0, const float* __restrict__ p1,
    float smax, float smin, int gx, int gy, int32_t* __restrict__ radii, float2* __restrict__ mean,
    float4* __restrict__ conic, int4* __restrict__ rect, uint32_t* __restrict__ touched,
    int* __restrict__ diff) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  radii[i] = 0;
  touched[i] = 0u;
  const float mx = fminf((float)(W - 1), fmaxf(0.f, xy[2 * i + 0]));
  const float my = fminf((float)(H - 1), fmaxf(0.f, xy[2 * i + 1]));
  float a, b, c, det;
  covariance<PARAM>(p0, p1, i, smax, smin, a, b, c, det);
  if (!(det > 0.f)) return;
  const float det_inv = 1.f / det;
  const float mid = 0.5f * (a + c);
  const float lam = mid + sqrtf(fmaxf(0.1f, mid * mid - det));
  const float rf = ceilf(3.f * sqrtf(lam));
  // NaN -> culled; beyond int range -> INT_MAX (the rect then spans the canvas)
  if (!(rf > 0.f)) return;
  const int r = rf >= 2147483520.f ? 2147483647 : (int)rf;
  mean[i] = make_float2(mx, my);
  conic[i] = make_float4(c * det_inv, -b * det_inv, a * det_inv, 0.f);
  radii[i] = r;
  const int4 rc = tile_rect(mx, my, r, gx, gy);
  rect[i] = rc;
  touched[i] = (uint32_t)((rc.w - rc.y) * (rc.z - rc.x));
  // the rect as +1 / -1 at its four corners of a (gy+1) x (gx+1) difference
  // array; its 2-D prefix sum is the per-tile pair count (four atomics per
  // Gaussian instead of one per covered tile)
  const int gx1 = gx + 1;
  atomicAdd(&diff[rc.y * gx1 + rc.x], 1);
  atomicAdd(&diff[rc.y * gx1 + rc.z], -1);
  atomicAdd(&diff[rc.w * gx1 + rc.x], -1);
  atomicAdd(&diff[rc.w * gx1 + rc.z], 1);
}

// count[ty][tx] = sum of diff over rows <= ty and columns <= tx.  The
// prefix kernels are defined after the scan helpers below.

// ------------------------------------------------------------------ scans
// Exclusive scan of kScanBlock elements per workgroup (256 lanes x 4); the
// workgroup total goes to sums[blockIdx.x].
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// inclusive scan over the workgroup (blockDim.x <= 1024), returns the lane's
// inclusive prefix and the workgroup total in *total
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_incl_scan(v);
  if (lane == 63) s_w[wave] = v;
  __syncthreads();
  if (wave == 0) {
    uint32_t t = lane < nw ? s_w[lane] : 0u;
    t = wave_incl_scan(t);
    if (lane < nw) s_w[lane] = t;
  }
  __syncthreads();
  if (wave > 0) v += s_w[wave - 1];
  *total = s_w[nw - 1];
  __syncthreads();
  return v;
}

__global__ __launch_bounds__(256) void k_scan_local(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                    uint32_t* __restrict__ sums, int64_t n) {
  __shared__ uint32_t s_w[16];
  const int64_t base = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * 4;
  uint32_t v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = base + k < n ? in[base + k] : 0u;
  const uint32_t t = v[0] + v[1] + v[2] + v[3];
  uint32_t tot;
  const uint32_t incl = block_incl_scan(t, s_w, &tot);
  uint32_t run = incl - t;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// One workgroup: exclusive scan of the nb workgroup totals in place, and the
// 64-bit grand total.
__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* __restrict__ sums, int nb,
                                                    unsigned long long* __restrict__ total) {
  __shared__ uint32_t s_w[16];
  unsigned long long carry = 0ull;
  for (int c = 0; c < nb; c += 1024) {
    const int i = c + threadIdx.x;
    const uint32_t v = i < nb ? sums[i] : 0u;
    uint32_t tot;
    const uint32_t incl = block_incl_scan(v, s_w, &tot);
    if (i < nb) sums[i] = (uint32_t)carry + (incl - v);
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* __restrict__ out, const uint32_t* __restrict__ sums,
                                                  int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += sums[i / kScanBlock];
}

// Inclusive prefix along x of every row of the (rows) x (gx1) difference
// array, in place; one workgroup per row.
__global__ __launch_bounds__(256) void k_diff_rows(uint32_t* __restrict__ diff, int gx1) {
  __shared__ uint32_t s_w[16];
  uint32_t* row = diff + (size_t)blockIdx.x * gx1;
  uint32_t carry = 0u;
  for (int c = 0; c < gx1; c += 256) {
    const int x = c + threadIdx.x;
    const uint32_t v = x < gx1 ? row[x] : 0u;
    uint32_t tot;
    const uint32_t incl = block_incl_scan(v, s_w, &tot);
    if (x < gx1) row[x] = carry + incl;
    carry += tot;
  }
}

// Inclusive prefix along y of the row-prefixed difference array: the pair
// count of every tile; one workgroup per tile column.  Integer sums wrap
// modulo 2^32, and every count is a non-negative total below 2^32.
__global__ __launch_bounds__(256) void k_diff_cols(const uint32_t* __restrict__ diff, int gx, int gy,
                                                   uint32_t* __restrict__ count) {
  __shared__ uint32_t s_w[16];
  const int x = blockIdx.x;
  uint32_t carry = 0u;
  for (int c = 0; c < gy; c += 256) {
    const int y = c + threadIdx.x;
    const uint32_t v = y < gy ? diff[(size_t)y * (gx + 1) + x] : 0u;
    uint32_t tot;
    const uint32_t incl = block_incl_scan(v, s_w, &tot);
    if (y < gy) count[(size_t)y * gx + x] = carry + incl;
    carry += tot;
  }
}

int exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t* sums, int nb, int64_t n,
                   unsigned long long* total, hipStream_t s) {
  if (n > 0) {
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(256), 0, s, in, out, sums, n);
  }
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, nb, total);
  if (n > 0) {
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, sums, n);
  }
  return check_launch("gaussian-field scan");
}

// ------------------------------------------------------------------ binning
__global__ __launch_bounds__(256) void k_gs_scatter(int n, int gx, const uint32_t* __restrict__ touched,
                                                    const int4* __restrict__ rect,
                                                    const uint32_t* __restrict__ start,
                                                    uint32_t* __restrict__ cursor, uint32_t* __restrict__ list_u) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || touched[i] == 0u) return;
  const int4 rc = rect[i];
  const int w = rc.z - rc.x, cnt = w * (rc.w - rc.y);
  // eight slot requests in flight before their results are used
  for (int base = 0; base < cnt; base += 8) {
    int tt[8];
    uint32_t slot[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = base + u;
      tt[u] = -1;
      if (k < cnt) {
        tt[u] = (rc.y + k / w) * gx + rc.x + k % w;
        slot[u] = atomicAdd(&cursor[tt[u]], 1u);
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (tt[u] >= 0) list_u[start[tt[u]] + slot[u]] = (uint32_t)i;
  }
}

// LDS binning (canvases of at most kLdsTiles tiles).  Chunk c holds Gaussians
// [c * chunk, (c + 1) * chunk); its workgroup counts the chunk's pairs per
// tile with LDS atomics into hist[c][t] (pass 1), a column scan over the
// chunks turns the counts into each chunk's offset inside every tile's list
// (pass 2), and pass 3 hands out slots from LDS counters.  No global atomics;
// the order inside a tile is settled by the sort that follows.
__global__ __launch_bounds__(256) void k_gs_chunk_hist(int n, int chunk, int tiles, int gx,
                                                       const uint32_t* __restrict__ touched,
                                                       const int4* __restrict__ rect, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kLdsTiles];
  for (int t = threadIdx.x; t < tiles; t += 256) h[t] = 0u;
  __syncthreads();
  const int g0 = blockIdx.x * chunk, g1 = min(n, g0 + chunk);
  for (int i = g0 + threadIdx.x; i < g1; i += 256) {
    if (touched[i] == 0u) continue;
    const int4 rc = rect[i];
    for (int ty = rc.y; ty < rc.w; ++ty)
      for (int tx = rc.x; tx < rc.z; ++tx) atomicAdd(&h[ty * gx + tx], 1u);
  }
  __syncthreads();
  uint32_t* out = hist + (size_t)blockIdx.x * tiles;
  for (int t = threadIdx.x; t < tiles; t += 256) out[t] = h[t];
}

// hist[c][t] <- start[t] + sum of hist[c'][t] over c' < c: one wavefront per
// tile, each lane holding up to eight consecutive chunks (n_chunks <= 512),
// a wave scan of the lanes' sums in between.
__global__ __launch_bounds__(256) void k_gs_chunk_prefix(int n_chunks, int tiles, const uint32_t* __restrict__ start,
                                                         uint32_t* __restrict__ hist) {
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= tiles) return;                      // uniform per wavefront
  const int lane = threadIdx.x & 63;
  const int k = (n_chunks + 63) / 64;
  uint32_t v[8];
  uint32_t sum = 0u;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = lane * k + j;
    v[j] = (j < k && c < n_chunks) ? hist[(size_t)c * tiles + t] : 0u;
    sum += v[j];
  }
  uint32_t run = start[t] + wave_incl_scan(sum) - sum;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = lane * k + j;
    if (j < k && c < n_chunks) {
      hist[(size_t)c * tiles + t] = run;
      run += v[j];
    }
  }
}

__global__ __launch_bounds__(256) void k_gs_chunk_scatter(int n, int chunk, int tiles, int gx,
                                                          const uint32_t* __restrict__ touched,
                                                          const int4* __restrict__ rect,
                                                          const uint32_t* __restrict__ base,
                                                          uint32_t* __restrict__ list_u) {
  __shared__ uint32_t cur[kLdsTiles];
  for (int t = threadIdx.x; t < tiles; t += 256) cur[t] = 0u;
  __syncthreads();
  const uint32_t* off = base + (size_t)blockIdx.x * tiles;
  const int g0 = blockIdx.x * chunk, g1 = min(n, g0 + chunk);
  for (int i = g0 + threadIdx.x; i < g1; i += 256) {
    if (touched[i] == 0u) continue;
    const int4 rc = rect[i];
    for (int ty = rc.y; ty < rc.w; ++ty)
      for (int tx = rc.x; tx < rc.z; ++tx) {
        const int t = ty * gx + tx;
        list_u[off[t] + atomicAdd(&cur[t], 1u)] = (uint32_t)i;
      }
  }
}

// Sort each tile's list of (unique) Gaussian ids ascending: bitonic in LDS up
// to kSortCap ids, else rank by counting (each id's rank is the number of
// smaller ids in the list), reading the list through LDS in chunks.
__global__ __launch_bounds__(256) void k_gs_sort(const uint32_t* __restrict__ start,
                                                 const uint32_t* __restrict__ count,
                                                 const uint32_t* __restrict__ list_u, uint32_t* __restrict__ list) {
  __shared__ uint32_t key[kSortCap];
  const int t = blockIdx.x;
  const uint32_t s = start[t], n = count[t];
  if (n == 0u) return;
  if (n == 1u) {
    if (threadIdx.x == 0) list[s] = list_u[s];
    return;
  }
  if (n <= 256u) {
    // one id per lane; its rank is the number of smaller ids (LDS broadcast reads)
    key[threadIdx.x] = threadIdx.x < n ? list_u[s + threadIdx.x] : 0xFFFFFFFFu;
    __syncthreads();
    if (threadIdx.x < n) {
      const uint32_t v = key[threadIdx.x];
      uint32_t r = 0;
      for (uint32_t j = 0; j < n; ++j) r += key[j] < v ? 1u : 0u;
      list[s + r] = v;
    }
    return;
  }
  if (n <= (uint32_t)kSortCap) {
    uint32_t S = 2;
    while (S < n) S <<= 1;
    for (uint32_t i = threadIdx.x; i < S; i += 256) key[i] = i < n ? list_u[s + i] : 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t k = 2; k <= S; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        for (uint32_t i = threadIdx.x; i < S; i += 256) {
          const uint32_t l = i ^ j;
          if (l > i) {
            const uint32_t a = key[i], b = key[l];
            const bool up = (i & k) == 0u;
            if ((a > b) == up) {
              key[i] = b;
              key[l] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    for (uint32_t i = threadIdx.x; i < n; i += 256) list[s + i] = key[i];
    return;
  }
  for (uint32_t base = 0; base < n; base += 256) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < n ? list_u[s + i] : 0xFFFFFFFFu;
    uint32_t r = 0;
    for (uint32_t c = 0; c < n; c += kSortCap) {
      const uint32_t cnt = min((uint32_t)kSortCap, n - c);
      __syncthreads();
      for (uint32_t j = threadIdx.x; j < cnt; j += 256) key[j] = list_u[s + c + j];
      __syncthreads();
      for (uint32_t j = 0; j < cnt; ++j) r += key[j] < v ? 1u : 0u;
    }
    if (i < n) list[s + r] = v;
  }
}

// ------------------------------------------------------------------ render
// power = -0.5 (cx dx^2 + cz dy^2) - cy dx dy (forward.cu:238), evaluated as
// power * log2(e) with the factors folded into the staged conic, so that
// G = exp2(.) is one hardware transcendental.
__device__ __forceinline__ float4 staged_conic(float4 c) {
  return make_float4(-0.5f * kLog2e * c.x, -kLog2e * c.y, -0.5f * kLog2e * c.z, 0.f);
}

__global__ __launch_bounds__(256) void k_gs_render(int W, int H, int gx, const uint32_t* __restrict__ start,
                                                   const uint32_t* __restrict__ count,
                                                   const uint32_t* __restrict__ list,
                                                   const float2* __restrict__ mean,
                                                   const float4* __restrict__ conic,
                                                   const float2* __restrict__ weights, float* __restrict__ out) {
  __shared__ float4 s_a[256];  // mean.x, mean.y, w.x, w.y
  __shared__ float4 s_b[256];  // staged conic
  const int t = blockIdx.x;
  const int tx = t % gx, ty = t / gx;
  const int px = tx * kTile + (threadIdx.x & 15), py = ty * kTile + (threadIdx.x >> 4);
  const float fx = (float)px, fy = (float)py;
  const uint32_t s = start[t], n = count[t];
  float acc0 = 0.f, acc1 = 0.f;
  for (uint32_t b = 0; b < n; b += 256) {
    __syncthreads();
    if (b + threadIdx.x < n) {
      const uint32_t g = list[s + b + threadIdx.x];
      const float2 m = mean[g], w = weights[g];
      s_a[threadIdx.x] = make_float4(m.x, m.y, w.x, w.y);
      s_b[threadIdx.x] = staged_conic(conic[g]);
    }
    __syncthreads();
    const int cnt = (int)min(256u, n - b);
#pragma unroll 4
    for (int k = 0; k < cnt; ++k) {
      const float4 A = s_a[k], B = s_b[k];
      const float dx = A.x - fx, dy = A.y - fy;
      const float p2 = B.x * dx * dx + B.z * dy * dy + B.y * dx * dy;
      const float G = __builtin_amdgcn_exp2f(p2);
      const bool keep = (p2 <= 0.f) & (G >= kGMin);
      acc0 += keep ? A.z * G : 0.f;
      acc1 += keep ? A.w * G : 0.f;
    }
  }
  if (px < W && py < H) {
    out[(size_t)py * W + px] = acc0;
    out[(size_t)H * W + (size_t)py * W + px] = acc1;
  }
}

// ------------------------------------------------------------------ backward
// Partial sums of entry (g, tile): (Σ G d0, Σ G d1, Σ q dx, Σ q dy) and
// (Σ q dx^2, Σ q dx dy, Σ q dy^2, 0) with q = G (w0 d0 + w1 d1), over the
// tile's pixels; the mean and conic gradients follow from them in the finish
// kernel: dmean = (-cx Sx - cy Sy, -cz Sy - cy Sx), dconic = (-Sxx/2, -Sxy,
// -Syy/2) (backward.cu:84-104 summed over pixels).
__global__ __launch_bounds__(256) void k_gs_render_bwd(
    int W, int H, int gx, const uint32_t* __restrict__ start, const uint32_t* __restrict__ count,
    const uint32_t* __restrict__ list, const float2* __restrict__ mean, const float4* __restrict__ conic,
    const float2* __restrict__ weights, const uint32_t* __restrict__ goff, const int4* __restrict__ rect,
    const float* __restrict__ dout, float4* __restrict__ partial) {
  __shared__ float2 s_d[kTilePix];
  const int t = blockIdx.x;
  const int tx = t % gx, ty = t / gx;
  {
    const int px = tx * kTile + (threadIdx.x & 15), py = ty * kTile + (threadIdx.x >> 4);
    float2 d = make_float2(0.f, 0.f);
    if (px < W && py < H) {
      const size_t o = (size_t)py * W + px;
      d = make_float2(dout[o], dout[(size_t)H * W + o]);
    }
    s_d[threadIdx.x] = d;
  }
  __syncthreads();
  const float fx0 = (float)(tx * kTile), fy0 = (float)(ty * kTile);
  const uint32_t s = start[t], n = count[t];
  for (uint32_t b = 0; b < n; b += 256) {
    const uint32_t m = min(256u, n - b);
    uint32_t mp = 1;
    while (mp < m) mp <<= 1;
    const int P = (int)min(64u, 256u / mp);     // lanes per entry (a power of two)
    const uint32_t e = threadIdx.x / P;
    const int part = threadIdx.x & (P - 1);
    float sw0 = 0.f, sw1 = 0.f, sx = 0.f, sy = 0.f, sxx = 0.f, sxy = 0.f, syy = 0.f;
    uint32_t g = 0;
    if (e < m) {
      g = list[s + b + e];
      const float2 mu = mean[g], w = weights[g];
      const float4 B = staged_conic(conic[g]);
      const float mx = mu.x - fx0, my = mu.y - fy0;
      for (int k = part; k < kTilePix; k += P) {
        const float2 d = s_d[k];
        const float dx = mx - (float)(k & 15), dy = my - (float)(k >> 4);
        const float p2 = B.x * dx * dx + B.z * dy * dy + B.y * dx * dy;
        const float G0 = __builtin_amdgcn_exp2f(p2);
        const float G = ((p2 <= 0.f) & (G0 >= kGMin)) ? G0 : 0.f;
        const float q = G * (w.x * d.x + w.y * d.y);
        sw0 += G * d.x;
        sw1 += G * d.y;
        const float qdx = q * dx;
        sx += qdx;
        sy += q * dy;
        sxx += qdx * dx;
        sxy += qdx * dy;
        syy += q * dy * dy;
      }
    }
    for (int o = P >> 1; o > 0; o >>= 1) {
      sw0 += __shfl_xor(sw0, o, 64);
      sw1 += __shfl_xor(sw1, o, 64);
      sx += __shfl_xor(sx, o, 64);
      sy += __shfl_xor(sy, o, 64);
      sxx += __shfl_xor(sxx, o, 64);
      sxy += __shfl_xor(sxy, o, 64);
      syy += __shfl_xor(syy, o, 64);
    }
    if (e < m && part == 0) {
      const int4 rc = rect[g];
      const uint32_t slot = goff[g] + (uint32_t)((ty - rc.y) * (rc.z - rc.x) + (tx - rc.x));
      partial[2 * (size_t)slot] = make_float4(sw0, sw1, sx, sy);
      partial[2 * (size_t)slot + 1] = make_float4(sxx, sxy, syy, 0.f);
    }
  }
}

// Sum of each Gaussian's partials in its rect order, then the preprocess
// adjoint (backward.cu:138-266).  Writes every output of every Gaussian.
template <int PARAM>
__global__ __launch_bounds__(256) void k_gs_finish(
    int n, const int32_t* __restrict__ radii, const uint32_t* __restrict__ goff,
    const uint32_t* __restrict__ touched, const float4* __restrict__ partial, const float4* __restrict__ conic,
    const float* __restrict__ p0, const float* __restrict__ p1, float smax, float smin,
    float* __restrict__ dxy, float* __restrict__ dp0, float* __restrict__ dp1, float* __restrict__ dweight) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (radii[i] <= 0) {
    dxy[2 * i] = 0.f; dxy[2 * i + 1] = 0.f;
    dp0[2 * i] = 0.f; dp0[2 * i + 1] = 0.f;
    dp1[i] = 0.f;
    dweight[2 * i] = 0.f; dweight[2 * i + 1] = 0.f;
    return;
  }
  float4 u = make_float4(0.f, 0.f, 0.f, 0.f), v = make_float4(0.f, 0.f, 0.f, 0.f);
  const uint32_t o = goff[i], cnt = touched[i];
  for (uint32_t k = 0; k < cnt; ++k) {
    const float4 a = partial[2 * (size_t)(o + k)], b = partial[2 * (size_t)(o + k) + 1];
    u.x += a.x; u.y += a.y; u.z += a.z; u.w += a.w;
    v.x += b.x; v.y += b.y; v.z += b.z;
  }
  dweight[2 * i] = u.x;
  dweight[2 * i + 1] = u.y;
  const float4 cc = conic[i];
  // the mean gradient passes through the position clamp (backward.cu:196-201)
  dxy[2 * i] = -cc.x * u.z - cc.y * u.w;
  dxy[2 * i + 1] = -cc.z * u.w - cc.y * u.z;
  const float dc0 = -0.5f * v.x, dc1 = -v.y, dc2 = -0.5f * v.z;
  if (PARAM == PTY_GS_RS) {
    const float th = p1[i];
    const float ct = cosf(th), st = sinf(th);
    const float sxr = expf(p0[2 * i]), syr = expf(p0[2 * i + 1]);
    const float sx = fminf(smax, fmaxf(smin, sxr)), sy = fminf(smax, fmaxf(smin, syr));
    const float sx2 = sx * sx, sy2 = sy * sy;
    const float a = ct * ct * sx2 + st * st * sy2;
    const float b = ct * st * (sx2 - sy2);
    const float c = st * st * sx2 + ct * ct * sy2;
    const float det = a * c - b * b;
    const float inv = 1.f / (det * det + 1e-7f);
    const float da = inv * (-c * c * dc0 + b * c * dc1 + (det - a * c) * dc2);
    const float db = inv * (2.f * b * c * dc0 - (det + 2.f * b * b) * dc1 + 2.f * a * b * dc2);
    const float dcc = inv * ((det - a * c) * dc0 + a * b * dc1 - a * a * dc2);
    const float dsx2 = ct * ct * da + ct * st * db + st * st * dcc;
    const float dsy2 = st * st * da - ct * st * db + ct * ct * dcc;
    const float dth = 2.f * ct * st * (sy2 - sx2) * da + (ct * ct - st * st) * (sx2 - sy2) * db +
                      2.f * st * ct * (sx2 - sy2) * dcc;
    const float inx = (sxr >= smin && sxr <= smax) ? 1.f : 0.f;
    const float iny = (syr >= smin && syr <= smax) ? 1.f : 0.f;
    dp0[2 * i] = dsx2 * 2.f * sx2 * inx;
    dp0[2 * i + 1] = dsy2 * 2.f * sy2 * iny;
    dp1[i] = dth;
  } else {
    const float r11 = expf(p0[2 * i]), r22 = expf(p0[2 * i + 1]);
    const float l11 = fminf(smax, fmaxf(smin, r11)), l22 = fminf(smax, fmaxf(smin, r22));
    const float l21 = p1[i];
    const float a = l11 * l11, b = l11 * l21, c = l21 * l21 + l22 * l22;
    const float det = a * l22 * l22;
    const float inv = 1.f / (det * det + 1e-7f);
    const float da = inv * (-c * c * dc0 + b * c * dc1 + (det - a * c) * dc2);
    const float db = inv * (2.f * b * c * dc0 - (det + 2.f * b * b) * dc1 + 2.f * a * b * dc2);
    const float dcc = inv * ((det - a * c) * dc0 + a * b * dc1 - a * a * dc2);
    const float dl11 = 2.f * l11 * da + l21 * db;
    const float dl21 = l11 * db + 2.f * l21 * dcc;
    const float dl22 = 2.f * l22 * dcc;
    const float i11 = (r11 >= smin && r11 <= smax) ? 1.f : 0.f;
    const float i22 = (r22 >= smin && r22 <= smax) ? 1.f : 0.f;
    dp0[2 * i] = dl11 * l11 * i11;
    dp0[2 * i + 1] = dl22 * l22 * i22;
    dp1[i] = dl21;
  }
}

int check_desc(const PtyGs* d) {
  PTY_REQUIRE(d != nullptr, PTY_EINVAL, "gaussian field: null descriptor");
  PTY_REQUIRE(d->n >= 0 && d->n <= (1 << 24), PTY_EUNSUP, "gaussian field: n = %d outside [0, 2^24]", d->n);
  PTY_REQUIRE(d->width >= 1 && d->width <= 16384 && d->height >= 1 && d->height <= 16384, PTY_EUNSUP,
              "gaussian field: canvas %d x %d outside [1, 16384]", d->width, d->height);
  PTY_REQUIRE(d->param == PTY_GS_RS || d->param == PTY_GS_CHOLESKY, PTY_EINVAL,
              "gaussian field: param %d is neither PTY_GS_RS nor PTY_GS_CHOLESKY", d->param);
  return PTY_OK;
}

unsigned grid1(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace pty

using namespace pty;

extern "C" size_t pty_gs_geom_bytes(const PtyGs* d) {
  if (check_desc(d) != PTY_OK) return 0;
  return geom_layout(d).bytes;
}

extern "C" size_t pty_gs_bin_bytes(const PtyGs* d, int64_t n_rendered, int32_t with_grad) {
  if (check_desc(d) != PTY_OK || n_rendered < 0) return 0;
  return bin_layout(n_rendered, with_grad).bytes;
}

extern "C" int pty_gs_preprocess(const PtyGs* d, const float* xy, const float* p0, const float* p1, int32_t* radii,
                                 void* geom, hipStream_t stream) {
  int rc = check_desc(d);
  if (rc != PTY_OK) return rc;
  PTY_REQUIRE(geom != nullptr, PTY_EINVAL, "gaussian field: null geometry buffer");
  PTY_REQUIRE(d->n == 0 || (xy && p0 && p1 && radii), PTY_EINVAL, "gaussian field: null input");
  const GeomLayout L = geom_layout(d);
  uint32_t* count = at<uint32_t>(geom, L.count);
  int* diff = at<int>(geom, L.diff);
  const int64_t n_diff = (int64_t)(L.tiles_x + 1) * (L.tiles_y + 1);
  hipLaunchKernelGGL(k_zero_u32, dim3(grid1(n_diff) < 4096 ? grid1(n_diff) : 4096), dim3(256), 0, stream,
                     reinterpret_cast<uint32_t*>(diff), n_diff);
  if (d->n > 0) {
    const float smax = (float)d->max_patch_radius, smin = d->min_scale;
    if (d->param == PTY_GS_RS)
      hipLaunchKernelGGL(k_gs_preprocess<PTY_GS_RS>, dim3(grid1(d->n)), dim3(256), 0, stream, d->n, d->width,
                         d->height, xy, p0, p1, smax, smin, L.tiles_x, L.tiles_y, radii, at<float2>(geom, L.mean),
                         at<float4>(geom, L.conic), at<int4>(geom, L.rect), at<uint32_t>(geom, L.touched), diff);
    else
      hipLaunchKernelGGL(k_gs_preprocess<PTY_GS_CHOLESKY>, dim3(grid1(d->n)), dim3(256), 0, stream, d->n,
                         d->width, d->height, xy, p0, p1, smax, smin, L.tiles_x, L.tiles_y, radii,
                         at<float2>(geom, L.mean), at<float4>(geom, L.conic), at<int4>(geom, L.rect),
                         at<uint32_t>(geom, L.touched), diff);
    rc = check_launch("pty_gs_preprocess");
    if (rc != PTY_OK) return rc;
  }
  hipLaunchKernelGGL(k_diff_rows, dim3(L.tiles_y + 1), dim3(256), 0, stream, reinterpret_cast<uint32_t*>(diff),
                     L.tiles_x + 1);
  hipLaunchKernelGGL(k_diff_cols, dim3(L.tiles_x), dim3(256), 0, stream, reinterpret_cast<const uint32_t*>(diff),
                     L.tiles_x, L.tiles_y, count);
  unsigned long long* total = at<unsigned long long>(geom, L.total);
  rc = exclusive_scan(at<uint32_t>(geom, L.touched), at<uint32_t>(geom, L.goff), at<uint32_t>(geom, L.gsums),
                      L.nb_g, d->n, total, stream);
  if (rc != PTY_OK) return rc;
  return exclusive_scan(count, at<uint32_t>(geom, L.start), at<uint32_t>(geom, L.tsums), L.nb_t, L.tiles,
                        total + 1, stream);
}

extern "C" int pty_gs_num_rendered(const PtyGs* d, const void* geom, int64_t* n_rendered, hipStream_t stream) {
  int rc = check_desc(d);
  if (rc != PTY_OK) return rc;
  PTY_REQUIRE(geom != nullptr && n_rendered != nullptr, PTY_EINVAL, "gaussian field: null argument");
  const GeomLayout L = geom_layout(d);
  unsigned long long tot[2] = {0ull, 0ull};
  hipError_t e = hipMemcpyAsync(tot, at<unsigned long long>(geom, L.total), sizeof(tot), hipMemcpyDeviceToHost,
                                stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  PTY_REQUIRE(e == hipSuccess, PTY_ELAUNCH, "pty_gs_num_rendered: %s", hipGetErrorString(e));
  PTY_REQUIRE(tot[0] == tot[1], PTY_ELAUNCH, "pty_gs_num_rendered: pair counts disagree (%llu per Gaussian, %llu per tile)",
              tot[0], tot[1]);
  PTY_REQUIRE(tot[0] < (1ull << 32), PTY_EUNSUP, "gaussian field: %llu (Gaussian, tile) pairs exceed 2^32 - 1",
              tot[0]);
  *n_rendered = (int64_t)tot[0];
  return PTY_OK;
}

extern "C" int pty_gs_render_fwd(const PtyGs* d, void* geom, int64_t n_rendered, const float* weights,
                                 void* bin, float* out, hipStream_t stream) {
  int rc = check_desc(d);
  if (rc != PTY_OK) return rc;
  PTY_REQUIRE(geom != nullptr && out != nullptr, PTY_EINVAL, "gaussian field: null argument");
  PTY_REQUIRE(n_rendered >= 0 && n_rendered < (int64_t)(1ull << 32), PTY_EUNSUP,
              "gaussian field: n_rendered %lld outside [0, 2^32)", (long long)n_rendered);
  PTY_REQUIRE(n_rendered == 0 || (bin != nullptr && weights != nullptr), PTY_EINVAL, "gaussian field: null argument");
  const GeomLayout L = geom_layout(d);
  const BinLayout B = bin_layout(n_rendered, 0);
  uint32_t* cursor = at<uint32_t>(geom, L.cursor);
  const uint32_t* start = at<uint32_t>(geom, L.start);
  const uint32_t* count = at<uint32_t>(geom, L.count);
  uint32_t* list_u = n_rendered ? at<uint32_t>(bin, B.list_u) : nullptr;
  uint32_t* list = n_rendered ? at<uint32_t>(bin, B.list) : nullptr;
  if (n_rendered > 0) {
    if (L.n_chunks == 0)
      hipLaunchKernelGGL(k_zero_u32, dim3(grid1(L.tiles) < 4096 ? grid1(L.tiles) : 4096), dim3(256), 0, stream,
                         cursor, (int64_t)L.tiles);
    if (L.n_chunks > 0) {
      uint32_t* hist = at<uint32_t>(geom, L.chunk);
      hipLaunchKernelGGL(k_gs_chunk_hist, dim3(L.n_chunks), dim3(256), 0, stream, d->n, L.chunk_size, L.tiles,
                         L.tiles_x, at<uint32_t>(geom, L.touched), at<int4>(geom, L.rect), hist);
      hipLaunchKernelGGL(k_gs_chunk_prefix, dim3((L.tiles + 3) / 4), dim3(256), 0, stream, L.n_chunks, L.tiles,
                         start, hist);
      hipLaunchKernelGGL(k_gs_chunk_scatter, dim3(L.n_chunks), dim3(256), 0, stream, d->n, L.chunk_size, L.tiles,
                         L.tiles_x, at<uint32_t>(geom, L.touched), at<int4>(geom, L.rect), hist, list_u);
    } else {
      hipLaunchKernelGGL(k_gs_scatter, dim3(grid1(d->n)), dim3(256), 0, stream, d->n, L.tiles_x,
                         at<uint32_t>(geom, L.touched), at<int4>(geom, L.rect), start, cursor, list_u);
    }
    hipLaunchKernelGGL(k_gs_sort, dim3(L.tiles), dim3(256), 0, stream, start, count, list_u, list);
  }
  hipLaunchKernelGGL(k_gs_render, dim3(L.tiles), dim3(256), 0, stream, d->width, d->height, L.tiles_x, start,
                     count, list, at<float2>(geom, L.mean), at<float4>(geom, L.conic),
                     reinterpret_cast<const float2*>(weights), out);
  return check_launch("pty_gs_render_fwd");
}

extern "C" int pty_gs_render_bwd(const PtyGs* d, const float* xy, const float* p0, const float* p1,
                                 const float* weights, const int32_t* radii, const void* geom, int64_t n_rendered,
                                 void* bin, const float* dout, float* dxy, float* dp0, float* dp1, float* dweight,
                                 hipStream_t stream) {
  (void)xy;
  int rc = check_desc(d);
  if (rc != PTY_OK) return rc;
  if (d->n == 0) return PTY_OK;
  PTY_REQUIRE(geom && p0 && p1 && weights && radii && dout && dxy && dp0 && dp1 && dweight, PTY_EINVAL,
              "gaussian field: null argument");
  PTY_REQUIRE(n_rendered >= 0 && n_rendered < (int64_t)(1ull << 32), PTY_EUNSUP,
              "gaussian field: n_rendered %lld outside [0, 2^32)", (long long)n_rendered);
  PTY_REQUIRE(n_rendered == 0 || bin != nullptr, PTY_EINVAL, "gaussian field: null bin buffer");
  const GeomLayout L = geom_layout(d);
  const BinLayout B = bin_layout(n_rendered, 1);
  float4* partial = n_rendered ? at<float4>(bin, B.partial) : nullptr;
  if (n_rendered > 0) {
    hipLaunchKernelGGL(k_gs_render_bwd, dim3(L.tiles), dim3(256), 0, stream, d->width, d->height, L.tiles_x,
                       at<uint32_t>(geom, L.start), at<uint32_t>(geom, L.count), at<uint32_t>(bin, B.list),
                       at<float2>(geom, L.mean), at<float4>(geom, L.conic),
                       reinterpret_cast<const float2*>(weights), at<uint32_t>(geom, L.goff), at<int4>(geom, L.rect),
                       dout, partial);
  }
  const float smax = (float)d->max_patch_radius, smin = d->min_scale;
  if (d->param == PTY_GS_RS)
    hipLaunchKernelGGL(k_gs_finish<PTY_GS_RS>, dim3(grid1(d->n)), dim3(256), 0, stream, d->n, radii,
                       at<uint32_t>(geom, L.goff), at<uint32_t>(geom, L.touched), partial, at<float4>(geom, L.conic),
                       p0, p1, smax, smin, dxy, dp0, dp1, dweight);
  else
    hipLaunchKernelGGL(k_gs_finish<PTY_GS_CHOLESKY>, dim3(grid1(d->n)), dim3(256), 0, stream, d->n, radii,
                       at<uint32_t>(geom, L.goff), at<uint32_t>(geom, L.touched), partial, at<float4>(geom, L.conic),
                       p0, p1, smax, smin, dxy, dp0, dp1, dweight);
  return check_launch("pty_gs_render_bwd");
}
