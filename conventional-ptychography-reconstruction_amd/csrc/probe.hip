// Probe centre-of-mass roll and energy normalisation (Recovery.py:262-285),
// forward and adjoint, with no host synchronisation.
//
// Forward: r_i = sum_j |p_ij|^2, c_j = sum_i |p_ij|^2, E = sum r_i.
//   The reference's cp = trunc([M,N]/2 - [M,N]*mean(cumsum(r or c))/E + 1)
//   equals trunc(S/E - size/2 + 1) with S = sum_i i*r_i (resp. sum_j j*c_j),
//   since mean_i cumsum(r)_i = E - S/M.  No roll when E == 0 (or not finite).
//   q = roll(p, -cp); n = s*q with s = sqrt(M*N/E) when E > 0, else n = q.
// Adjoint (cp is a constant, s depends on p through E):
//   g_q = s*g_n - (s/E) * Re<q, g_n> * q;  g_p = roll(g_q, +cp).
#include "common.h"

namespace pty {

constexpr int kProbeBlocks = 256;

// i mod n for i in [0, 2n) (a rolled index: i < n plus a shift in [0, n))
__device__ __forceinline__ int wrap(int i, int n) { return i < n ? i : i - n; }

// The pixel loops are grid-strided with kPix pixels per thread and batch:
// the batch's loads are issued together, then its pixels are processed in
// order (the grids give at most kPix pixels per thread up to 512 x 512).
constexpr int kPix = 4;

__global__ __launch_bounds__(256) void k_probe_stats(const float2* __restrict__ p, int M, int N,
                                                     double* __restrict__ part) {
  __shared__ double red[3][256];
  double e = 0.0, sy = 0.0, sx = 0.0;
  const int total = M * N, stride = gridDim.x * blockDim.x;
  for (int t0 = blockIdx.x * blockDim.x + threadIdx.x; t0 < total; t0 += kPix * stride) {
    float2 v[kPix];
#pragma unroll
    for (int u = 0; u < kPix; ++u) v[u] = p[min(t0 + u * stride, total - 1)];
#pragma unroll
    for (int u = 0; u < kPix; ++u) {
      const int t = t0 + u * stride;
      if (t < total) {
        const double a = (double)(v[u].x * v[u].x + v[u].y * v[u].y);
        const int i = t / N, j = t - i * N;
        e += a;
        sy += a * i;
        sx += a * j;
      }
    }
  }
  red[0][threadIdx.x] = e;
  red[1][threadIdx.x] = sy;
  red[2][threadIdx.x] = sx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[k * gridDim.x + blockIdx.x] = red[k][0];
}

struct ProbeNorm {
  int cp0, cp1;
  float s, E;
  bool roll, scale;
};

// Sum of the nb (<= kProbeBlocks) partials part[0..nb) by the first
// wavefront: lane l adds partials l, l + 64, ... (its loads all in flight
// together), then a fixed xor butterfly; the total is returned on every lane
// of that wavefront.  The order is fixed, so the sum is deterministic.
__device__ __forceinline__ double wave_sum_partials(const double* part, int nb) {
  const int lane = threadIdx.x & 63;
  double v[kProbeBlocks / 64];
#pragma unroll
  for (int u = 0; u < kProbeBlocks / 64; ++u) {
    const int b = lane + 64 * u;
    v[u] = b < nb ? part[b] : 0.0;
  }
  double acc = v[0];
#pragma unroll
  for (int u = 1; u < kProbeBlocks / 64; ++u) acc += v[u];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  return acc;
}

// Every block reduces the partials itself (a few hundred doubles) so no
// extra launch is needed.
__device__ ProbeNorm probe_norm_from(const double* part, int nb, int M, int N) {
  __shared__ double tot[3];
  if (threadIdx.x < 64) {
    const double e = wave_sum_partials(part, nb), sy = wave_sum_partials(part + nb, nb),
                 sx = wave_sum_partials(part + 2 * nb, nb);
    if (threadIdx.x == 0) {
      tot[0] = e;
      tot[1] = sy;
      tot[2] = sx;
    }
  }
  __syncthreads();
  ProbeNorm r;
  const double E = tot[0];
  r.E = (float)E;
  r.cp0 = r.cp1 = 0;
  r.roll = false;
  r.scale = false;
  r.s = 1.f;
  if (E > 0.0 && isfinite(E)) {
    const double c0 = tot[1] / E - 0.5 * M + 1.0;
    const double c1 = tot[2] / E - 0.5 * N + 1.0;
    const long long t0 = (long long)trunc(c0), t1 = (long long)trunc(c1);
    // roll is circular: reduce to [0, size)
    r.cp0 = (int)(((t0 % M) + M) % M);
    r.cp1 = (int)(((t1 % N) + N) % N);
    r.roll = (t0 != 0 || t1 != 0);
    r.scale = true;
    r.s = (float)sqrt((double)M * (double)N / E);
  }
  return r;
}

__global__ __launch_bounds__(256) void k_probe_apply(const float2* __restrict__ p, int M, int N,
                                                     const double* __restrict__ part, int nb,
                                                     float2* __restrict__ out, float* __restrict__ stats) {
  const ProbeNorm pn = probe_norm_from(part, nb, M, N);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    stats[0] = (float)pn.cp0;
    stats[1] = (float)pn.cp1;
    stats[2] = pn.s;
    stats[3] = pn.E;
    stats[4] = pn.roll ? 1.f : 0.f;
    stats[5] = pn.scale ? 1.f : 0.f;
    stats[6] = 0.f;
    stats[7] = 0.f;
  }
  const int total = M * N, stride = gridDim.x * blockDim.x;
  for (int t0 = blockIdx.x * blockDim.x + threadIdx.x; t0 < total; t0 += kPix * stride) {
    float2 v[kPix];
#pragma unroll
    for (int u = 0; u < kPix; ++u) {
      const int t = min(t0 + u * stride, total - 1);
      const int i = t / N, j = t - i * N;
      v[u] = p[wrap(i + pn.cp0, M) * N + wrap(j + pn.cp1, N)];
    }
#pragma unroll
    for (int u = 0; u < kPix; ++u) {
      const int t = t0 + u * stride;
      if (t < total) out[t] = cscale(v[u], pn.s);
    }
  }
}

// partial of Re<q, g_n> = sum Re(conj(g_n) * q), q = roll(p, -cp)
__global__ __launch_bounds__(256) void k_probe_bwd_dot(const float2* __restrict__ p, const float2* __restrict__ gn,
                                                       int M, int N, const float* __restrict__ stats,
                                                       double* __restrict__ part) {
  __shared__ double red[256];
  const int cp0 = (int)stats[0], cp1 = (int)stats[1];
  double acc = 0.0;
  const int total = M * N, stride = gridDim.x * blockDim.x;
  for (int t0 = blockIdx.x * blockDim.x + threadIdx.x; t0 < total; t0 += kPix * stride) {
    float2 q[kPix], g[kPix];
#pragma unroll
    for (int u = 0; u < kPix; ++u) {
      const int t = min(t0 + u * stride, total - 1);
      const int i = t / N, j = t - i * N;
      q[u] = p[wrap(i + cp0, M) * N + wrap(j + cp1, N)];
      g[u] = gn[t];
    }
#pragma unroll
    for (int u = 0; u < kPix; ++u)
      if (t0 + u * stride < total) acc += (double)(g[u].x * q[u].x + g[u].y * q[u].y);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_probe_bwd_apply(const float2* __restrict__ p, const float2* __restrict__ gn,
                                                         int M, int N, const float* __restrict__ stats,
                                                         const double* __restrict__ part, int nb,
                                                         float2* __restrict__ gp) {
  __shared__ double dot;
  if (threadIdx.x < 64) {
    const double acc = wave_sum_partials(part, nb);
    if (threadIdx.x == 0) dot = acc;
  }
  __syncthreads();
  const int cp0 = (int)stats[0], cp1 = (int)stats[1];
  const bool scale = stats[5] != 0.f;
  const float s = stats[2], E = stats[3];
  const float coef = scale ? (float)(dot * (double)s / (double)E) : 0.f;
  const int total = M * N, stride = gridDim.x * blockDim.x;
  for (int t0 = blockIdx.x * blockDim.x + threadIdx.x; t0 < total; t0 += kPix * stride) {
    float2 q[kPix], g[kPix];
    int src[kPix];
#pragma unroll
    for (int u = 0; u < kPix; ++u) {
      const int t = min(t0 + u * stride, total - 1);
      const int i = t / N, j = t - i * N;
      src[u] = wrap(i + cp0, M) * N + wrap(j + cp1, N);
      q[u] = p[src[u]];
      g[u] = gn[t];
    }
#pragma unroll
    for (int u = 0; u < kPix; ++u) {
      if (t0 + u * stride >= total) continue;
      float2 r;
      if (scale) r = make_float2(s * g[u].x - coef * q[u].x, s * g[u].y - coef * q[u].y);
      else r = g[u];
      gp[src[u]] = r;
    }
  }
}

static int probe_blocks(int M, int N) {
  int b = (M * N + 1023) / 1024;
  return b > kProbeBlocks ? kProbeBlocks : (b < 1 ? 1 : b);
}

static int check_probe(int M, int N) {
  PTY_REQUIRE(M >= 2 && N >= 2 && M <= 4096 && N <= 4096, PTY_EUNSUP, "probe: %dx%d outside [2, 4096]", M, N);
  return PTY_OK;
}

}  // namespace pty

using namespace pty;

extern "C" size_t pty_probe_workspace(int32_t M, int32_t N) {
  return sizeof(double) * 3 * (size_t)probe_blocks(M, N);
}

extern "C" int pty_probe_normalize_fwd(const float* probe, int32_t M, int32_t N, float* probe_n, float* stats,
                                       void* work, size_t work_bytes, hipStream_t stream) {
  int rc = check_probe(M, N);
  if (rc) return rc;
  PTY_REQUIRE(probe && probe_n && stats, PTY_EINVAL, "probe_normalize_fwd: null pointer");
  PTY_REQUIRE(work && work_bytes >= pty_probe_workspace(M, N), PTY_EWORK, "probe_normalize_fwd: workspace too small");
  const int nb = probe_blocks(M, N);
  double* part = (double*)work;
  hipLaunchKernelGGL(k_probe_stats, dim3(nb), dim3(256), 0, stream, (const float2*)probe, M, N, part);
  hipLaunchKernelGGL(k_probe_apply, dim3(nb), dim3(256), 0, stream, (const float2*)probe, M, N,
                     (const double*)part, nb, (float2*)probe_n, stats);
  return check_launch("probe_normalize_fwd");
}

extern "C" int pty_probe_normalize_bwd(const float* probe, const float* gprobe_n, int32_t M, int32_t N,
                                       const float* stats, float* gprobe, void* work, size_t work_bytes,
                                       hipStream_t stream) {
  int rc = check_probe(M, N);
  if (rc) return rc;
  PTY_REQUIRE(probe && gprobe_n && stats && gprobe, PTY_EINVAL, "probe_normalize_bwd: null pointer");
  PTY_REQUIRE(work && work_bytes >= pty_probe_workspace(M, N), PTY_EWORK, "probe_normalize_bwd: workspace too small");
  const int nb = probe_blocks(M, N);
  double* part = (double*)work;
  hipLaunchKernelGGL(k_probe_bwd_dot, dim3(nb), dim3(256), 0, stream, (const float2*)probe,
                     (const float2*)gprobe_n, M, N, stats, part);
  hipLaunchKernelGGL(k_probe_bwd_apply, dim3(nb), dim3(256), 0, stream, (const float2*)probe,
                     (const float2*)gprobe_n, M, N, stats, (const double*)part, nb, (float2*)gprobe);
  return check_launch("probe_normalize_bwd");
}
