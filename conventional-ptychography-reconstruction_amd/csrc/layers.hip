// Single layers of ComplexMLP.net, for callers that run the network layer by
// layer (models/complex_inr.py:6-40 ComplexGaborLayer2D.forward, :73-78 the
// final complex nn.Linear, :308-314 _get_complex_output's model.net(coords)).
// The trained path evaluates the whole network in one fused kernel
// (inr_f32.hip); these kernels serve the module-level API.
//
//   activation = 1: z = exp(1j * omega * (W x + b))   (ComplexGaborLayer2D)
//   activation = 0: o = W x + b                       (nn.Linear)
//   complex_io = 0: real x, W, b (the first Gabor layer, output complex with
//                   the activation, real without it); 1: complex everything.
//
// Forward: one thread per (point, output unit).  Backward: d lin per (point,
// unit) into the workspace (the pre-activation is recomputed), d x from it,
// and the parameter gradients as per-workgroup partial sums over contiguous
// point ranges, totalled in a fixed order: deterministic, no atomics.
// Gradient convention: PyTorch's for complex tensors (grad = dL/dRe + i dL/dIm),
//   d lin = -i omega conj(z) g  (complex lin), omega Im(conj(z) g) (real lin),
//   d omega = sum Im(conj(z) g conj(lin)),  d x = W^H d lin,
//   d W = sum_p d lin x^H,  d b = sum_p d lin.
#include "common.h"

namespace pty {

constexpr int kLayerThreads = 256;
constexpr int kLayerBlocks = 256;   // partial-sum workgroups of the backward

struct LayerArgs {
  int n_in, n_out, cplx, act;
  const float* w;
  const float* b;
  const float* om;
};

static LayerArgs layer_args(const PtyLayer* d) {
  LayerArgs a;
  a.n_in = d->n_in; a.n_out = d->n_out; a.cplx = d->complex_io; a.act = d->activation;
  a.w = d->weight; a.b = d->bias; a.om = d->omega;
  return a;
}

// lin = W x + b for point p, unit o (imaginary part 0 for real layers)
__device__ __forceinline__ float2 layer_lin(const LayerArgs& a, const float* __restrict__ x, int64_t p, int o) {
  float2 s = make_float2(0.f, 0.f);
  if (a.cplx) {
    const float2* xp = (const float2*)x + p * a.n_in;
    const float2* wo = (const float2*)a.w + (int64_t)o * a.n_in;
    for (int k = 0; k < a.n_in; ++k) s = cadd(s, cmul(wo[k], xp[k]));
    if (a.b) s = cadd(s, ((const float2*)a.b)[o]);
  } else {
    const float* xp = x + p * a.n_in;
    const float* wo = a.w + (int64_t)o * a.n_in;
    for (int k = 0; k < a.n_in; ++k) s.x += wo[k] * xp[k];
    if (a.b) s.x += a.b[o];
  }
  return s;
}

// exp(1j * w * lin) = e^{-w Im lin} (cos(w Re lin), sin(w Re lin))
__device__ __forceinline__ float2 layer_act(float w, float2 lin) {
  float sn, cs;
  sincosf(w * lin.x, &sn, &cs);
  const float m = expf(-w * lin.y);
  return make_float2(m * cs, m * sn);
}

__global__ __launch_bounds__(kLayerThreads) void k_layer_fwd(LayerArgs a, const float* __restrict__ x, int64_t n,
                                                             float* __restrict__ out) {
  const float w = a.act ? *a.om : 0.f;
  const int64_t total = n * a.n_out;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = t / a.n_out;
    const int o = (int)(t - p * a.n_out);
    const float2 lin = layer_lin(a, x, p, o);
    if (a.act) ((float2*)out)[t] = layer_act(w, lin);
    else if (a.cplx) ((float2*)out)[t] = lin;
    else out[t] = lin.x;
  }
}

// D[t] = d lin (real layers: (d lin, 0)); E[t] = d omega term
__global__ __launch_bounds__(kLayerThreads) void k_layer_dlin(LayerArgs a, const float* __restrict__ x, int64_t n,
                                                              const float* __restrict__ gout, float2* __restrict__ D,
                                                              float* __restrict__ E) {
  const float w = a.act ? *a.om : 0.f;
  const int64_t total = n * a.n_out;
  const bool cout = a.act || a.cplx;   // complex output
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const float2 g = cout ? ((const float2*)gout)[t] : make_float2(gout[t], 0.f);
    if (!a.act) {
      D[t] = g;
      continue;
    }
    const int64_t p = t / a.n_out;
    const int o = (int)(t - p * a.n_out);
    const float2 lin = layer_lin(a, x, p, o);
    const float2 du = cmulc(g, layer_act(w, lin));   // g conj(z)
    D[t] = a.cplx ? make_float2(w * du.y, -w * du.x) : make_float2(w * du.y, 0.f);
    E[t] = du.y * lin.x - du.x * lin.y;                // Im(du conj(lin))
  }
}

__global__ __launch_bounds__(kLayerThreads) void k_layer_dx(LayerArgs a, int64_t n, const float2* __restrict__ D,
                                                            float* __restrict__ gx) {
  const int64_t total = n * a.n_in;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = t / a.n_in;
    const int k = (int)(t - p * a.n_in);
    const float2* dp = D + p * a.n_out;
    if (a.cplx) {
      float2 s = make_float2(0.f, 0.f);
      for (int o = 0; o < a.n_out; ++o) s = cadd(s, cmulc(dp[o], ((const float2*)a.w)[(int64_t)o * a.n_in + k]));
      ((float2*)gx)[t] = s;
    } else {
      float s = 0.f;
      for (int o = 0; o < a.n_out; ++o) s += a.w[(int64_t)o * a.n_in + k] * dp[o].x;
      gx[t] = s;
    }
  }
}

// Partial sums of workgroup b over points [b*chunk, (b+1)*chunk): slots
// [dW (n_out*n_in, x2 complex)] [db (n_out, x2 complex)] [d omega]
__global__ __launch_bounds__(kLayerThreads) void k_layer_partials(LayerArgs a, const float* __restrict__ x, int64_t n,
                                                                  int64_t chunk, const float2* __restrict__ D,
                                                                  const float* __restrict__ E,
                                                                  float* __restrict__ partial, int slots) {
  const int64_t p0 = (int64_t)blockIdx.x * chunk;
  const int64_t p1 = p0 + chunk < n ? p0 + chunk : n;
  float* out = partial + (int64_t)blockIdx.x * slots;
  const int nw = a.n_out * a.n_in;
  for (int s = threadIdx.x; s < nw + a.n_out + 1; s += blockDim.x) {
    float2 acc = make_float2(0.f, 0.f);
    if (s < nw) {
      const int o = s / a.n_in, k = s - o * a.n_in;
      if (a.cplx) {
        for (int64_t p = p0; p < p1; ++p) acc = cadd(acc, cmulc(D[p * a.n_out + o], ((const float2*)x)[p * a.n_in + k]));
      } else {
        for (int64_t p = p0; p < p1; ++p) acc.x += D[p * a.n_out + o].x * x[p * a.n_in + k];
      }
      out[2 * s] = acc.x;
      out[2 * s + 1] = acc.y;
    } else if (s < nw + a.n_out) {
      const int o = s - nw;
      for (int64_t p = p0; p < p1; ++p) acc = cadd(acc, D[p * a.n_out + o]);
      out[2 * s] = acc.x;
      out[2 * s + 1] = acc.y;
    } else {
      float e = 0.f;
      if (a.act)
        for (int64_t t = p0 * a.n_out; t < p1 * a.n_out; ++t) e += E[t];
      out[2 * s] = e;
      out[2 * s + 1] = 0.f;
    }
  }
}

// Totals over the workgroups in order; writes the parameter gradients.
__global__ void k_layer_finish(LayerArgs a, const float* __restrict__ partial, int blocks, int slots,
                               float* __restrict__ gw, float* __restrict__ gb, float* __restrict__ gom) {
  const int nw = a.n_out * a.n_in;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nw + a.n_out + 1; s += gridDim.x * blockDim.x) {
    float re = 0.f, im = 0.f;
    for (int b = 0; b < blocks; ++b) {
      re += partial[(int64_t)b * slots + 2 * s];
      im += partial[(int64_t)b * slots + 2 * s + 1];
    }
    if (s < nw) {
      if (!gw) continue;
      if (a.cplx) { gw[2 * s] = re; gw[2 * s + 1] = im; } else gw[s] = re;
    } else if (s < nw + a.n_out) {
      if (!gb) continue;
      const int o = s - nw;
      if (a.cplx) { gb[2 * o] = re; gb[2 * o + 1] = im; } else gb[o] = re;
    } else if (gom) {
      gom[0] = re;
    }
  }
}

static int check_layer(const PtyLayer* d, const char* what) {
  PTY_REQUIRE(d, PTY_EINVAL, "%s: null descriptor", what);
  PTY_REQUIRE(d->n_in >= 1 && d->n_in <= 4096 && d->n_out >= 1 && d->n_out <= 4096, PTY_EUNSUP,
              "%s: %d -> %d features (1..4096 each)", what, d->n_in, d->n_out);
  PTY_REQUIRE(d->weight, PTY_EINVAL, "%s: null weight", what);
  PTY_REQUIRE(!d->activation || d->omega, PTY_EINVAL, "%s: the Gabor activation needs omega", what);
  return PTY_OK;
}

struct LayerWork {
  size_t D, E, partial, total;
  int64_t chunk;
  int blocks, slots;
};

static LayerWork layer_work(const PtyLayer* d, int64_t n) {
  LayerWork L;
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
  L.blocks = (int)(n < kLayerBlocks ? (n > 0 ? n : 1) : kLayerBlocks);
  L.chunk = (n + L.blocks - 1) / L.blocks;
  L.slots = 2 * (d->n_out * d->n_in + d->n_out + 1);
  L.D = take((size_t)n * d->n_out * sizeof(float2));
  L.E = take((size_t)n * d->n_out * sizeof(float));
  L.partial = take((size_t)L.blocks * L.slots * sizeof(float));
  L.total = off;
  return L;
}

static unsigned grid_for(int64_t total) {
  int64_t b = (total + kLayerThreads - 1) / kLayerThreads;
  if (b > 8192) b = 8192;
  return (unsigned)(b > 0 ? b : 1);
}

}  // namespace pty

using namespace pty;

extern "C" int pty_layer_fwd(const PtyLayer* d, const float* x, int64_t n, float* out, hipStream_t stream) {
  int rc = check_layer(d, "layer_fwd");
  if (rc) return rc;
  PTY_REQUIRE(n >= 0, PTY_EINVAL, "layer_fwd: n < 0");
  if (n == 0) return PTY_OK;
  PTY_REQUIRE(x && out, PTY_EINVAL, "layer_fwd: null pointer");
  const LayerArgs a = layer_args(d);
  hipLaunchKernelGGL(k_layer_fwd, dim3(grid_for(n * d->n_out)), dim3(kLayerThreads), 0, stream, a, x, n, out);
  return check_launch("layer_fwd");
}

extern "C" size_t pty_layer_bwd_workspace(const PtyLayer* d, int64_t n) {
  if (check_layer(d, "layer_bwd_workspace") || n < 0) return 0;
  return layer_work(d, n).total;
}

extern "C" int pty_layer_bwd(const PtyLayer* d, const float* x, int64_t n, const float* gout, float* gx,
                             float* gweight, float* gbias, float* gomega, void* work, size_t work_bytes,
                             hipStream_t stream) {
  int rc = check_layer(d, "layer_bwd");
  if (rc) return rc;
  PTY_REQUIRE(n >= 1, PTY_EINVAL, "layer_bwd: n = %lld", (long long)n);
  PTY_REQUIRE(x && gout, PTY_EINVAL, "layer_bwd: null pointer");
  const LayerWork L = layer_work(d, n);
  PTY_REQUIRE(work && work_bytes >= L.total, PTY_EWORK, "layer_bwd: workspace %zu < %zu", work_bytes, L.total);
  const LayerArgs a = layer_args(d);
  char* w = (char*)work;
  float2* D = (float2*)(w + L.D);
  float* E = (float*)(w + L.E);
  float* partial = (float*)(w + L.partial);
  hipLaunchKernelGGL(k_layer_dlin, dim3(grid_for(n * d->n_out)), dim3(kLayerThreads), 0, stream, a, x, n, gout, D, E);
  if (gx)
    hipLaunchKernelGGL(k_layer_dx, dim3(grid_for(n * d->n_in)), dim3(kLayerThreads), 0, stream, a, n, D, gx);
  if (gweight || gbias || gomega) {
    hipLaunchKernelGGL(k_layer_partials, dim3(L.blocks), dim3(kLayerThreads), 0, stream, a, x, n, L.chunk, D, E,
                       partial, L.slots);
    const int nslots = d->n_out * d->n_in + d->n_out + 1;
    hipLaunchKernelGGL(k_layer_finish, dim3((nslots + 255) / 256), dim3(256), 0, stream, a, partial, L.blocks,
                       L.slots, gweight, gbias, gomega);
  }
  return check_launch("layer_bwd");
}
