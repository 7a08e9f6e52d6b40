// Haar wavelet sparsity of the complex object canvas and its adjoint.
//
// Reference: haar_wavelet_sparsity_loss (utils/loss.py:60-83), called once per
// step on the whole object at Recovery.py:309-310 with reduction='sum' and
// weighted by sparsity_weight.  For every 2x2 block
//     a b
//     c d          (rows 2r, 2r+1; columns 2s, 2s+1)
// of the real part and of the imaginary part:
//     lh = a - b + c - d,  hl = a + b - c - d,  hh = a - b - c + d
//     loss = sum over blocks and parts of |lh| + |hl| + |hh|
// The gradient of |x| is sign(x) with sign(0) = 0 (torch.abs backward), so
//     dL/da =  s_lh + s_hl + s_hh      dL/db = -s_lh + s_hl - s_hh
//     dL/dc =  s_lh - s_hl - s_hh      dL/dd = -s_lh - s_hl + s_hh
// and the complex gradient is dL/dRe + i dL/dIm (PyTorch's convention).
//
// One thread per block: two 16-byte loads (a b, c d) per part pair, the
// gradient added to the caller's object gradient with two 16-byte
// read-modify-writes.  The loss is summed per workgroup (compensated per
// thread, fixed tree per workgroup) and the workgroup partials are summed in
// a fixed order by one workgroup: the result is deterministic.
#include "common.h"

namespace pty {

constexpr int kHaarThreads = 256;
constexpr int kHaarMaxBlocks = 2048;

__device__ __forceinline__ float sgnf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }

// one part (real or imaginary) of one 2x2 block: adds |lh|+|hl|+|hh| to acc and
// returns the four gradient coefficients
__device__ __forceinline__ void haar_part(float a, float b, float c, float d, KSum& acc, float& ga, float& gb,
                                          float& gc, float& gd) {
  const float lh = a - b + c - d;
  const float hl = a + b - c - d;
  const float hh = a - b - c + d;
  acc.add(fabsf(lh));
  acc.add(fabsf(hl));
  acc.add(fabsf(hh));
  const float slh = sgnf(lh), shl = sgnf(hl), shh = sgnf(hh);
  ga = slh + shl + shh;
  gb = -slh + shl - shh;
  gc = slh - shl - shh;
  gd = -slh - shl + shh;
}

// obj [rows, W] c64 with rows, W even; partial[blockIdx.x] = sum of the
// block's terms; when dobj != NULL, dobj += wgrad * dL/dobj.
__global__ __launch_bounds__(kHaarThreads) void k_haar(const float4* __restrict__ obj, int64_t rows, int W,
                                                       float wgrad, float4* __restrict__ dobj,
                                                       float* __restrict__ partial) {
  const int64_t bw = W / 2;                  // blocks per block-row
  const int64_t nblk = (rows / 2) * bw;
  const int64_t row4 = W / 2;                // float4 (two complex) per canvas row
  KSum acc;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nblk; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / bw, s = t - r * bw;
    const int64_t i0 = (2 * r) * row4 + s;   // (a, b) of row 2r
    const int64_t i1 = i0 + row4;            // (c, d) of row 2r+1
    const float4 ab = obj[i0], cd = obj[i1];
    float ra, rb, rc, rd, ia, ib, ic, id;
    haar_part(ab.x, ab.z, cd.x, cd.z, acc, ra, rb, rc, rd);
    haar_part(ab.y, ab.w, cd.y, cd.w, acc, ia, ib, ic, id);
    if (dobj) {
      float4 g0 = dobj[i0], g1 = dobj[i1];
      g0.x += wgrad * ra;
      g0.y += wgrad * ia;
      g0.z += wgrad * rb;
      g0.w += wgrad * ib;
      g1.x += wgrad * rc;
      g1.y += wgrad * ic;
      g1.z += wgrad * rd;
      g1.w += wgrad * id;
      dobj[i0] = g0;
      dobj[i1] = g1;
    }
  }
  __shared__ float red[kHaarThreads];
  red[threadIdx.x] = acc.value();
  __syncthreads();
  for (int st = kHaarThreads / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// loss[0] = scale * sum(partial[0..m)) in a fixed order.
__global__ __launch_bounds__(kHaarThreads) void k_haar_final(const float* __restrict__ partial, int m, float scale,
                                                             float* __restrict__ loss) {
  __shared__ float red[kHaarThreads];
  KSum acc;
  for (int i = threadIdx.x; i < m; i += blockDim.x) acc.add(partial[i]);
  red[threadIdx.x] = acc.value();
  __syncthreads();
  for (int st = kHaarThreads / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = scale * red[0];
}

static int haar_blocks(int64_t rows, int64_t W) {
  const int64_t nblk = (rows / 2) * (W / 2);
  int64_t b = (nblk + kHaarThreads - 1) / kHaarThreads;
  if (b > kHaarMaxBlocks) b = kHaarMaxBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace pty

using namespace pty;

extern "C" size_t pty_haar_workspace(int64_t rows, int32_t W) {
  if (rows <= 0 || W <= 0) return 0;
  return sizeof(float) * (size_t)haar_blocks(rows, W);
}

extern "C" int pty_haar_sparsity(const float* object, int64_t rows, int32_t W, float scale, float grad_weight,
                                 float* loss, float* dobject, void* work, size_t work_bytes, hipStream_t stream) {
  PTY_REQUIRE(object && loss && work, PTY_EINVAL, "haar_sparsity: null pointer");
  PTY_REQUIRE(rows > 0 && W > 0, PTY_EINVAL, "haar_sparsity: empty image %lld x %d", (long long)rows, W);
  PTY_REQUIRE(rows % 2 == 0 && W % 2 == 0, PTY_EUNSUP,
              "haar_sparsity: %lld x %d has an odd side (the 2x2 sub-bands need even sizes)", (long long)rows, W);
  PTY_REQUIRE(((uintptr_t)object & 15u) == 0 && (!dobject || ((uintptr_t)dobject & 15u) == 0), PTY_EINVAL,
              "haar_sparsity: object / dobject must be 16-byte aligned");
  const int blocks = haar_blocks(rows, W);
  PTY_REQUIRE(work_bytes >= sizeof(float) * (size_t)blocks, PTY_EWORK, "haar_sparsity: workspace %zu < %zu",
              work_bytes, sizeof(float) * (size_t)blocks);
  float* partial = (float*)work;
  hipLaunchKernelGGL(k_haar, dim3(blocks), dim3(kHaarThreads), 0, stream, (const float4*)object, rows, (int)W,
                     grad_weight * scale, (float4*)dobject, partial);
  hipLaunchKernelGGL(k_haar_final, dim3(1), dim3(kHaarThreads), 0, stream, (const float*)partial, blocks, scale,
                     loss);
  return check_launch("haar_sparsity");
}
