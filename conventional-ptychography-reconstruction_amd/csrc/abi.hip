// Error reporting, fill, residual and Adam entry points of the C ABI.
#include "common.h"

#include <string>

namespace pty {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(PTY_ELAUNCH, "%s: %s", what, hipGetErrorString(e));
  return PTY_OK;
}

int device_cu_count() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
  return n;
}

__global__ void k_fill(float* __restrict__ p, int64_t n, float v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// dst = half(src), round to nearest even; four values per thread when the
// pointers allow 16-byte loads and 8-byte stores, else one.
__global__ void k_cast_f16_vec(const float4* __restrict__ src, int64_t n4, uint2* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x) {
    const float4 v = src[i];
    __half2 a = __floats2half2_rn(v.x, v.y), b = __floats2half2_rn(v.z, v.w);
    uint2 o;
    o.x = *reinterpret_cast<uint32_t*>(&a);
    o.y = *reinterpret_cast<uint32_t*>(&b);
    dst[i] = o;
  }
}

__global__ void k_cast_f16(const float* __restrict__ src, int64_t lo, int64_t n, __half* __restrict__ dst) {
  for (int64_t i = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = __float2half_rn(src[i]);
}

// out = k * x + initial  (k is a device scalar; complex x, initial, out)
__global__ void k_residual_fwd(const float* __restrict__ k, const float2* __restrict__ x,
                               const float2* __restrict__ init, int64_t n, float2* __restrict__ out) {
  const float kk = *k;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float2 a = x[i], b = init[i];
    out[i] = make_float2(kk * a.x + b.x, kk * a.y + b.y);
  }
}

// gx = k * gout; partial[block] = sum Re(gout * conj(x))
__global__ __launch_bounds__(256) void k_residual_bwd(const float* __restrict__ k, const float2* __restrict__ x,
                                                      const float2* __restrict__ g, int64_t n,
                                                      float2* __restrict__ gx, float* __restrict__ partial) {
  const float kk = *k;
  KSum acc;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float2 a = x[i], b = g[i];
    gx[i] = make_float2(kk * b.x, kk * b.y);
    acc.add(b.x * a.x + b.y * a.y);
  }
  __shared__ float red[256];
  red[threadIdx.x] = acc.value();
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[0] += sum(partial[0..m)) in a fixed order (deterministic).
__global__ void k_sum_into(const float* __restrict__ partial, int m, float* __restrict__ out) {
  __shared__ float red[256];
  KSum acc;
  for (int i = threadIdx.x; i < m; i += blockDim.x) acc.add(partial[i]);
  red[threadIdx.x] = acc.value();
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] += red[0];
}

// torch.optim.Adam over real views (adam_elem in common.h); hyper = [lr, step]
// on the device; pty_adam_tick advances step.  float4
// lanes over n4 = n/4 groups (the flat buffer is padded to 4 floats); the
// consumed gradient is cleared for the next step when clear != 0.
__global__ void k_adam(float4* __restrict__ p, float4* __restrict__ g, float4* __restrict__ m,
                       float4* __restrict__ v, int64_t n4, const float* __restrict__ hyper, float b1, float b2,
                       float eps, int clear) {
  const AdamCoef c = adam_coef(hyper, b1, b2, eps);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x) {
    const float4 gi = g[i];
    float4 mi = m[i], vi = v[i], pi = p[i];
    adam_elem(pi.x, mi.x, vi.x, gi.x, c);
    adam_elem(pi.y, mi.y, vi.y, gi.y, c);
    adam_elem(pi.z, mi.z, vi.z, gi.z, c);
    adam_elem(pi.w, mi.w, vi.w, gi.w, c);
    m[i] = mi;
    v[i] = vi;
    p[i] = pi;
    if (clear) g[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// scalar tail (n not a multiple of 4)
__global__ void k_adam_tail(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t i0, int64_t n, const float* __restrict__ hyper, float b1,
                            float b2, float eps, int clear) {
  const AdamCoef c = adam_coef(hyper, b1, b2, eps);
  const int64_t i = i0 + threadIdx.x;
  if (i >= n) return;
  float pi = p[i], mi = m[i], vi = v[i];
  adam_elem(pi, mi, vi, g[i], c);
  m[i] = mi;
  v[i] = vi;
  p[i] = pi;
  if (clear) g[i] = 0.f;
}

// [M, N, F] -> [F, M, N] with ifftshift over (M, N): out[f, k0, k1] =
// in[(k0 + M/2) % M, (k1 + N/2) % N, f]  (natural FFT order of the shifted
// measurement, so that fftshift(fft2(.)) compares with it pixel by pixel).
__global__ void k_stack_prepare(const float* __restrict__ in, int M, int N, int F, float* __restrict__ out) {
  const int64_t total = (int64_t)M * N * F;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    // t indexes the input contiguously: t = (y*N + x)*F + f
    const int f = (int)(t % F);
    const int64_t yx = t / F;
    const int y = (int)(yx / N), x = (int)(yx % N);
    const int k0 = (y + M - M / 2) % M;
    const int k1 = (x + N - N / 2) % N;
    out[((int64_t)f * M + k0) * N + k1] = in[t];
  }
}

__global__ void k_adam_tick(float* hyper) { hyper[1] += 1.f; }

// Next batch of an epoch's position order (one workgroup): with c = cursor[0]
// and start = c * batch (0, and c = 0, once start + batch would pass n_order),
// idx[k] = order[start + k]; the previous batch's per-pattern losses
// loss[0..batch) go to hist[(c - 1) * batch + k] when c >= 1 and that range
// lies inside hist; then cursor[0] = c + 1 and, with hyper, hyper[1] += 1.
__global__ void k_batch_next(const int32_t* __restrict__ order, int32_t n_order, int32_t* cursor, int32_t batch,
                             int32_t* __restrict__ idx, float* hyper, const float* __restrict__ loss,
                             float* __restrict__ hist, int64_t n_hist) {
  __shared__ int s_c;
  if (threadIdx.x == 0) {
    int c = cursor[0];
    if (c < 0 || (int64_t)c * batch + batch > n_order) c = 0;
    s_c = c;
  }
  __syncthreads();
  const int c = s_c;
  const int64_t start = (int64_t)c * batch;
  for (int k = threadIdx.x; k < batch; k += blockDim.x) idx[k] = order[start + k];
  if (loss && hist && c >= 1 && (int64_t)c * batch <= n_hist)
    for (int k = threadIdx.x; k < batch; k += blockDim.x) hist[(int64_t)(c - 1) * batch + k] = loss[k];
  if (threadIdx.x == 0) {
    cursor[0] = c + 1;
    if (hyper) hyper[1] += 1.f;
  }
}

static unsigned grid_for(int64_t n, int block) {
  int64_t b = (n + block - 1) / block;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace pty

using namespace pty;

extern "C" int pty_abi_version(void) { return PTY_ABI_VERSION; }
extern "C" const char* pty_last_error(void) { return g_last_error.c_str(); }

extern "C" int pty_fill(float* p, int64_t n, float value, hipStream_t stream) {
  PTY_REQUIRE(n >= 0 && (n == 0 || p), PTY_EINVAL, "fill: null pointer");
  if (n == 0) return PTY_OK;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256)), dim3(256), 0, stream, p, n, value);
  return check_launch("fill");
}

extern "C" int pty_cast_f16(const float* src, int64_t n, void* dst, hipStream_t stream) {
  PTY_REQUIRE(n >= 0 && (n == 0 || (src && dst)), PTY_EINVAL, "cast_f16: null pointer");
  if (n == 0) return PTY_OK;
  int64_t done = 0;
  if (((uintptr_t)src & 15) == 0 && ((uintptr_t)dst & 7) == 0 && n >= 4) {
    const int64_t n4 = n / 4;
    hipLaunchKernelGGL(k_cast_f16_vec, dim3(grid_for(n4, 256)), dim3(256), 0, stream, (const float4*)src, n4,
                       (uint2*)dst);
    done = n4 * 4;
  }
  if (done < n)
    hipLaunchKernelGGL(k_cast_f16, dim3(grid_for(n - done, 256)), dim3(256), 0, stream, src, done, n,
                       (__half*)dst);
  return check_launch("cast_f16");
}

extern "C" int pty_residual_fwd(const float* k, const float* x, const float* initial, int64_t n, float* out,
                                hipStream_t stream) {
  PTY_REQUIRE(k && x && initial && out && n > 0, PTY_EINVAL, "residual_fwd: bad arguments");
  hipLaunchKernelGGL(k_residual_fwd, dim3(grid_for(n, 256)), dim3(256), 0, stream, k, (const float2*)x,
                     (const float2*)initial, n, (float2*)out);
  return check_launch("residual_fwd");
}

static int residual_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b > 1024 ? 1024 : (b < 1 ? 1 : b));
}

extern "C" size_t pty_residual_workspace(int64_t n) { return sizeof(float) * (size_t)residual_blocks(n); }

extern "C" int pty_residual_bwd(const float* k, const float* x, const float* gout, int64_t n, float* gx,
                                float* gk, void* work, size_t work_bytes, hipStream_t stream) {
  PTY_REQUIRE(k && x && gout && gx && gk && n > 0, PTY_EINVAL, "residual_bwd: bad arguments");
  const int blocks = residual_blocks(n);
  PTY_REQUIRE(work && work_bytes >= pty_residual_workspace(n), PTY_EWORK, "residual_bwd: workspace too small");
  float* partial = (float*)work;
  hipLaunchKernelGGL(k_residual_bwd, dim3(blocks), dim3(256), 0, stream, k, (const float2*)x,
                     (const float2*)gout, n, (float2*)gx, partial);
  hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(256), 0, stream, partial, blocks, gk);
  return check_launch("residual_bwd");
}

extern "C" int pty_adam(float* param, float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                        const float* hyper, float beta1, float beta2, float eps, int32_t clear_grad,
                        hipStream_t stream) {
  PTY_REQUIRE(param && grad && exp_avg && exp_avg_sq && hyper && n >= 0, PTY_EINVAL, "adam: bad arguments");
  if (n == 0) return PTY_OK;
  const bool aligned = ((((uintptr_t)param) | ((uintptr_t)grad) | ((uintptr_t)exp_avg) | ((uintptr_t)exp_avg_sq)) &
                        15) == 0;
  const int64_t n4 = aligned ? n / 4 : 0;
  if (n4 > 0)
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n4, 256)), dim3(256), 0, stream, (float4*)param, (float4*)grad,
                       (float4*)exp_avg, (float4*)exp_avg_sq, n4, hyper, beta1, beta2, eps, (int)clear_grad);
  for (int64_t i0 = n4 * 4; i0 < n; i0 += 1024)
    hipLaunchKernelGGL(k_adam_tail, dim3(1), dim3(1024), 0, stream, param, grad, exp_avg, exp_avg_sq, i0, n, hyper,
                       beta1, beta2, eps, (int)clear_grad);
  return check_launch("adam");
}

extern "C" int pty_adam_tick(float* hyper, hipStream_t stream) {
  PTY_REQUIRE(hyper, PTY_EINVAL, "adam_tick: null pointer");
  hipLaunchKernelGGL(k_adam_tick, dim3(1), dim3(1), 0, stream, hyper);
  return check_launch("adam_tick");
}

extern "C" int pty_batch_next(const int32_t* order, int32_t n_order, int32_t* cursor, int32_t batch, int32_t* idx,
                              float* hyper, const float* loss, float* hist, int64_t n_hist, hipStream_t stream) {
  PTY_REQUIRE(order && cursor && idx, PTY_EINVAL, "batch_next: null pointer");
  PTY_REQUIRE(batch > 0 && n_order >= batch, PTY_EINVAL, "batch_next: batch %d with an order of %d positions",
              batch, n_order);
  PTY_REQUIRE(n_hist >= 0 && (n_hist == 0 || hist), PTY_EINVAL, "batch_next: bad loss history");
  hipLaunchKernelGGL(k_batch_next, dim3(1), dim3(256), 0, stream, order, n_order, cursor, batch, idx, hyper,
                     hist ? loss : nullptr, hist, n_hist);
  return check_launch("batch_next");
}

extern "C" int pty_stack_prepare(const float* stack_mnf, int32_t M, int32_t N, int32_t frames, float* out,
                                 hipStream_t stream) {
  PTY_REQUIRE(stack_mnf && out && M > 0 && N > 0 && frames > 0, PTY_EINVAL, "stack_prepare: bad arguments");
  const int64_t total = (int64_t)M * N * frames;
  hipLaunchKernelGGL(k_stack_prepare, dim3(grid_for(total, 256)), dim3(256), 0, stream, stack_mnf, M, N, frames, out);
  return check_launch("stack_prepare");
}
