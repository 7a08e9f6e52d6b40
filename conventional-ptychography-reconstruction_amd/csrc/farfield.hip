// Far-field forward model, the three losses and their adjoint for one batch
// of scan positions (reference: Recovery.py:287-307; same math as
// utils/optics.py:102-112).
//
// For position p = batch_idx[b] with window (tlY, tlX):
//   psi = probe_n * object[tlY:tlY+M, tlX:tlX+N]
//   Psi = fft2(psi)   (natural order; the measured stack is stored
//                      ifftshifted, so fftshift never has to move data)
//   I = |Psi|^2,  loss_b = mean over pixels of l(I, meas_p)
// Adjoint of sum_b loss_b:
//   dPsi = 2 dI Psi,  dpsi = conj-DFT(dPsi) (= M*N*ifft2),
//   dprobe += dpsi * conj(patch),  dobject[window] += dpsi * conj(probe_n)
// Pass structure (LDS Stockham FFTs):
//   F1 rows  : Psi1[b, r, :] = FFT_N(probe_n[r,:] * patch[r,:])
//   F2 cols  : FFT_M; smooth-L1 / Poisson: loss, dI, dPsi, IFFT_M in place
//              FD: store Psi and I, then FD1 (loss, gx, gy), FD2 (dPsi), F2i (IFFT_M)
//   F3 rows  : per (row block, b): dpsi = IFFT_N(dPsi rows), in place; the
//              first row block of each pattern also sums its loss partials
//   F4       : dprobe = sum_b dpsi_b * conj(patch_b)             (fixed order)
//   F5       : dobject[px] = sum over windows covering px of dpsi_b * conj(probe_n)
//              (a per-pixel gather in batch order: no atomics, deterministic)
//              F4 and F5 share one launch.
#include "common.h"

#include <mutex>

namespace pty {

struct FarArgs {
  int M, N, lgM, lgN, H, W, B, F, loss;
  const float2* obj;
  const float2* probe;
  const float* stack;     // [F, M, N] natural order
  const int* bidx;        // [B]
  const int* tl;          // [F, 2] (tlY, tlX)
  float2* psi;            // [B, M, N] workspace
  float* inten;           // [B, M, N] (FD / intensity output)
  float* gx;              // [B, M, N] (FD)
  float* gy;              // [B, M, N] (FD)
  float* partial;         // [B, P]
  int P;                  // partial sums per pattern
  float* loss_out;        // [B]
  float2* dobj;
  float2* dprobe;
};

// Window of batch entry b; returns false when the entry is invalid (index or
// window out of range): such entries contribute nothing.
__device__ __forceinline__ bool window_of(const FarArgs& a, int b, int& ty, int& tx, int& p) {
  p = a.bidx[b];
  if (p < 0 || p >= a.F) return false;
  ty = a.tl[2 * p];
  tx = a.tl[2 * p + 1];
  return ty >= 0 && tx >= 0 && ty + a.M <= a.H && tx + a.N <= a.W;
}

// F1: rows.  lines rows of the flattened (b, r) index per workgroup.
__global__ __launch_bounds__(256) void k_far_rows_fwd(FarArgs a, int lines) {
  extern __shared__ float2 smem[];
  const int N = a.N;
  float2* tw = smem;
  float2* x = tw + N;
  float2* y = x + lines * N;
  const int tid = threadIdx.x, nt = blockDim.x;
  fill_twiddles(tw, N, tid, nt);
  const int row0 = blockIdx.x * lines;
  const int total_rows = a.B * a.M;
  const int lgN = a.lgN;
  for (int t = tid; t < lines * N; t += nt) {
    const int li = t >> lgN, c = t & (N - 1);
    const int row = row0 + li;
    float2 v = make_float2(0.f, 0.f);
    if (row < total_rows) {
      const int b = row / a.M, r = row - b * a.M;
      int ty, tx, p;
      if (window_of(a, b, ty, tx, p))
        v = cmul(a.probe[r * N + c], a.obj[(int64_t)(ty + r) * a.W + tx + c]);
    }
    x[t] = v;
  }
  __syncthreads();
  float2* res = fft_lines<-1>(x, y, N, a.lgN, lines, N, tw, 1, tid, nt);
  for (int t = tid; t < lines * N; t += nt) {
    const int row = row0 + (t >> lgN);
    if (row < total_rows) a.psi[(int64_t)row * N + (t & (N - 1))] = res[t];
  }
}

__device__ __forceinline__ float smooth_l1(float d, float& g) {
  const float ad = fabsf(d);
  if (ad < 1.f) { g = d; return 0.5f * d * d; }
  g = d > 0.f ? 1.f : -1.f;
  return ad - 0.5f;
}

template <int NT>
__device__ __forceinline__ float block_sum(float v, float* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  float r = red[0];
  __syncthreads();
  return r;
}

// F2: columns of pattern b = blockIdx.y, C columns per workgroup of
// kColThreads threads.
constexpr int kColThreads = 512;
// MODE 0: pointwise loss (smooth-L1 / Poisson) fused with dPsi and IFFT_M.
// MODE 1: forward only, store natural-order Psi (in place) and I.
// MODE 2: inverse only (FD path): IFFT_M of the stored dPsi, in place.
template <int MODE>
__global__ __launch_bounds__(kColThreads) void k_far_cols(FarArgs a, int C) {
  extern __shared__ float2 smem[];
  __shared__ float red[kColThreads];
  const int M = a.M, N = a.N, S = M + 1;
  float2* tw = smem;
  float2* x = tw + M;
  float2* y = x + C * S;
  const int tid = threadIdx.x, nt = blockDim.x;
  fill_twiddles(tw, M, tid, nt);
  const int b = blockIdx.y;
  const int c0 = blockIdx.x * C;
  const int lgC = __builtin_ctz(C);   // C is a power of two
  int ty, tx, p;
  const bool valid = window_of(a, b, ty, tx, p);
  float2* base = a.psi + (int64_t)b * M * N;
  for (int t = tid; t < M * C; t += nt) {
    const int r = t >> lgC, c = t & (C - 1);
    x[c * S + r] = base[(int64_t)r * N + c0 + c];
  }
  __syncthreads();
  if (MODE == 2) {
    float2* res = fft_lines<+1>(x, y, M, a.lgM, C, S, tw, 1, tid, nt);
    for (int t = tid; t < M * C; t += nt) {
      const int r = t >> lgC, c = t & (C - 1);
      base[(int64_t)r * N + c0 + c] = res[c * S + r];
    }
    return;
  }
  float2* f = fft_lines<-1>(x, y, M, a.lgM, C, S, tw, 1, tid, nt);
  if (MODE == 1) {
    for (int t = tid; t < M * C; t += nt) {
      const int r = t >> lgC, c = t & (C - 1);
      const float2 v = f[c * S + r];
      base[(int64_t)r * N + c0 + c] = v;
      a.inten[(int64_t)b * M * N + (int64_t)r * N + c0 + c] = v.x * v.x + v.y * v.y;
    }
    return;
  }
  // MODE 0
  const float inv_mn = 1.f / ((float)M * (float)N);
  const float* meas = a.stack + (int64_t)(valid ? p : 0) * M * N;
  KSum acc;
  for (int t = tid; t < M * C; t += nt) {
    const int r = t >> lgC, c = t & (C - 1);   // consecutive threads -> consecutive columns
    float2 v = f[c * S + r];
    float2 gv = make_float2(0.f, 0.f);
    if (valid) {
      const float I = v.x * v.x + v.y * v.y;
      const float m = meas[(int64_t)r * N + c0 + c];
      float g;
      if (a.loss == PTY_LOSS_SMOOTH_L1) {
        acc.add(smooth_l1(I - m, g));
      } else {  // Poisson
        const float e = I + 1e-8f;
        acc.add(I - m * logf(e));
        g = 1.f - m / e;
      }
      gv = cscale(v, 2.f * g * inv_mn);
    }
    f[c * S + r] = gv;
  }
  __syncthreads();
  const float s = block_sum<kColThreads>(acc.value(), red);
  if (tid == 0) a.partial[b * a.P + blockIdx.x] = s;
  float2* g2 = (f == x) ? y : x;
  float2* res = fft_lines<+1>(f, g2, M, a.lgM, C, S, tw, 1, tid, nt);
  for (int t = tid; t < M * C; t += nt) {
    const int r = t >> lgC, c = t & (C - 1);
    base[(int64_t)r * N + c0 + c] = res[c * S + r];
  }
}

// FD loss (Recovery.py:296-302) on natural-order pixels: the circular
// neighbours of the shifted array are the circular neighbours here too.
// FD1: dW = sqrt(m+eps) - sqrt(I+eps); Rx = dW[., x+1] - dW; Ry = dW[y+1, .] - dW;
//      G = sqrt(Rx^2 + Ry^2 + eps); loss += G; gx = Rx/(G MN); gy = Ry/(G MN).
__global__ __launch_bounds__(256) void k_far_fd1(FarArgs a) {
  __shared__ float red[256];
  const int M = a.M, N = a.N, b = blockIdx.y;
  int ty, tx, p;
  const bool valid = window_of(a, b, ty, tx, p);
  const float eps = 1e-8f;
  const float inv_mn = 1.f / ((float)M * (float)N);
  const float* I = a.inten + (int64_t)b * M * N;
  const float* meas = a.stack + (int64_t)(valid ? p : 0) * M * N;
  const int per = (M * N + a.P - 1) / a.P;
  const int q0 = blockIdx.x * per, q1 = min(M * N, q0 + per);
  KSum acc;
  for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    const int r = q >> a.lgN, c = q & (N - 1);
    float gxv = 0.f, gyv = 0.f;
    if (valid) {
      const int qx = r * N + ((c + 1) & (N - 1));
      const int qy = ((r + 1) & (M - 1)) * N + c;
      const float w0 = sqrtf(meas[q] + eps) - sqrtf(I[q] + eps);
      const float wx = sqrtf(meas[qx] + eps) - sqrtf(I[qx] + eps);
      const float wy = sqrtf(meas[qy] + eps) - sqrtf(I[qy] + eps);
      const float rx = wx - w0, ry = wy - w0;
      const float G = sqrtf(rx * rx + ry * ry + eps);
      acc.add(G);
      const float s = inv_mn / G;
      gxv = rx * s;
      gyv = ry * s;
    }
    a.gx[(int64_t)b * M * N + q] = gxv;
    a.gy[(int64_t)b * M * N + q] = gyv;
  }
  const float s = block_sum<256>(acc.value(), red);
  if (threadIdx.x == 0) a.partial[b * a.P + blockIdx.x] = s;
}

// FD2: d(dW)_q = -gx_q - gy_q + gx_{q-x} + gy_{q-y}; dI = -d(dW) / (2 sqrt(I+eps));
//      dPsi = 2 dI Psi (in place over Psi).
__global__ __launch_bounds__(256) void k_far_fd2(FarArgs a) {
  const int M = a.M, N = a.N, b = blockIdx.y;
  const float eps = 1e-8f;
  const int64_t off = (int64_t)b * M * N;
  const int per = (M * N + a.P - 1) / a.P;
  const int q0 = blockIdx.x * per, q1 = min(M * N, q0 + per);
  for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    const int r = q >> a.lgN, c = q & (N - 1);
    const int qxm = r * N + ((c - 1) & (N - 1));
    const int qym = ((r - 1) & (M - 1)) * N + c;
    const float ddw = -a.gx[off + q] - a.gy[off + q] + a.gx[off + qxm] + a.gy[off + qym];
    const float dI = -ddw / (2.f * sqrtf(a.inten[off + q] + eps));
    a.psi[off + q] = cscale(a.psi[off + q], 2.f * dI);
  }
}

// F3: rows, inverse.  Workgroup (row block, pattern b): dpsi = IFFT_N of the
// dPsi rows (unnormalised), written back in place.  Workgroup (0, b) also
// forms loss_b = sum of the pattern's P loss partials (in order) / (M N).
__global__ __launch_bounds__(256) void k_far_rows_bwd(FarArgs a, int lines) {
  extern __shared__ float2 smem[];
  const int N = a.N;
  float2* tw = smem;
  float2* x = tw + N;
  float2* y = x + lines * N;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int b = blockIdx.y;
  if (blockIdx.x == 0 && tid == 0 && a.loss_out) {
    KSum acc;
    for (int i = 0; i < a.P; ++i) acc.add(a.partial[b * a.P + i]);
    a.loss_out[b] = acc.value() / ((float)a.M * (float)N);
  }
  int ty, tx, p;
  if (!window_of(a, b, ty, tx, p)) return;  // uniform across the block
  fill_twiddles(tw, N, tid, nt);
  const int r0 = blockIdx.x * lines;
  float2* base = a.psi + (int64_t)b * a.M * N;
  // the lines of the block are contiguous rows: base + r0 N + t
  for (int t = tid; t < lines * N; t += nt) x[t] = base[(int64_t)r0 * N + t];
  __syncthreads();
  float2* res = fft_lines<+1>(x, y, N, a.lgN, lines, N, tw, 1, tid, nt);
  for (int t = tid; t < lines * N; t += nt) base[(int64_t)r0 * N + t] = res[t];
}

// F4: dprobe = sum_b dpsi_b * conj(patch_b) over valid batch entries (fixed
// order); workgroup blk of nblk strides over the probe pixels.
__device__ __forceinline__ void far_dprobe(const FarArgs& a, int blk, int nblk) {
  const int MN = a.M * a.N;
  for (int q = blk * 256 + (int)threadIdx.x; q < MN; q += nblk * 256) {
    const int r = q >> a.lgN, c = q & (a.N - 1);
    float2 acc = make_float2(0.f, 0.f);
    for (int b = 0; b < a.B; ++b) {
      int ty, tx, p;
      if (!window_of(a, b, ty, tx, p)) continue;
      const float2 v = cmulc(a.psi[(int64_t)b * MN + q], a.obj[(int64_t)(ty + r) * a.W + tx + c]);
      acc.x += v.x;
      acc.y += v.y;
    }
    a.dprobe[q] = acc;
  }
}

// F5: dobject (overwritten, whole canvas) = sum over the windows covering each
// pixel of dpsi_b * conj(probe_n), in batch order: a gather, no atomics.
// Workgroup = one kTile x kTile canvas tile: wave 0 first lists, in batch
// order, the batch windows that intersect the tile (a ballot scan), so each
// pixel only tests the few windows near it instead of all B.
constexpr int kMaxWinLds = 1024;
constexpr int kTile = 32;

__device__ __forceinline__ void far_dobject(const FarArgs& a, int tile) {
  __shared__ int3 list[kMaxWinLds];
  __shared__ int n_list;
  const int tiles_x = (a.W + kTile - 1) / kTile;
  const int Y0 = (tile / tiles_x) * kTile, X0 = (tile % tiles_x) * kTile;
  const int MN = a.M * a.N;
  constexpr int kPer = kTile * kTile / 256;   // pixels per thread
  float2 acc[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) acc[i] = make_float2(0.f, 0.f);
  // batch entries in chunks of kMaxWinLds (one chunk for every practical batch)
  for (int c0 = 0; c0 < a.B; c0 += kMaxWinLds) {
    const int c1 = min(a.B, c0 + kMaxWinLds);
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x;
      int count = 0;
      for (int b0 = c0; b0 < c1; b0 += 64) {
        const int b = b0 + lane;
        int ty = 0, tx = 0, p;
        bool hit = false;
        if (b < c1 && window_of(a, b, ty, tx, p))
          hit = ty < Y0 + kTile && ty + a.M > Y0 && tx < X0 + kTile && tx + a.N > X0;
        const uint64_t mask = __ballot(hit);
        const int pos = count + __popcll(mask & ((1ull << lane) - 1ull));
        if (hit) list[pos] = make_int3(b, ty, tx);
        count += __popcll(mask);
      }
      if (lane == 0) n_list = count;
    }
    __syncthreads();
    const int nl = n_list;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int t = threadIdx.x + 256 * i;
      const int yy = Y0 + t / kTile, xx = X0 + t % kTile;
      for (int k = 0; k < nl; ++k) {
        const int3 w = list[k];
        const int r = yy - w.y, c = xx - w.z;
        if ((unsigned)r < (unsigned)a.M && (unsigned)c < (unsigned)a.N) {
          const float2 v = cmulc(a.psi[(int64_t)w.x * MN + r * a.N + c], a.probe[r * a.N + c]);
          acc[i].x += v.x;
          acc[i].y += v.y;
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int t = threadIdx.x + 256 * i;
    const int yy = Y0 + t / kTile, xx = X0 + t % kTile;
    if (yy < a.H && xx < a.W) a.dobj[(int64_t)yy * a.W + xx] = acc[i];
  }
}

// F4 + F5 in one launch: the first n_probe_blocks workgroups form dprobe, the
// rest one canvas tile of dobject each (the two read the same dpsi and write
// disjoint outputs).
__global__ __launch_bounds__(256) void k_far_dgrad(FarArgs a, int n_probe_blocks) {
  if ((int)blockIdx.x < n_probe_blocks) far_dprobe(a, blockIdx.x, n_probe_blocks);
  else far_dobject(a, blockIdx.x - n_probe_blocks);
}

// Columns per workgroup of the column pass: up to 8 (measured best for
// probes of 256 and 512), within 80 KB of LDS (two workgroups per CU).
static int far_col_tile(int M, int N) {
  int C = 8;
  while (C > 1 && (size_t)(M + 2 * C * (M + 1)) * sizeof(float2) > 81920) C >>= 1;
  while (C > 1 && N % C) C >>= 1;
  return C;
}

static int check_far(const PtyFar* d) {
  PTY_REQUIRE(d, PTY_EINVAL, "farfield: descriptor is NULL");
  PTY_REQUIRE(is_pow2(d->M) && is_pow2(d->N) && d->M >= 16 && d->N >= 16 && d->M <= 1024 && d->N <= 1024,
              PTY_EUNSUP, "farfield: probe %dx%d must be powers of two in [16, 1024]", d->M, d->N);
  PTY_REQUIRE(d->canvas_h >= d->M && d->canvas_w >= d->N, PTY_EINVAL, "farfield: canvas smaller than probe");
  PTY_REQUIRE(d->n_batch >= 1 && d->n_frames >= 1, PTY_EINVAL, "farfield: empty batch or stack");
  PTY_REQUIRE(d->loss_type >= 0 && d->loss_type <= 2, PTY_EINVAL, "farfield: bad loss_type %d", d->loss_type);
  return PTY_OK;
}

struct FarLayout {
  size_t psi, inten, gx, gy, partial, total;
  int P, C;
};

static FarLayout far_layout(const PtyFar* d) {
  FarLayout L;
  const size_t px = (size_t)d->n_batch * d->M * d->N;
  L.C = far_col_tile(d->M, d->N);
  L.P = d->N / L.C;
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
  L.psi = take(px * sizeof(float2));
  const bool fd = d->loss_type == PTY_LOSS_FD;
  L.inten = take(px * sizeof(float));
  L.gx = fd ? take(px * sizeof(float)) : 0;
  L.gy = fd ? take(px * sizeof(float)) : 0;
  L.partial = take((size_t)d->n_batch * L.P * sizeof(float));
  L.total = off;
  return L;
}

static FarArgs far_args(const PtyFar* d, const FarLayout& L, const float* object, const float* probe_n,
                        const float* stack, const int32_t* bidx, const int32_t* tl, float* loss, float* dobject,
                        float* dprobe, void* work) {
  FarArgs a;
  a.M = d->M; a.N = d->N; a.lgM = ilog2(d->M); a.lgN = ilog2(d->N);
  a.H = d->canvas_h; a.W = d->canvas_w; a.B = d->n_batch; a.F = d->n_frames; a.loss = d->loss_type;
  a.obj = (const float2*)object; a.probe = (const float2*)probe_n; a.stack = stack;
  a.bidx = bidx; a.tl = tl;
  char* w = (char*)work;
  a.psi = (float2*)(w + L.psi);
  a.inten = (float*)(w + L.inten);
  a.gx = d->loss_type == PTY_LOSS_FD ? (float*)(w + L.gx) : nullptr;
  a.gy = d->loss_type == PTY_LOSS_FD ? (float*)(w + L.gy) : nullptr;
  a.partial = (float*)(w + L.partial);
  a.P = L.P;
  a.loss_out = loss; a.dobj = (float2*)dobject; a.dprobe = (float2*)dprobe;
  return a;
}

// Row passes: two rows per workgroup below N = 1024 (measured: 2 beats 4 at
// N = 256 and beats 4 or 1 at N = 512), one from 1024 on.
static int far_row_lines(int N, int M) {
  const int lines = N >= 1024 ? 1 : 2;
  return lines < M ? lines : M;
}

static void launch_rows_fwd(const FarArgs& a, hipStream_t s) {
  const int lines = far_row_lines(a.N, a.M);
  const size_t shm = sizeof(float2) * (size_t)(a.N + 2 * lines * a.N);
  const int blocks = (a.B * a.M + lines - 1) / lines;
  hipLaunchKernelGGL(k_far_rows_fwd, dim3(blocks), dim3(256), shm, s, a, lines);
}

template <int MODE>
static void launch_cols(const FarArgs& a, int C, hipStream_t s) {
  const size_t shm = sizeof(float2) * (size_t)(a.M + 2 * C * (a.M + 1));
  hipLaunchKernelGGL(k_far_cols<MODE>, dim3(a.N / C, a.B), dim3(kColThreads), shm, s, a, C);
}

}  // namespace pty

using namespace pty;

extern "C" size_t pty_farfield_workspace(const PtyFar* d) {
  if (check_far(d)) return 0;
  return far_layout(d).total;
}

namespace pty {

// The far-field launch sequence over a given workspace (pty_farfield_fwd_bwd
// and pty_farfield_fwd_bwd_noalloc).
static int far_fwd_bwd_launch(const PtyFar* d, const FarLayout& L, const float* object, const float* probe_n,
                              const float* stack, const int32_t* batch_idx, const int32_t* tl, float* loss,
                              float* dobject, float* dprobe, void* work, hipStream_t stream) {
  FarArgs a = far_args(d, L, object, probe_n, stack, batch_idx, tl, loss, dobject, dprobe, work);
  launch_rows_fwd(a, stream);
  if (d->loss_type == PTY_LOSS_FD) {
    launch_cols<1>(a, L.C, stream);
    hipLaunchKernelGGL(k_far_fd1, dim3(a.P, a.B), dim3(256), 0, stream, a);
    hipLaunchKernelGGL(k_far_fd2, dim3(a.P, a.B), dim3(256), 0, stream, a);
    launch_cols<2>(a, L.C, stream);
  } else {
    launch_cols<0>(a, L.C, stream);
  }
  const int lines = far_row_lines(a.N, a.M);
  const size_t shm = sizeof(float2) * (size_t)(a.N + 2 * lines * a.N);
  hipLaunchKernelGGL(k_far_rows_bwd, dim3(a.M / lines, a.B), dim3(256), shm, stream, a, lines);
  const int mn = a.M * a.N;
  const int probe_blocks = (mn + 255) / 256 < 1024 ? (mn + 255) / 256 : 1024;
  const int64_t tiles = (int64_t)((a.H + kTile - 1) / kTile) * ((a.W + kTile - 1) / kTile);
  hipLaunchKernelGGL(k_far_dgrad, dim3((unsigned)(probe_blocks + tiles)), dim3(256), 0, stream, a, probe_blocks);
  return check_launch("farfield_fwd_bwd");
}

// Scratch of pty_farfield_fwd_bwd_noalloc: a file-scope device array, one
// copy per GPU (each holds it until the process ends).  Host code never
// passes the array's name to a kernel (on the host it is a stand-in, not a
// device address): its device address is looked up once per device with
// hipGetSymbolAddress, with that device current, and kept here.
constexpr int kMaxDevices = 64;
static __device__ __align__(256) char far_scratch[PTY_FAR_SCRATCH_BYTES];
static void* g_far_scratch[kMaxDevices];
static std::mutex g_far_scratch_mu;

static int far_scratch_ptr(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(PTY_ELAUNCH, "farfield_fwd_bwd_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  PTY_REQUIRE(dev >= 0 && dev < kMaxDevices, PTY_EUNSUP, "farfield_fwd_bwd_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lock(g_far_scratch_mu);
  if (!g_far_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(far_scratch));
    if (e != hipSuccess || !p)
      return set_error(PTY_ELAUNCH, "farfield_fwd_bwd_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
    g_far_scratch[dev] = p;
  }
  *out = g_far_scratch[dev];
  return PTY_OK;
}

}  // namespace pty

extern "C" int pty_farfield_fwd_bwd(const PtyFar* d, const float* object, const float* probe_n, const float* stack,
                                    const int32_t* batch_idx, const int32_t* tl, float* loss, float* dobject,
                                    float* dprobe, void* work, size_t work_bytes, hipStream_t stream) {
  int rc = check_far(d);
  if (rc) return rc;
  PTY_REQUIRE(object && probe_n && stack && batch_idx && tl && loss && dobject && dprobe, PTY_EINVAL,
              "farfield_fwd_bwd: null pointer");
  const FarLayout L = far_layout(d);
  PTY_REQUIRE(work && work_bytes >= L.total, PTY_EWORK, "farfield_fwd_bwd: workspace %zu < %zu", work_bytes, L.total);
  return far_fwd_bwd_launch(d, L, object, probe_n, stack, batch_idx, tl, loss, dobject, dprobe, work, stream);
}

extern "C" int pty_farfield_fwd_bwd_noalloc(const PtyFar* d, const float* object, const float* probe_n,
                                            const float* stack, const int32_t* batch_idx, const int32_t* tl,
                                            float* loss, float* dobject, float* dprobe, hipStream_t stream) {
  int rc = check_far(d);
  if (rc) return rc;
  PTY_REQUIRE(object && probe_n && stack && batch_idx && tl && loss && dobject && dprobe, PTY_EINVAL,
              "farfield_fwd_bwd_noalloc: null pointer");
  const FarLayout L = far_layout(d);
  PTY_REQUIRE(L.total <= (size_t)PTY_FAR_SCRATCH_BYTES, PTY_EWORK,
              "farfield_fwd_bwd_noalloc: this batch needs %zu bytes of scratch, more than the %zu of the static "
              "scratch (use pty_farfield_fwd_bwd with a workspace, or a smaller batch)",
              L.total, (size_t)PTY_FAR_SCRATCH_BYTES);
  void* scratch = nullptr;
  rc = far_scratch_ptr(&scratch);
  if (rc) return rc;
  return far_fwd_bwd_launch(d, L, object, probe_n, stack, batch_idx, tl, loss, dobject, dprobe, scratch, stream);
}

// Intensity output in the reference's fftshifted order.
__global__ void k_far_shift_out(FarArgs a, float* out) {
  const int M = a.M, N = a.N;
  const int64_t total = (int64_t)a.B * M * N;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(t / ((int64_t)M * N));
    const int q = (int)(t - (int64_t)b * M * N);
    const int k0 = q / N, k1 = q - k0 * N;
    const int y = (k0 + M / 2) & (M - 1), x = (k1 + N / 2) & (N - 1);
    out[(int64_t)b * M * N + (int64_t)y * N + x] = a.inten[t];
  }
}

extern "C" int pty_farfield_intensity(const PtyFar* d, const float* object, const float* probe_n,
                                      const int32_t* batch_idx, const int32_t* tl, float* intensity, void* work,
                                      size_t work_bytes, hipStream_t stream) {
  int rc = check_far(d);
  if (rc) return rc;
  PTY_REQUIRE(object && probe_n && batch_idx && tl && intensity, PTY_EINVAL, "farfield_intensity: null pointer");
  const FarLayout L = far_layout(d);
  PTY_REQUIRE(work && work_bytes >= L.total, PTY_EWORK, "farfield_intensity: workspace too small");
  FarArgs a = far_args(d, L, object, probe_n, nullptr, batch_idx, tl, nullptr, nullptr, nullptr, work);
  launch_rows_fwd(a, stream);
  launch_cols<1>(a, L.C, stream);
  const int64_t total = (int64_t)a.B * a.M * a.N;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_far_shift_out, dim3((unsigned)blocks), dim3(256), 0, stream, a, intensity);
  return check_launch("farfield_intensity");
}
