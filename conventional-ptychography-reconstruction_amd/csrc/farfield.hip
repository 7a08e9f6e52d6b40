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
// Any M x N whose prime factors are at most 64 (mixed-radix LDS Stockham
// FFTs, common.h); powers of two take the register / radix-8 paths.
// Pass structure:
//   F1 rows  : Psi1[b, r, :] = FFT_N(probe_n[r,:] * patch[r,:]) (register FFT of
//              one wavefront for M, N powers of two and N = 64 .. 512, fft_wave.h;
//              LDS Stockham otherwise)
//   F2 cols  : FFT_M; smooth-L1 / Poisson: loss, dI, dPsi, IFFT_M in place
//              FD: store Psi and I, then one pass per column block: loss,
//              d(dW) from the circular neighbours (one halo column each
//              side), dPsi, IFFT_M in place
//   F3 rows  : dpsi = IFFT_N(dPsi rows), in place (register FFT as F1); the
//              first row block of each pattern also sums its loss partials
//   F4       : dprobe = sum_b dpsi_b * conj(patch_b)             (fixed order)
//   F5       : dobject[px] = sum over windows covering px of dpsi_b * conj(probe_n)
//              (a per-pixel gather in batch order: no atomics, deterministic)
//              F4 and F5 share one launch.
#include "common.h"
#include "fft_wave.h"

#include <mutex>

namespace pty {

struct FarArgs {
  int M, N, lgM, lgN, H, W, B, F, loss;
  FftPlan pm, pn;         // radix plans of M and N
  const float2* obj;
  const float2* probe;
  const float* stack;     // [F, M, N] natural order
  const int* bidx;        // [B]
  const int* tl;          // [F, 2] (tlY, tlX)
  float2* psi;            // [B, M, N] workspace
  float* inten;           // [B, M, N] (FD / intensity output)
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

// Column block (cb) and pattern (b) of a column-pass workgroup on a 1-D grid
// of nb * B workgroups.  Workgroups are dispatched round-robin to the 8 XCDs
// (each with its own L2), so consecutive grid ids land on different XCDs;
// when nb * B is a multiple of 8 the ids are remapped so that each XCD takes
// a contiguous run of column blocks of one pattern.  Neighbouring blocks then
// share their L2 (the other half of every 128-byte row segment, and the halo
// columns of the FD pass).
__device__ __forceinline__ void col_block(int nb, int B, int& cb, int& b) {
  const int total = nb * B;
  int id = blockIdx.x;
  if ((total & 7) == 0) id = (id & 7) * (total >> 3) + (id >> 3);
  b = id / nb;
  cb = id - b * nb;
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
  for (int t = tid; t < lines * N; t += nt) {
    const int li = t / N, c = t - li * N;
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
  float2* res = fft_any<-1>(x, y, a.pn, lines, N, tw, 1, tid, nt);
  for (int t = tid; t < lines * N; t += nt) {
    if (t < (total_rows - row0) * N) a.psi[(int64_t)row0 * N + t] = res[t];
  }
}

// F1 for N = 64 E (E in 1, 2, 4, 8): each wavefront transforms L = 8 / E rows
// per job in registers (fft_wave.h), jobs strided over the grid's
// wavefronts; the lane's twiddles are loaded once.
template <int E>
__global__ __launch_bounds__(256) void k_far_rows_fwd_w(FarArgs a, int jobs) {
  using F = WaveFft<E>;
  __shared__ float2 scratch[4][kWaveFftScratch];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  float2* s = scratch[wv];
  F fft;
  fft.init(lane);
  const int N = F::N, M = a.M;
  const int nwaves = gridDim.x * 4;
  for (int j = blockIdx.x * 4 + wv; j < jobs; j += nwaves) {
    float2 v[8];
#pragma unroll
    for (int l = 0; l < F::L; ++l) {
      const int row = j * F::L + l;
      const int b = row >> a.lgM, r = row & (M - 1);
      int ty = 0, tx = 0, p;
      const bool valid = window_of(a, b, ty, tx, p);
      const float2* prow = a.probe + r * N;
      const float2* orow = a.obj + (int64_t)(ty + r) * a.W + tx;
#pragma unroll
      for (int n2 = 0; n2 < E; ++n2) {
        const int c = lane + 64 * n2;
        v[l * E + n2] = valid ? cmul(prow[c], orow[c]) : make_float2(0.f, 0.f);
      }
    }
    fft.template run<-1>(v, s, lane);
    float2* out = a.psi + (int64_t)(j * F::L + F::out_line(lane)) * N + F::out_first(lane);
#pragma unroll
    for (int d = 0; d < 8; ++d) out[8 * E * d] = v[d];
  }
}

// F3 for N = 64 E: dpsi rows = IFFT_N of the dPsi rows, in place; the job
// holding a pattern's first row also forms loss_b (as k_far_rows_bwd).
template <int E>
__global__ __launch_bounds__(256) void k_far_rows_bwd_w(FarArgs a, int jobs) {
  using F = WaveFft<E>;
  __shared__ float2 scratch[4][kWaveFftScratch];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  float2* s = scratch[wv];
  F fft;
  fft.init(lane);
  const int N = F::N, M = a.M;
  const int nwaves = gridDim.x * 4;
  for (int j = blockIdx.x * 4 + wv; j < jobs; j += nwaves) {
    const int row0 = j * F::L;            // the job's rows lie in one pattern (L <= M)
    const int b = row0 >> a.lgM;
    if ((row0 & (M - 1)) == 0 && a.loss_out) {
      // the pattern's loss partials, summed across the wavefront in a fixed order
      float v = 0.f;
      for (int i = lane; i < a.P; i += 64) v += a.partial[b * a.P + i];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) a.loss_out[b] = v / ((float)M * (float)N);
    }
    int ty, tx, p;
    if (!window_of(a, b, ty, tx, p)) continue;   // uniform across the wavefront
    float2 v[8];
    const float2* in = a.psi + (int64_t)row0 * N;
#pragma unroll
    for (int l = 0; l < F::L; ++l)
#pragma unroll
      for (int n2 = 0; n2 < E; ++n2) v[l * E + n2] = in[(int64_t)l * N + lane + 64 * n2];
    fft.template run<+1>(v, s, lane);
    float2* out = a.psi + (int64_t)(row0 + F::out_line(lane)) * N + F::out_first(lane);
#pragma unroll
    for (int d = 0; d < 8; ++d) out[8 * E * d] = v[d];
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

// F2: columns of pattern b (col_block), C columns per workgroup of
// kColThreads threads.
constexpr int kColThreads = 512;
// MODE 0: pointwise loss (smooth-L1 / Poisson) fused with dPsi and IFFT_M.
// MODE 1: forward only, store natural-order Psi (in place) and I (the FD
//         path, finished by k_far_fd_cols).
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
  int cb, b;
  col_block(N / C, a.B, cb, b);
  const int c0 = cb * C;
  const int lgC = __builtin_ctz(C);   // C is a power of two
  int ty, tx, p;
  const bool valid = window_of(a, b, ty, tx, p);
  float2* base = a.psi + (int64_t)b * M * N;
  for (int t = tid; t < M * C; t += nt) {
    const int r = t >> lgC, c = t & (C - 1);
    x[c * S + r] = base[(int64_t)r * N + c0 + c];
  }
  __syncthreads();
  float2* f = fft_any<-1>(x, y, a.pm, C, S, tw, 1, tid, nt);
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
  if (tid == 0) a.partial[b * a.P + cb] = s;
  float2* g2 = (f == x) ? y : x;
  float2* res = fft_any<+1>(f, g2, a.pm, C, S, tw, 1, tid, nt);
  for (int t = tid; t < M * C; t += nt) {
    const int r = t >> lgC, c = t & (C - 1);
    base[(int64_t)r * N + c0 + c] = res[c * S + r];
  }
}

// FD adjoint at pixel (r, column cc) of a dW block in LDS (column stride S,
// rows circular with ru = r + 1, rd = r - 1): G, gx, gy at q, at q - x and at
// q - y, and d(dW)_q = -gx_q - gy_q + gx_{q-x} + gy_{q-y} with
// g = R / (G M N); G_q is returned for the loss.  1 / G by rsqrt.
__device__ __forceinline__ float fd_ddw(const float* dW, int S, int cc, int r, int ru, int rd, float inv_mn,
                                        float& Gq) {
  const float eps = 1e-8f;
  const float w = dW[cc * S + r];
  const float rx = dW[(cc + 1) * S + r] - w, ry = dW[cc * S + ru] - w;
  const float s0 = rx * rx + ry * ry + eps;
  const float i0 = rsqrtf(s0);
  Gq = s0 * i0;
  const float wl = dW[(cc - 1) * S + r];
  const float rxl = w - wl, ryl = dW[(cc - 1) * S + ru] - wl;
  const float il = rsqrtf(rxl * rxl + ryl * ryl + eps);
  const float wd = dW[cc * S + rd];
  const float rxd = dW[(cc + 1) * S + rd] - wd, ryd = w - wd;
  const float id = rsqrtf(rxd * rxd + ryd * ryd + eps);
  return inv_mn * (-(rx + ry) * i0 + rxl * il + ryd * id);
}

// Sum over the workgroup (blockDim a multiple of 64) in a fixed order.
__device__ __forceinline__ float wg_sum(float v, float* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  float t = 0.f;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
  __syncthreads();
  return t;
}

// F2 (MODE 0: smooth-L1 / Poisson loss, dPsi, inverse; MODE 1: store Psi
// and I) for M = 64 E, as a persistent pipeline over jobs of 8 columns of
// one pattern (XCD-grouped as col_block): while the wavefronts transform the
// current job's columns in LDS, the next job's tile (and measurement values)
// is already being loaded into registers.  Each thread owns the same 8 tile
// positions in every job (column tid & 7, rows (tid >> 3) + k (64 E / 8)), so
// the loss phase reads the measurement values it prefetched.  The LDS column
// stride 584 keeps the tile writes 2-way bank conflicted and leaves room for
// a wavefront's exchange scratch inside its own column slot.
constexpr int kColStrideW = 584;

template <int E, int MODE>
__global__ __launch_bounds__(64 * E) void k_far_cols_wp(FarArgs a, int jobs) {
  constexpr int C = 8, S = kColStrideW, NT = 64 * E, RSTEP = NT / 8;
  using Ff = WaveFft<E>;
  using Fi = WaveFft<E>;
  __shared__ float2 x[C * S];
  __shared__ float red[8];
  const int M = a.M, N = a.N, nb = N / C;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int c = tid & 7, r0 = tid >> 3;
  const float inv_mn = 1.f / ((float)M * (float)N);
  Ff fft;
  fft.init(lane);
  float2* wscr = x + wv * Ff::L * S;        // the wavefront's first column slot
  auto job_of = [&](int j, int& cb, int& b) {
    const int id = (jobs & 7) == 0 ? (j & 7) * (jobs >> 3) + (j >> 3) : j;
    b = id / nb;
    cb = id - b * nb;
  };
  float2 pf[8];
  float pm[8];
  auto prefetch = [&](int j) {
    if (j >= jobs) return;
    int cb, b, ty, tx, p;
    job_of(j, cb, b);
    const float2* src = a.psi + (int64_t)b * M * N + cb * C + c;
#pragma unroll
    for (int k = 0; k < 8; ++k) pf[k] = src[(int64_t)(r0 + k * RSTEP) * N];
    if (MODE == 0) {
      const bool valid = window_of(a, b, ty, tx, p);
      const float* ms = a.stack + (int64_t)(valid ? p : 0) * M * N + cb * C + c;
#pragma unroll
      for (int k = 0; k < 8; ++k) pm[k] = valid ? ms[(int64_t)(r0 + k * RSTEP) * N] : 0.f;
    }
  };
  prefetch(blockIdx.x);
  for (int j = blockIdx.x; j < jobs; j += gridDim.x) {
    int cb, b, ty, tx, p;
    job_of(j, cb, b);
    const bool valid = window_of(a, b, ty, tx, p);
    float mc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      x[c * S + r0 + k * RSTEP] = pf[k];
      mc[k] = pm[k];
    }
    __syncthreads();
    prefetch(j + gridDim.x);
    {
      float2 v[8];
#pragma unroll
      for (int l = 0; l < Ff::L; ++l)
#pragma unroll
        for (int n2 = 0; n2 < E; ++n2) v[l * E + n2] = x[(wv * Ff::L + l) * S + lane + 64 * n2];
      fft.template run<-1>(v, wscr, lane);
      float2* col = x + (wv * Ff::L + Ff::out_line(lane)) * S + Ff::out_first(lane);
#pragma unroll
      for (int d = 0; d < 8; ++d) col[8 * E * d] = v[d];
    }
    __syncthreads();
    float2* dst = a.psi + (int64_t)b * M * N + cb * C + c;
    if (MODE == 1) {
      float* idst = a.inten + (int64_t)b * M * N + cb * C + c;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const float2 v = x[c * S + r0 + k * RSTEP];
        dst[(int64_t)(r0 + k * RSTEP) * N] = v;
        idst[(int64_t)(r0 + k * RSTEP) * N] = v.x * v.x + v.y * v.y;
      }
      __syncthreads();
      continue;
    }
    KSum acc;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int q = c * S + r0 + k * RSTEP;
      const float2 v = x[q];
      float2 gv = make_float2(0.f, 0.f);
      if (valid) {
        const float I = v.x * v.x + v.y * v.y;
        const float m = mc[k];
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
      x[q] = gv;
    }
    __syncthreads();
    const float sum = wg_sum(acc.value(), red);
    if (tid == 0) a.partial[b * a.P + cb] = sum;
    {
      float2 v[8];
#pragma unroll
      for (int l = 0; l < Fi::L; ++l)
#pragma unroll
        for (int n2 = 0; n2 < E; ++n2) v[l * E + n2] = x[(wv * Fi::L + l) * S + lane + 64 * n2];
      fft.template run<+1>(v, wscr, lane);
      float2* col = x + (wv * Fi::L + Fi::out_line(lane)) * S + Fi::out_first(lane);
#pragma unroll
      for (int d = 0; d < 8; ++d) col[8 * E * d] = v[d];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[(int64_t)(r0 + k * RSTEP) * N] = x[c * S + r0 + k * RSTEP];
    __syncthreads();
  }
}

// The FD loss, its adjoint and the column IFFT for M = 64 E (after the
// forward column pass stored Psi and I), pipelined like k_far_cols_wp: the
// next job's I, the measurement (own columns and one halo column each side,
// circular) and Psi (own columns, kept in registers for dPsi) are loaded
// while the current job finishes.  Per job: dW = sqrt(m + eps) - sqrt(I + eps)
// for the C + 2 columns into LDS, then for the own pixels G, d(dW) (fd_ddw),
// dI, dPsi into the tile, the loss partial, the inverse transforms and the
// store (in place: the halo columns are read from I, which stays untouched).
template <int E>
__global__ __launch_bounds__(64 * E) void k_far_fd_wp(FarArgs a, int jobs) {
  constexpr int C = 8, S = kColStrideW, NT = 64 * E, RSTEP = NT / 8;
  using F = WaveFft<E>;
  __shared__ float2 x[C * S];
  __shared__ float dW[(C + 2) * (64 * E + 8)];
  __shared__ float red[8];
  // dW column stride M + 8 (= 8 mod 64): a wavefront's 8 columns x 8 rows hit
  // 64 distinct banks
  const int M = a.M, N = a.N, nb = N / C, SW = M + 8;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int c = tid & 7, r0 = tid >> 3;
  const float eps = 1e-8f;
  const float inv_mn = 1.f / ((float)M * (float)N);
  F fft;
  fft.init(lane);
  float2* wscr = x + wv * F::L * S;
  auto job_of = [&](int j, int& cb, int& b) {
    const int id = (jobs & 7) == 0 ? (j & 7) * (jobs >> 3) + (j >> 3) : j;
    b = id / nb;
    cb = id - b * nb;
  };
  // prefetched: own-column Psi, I, m; halo I, m (entries tid and tid + NT of
  // the 2 M halo values: side = t & 1, row = t >> 1)
  float2 pf[8];
  float pi[8], pm[8], hi[2], hm[2];
  auto prefetch = [&](int j) {
    if (j >= jobs) return;
    int cb, b, ty, tx, p;
    job_of(j, cb, b);
    const bool valid = window_of(a, b, ty, tx, p);
    const int64_t off = (int64_t)b * M * N;
    const float* Ib = a.inten + off;
    const float* ms = a.stack + (int64_t)(valid ? p : 0) * M * N;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int64_t q = (int64_t)(r0 + k * RSTEP) * N + cb * C + c;
      pf[k] = a.psi[off + q];
      pi[k] = Ib[q];
      pm[k] = valid ? ms[q] : 0.f;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int t = tid + h * NT, side = t & 1, r = t >> 1;
      const int col = side ? (cb * C + C == N ? 0 : cb * C + C) : (cb == 0 ? N - 1 : cb * C - 1);
      const int64_t q = (int64_t)r * N + col;
      hi[h] = Ib[q];
      hm[h] = valid ? ms[q] : 0.f;
    }
  };
  prefetch(blockIdx.x);
  for (int j = blockIdx.x; j < jobs; j += gridDim.x) {
    int cb, b, ty, tx, p;
    job_of(j, cb, b);
    const bool valid = window_of(a, b, ty, tx, p);
    float2 psi[8];
    float h2[8];     // 1 / (2 sqrt(I + eps)) of the own pixels
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      // hardware rsqrt / sqrt (about 1 ulp): sqrt(I + eps) = (I + eps) rsqrt(I + eps)
      const float ie = pi[k] + eps, ri = __builtin_amdgcn_rsqf(ie);
      dW[(c + 1) * SW + r0 + k * RSTEP] = valid ? __builtin_amdgcn_sqrtf(pm[k] + eps) - ie * ri : 0.f;
      h2[k] = 0.5f * ri;
      psi[k] = pf[k];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int t = tid + h * NT, side = t & 1, r = t >> 1;
      const float ie = hi[h] + eps;   // the same expression as the owning block's
      dW[(side ? C + 1 : 0) * SW + r] =
          valid ? __builtin_amdgcn_sqrtf(hm[h] + eps) - ie * __builtin_amdgcn_rsqf(ie) : 0.f;
    }
    __syncthreads();
    prefetch(j + gridDim.x);
    KSum acc;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int r = r0 + k * RSTEP;
      float2 g = make_float2(0.f, 0.f);
      if (valid) {
        float G;
        const float ddw = fd_ddw(dW, SW, c + 1, r, (r + 1) & (M - 1), (r - 1) & (M - 1), inv_mn, G);
        acc.add(G);
        g = cscale(psi[k], -2.f * ddw * h2[k]);
      }
      x[c * S + r] = g;
    }
    __syncthreads();
    const float sum = wg_sum(acc.value(), red);
    if (tid == 0) a.partial[b * a.P + cb] = sum;
    {
      float2 v[8];
#pragma unroll
      for (int l = 0; l < F::L; ++l)
#pragma unroll
        for (int n2 = 0; n2 < E; ++n2) v[l * E + n2] = x[(wv * F::L + l) * S + lane + 64 * n2];
      fft.template run<+1>(v, wscr, lane);
      float2* col = x + (wv * F::L + F::out_line(lane)) * S + F::out_first(lane);
#pragma unroll
      for (int d = 0; d < 8; ++d) col[8 * E * d] = v[d];
    }
    __syncthreads();
    float2* dst = a.psi + (int64_t)b * M * N + cb * C + c;
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[(int64_t)(r0 + k * RSTEP) * N] = x[c * S + r0 + k * RSTEP];
    __syncthreads();
  }
}

// FD loss and adjoint fused with the column IFFT (Recovery.py:296-302): one
// workgroup per (C columns, pattern b) after k_far_cols<1> stored Psi and I.
// Phase A: dW = sqrt(m + eps) - sqrt(I + eps) for the C columns and one halo
// column on each side (circular), in LDS.  Phase B, own pixels q: G, gx, gy
// at q, at q - x and at q - y (recomputed from dW, so no gradient arrays),
// d(dW) = -gx_q - gy_q + gx_{q-x} + gy_{q-y}, dI = -d(dW) / (2 sqrt(I + eps)),
// dPsi = 2 dI Psi; loss partial = sum of G over own pixels.  Then IFFT_M of
// the C columns in place.  Psi, I and the measurement are read once and the
// gradients of the loss never leave the workgroup.
__global__ __launch_bounds__(kColThreads) void k_far_fd_cols(FarArgs a, int C) {
  extern __shared__ float2 smem[];
  __shared__ float red[kColThreads];
  const int M = a.M, N = a.N, S = M + 1, lgC = __builtin_ctz(C);
  float2* tw = smem;
  float2* x = tw + M;
  float2* y = x + C * S;
  float* dW = reinterpret_cast<float*>(y);     // (C + 2) x S floats, column cc = c + 1 (halos 0, C + 1);
                                               // the launch sizes y for both uses
  const int tid = threadIdx.x, nt = blockDim.x;
  fill_twiddles(tw, M, tid, nt);
  int cb, b;
  col_block(N / C, a.B, cb, b);
  const int c0 = cb * C;
  int ty, tx, p;
  const bool valid = window_of(a, b, ty, tx, p);
  const float eps = 1e-8f;
  const float inv_mn = 1.f / ((float)M * (float)N);
  const int64_t off = (int64_t)b * M * N;
  const float* I = a.inten + off;
  const float* meas = a.stack + (int64_t)(valid ? p : 0) * M * N;
  float2* base = a.psi + off;
  // phase A: own columns (coalesced over c), keeping sqrt(I + eps) in x
  for (int t = tid; t < M * C; t += nt) {
    const int r = t >> lgC, c = t & (C - 1);
    const int64_t q = (int64_t)r * N + c0 + c;
    const float si = sqrtf(I[q] + eps);
    dW[(c + 1) * S + r] = valid ? sqrtf(meas[q] + eps) - si : 0.f;
    x[c * S + r] = make_float2(0.5f / si, 0.f);
  }
  for (int t = tid; t < 2 * M; t += nt) {
    const int r = t >> 1, side = t & 1;
    const int col = side ? (c0 + C == N ? 0 : c0 + C) : (c0 == 0 ? N - 1 : c0 - 1);
    const int64_t q = (int64_t)r * N + col;
    dW[(side ? C + 1 : 0) * S + r] = valid ? sqrtf(meas[q] + eps) - sqrtf(I[q] + eps) : 0.f;
  }
  __syncthreads();
  KSum acc;
  for (int t = tid; t < M * C; t += nt) {
    const int r = t >> lgC, c = t & (C - 1);
    const int cc = c + 1, ru = r + 1 == M ? 0 : r + 1, rd = r == 0 ? M - 1 : r - 1;
    float2 g = make_float2(0.f, 0.f);
    if (valid) {
      float G;
      const float ddw = fd_ddw(dW, S, cc, r, ru, rd, inv_mn, G);
      acc.add(G);
      const float dI = -ddw * x[c * S + r].x;      // x holds 1 / (2 sqrt(I + eps))
      g = cscale(base[(int64_t)r * N + c0 + c], 2.f * dI);
    }
    x[c * S + r] = g;
  }
  __syncthreads();
  const float sum = block_sum<kColThreads>(acc.value(), red);
  if (tid == 0) a.partial[b * a.P + cb] = sum;
  float2* res = fft_any<+1>(x, y, a.pm, C, S, tw, 1, tid, nt);
  for (int t = tid; t < M * C; t += nt) {
    const int r = t >> lgC, c = t & (C - 1);
    base[(int64_t)r * N + c0 + c] = res[c * S + r];
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
  float2* res = fft_any<+1>(x, y, a.pn, lines, N, tw, 1, tid, nt);
  for (int t = tid; t < lines * N; t += nt) base[(int64_t)r0 * N + t] = res[t];
}

// F4 + F5 share a list of batch windows in LDS: wave 0 lists, in batch order,
// the valid entries of batch[c0, c1) (and, for a canvas tile, only those that
// intersect it) by a ballot scan; returns the count.
constexpr int kMaxWinLds = 1024;
constexpr int kTile = 32;
constexpr int kWinBatch = 8;   // dprobe: windows whose loads are in flight together
constexpr int kWinBatchObj = 4;   // dobject: the same per pixel (four pixels per thread)

__device__ __forceinline__ int list_windows(const FarArgs& a, int c0, int c1, bool tile_test, int Y0, int X0,
                                            int3* list, int* n_list) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    int count = 0;
    for (int b0 = c0; b0 < c1; b0 += 64) {
      const int b = b0 + lane;
      int ty = 0, tx = 0, p;
      bool hit = false;
      if (b < c1 && window_of(a, b, ty, tx, p))
        hit = !tile_test || (ty < Y0 + kTile && ty + a.M > Y0 && tx < X0 + kTile && tx + a.N > X0);
      const uint64_t mask = __ballot(hit);
      const int pos = count + __popcll(mask & ((1ull << lane) - 1ull));
      if (hit) list[pos] = make_int3(b, ty, tx);
      count += __popcll(mask);
    }
    if (lane == 0) *n_list = count;
  }
  __syncthreads();
  const int n = *n_list;
  return n;
}

// F4: dprobe = sum_b dpsi_b * conj(patch_b) over valid batch entries (fixed
// order); workgroup blk of nblk strides over the probe pixels.  The windows
// come from the LDS list and the loads of kWinBatch of them are issued
// together (one round trip per batch instead of three per entry).
__device__ __forceinline__ void far_dprobe(const FarArgs& a, int blk, int nblk, int3* list, int* n_list) {
  const int MN = a.M * a.N;
  const int stride = nblk * 256;
  const int q0 = blk * 256 + (int)threadIdx.x;
  if (a.B > kMaxWinLds) {   // uniform: batches longer than the list, one entry at a time
    for (int q = q0; q < MN; q += stride) {
      const int r = q / a.N, c = q - r * a.N;
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
    return;
  }
  const int nl = list_windows(a, 0, a.B, false, 0, 0, list, n_list);
  for (int q = q0; q < MN; q += stride) {
    const int r = q / a.N, c = q - r * a.N;
    float2 acc = make_float2(0.f, 0.f);
    for (int k0 = 0; k0 < nl; k0 += kWinBatch) {
      float2 pv[kWinBatch], ov[kWinBatch];
#pragma unroll
      for (int u = 0; u < kWinBatch; ++u) {
        const int3 w = list[min(k0 + u, nl - 1)];
        pv[u] = a.psi[(int64_t)w.x * MN + q];
        ov[u] = a.obj[(int64_t)(w.y + r) * a.W + w.z + c];
      }
#pragma unroll
      for (int u = 0; u < kWinBatch; ++u)
        if (k0 + u < nl) {
          const float2 v = cmulc(pv[u], ov[u]);
          acc.x += v.x;
          acc.y += v.y;
        }
    }
    a.dprobe[q] = acc;
  }
}

// F5: dobject (overwritten, whole canvas) = sum over the windows covering each
// pixel of dpsi_b * conj(probe_n), in batch order: a gather, no atomics.
// Workgroup = one kTile x kTile canvas tile: each pixel only tests the few
// listed windows that intersect the tile; a thread's loads for its kPer
// pixels and kWinBatch windows are issued together.
__device__ __forceinline__ void far_dobject(const FarArgs& a, int tile, int3* list, int* n_list) {
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
    const int nl = list_windows(a, c0, c1, true, Y0, X0, list, n_list);
    for (int k0 = 0; k0 < nl; k0 += kWinBatchObj) {
      float2 pv[kPer][kWinBatchObj], qv[kPer][kWinBatchObj];
      bool in[kPer][kWinBatchObj];
#pragma unroll
      for (int i = 0; i < kPer; ++i) {
        const int t = threadIdx.x + 256 * i;
        const int yy = Y0 + t / kTile, xx = X0 + t % kTile;
#pragma unroll
        for (int u = 0; u < kWinBatchObj; ++u) {
          const int3 w = list[min(k0 + u, nl - 1)];
          const int r = yy - w.y, c = xx - w.z;
          in[i][u] = k0 + u < nl && (unsigned)r < (unsigned)a.M && (unsigned)c < (unsigned)a.N;
          pv[i][u] = in[i][u] ? a.psi[(int64_t)w.x * MN + r * a.N + c] : make_float2(0.f, 0.f);
          qv[i][u] = in[i][u] ? a.probe[r * a.N + c] : make_float2(0.f, 0.f);
        }
      }
#pragma unroll
      for (int i = 0; i < kPer; ++i)
#pragma unroll
        for (int u = 0; u < kWinBatchObj; ++u)
          if (in[i][u]) {
            const float2 v = cmulc(pv[i][u], qv[i][u]);
            acc[i].x += v.x;
            acc[i].y += v.y;
          }
    }
    __syncthreads();   // the next chunk's list overwrites this one
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
  __shared__ int3 list[kMaxWinLds];
  __shared__ int n_list;
  if ((int)blockIdx.x < n_probe_blocks) far_dprobe(a, blockIdx.x, n_probe_blocks, list, &n_list);
  else far_dobject(a, blockIdx.x - n_probe_blocks, list, &n_list);
}

// Dynamic LDS of the LDS-FFT column passes (k_far_cols, k_far_fd_cols):
// twiddles, then two C x (M + 1) buffers; the FD pass keeps (C + 2) x (M + 1)
// floats of dW in the second, which needs more room only at C = 1.
static size_t far_col_lds(int M, int C, bool fd) {
  const size_t S = (size_t)M + 1;
  const size_t ybuf = (fd && C == 1) ? (3 * S + 1) / 2 : (size_t)C * S;
  return ((size_t)M + (size_t)C * S + ybuf) * sizeof(float2);
}
// Dynamic LDS of the LDS-FFT row passes (k_far_rows_fwd, k_far_rows_bwd).
static size_t far_row_lds(int N, int lines) { return sizeof(float2) * (size_t)(N + 2 * lines * N); }
constexpr size_t kFarMaxLds = 144 * 1024;   // of gfx950's 160 KB per workgroup

// Columns per workgroup of the column pass: up to 8 (measured best for
// probes of 256 and 512), a power of two dividing N, within 80 KB of LDS
// (two workgroups per CU) where one column allows it.
static int far_col_tile(int M, int N) {
  int C = 8;
  while (C > 1 && far_col_lds(M, C, true) > 81920) C >>= 1;
  while (C > 1 && N % C) C >>= 1;
  return C;
}

// Row passes: two rows per workgroup below N = 1024 (measured: 2 beats 4 at
// N = 256 and beats 4 or 1 at N = 512), one from 1024 on; rows per workgroup
// divide M.
static int far_row_lines(int N, int M) { return (N >= 1024 || M % 2) ? 1 : 2; }

static int check_far(const PtyFar* d) {
  PTY_REQUIRE(d, PTY_EINVAL, "farfield: descriptor is NULL");
  PTY_REQUIRE(d->M >= 8 && d->N >= 8 && d->M <= 4096 && d->N <= 4096, PTY_EUNSUP,
              "farfield: probe %dx%d outside [8, 4096]", d->M, d->N);
  FftPlan pm, pn;
  PTY_REQUIRE(make_fft_plan(d->M, pm) == PTY_OK && make_fft_plan(d->N, pn) == PTY_OK, PTY_EUNSUP,
              "farfield: probe %dx%d has a prime factor above %d", d->M, d->N, kMaxRadix);
  PTY_REQUIRE(far_col_lds(d->M, far_col_tile(d->M, d->N), true) <= kFarMaxLds &&
                  far_row_lds(d->N, far_row_lines(d->N, d->M)) <= kFarMaxLds,
              PTY_EUNSUP, "farfield: probe %dx%d needs more LDS than a workgroup has", d->M, d->N);
  PTY_REQUIRE(d->canvas_h >= d->M && d->canvas_w >= d->N, PTY_EINVAL, "farfield: canvas smaller than probe");
  PTY_REQUIRE(d->n_batch >= 1 && d->n_frames >= 1, PTY_EINVAL, "farfield: empty batch or stack");
  PTY_REQUIRE(d->loss_type >= 0 && d->loss_type <= 2, PTY_EINVAL, "farfield: bad loss_type %d", d->loss_type);
  return PTY_OK;
}

struct FarLayout {
  size_t psi, inten, partial, total;
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
  L.inten = take(px * sizeof(float));
  L.partial = take((size_t)d->n_batch * L.P * sizeof(float));
  L.total = off;
  return L;
}

static FarArgs far_args(const PtyFar* d, const FarLayout& L, const float* object, const float* probe_n,
                        const float* stack, const int32_t* bidx, const int32_t* tl, float* loss, float* dobject,
                        float* dprobe, void* work) {
  FarArgs a;
  a.M = d->M; a.N = d->N; a.lgM = ilog2(d->M); a.lgN = ilog2(d->N);
  make_fft_plan(d->M, a.pm);   // checked by check_far
  make_fft_plan(d->N, a.pn);
  a.H = d->canvas_h; a.W = d->canvas_w; a.B = d->n_batch; a.F = d->n_frames; a.loss = d->loss_type;
  a.obj = (const float2*)object; a.probe = (const float2*)probe_n; a.stack = stack;
  a.bidx = bidx; a.tl = tl;
  char* w = (char*)work;
  a.psi = (float2*)(w + L.psi);
  a.inten = (float*)(w + L.inten);
  a.partial = (float*)(w + L.partial);
  a.P = L.P;
  a.loss_out = loss; a.dobj = (float2*)dobject; a.dprobe = (float2*)dprobe;
  return a;
}

// Wavefronts of the register-FFT row passes: jobs strided over at most
// 32 wavefronts per CU.
static int far_wave_blocks(int jobs) {
  const int cap = device_cu_count() * 8;   // 8 workgroups of 4 wavefronts per CU
  const int need = (jobs + 3) / 4;
  return need < cap ? need : cap;
}

// Row passes on the register FFT: M and N powers of two, N = 64 .. 512.
static bool far_wave_rows(const FarArgs& a) { return is_pow2(a.M) && is_pow2(a.N) && a.N >= 64 && a.N <= 512; }

static void launch_rows_fwd(const FarArgs& a, hipStream_t s) {
  if (far_wave_rows(a)) {
    const int E = a.N / 64, jobs = a.B * a.M / (8 / E);
    const dim3 g(far_wave_blocks(jobs)), blk(256);
    switch (E) {
      case 1: hipLaunchKernelGGL(k_far_rows_fwd_w<1>, g, blk, 0, s, a, jobs); break;
      case 2: hipLaunchKernelGGL(k_far_rows_fwd_w<2>, g, blk, 0, s, a, jobs); break;
      case 4: hipLaunchKernelGGL(k_far_rows_fwd_w<4>, g, blk, 0, s, a, jobs); break;
      default: hipLaunchKernelGGL(k_far_rows_fwd_w<8>, g, blk, 0, s, a, jobs); break;
    }
    return;
  }
  const int lines = far_row_lines(a.N, a.M);
  const size_t shm = far_row_lds(a.N, lines);
  const int blocks = (a.B * a.M + lines - 1) / lines;
  hipLaunchKernelGGL(k_far_rows_fwd, dim3(blocks), dim3(256), shm, s, a, lines);
}

// Column passes on the register FFT (M a power of two in 64 .. 512, 8
// columns per workgroup: C / L = E wavefronts) when the layout's column
// block is 8 wide.
static bool far_wave_cols(const FarArgs& a, int C) { return is_pow2(a.M) && a.M >= 64 && a.M <= 512 && C == 8; }

// MODE 0: pointwise losses; 1: forward store (Psi, I); 3: FD loss + adjoint + inverse
template <int MODE>
static void launch_cols(const FarArgs& a, int C, hipStream_t s) {
  if (far_wave_cols(a, C)) {
    const int jobs = (a.N / C) * a.B;
    // resident workgroups per CU: LDS allows 4 of the pointwise-loss / store
    // passes (37 KB) and 2 of the FD pass (58 KB)
    const int cap = device_cu_count() * (MODE == 3 ? 2 : 4);
    const dim3 g(jobs < cap ? jobs : cap);
    if (MODE == 3) {
      switch (a.M) {
        case 64: hipLaunchKernelGGL(k_far_fd_wp<1>, g, dim3(64), 0, s, a, jobs); break;
        case 128: hipLaunchKernelGGL(k_far_fd_wp<2>, g, dim3(128), 0, s, a, jobs); break;
        case 256: hipLaunchKernelGGL(k_far_fd_wp<4>, g, dim3(256), 0, s, a, jobs); break;
        default: hipLaunchKernelGGL(k_far_fd_wp<8>, g, dim3(512), 0, s, a, jobs); break;
      }
      return;
    }
    constexpr int M1 = MODE == 1 ? 1 : 0;
    switch (a.M) {
      case 64: hipLaunchKernelGGL((k_far_cols_wp<1, M1>), g, dim3(64), 0, s, a, jobs); break;
      case 128: hipLaunchKernelGGL((k_far_cols_wp<2, M1>), g, dim3(128), 0, s, a, jobs); break;
      case 256: hipLaunchKernelGGL((k_far_cols_wp<4, M1>), g, dim3(256), 0, s, a, jobs); break;
      default: hipLaunchKernelGGL((k_far_cols_wp<8, M1>), g, dim3(512), 0, s, a, jobs); break;
    }
    return;
  }
  if (MODE == 3) {
    hipLaunchKernelGGL(k_far_fd_cols, dim3((a.N / C) * a.B), dim3(kColThreads), far_col_lds(a.M, C, true), s, a, C);
    return;
  }
  const size_t shm = far_col_lds(a.M, C, false);
  hipLaunchKernelGGL(k_far_cols<MODE == 1 ? 1 : 0>, dim3((a.N / C) * a.B), dim3(kColThreads), shm, s, a, C);
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
    launch_cols<3>(a, L.C, stream);
  } else {
    launch_cols<0>(a, L.C, stream);
  }
  if (far_wave_rows(a)) {
    const int E = a.N / 64, jobs = a.B * a.M / (8 / E);
    const dim3 g(far_wave_blocks(jobs)), blk(256);
    switch (E) {
      case 1: hipLaunchKernelGGL(k_far_rows_bwd_w<1>, g, blk, 0, stream, a, jobs); break;
      case 2: hipLaunchKernelGGL(k_far_rows_bwd_w<2>, g, blk, 0, stream, a, jobs); break;
      case 4: hipLaunchKernelGGL(k_far_rows_bwd_w<4>, g, blk, 0, stream, a, jobs); break;
      default: hipLaunchKernelGGL(k_far_rows_bwd_w<8>, g, blk, 0, stream, a, jobs); break;
    }
  } else {
    const int lines = far_row_lines(a.N, a.M);
    hipLaunchKernelGGL(k_far_rows_bwd, dim3(a.M / lines, a.B), dim3(256), far_row_lds(a.N, lines), stream, a,
                       lines);
  }
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
    const int y = k0 + M / 2 < M ? k0 + M / 2 : k0 + M / 2 - M;
    const int x = k1 + N / 2 < N ? k1 + N / 2 : k1 + N / 2 - N;
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
