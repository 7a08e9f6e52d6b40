// Low-precision variants of the fused INR kernels on gfx950 MFMA with FP32
// accumulation, templated over the operand format:
//   FmtBf16: v_mfma_f32_32x32x16_bf16;
//   FmtFp8 : v_mfma_f32_32x32x16_fp8_fp8 on OCP e4m3fn (max 448) operands
//            with per-tensor scales: weights by their exact amax (each
//            workgroup reduces it while staging the operands), activations
//            and gradients by the amax recorded in the previous step
//            (delayed scaling; PtyMlp.fp8_state).  Values are clamped to
//            +-448 before conversion (out-of-range e4m3 conversions give
//            NaN).  A GEMM result is the MFMA sum times 1 / (s_a s_b).
// Same tile layout and chain rule as the FP32 kernels in inr.hip; per MFMA
// step a lane supplies 8 consecutive K values, lane (x, h) holding K = 8h + t
// of the step (for A and for B alike), so the kappa maps are applied to value
// index 8q + t.
//
// * Weight operands are converted once per workgroup into LDS in operand
//   order: one LDS read per MFMA operand.
// * Activations are packed in registers for the B operand; the elementwise
//   Gabor activation and its adjoint stay FP32.
// * The weight-gradient GEMMs take their point-major operands from LDS
//   staging rows of 32 points (bf16: 80-B stride, fp8: 40-B stride; aligned
//   and conflict-free).
// * Bias gradients are FP32 sums of the staged (rounded) values.
#include "inr_common.h"

namespace pty {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------- formats
struct FmtBf16 {
  using op_t = uint4;      // 8 bf16
  using el_t = uint16_t;
  static constexpr bool kScaled = false;
  static constexpr int kStride = 40;   // staging row stride in elements (80 B)

  static __device__ __forceinline__ op_t pack8(const float* x, float) {
    bf16x8 r;
#pragma unroll
    for (int t = 0; t < 8; ++t) r[t] = (__bf16)x[t];
    return __builtin_bit_cast(uint4, r);
  }
  static __device__ __forceinline__ v16f mfma(op_t a, op_t b, v16f c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                    0, 0, 0);
  }
  static __device__ __forceinline__ op_t neg(op_t a) {
    a.x ^= 0x80008000u;
    a.y ^= 0x80008000u;
    a.z ^= 0x80008000u;
    a.w ^= 0x80008000u;
    return a;
  }
  static __device__ __forceinline__ float val(op_t a, int t) {
    const uint32_t word = (t < 2) ? a.x : (t < 4) ? a.y : (t < 6) ? a.z : a.w;
    return (t & 1) ? __uint_as_float(word & 0xFFFF0000u) : __uint_as_float(word << 16);
  }
  static __device__ __forceinline__ float sum8(op_t a) {
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < 8; ++t) s += val(a, t);
    return s;
  }
  static __device__ __forceinline__ el_t el(float x, float) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
  static __device__ __forceinline__ op_t row8(const el_t* base, int row, int col0) {
    return *(const uint4*)(base + row * kStride + col0);
  }
};

struct FmtFp8 {
  using op_t = uint2;      // 8 e4m3
  using el_t = uint8_t;
  static constexpr bool kScaled = true;
  static constexpr int kStride = 40;   // staging row stride in bytes

  static __device__ __forceinline__ float clamp448(float x) { return __builtin_amdgcn_fmed3f(x, -448.f, 448.f); }
  static __device__ __forceinline__ uint32_t pack4(float a, float b, float c, float d) {
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(clamp448(a), clamp448(b), 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(clamp448(c), clamp448(d), w, true);
    return (uint32_t)w;
  }
  static __device__ __forceinline__ op_t pack8(const float* x, float s) {
    op_t r;
    r.x = pack4(x[0] * s, x[1] * s, x[2] * s, x[3] * s);
    r.y = pack4(x[4] * s, x[5] * s, x[6] * s, x[7] * s);
    return r;
  }
  static __device__ __forceinline__ v16f mfma(op_t a, op_t b, v16f c) {
    const long la = (long)(((uint64_t)a.y << 32) | a.x), lb = (long)(((uint64_t)b.y << 32) | b.x);
    return __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(la, lb, c, 0, 0, 0);
  }
  static __device__ __forceinline__ op_t neg(op_t a) {
    a.x ^= 0x80808080u;
    a.y ^= 0x80808080u;
    return a;
  }
  static __device__ __forceinline__ float val(op_t a, int t) {
    const uint32_t word = (t < 4) ? a.x : a.y;
    typedef float f2v __attribute__((ext_vector_type(2)));
    const f2v f = ((t >> 1) & 1) ? __builtin_amdgcn_cvt_pk_f32_fp8((int)word, true)
                                 : __builtin_amdgcn_cvt_pk_f32_fp8((int)word, false);
    return (t & 1) ? f[1] : f[0];
  }
  static __device__ __forceinline__ float sum8(op_t a) {
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < 8; ++t) s += val(a, t);
    return s;
  }
  static __device__ __forceinline__ el_t el(float x, float s) {
    return (el_t)(__builtin_amdgcn_cvt_pk_fp8_f32(clamp448(x * s), 0.f, 0, false) & 0xFF);
  }
  static __device__ __forceinline__ op_t row8(const el_t* base, int row, int col0) {
    return *(const uint2*)(base + row * kStride + col0);
  }
};

// fp8_state (see PtyMlp.fp8_state): scales [0, 8), then the amax recorded in
// this step, slot i spread over kAmaxSubs sub-slots (workgroup b writes
// sub-slot b % kAmaxSubs, so no address sees more than a few atomics)
constexpr int kAmaxSubs = 64;
static_assert(8 + 8 * kAmaxSubs == PTY_FP8_STATE_FLOATS, "fp8 state layout");
__device__ __forceinline__ float* amax_slot(float* state, int slot) {
  return state + 8 + slot * kAmaxSubs + (blockIdx.x % kAmaxSubs);
}
constexpr int kSlotH = 0;                                   // layer-0 input (features)
__device__ __forceinline__ int slot_z(int k) { return 1 + k; }      // input of hidden layer k
// layer-0 gradient Im(g_z conj(z0)) (the pre-activation gradient / w0)
constexpr int kSlotGu0 = 3;
// hidden layer k: g_w = g_z conj(z) (the pre-activation gradient / (-i w))
__device__ __forceinline__ int slot_gv(int k) { return 4 + k; }

__device__ __forceinline__ float amax8(const float* x) {
  float m = 0.f;
#pragma unroll
  for (int t = 0; t < 8; ++t) m = fmaxf(m, fabsf(x[t]));
  return m;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}


// ---------------------------------------------------------------- LDS
template <class F, int H>
struct LowpFwd {
  typename F::op_t w0[2][64];         // layer 0: lane (j,h), step q: W0[j][kappa0(8q+t, h)]
  typename F::op_t a[H][2][4][64];    // hidden k, M block (Re rows, Im rows), step q
};

template <class F, int H>
struct LowpBwd {
  typename F::op_t w0t[2][64];        // layer-0 transpose: W0[crow(8q+t, h)][phi0(j)]
  typename F::op_t at[H][2][4][64];   // hidden transpose, block = (Re inputs, Im inputs)
};

template <int H>
struct SmallLds {
  float b0[32];
  float cr[H][32];
  float ci[H][32];
  float om[H + 1];
  float wor[32], woi[32];
  float bo[2];
  float ws[H + 1];       // weight scales (W0, A_k)
  float xs[8];           // activation / gradient scales (fp8_state slots)
  unsigned int amax[H + 1];
  unsigned int rec[8];   // this workgroup's amax per fp8_state slot
};

// Workgroup amax of per-lane values v[k] (k <= H) into slots slot_of(k):
// LDS max, then one global atomic per slot.  Contains __syncthreads().
template <int H, class SlotOf>
__device__ void record_block_amax(SmallLds<H>& s, float* state, const float v[H + 1], SlotOf slot_of) {
  if (threadIdx.x < 8) s.rec[threadIdx.x] = 0u;
  __syncthreads();
#pragma unroll
  for (int k = 0; k <= H; ++k) {
    const float a = wave_max(v[k]);
    if ((threadIdx.x & 63) == 0 && a > 0.f && a < 3.0e38f) atomicMax(&s.rec[slot_of(k)], __float_as_uint(a));
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k <= H; ++k)
    if (threadIdx.x == 0 && s.rec[slot_of(k)] != 0u)
      atomicMax((unsigned int*)amax_slot(state, slot_of(k)), s.rec[slot_of(k)]);
}

template <int H>
__device__ void load_small(SmallLds<H>& s, const PtyMlp& m) {
  for (int t = threadIdx.x; t < 32; t += blockDim.x) {
    s.b0[t] = m.b0[t];
    for (int k = 0; k < H; ++k) {
      s.cr[k][t] = m.bh[k][2 * t];
      s.ci[k][t] = m.bh[k][2 * t + 1];
    }
    s.wor[t] = m.wo[2 * t];
    s.woi[t] = m.wo[2 * t + 1];
  }
  if (threadIdx.x <= H) s.om[threadIdx.x] = *m.omega[threadIdx.x];
  if (threadIdx.x == 0) { s.bo[0] = m.bo[0]; s.bo[1] = m.bo[1]; }
  if (threadIdx.x < 8) s.xs[threadIdx.x] = m.fp8_state ? m.fp8_state[threadIdx.x] : 1.f;
}

// Per-tensor weight scales 448 / amax (1 for unscaled formats).  Contains
// __syncthreads(); every thread of the workgroup calls it.
template <class F, int H>
__device__ void weight_scales(SmallLds<H>& s, const PtyMlp& m) {
  if (!F::kScaled) {
    if (threadIdx.x <= H) s.ws[threadIdx.x] = 1.f;
    __syncthreads();
    return;
  }
  if (threadIdx.x <= H) s.amax[threadIdx.x] = 0u;
  __syncthreads();
  float a0 = 0.f;
  for (int t = threadIdx.x; t < 32 * m.n_in; t += blockDim.x) a0 = fmaxf(a0, fabsf(m.w0[t]));
  atomicMax(&s.amax[0], __float_as_uint(a0));
  for (int k = 0; k < H; ++k) {
    float ak = 0.f;
    for (int t = threadIdx.x; t < 2048; t += blockDim.x) ak = fmaxf(ak, fabsf(m.wh[k][t]));
    atomicMax(&s.amax[1 + k], __float_as_uint(ak));
  }
  __syncthreads();
  if (threadIdx.x <= H) {
    const float a = __uint_as_float(s.amax[threadIdx.x]);
    s.ws[threadIdx.x] = (a > 0.f && a < 3.0e38f) ? 448.f / a : 1.f;
  }
  __syncthreads();
}

template <class F, int H>
__device__ void build_fwd_ops(LowpFwd<F, H>& o, const SmallLds<H>& s, const PtyMlp& m) {
  for (int idx = threadIdx.x; idx < 2 * 64; idx += blockDim.x) {
    const int q = idx >> 6, lane = idx & 63, j = lane & 31, h = lane >> 5;
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int f = kappa0(8 * q + t, h);
      v[t] = f < m.n_in ? m.w0[j * m.n_in + f] : 0.f;
    }
    o.w0[q][lane] = F::pack8(v, s.ws[0]);
  }
  for (int idx = threadIdx.x; idx < H * 2 * 4 * 64; idx += blockDim.x) {
    const int lane = idx & 63, q = (idx >> 6) & 3, mb = (idx >> 8) & 1, k = idx >> 9;
    const int j = lane & 31, h = lane >> 5;
    const float2* A = (const float2*)m.wh[k];
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int sidx = 8 * q + t;
      const float2 w = A[j * 32 + crow(sidx & 15, h)];   // A[out j][in u]
      v[t] = mb == 0 ? (sidx < 16 ? w.x : -w.y) : (sidx < 16 ? w.y : w.x);
    }
    o.a[k][mb][q][lane] = F::pack8(v, s.ws[1 + k]);
  }
}

template <class F, int H>
__device__ void build_bwd_ops(LowpBwd<F, H>& o, const SmallLds<H>& s, const PtyMlp& m) {
  for (int idx = threadIdx.x; idx < 2 * 64; idx += blockDim.x) {
    const int q = idx >> 6, lane = idx & 63, j = lane & 31, h = lane >> 5;
    const int f = phi0(j);
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = f < m.n_in ? m.w0[crow(8 * q + t, h) * m.n_in + f] : 0.f;
    o.w0t[q][lane] = F::pack8(v, s.ws[0]);
  }
  for (int idx = threadIdx.x; idx < H * 2 * 4 * 64; idx += blockDim.x) {
    const int lane = idx & 63, q = (idx >> 6) & 3, mb = (idx >> 8) & 1, k = idx >> 9;
    const int j = lane & 31, h = lane >> 5;
    const float2* A = (const float2*)m.wh[k];
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int sidx = 8 * q + t;
      const float2 w = A[crow(sidx & 15, h) * 32 + j];   // A[out u][in j]
      v[t] = mb == 0 ? (sidx < 16 ? w.x : w.y) : (sidx < 16 ? -w.y : w.x);
    }
    o.at[k][mb][q][lane] = F::pack8(v, s.ws[1 + k]);
  }
}

// u0 = b0 + W0 feat from the packed features
template <class F, int H>
__device__ __forceinline__ v16f layer0_lp(const LowpFwd<F, H>& o, const SmallLds<H>& s,
                                          const typename F::op_t pf[2], int lane, int h) {
  v16f acc;
  if (!F::kScaled) {
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = s.b0[crow(v, h)];
    acc = F::mfma(o.w0[0][lane], pf[0], acc);
    acc = F::mfma(o.w0[1][lane], pf[1], acc);
  } else {
    // the bias enters the accumulator in the scaled domain
    const float sc = s.ws[0] * s.xs[kSlotH], inv = 1.f / sc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = s.b0[crow(v, h)] * sc;
    acc = F::mfma(o.w0[0][lane], pf[0], acc);
    acc = F::mfma(o.w0[1][lane], pf[1], acc);
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] *= inv;
  }
  return acc;
}

// Hidden layer k from its packed input (4 operands in kappa1 order).
template <class F, int H>
__device__ __forceinline__ void hidden_lp(const LowpFwd<F, H>& o, const SmallLds<H>& s, int k,
                                          const typename F::op_t b[4], int lane, int h, v16f& vr, v16f& vi) {
  if (!F::kScaled) {
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      vr[v] = s.cr[k][crow(v, h)];
      vi[v] = s.ci[k][crow(v, h)];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      vr = F::mfma(o.a[k][0][q][lane], b[q], vr);
      vi = F::mfma(o.a[k][1][q][lane], b[q], vi);
    }
  } else {
    const float sc = s.ws[1 + k] * s.xs[slot_z(k)], inv = 1.f / sc;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      vr[v] = s.cr[k][crow(v, h)] * sc;
      vi[v] = s.ci[k][crow(v, h)] * sc;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      vr = F::mfma(o.a[k][0][q][lane], b[q], vr);
      vi = F::mfma(o.a[k][1][q][lane], b[q], vi);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      vr[v] *= inv;
      vi[v] *= inv;
    }
  }
}

// gabor_act that also returns max |z| over the lane's units (a bound of the
// components' amax)
__device__ __forceinline__ float gabor_act_mag(float w, const v16f& vr, const v16f& vi, float zr[16],
                                               float zi[16]) {
  float mm = 0.f;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    float c, sn;
    cis(w * vr[v], c, sn);
    const float mag = exp_f(-w * vi[v]);
    zr[v] = mag * c;
    zi[v] = mag * sn;
    mm = fmaxf(mm, mag);
  }
  return mm;
}

template <class F>
__device__ __forceinline__ void pack_z(const float zr[16], const float zi[16], float sc, typename F::op_t b[4]) {
  b[0] = F::pack8(zr, sc);
  b[1] = F::pack8(zr + 8, sc);
  b[2] = F::pack8(zi, sc);
  b[3] = F::pack8(zi + 8, sc);
}

// ============================================================== forward
template <class F, int H>
__global__ __launch_bounds__(kThreads) void k_inr_fwd_lp(PtyGrid g, PtyMlp m, const float2* __restrict__ table,
                                                         PtyPoints pts, float2* __restrict__ out) {
  __shared__ LowpFwd<F, H> o;
  __shared__ SmallLds<H> s;
  load_small<H>(s, m);
  weight_scales<F, H>(s, m);
  build_fwd_ops<F, H>(o, s, m);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  float mx[H + 1];
#pragma unroll
  for (int k = 0; k <= H; ++k) mx[k] = 0.f;
  const int64_t ntiles = (pts.n + 31) / 32;
  for (int64_t tile = (int64_t)blockIdx.x * kWaves + wave; tile < ntiles; tile += (int64_t)gridDim.x * kWaves) {
    asm volatile("" ::: "memory");
    const int64_t p = tile * 32 + j;
    const bool valid = p < pts.n;
    float x0 = 0.5f, x1 = 0.5f;
    if (valid) point_coords(pts, p, x0, x1);
    float feat[16];
    encode_lane(g, x0, x1, h, table, feat);
    if (pts.features) store_features32(pts.features, tile, lane, feat);
    typename F::op_t pf[2] = {F::pack8(feat, s.xs[kSlotH]), F::pack8(feat + 8, s.xs[kSlotH])};
    if (F::kScaled) mx[0] = fmaxf(mx[0], fmaxf(amax8(feat), amax8(feat + 8)));
    const v16f u0 = layer0_lp<F, H>(o, s, pf, lane, h);
    float zr[16], zi[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) cis(s.om[0] * u0[v], zr[v], zi[v]);
    // amax of the hidden-layer inputs, bounded by |z|: 1 after the real
    // layer 0, exp(-omega Im v) after a complex layer
    if (F::kScaled && H > 0) mx[1] = 1.f;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      typename F::op_t b[4];
      pack_z<F>(zr, zi, s.xs[slot_z(k)], b);
      v16f vr, vi;
      hidden_lp<F, H>(o, s, k, b, lane, h, vr, vi);
      if (F::kScaled && k + 1 < H) mx[2 + k] = fmaxf(mx[2 + k], gabor_act_mag(s.om[k + 1], vr, vi, zr, zi));
      else gabor_act(s.om[k + 1], vr, vi, zr, zi);
    }
    float orr = 0.f, oi = 0.f;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int u = crow(v, h);
      orr += s.wor[u] * zr[v] - s.woi[u] * zi[v];
      oi += s.wor[u] * zi[v] + s.woi[u] * zr[v];
    }
    orr += __shfl_xor(orr, 32);
    oi += __shfl_xor(oi, 32);
    if (valid && h == 0) out[p] = make_float2(orr + s.bo[0], oi + s.bo[1]);
  }
  if (F::kScaled && m.fp8_state)
    record_block_amax<H>(s, m.fp8_state, mx, [](int k) { return k == 0 ? kSlotH : slot_z(k - 1); });
}

// ============================================================== backward
template <class F, int H>
struct alignas(16) LowpStage {
  typename F::el_t X[H][64 * F::kStride];  // input z of hidden layer k, o-major, column = point
  typename F::el_t F_[32 * F::kStride];    // features, feature-major
  // gradient staging: rows of kStride elements of the format, or (output
  // layer) bf16 rows of 40
  alignas(16) uint16_t G16[64 * FmtBf16::kStride];
};

template <class F, int H>
__global__ __launch_bounds__(kThreads) void k_inr_bwd_lp(PtyGrid g, PtyMlp m, const float2* __restrict__ table,
                                                         float2* __restrict__ gfeat, PtyPoints pts,
                                                         const float2* __restrict__ dout,
                                                         float* __restrict__ partial) {
  using op_t = typename F::op_t;
  constexpr int BS = F::kStride;
  __shared__ LowpFwd<F, H> o;
  __shared__ LowpBwd<F, H> ot;
  __shared__ SmallLds<H> s;
  __shared__ LowpStage<F, H> st[kWaves];
  load_small<H>(s, m);
  weight_scales<F, H>(s, m);
  build_fwd_ops<F, H>(o, s, m);
  build_bwd_ops<F, H>(ot, s, m);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  LowpStage<F, H>& S = st[wave];
  typename F::el_t* SG = (typename F::el_t*)S.G16;
  const float sh = s.xs[kSlotH], sgu = s.xs[kSlotGu0];

  v16f dWr[H], dWi[H];
  v16f dW0;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    dW0[v] = 0.f;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      dWr[k][v] = 0.f;
      dWi[k][v] = 0.f;
    }
  }
  float acc_c[H][2];
  float mg[H + 1];
#pragma unroll
  for (int k = 0; k < H; ++k) acc_c[k][0] = acc_c[k][1] = 0.f;
#pragma unroll
  for (int k = 0; k <= H; ++k) mg[k] = 0.f;
  float acc_b0 = 0.f, acc_wo = 0.f, acc_bor = 0.f, acc_boi = 0.f;

  const int64_t ntiles = (pts.n + 31) / 32;
  for (int64_t tile = (int64_t)blockIdx.x * kWaves + wave; tile < ntiles; tile += (int64_t)gridDim.x * kWaves) {
    asm volatile("" ::: "memory");
    const int64_t p = tile * 32 + j;
    const bool valid = p < pts.n;
    // ---------------- recompute forward; keep only the packed inputs of
    // every layer (the pre-activations are not needed by the backward); the
    // features come from the forward's cache when there is one
    op_t pf[2];
    {
      float feat[16];
      if (pts.features) {
        load_features32(pts.features, tile, lane, feat);
      } else {
        float x0 = 0.5f, x1 = 0.5f;
        if (valid) point_coords(pts, p, x0, x1);
        encode_lane_batched(g, x0, x1, h, table, feat);
      }
#pragma unroll
      for (int v = 0; v < 16; ++v) S.F_[kappa0(v, h) * BS + j] = F::el(feat[v], sh);
      pf[0] = F::pack8(feat, sh);
      pf[1] = F::pack8(feat + 8, sh);
    }
    float zr[16], zi[16];
    {
      const v16f u0 = layer0_lp<F, H>(o, s, pf, lane, h);
#pragma unroll
      for (int v = 0; v < 16; ++v) cis(s.om[0] * u0[v], zr[v], zi[v]);
    }
    op_t px[H][4];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const float sz = s.xs[slot_z(k)];
      pack_z<F>(zr, zi, sz, px[k]);
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        S.X[k][crow(v, h) * BS + j] = F::el(zr[v], sz);
        S.X[k][(32 + crow(v, h)) * BS + j] = F::el(zi[v], sz);
      }
      v16f vr, vi;
      hidden_lp<F, H>(o, s, k, px[k], lane, h, vr, vi);
      gabor_act(s.om[k + 1], vr, vi, zr, zi);
    }
    // ---------------- output layer (FP32)
    float2 go = make_float2(0.f, 0.f);
    if (valid) go = dout[p];
    if (h == 0) { acc_bor += go.x; acc_boi += go.y; }
    float gzr[16], gzi[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int u = crow(v, h);
      // g_wo[u] += g_o * conj(z_u) (bf16 staging); g_z = conj(wo_u) * g_o
      S.G16[u * FmtBf16::kStride + j] = FmtBf16::el(go.x * zr[v] + go.y * zi[v], 1.f);
      S.G16[(32 + u) * FmtBf16::kStride + j] = FmtBf16::el(go.y * zr[v] - go.x * zi[v], 1.f);
      gzr[v] = s.wor[u] * go.x + s.woi[u] * go.y;
      gzi[v] = s.wor[u] * go.y - s.woi[u] * go.x;
    }
    wave_sync();
    {
      float t = 0.f;
#pragma unroll
      for (int c = 0; c < 32; c += 8) t += FmtBf16::sum8(FmtBf16::row8(S.G16, lane, c));
      acc_wo += t;
    }
    wave_sync();
    // ---------------- hidden layers, last to first (the staged gradient is
    // g_w = g_z conj(z); the factor -i w of g_v is applied to g_x here and
    // to the weight gradients by the finish kernel; dw comes from omega_terms)
#pragma unroll
    for (int k = H - 1; k >= 0; --k) {
      const float w = s.om[k + 1];
      const float sg = s.xs[slot_gv(k)];
      float gwr[16], gwi[16];
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        gwr[v] = gzr[v] * zr[v] + gzi[v] * zi[v];
        gwi[v] = gzi[v] * zr[v] - gzr[v] * zi[v];
        if (F::kScaled) mg[1 + k] = fmaxf(mg[1 + k], fmaxf(fabsf(gwr[v]), fabsf(gwi[v])));
        const int u = crow(v, h);
        SG[u * BS + j] = F::el(gwr[v], sg);
        SG[(32 + u) * BS + j] = F::el(gwi[v], sg);
      }
      wave_sync();
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int c0 = 16 * q + 8 * h;
        const op_t a0 = F::row8(SG, j, c0), a1 = F::row8(SG, 32 + j, c0);
        const op_t b0 = F::row8(S.X[k], j, c0), b1 = F::row8(S.X[k], 32 + j, c0);
        dWr[k] = F::mfma(a0, b0, dWr[k]);
        dWr[k] = F::mfma(a1, b1, dWr[k]);
        dWi[k] = F::mfma(a1, b0, dWi[k]);
        dWi[k] = F::mfma(F::neg(a0), b1, dWi[k]);
        acc_c[k][0] += F::sum8(a0);
        acc_c[k][1] += F::sum8(a1);
      }
      // t = A^H g_w (MFMA), g_x = -i w t
      v16f txr, txi;
#pragma unroll
      for (int v = 0; v < 16; ++v) { txr[v] = 0.f; txi[v] = 0.f; }
      const op_t bq[4] = {F::pack8(gwr, sg), F::pack8(gwr + 8, sg), F::pack8(gwi, sg), F::pack8(gwi + 8, sg)};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        txr = F::mfma(ot.at[k][0][q][lane], bq[q], txr);
        txi = F::mfma(ot.at[k][1][q][lane], bq[q], txi);
      }
      const float ginv = (F::kScaled ? 1.f / (s.ws[1 + k] * sg) : 1.f) * w;
      const float zinv = F::kScaled ? 1.f / s.xs[slot_z(k)] : 1.f;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        gzr[v] = txi[v] * ginv;
        gzi[v] = -txr[v] * ginv;
        zr[v] = F::val(px[k][v >> 3], v & 7) * zinv;
        zi[v] = F::val(px[k][2 + (v >> 3)], v & 7) * zinv;
      }
      wave_sync();
    }
    // ---------------- layer 0 (real): staged Im(g_z conj(z0)) = g_u0 / w0
    {
      const float w = s.om[0];
      float gu[16];
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        gu[v] = gzi[v] * zr[v] - gzr[v] * zi[v];
        if (F::kScaled) mg[0] = fmaxf(mg[0], fabsf(gu[v]));
        SG[crow(v, h) * BS + j] = F::el(gu[v], sgu);
      }
      wave_sync();
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int c0 = 16 * q + 8 * h;
        const op_t a = F::row8(SG, j, c0);
        dW0 = F::mfma(a, F::row8(S.F_, j, c0), dW0);
        acc_b0 += F::sum8(a);
      }
      v16f gh;
#pragma unroll
      for (int v = 0; v < 16; ++v) gh[v] = 0.f;
      gh = F::mfma(ot.w0t[0][lane], F::pack8(gu, sgu), gh);
      gh = F::mfma(ot.w0t[1][lane], F::pack8(gu + 8, sgu), gh);
      const float hinv = (F::kScaled ? 1.f / (s.ws[0] * sgu) : 1.f) * w;
      wave_sync();
      if (valid) {
#pragma unroll
        for (int mm = 0; mm < 8; ++mm) {
          const int l = 2 * mm + h;
          if (l < g.n_levels)
            gfeat[(int64_t)l * pts.n + p] = make_float2(gh[2 * mm] * hinv, gh[2 * mm + 1] * hinv);
        }
      }
    }
  }
  if (F::kScaled && m.fp8_state)
    record_block_amax<H>(s, m.fp8_state, mg, [](int k) { return k == 0 ? kSlotGu0 : slot_gv(k - 1); });

  // ---------------- workgroup reduction into partial[blockIdx.x][GradLayout::total]
  // (scaled formats: the MFMA sums of staged operands are multiplied by
  // 1 / (s_G s_X), the bias sums by 1 / s_G)
  // Waves 0-1 store their sums into two LDS regions, waves 2-3 add theirs
  // into region wave - 2, then every thread totals its slots over the two
  // regions: two barriers instead of one per wave.
  using GL = GradLayout<H>;
  constexpr int kRegions = kWaves / 2;
  static_assert(kRegions == 2, "the summation below assumes two regions");
  static_assert(sizeof(st) / (sizeof(float)) >= (size_t)kRegions * GL::total, "reduction buffer does not fit");
  float dwi0 = 1.f, dwik[H], cik[H], b0i = 1.f;
#pragma unroll
  for (int k = 0; k < H; ++k) dwik[k] = cik[k] = 1.f;
  if (F::kScaled) {
    dwi0 = 1.f / (sgu * sh);
    b0i = 1.f / sgu;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      dwik[k] = 1.f / (s.xs[slot_gv(k)] * s.xs[slot_z(k)]);
      cik[k] = 1.f / s.xs[slot_gv(k)];
    }
  }
  // cross-lane sums first, with every lane of the wave taking part
  float cr[H], ci[H];
#pragma unroll
  for (int k = 0; k < H; ++k) {
    cr[k] = acc_c[k][0] + __shfl_xor(acc_c[k][0], 32);
    ci[k] = acc_c[k][1] + __shfl_xor(acc_c[k][1], 32);
  }
  const float b0s = acc_b0 + __shfl_xor(acc_b0, 32);
  const float br = wave_sum(acc_bor), bi = wave_sum(acc_boi);
  __syncthreads();
  float* red = (float*)&st[0];
  float* R = red + (wave % kRegions) * GL::total;
  for (int phase = 0; phase < 2; ++phase) {
    if (wave / kRegions == phase) {
      const bool add = phase == 1;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int oo = crow(v, h);
        float* r0 = &R[GL::w0 + oo * 32 + j];
        *r0 = (add ? *r0 : 0.f) + dW0[v] * dwi0;
#pragma unroll
        for (int k = 0; k < H; ++k) {
          float* rr = &R[GL::A(k) + 2 * (oo * 32 + j)];
          rr[0] = (add ? rr[0] : 0.f) + dWr[k][v] * dwik[k];
          rr[1] = (add ? rr[1] : 0.f) + dWi[k][v] * dwik[k];
        }
      }
      if (h == 0) {
#pragma unroll
        for (int k = 0; k < H; ++k) {
          float* rc = &R[GL::c(k) + 2 * j];
          rc[0] = (add ? rc[0] : 0.f) + cr[k] * cik[k];
          rc[1] = (add ? rc[1] : 0.f) + ci[k] * cik[k];
        }
        R[GL::b0 + j] = (add ? R[GL::b0 + j] : 0.f) + b0s * b0i;
      }
      float* rw = &R[GL::wo + 2 * (lane & 31) + (lane >> 5)];
      *rw = (add ? *rw : 0.f) + acc_wo;
      if (lane == 0) {
        R[GL::bo] = (add ? R[GL::bo] : 0.f) + br;
        R[GL::bo + 1] = (add ? R[GL::bo + 1] : 0.f) + bi;
      }
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < GL::total; t += blockDim.x)
    if (t < GL::om || t >= GL::wo) red[t] += red[GL::total + t];
  __syncthreads();
  omega_terms<H>(red, m);
  for (int t = threadIdx.x; t < GL::total; t += blockDim.x) partial[(int64_t)blockIdx.x * GL::total + t] = red[t];
}

// Scales of the next step from the amax recorded in this one: thread i < 8
// reduces slot i over its sub-slots in a fixed order and clears them;
// scale = 224 / amax.
__global__ void k_fp8_roll(float* state) {
  const int i = threadIdx.x;
  if (blockIdx.x != 0 || i >= 8) return;
  float a = 0.f;
  float* sub = state + 8 + i * kAmaxSubs;
  for (int r = 0; r < kAmaxSubs; ++r) {
    a = fmaxf(a, sub[r]);
    sub[r] = 0.f;
  }
  // one binade of headroom: a tensor may grow 2x between steps before its
  // largest values saturate
  if (a > 0.f && a < 3.0e38f) state[i] = 224.f / a;
}

template <class F>
static void launch_fwd(const PtyGrid& g, const PtyMlp& m, const float2* table, const PtyPoints& pts, float2* out,
                       int blocks, hipStream_t stream) {
  if (m.n_hidden == 1)
    hipLaunchKernelGGL((k_inr_fwd_lp<F, 1>), dim3(blocks), dim3(kThreads), 0, stream, g, m, table, pts, out);
  else
    hipLaunchKernelGGL((k_inr_fwd_lp<F, 2>), dim3(blocks), dim3(kThreads), 0, stream, g, m, table, pts, out);
}

template <class F>
static void launch_bwd(const PtyGrid& g, const PtyMlp& m, const float2* table, float2* gfeat, const PtyPoints& pts,
                       const float2* dout, float* partial, int blocks, hipStream_t stream) {
  if (m.n_hidden == 1)
    hipLaunchKernelGGL((k_inr_bwd_lp<F, 1>), dim3(blocks), dim3(kThreads), 0, stream, g, m, table, gfeat, pts, dout,
                       partial);
  else
    hipLaunchKernelGGL((k_inr_bwd_lp<F, 2>), dim3(blocks), dim3(kThreads), 0, stream, g, m, table, gfeat, pts, dout,
                       partial);
}

void launch_inr_fwd_lowp(const PtyGrid& g, const PtyMlp& m, const float2* table, const PtyPoints& pts, float2* out,
                         int blocks, hipStream_t stream) {
  if (m.precision == PTY_PREC_FP8) launch_fwd<FmtFp8>(g, m, table, pts, out, blocks, stream);
  else launch_fwd<FmtBf16>(g, m, table, pts, out, blocks, stream);
}

void launch_inr_bwd_lowp(const PtyGrid& g, const PtyMlp& m, const float2* table, float2* gfeat, const PtyPoints& pts,
                         const float2* dout, float* partial, int blocks, hipStream_t stream) {
  if (m.precision == PTY_PREC_FP8) launch_bwd<FmtFp8>(g, m, table, gfeat, pts, dout, partial, blocks, stream);
  else launch_bwd<FmtBf16>(g, m, table, gfeat, pts, dout, partial, blocks, stream);
}

}  // namespace pty

using namespace pty;

extern "C" int pty_fp8_roll(float* state, hipStream_t stream) {
  PTY_REQUIRE(state, PTY_EINVAL, "fp8_roll: null state");
  hipLaunchKernelGGL(k_fp8_roll, dim3(1), dim3(64), 0, stream, state);
  return check_launch("fp8_roll");
}
