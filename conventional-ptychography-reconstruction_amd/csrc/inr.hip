// Entry points of the fused hash-grid encode + complex-exponential MLP
// (gfx950): argument checks, workspace layout, the fixed-order reduction of
// the per-workgroup gradient partials, and the MLP over given features.
//
// Reference: ComplexINRModel2D.forward -> HashGrid (models/complex_inr.py:279,
// 294), ComplexMLP.net (models/complex_inr.py:57-78, 313):
//   u0 = W0 h + b0            z0 = exp(i w0 u0)                 (real first layer)
//   v_k = A_k z_{k-1} + c_k   z_k = exp(i w_k v_k) = e^{-w Im v}(cos w Re v, sin w Re v)
//   o = wo . z_H + bo
// The fused kernels are in inr_f32.hip (FP32, 16-point wave tiles on
// v_mfma_f32_16x16x4f32) and inr_lowp.hip (BF16 / FP8, 32-point tiles).  The
// MLP over given features (k_mlp_fwd) uses the 32-point FP32 layout of
// inr_common.h.
//
// The backward kernels accumulate, over the points, P = sum g_w x^H and
// q = sum g_w per hidden layer (g_w = g_z conj(z), the pre-activation
// gradient g_v = -i w g_w without its factor) and P0 = sum Im(g_z conj(z0)) h^T,
// q0 = sum Im(g_z conj(z0)) for the first layer.  The finish kernel turns the
// totals into the parameter gradients, dA = -i w P, dc = -i w q,
// dW0 = w0 P0, db0 = w0 q0.  The frequency gradients follow from the
// homogeneity of each layer in its weights (v = A x + c, so
// sum g_w conj(v) = <conj(A), P> + <conj(c), q>):
//   dw = Im(sum conj(A) P + sum conj(c) q),  dw0 = sum W0 P0 + sum b0 q0,
// which equals sum Im(g_w conj(v)) over the points without keeping the
// pre-activations and holds for w = 0 as well; being linear in P and q, each
// backward workgroup forms it for its own share (omega_terms) and the finish
// totals the shares.
#include "inr_common.h"

namespace pty {
// ComplexMLP.net over precomputed features [n, n_in] (models/complex_inr.py:313).
template <int H>
__global__ __launch_bounds__(kThreads) void k_mlp_fwd(PtyMlp m, const float* __restrict__ feats, int64_t n,
                                                      float2* __restrict__ out) {
  __shared__ MlpLds<H> s;
  load_weights<H>(s, m);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  const int64_t ntiles = (n + 31) / 32;
  for (int64_t tile = (int64_t)blockIdx.x * kWaves + wave; tile < ntiles; tile += (int64_t)gridDim.x * kWaves) {
    // weights are re-read from LDS at each use instead of being hoisted into
    // ~270 registers for the whole loop (which forces one wave per SIMD)
    asm volatile("" ::: "memory");
    const int64_t p = tile * 32 + j;
    const bool valid = p < n;
    float feat[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int f = kappa0(v, h);
      feat[v] = (valid && f < m.n_in) ? feats[p * m.n_in + f] : 0.f;
    }
    v16f u0 = layer0<H>(s, feat, j, h);
    float zr[16], zi[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) cis(s.om[0] * u0[v], zr[v], zi[v]);
#pragma unroll
    for (int k = 0; k < H; ++k) {
      v16f vr, vi;
      hidden_fwd<H>(s, k, zr, zi, j, h, vr, vi);
      gabor_act(s.om[k + 1], vr, vi, zr, zi);
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
}

// Fixed-order reduction of the workgroup partials into the parameter
// gradients, one launch: workgroup c covers the 32 gradient slots
// [32c, 32c + 32); its 8 row groups of 32 threads each sum the partials of
// workgroups rg, rg + 8, ... (in order), then the 8 group sums are added in
// group order.  With total t of a slot:
//   w0 / b0 : += w0 t
//   A_k / c_k (complex pairs (re, im)): += -i w_k t, i.e. (w Im t, -w Re t)
//   omega_k : += t   (each workgroup stored its frequency terms in these slots)
//   wo / bo : += t
constexpr int kFinishCols = 32;
constexpr int kFinishGroups = 8;
constexpr int kFinishBatch = 32;

template <int H>
__global__ __launch_bounds__(256) void k_inr_grad_finish(const float* __restrict__ partial, int nblocks, PtyMlp m,
                                                         PtyMlp gm) {
  using GL = GradLayout<H>;
  __shared__ float sums[kFinishGroups][kFinishCols];
  __shared__ float tot[kFinishCols];
  const int c = threadIdx.x & (kFinishCols - 1), rg = threadIdx.x / kFinishCols;
  const int t = blockIdx.x * kFinishCols + c;
  float acc = 0.f;
  if (t < GL::total) {
    // batches of kFinishBatch partials with their loads in flight together,
    // added in workgroup order (a missing one adds 0)
    for (int b0 = rg; b0 < nblocks; b0 += kFinishBatch * kFinishGroups) {
      float v[kFinishBatch];
#pragma unroll
      for (int u = 0; u < kFinishBatch; ++u) {
        const int b = b0 + u * kFinishGroups;
        v[u] = b < nblocks ? partial[(int64_t)b * GL::total + t] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < kFinishBatch; ++u) acc += v[u];
    }
  }
  sums[rg][c] = acc;
  __syncthreads();
  if (rg == 0) {
    float s = sums[0][c];
#pragma unroll
    for (int r = 1; r < kFinishGroups; ++r) s += sums[r][c];
    tot[c] = s;
  }
  __syncthreads();
  if (rg != 0 || t >= GL::total) return;
  const float a = tot[c];
  if (t < GL::b0) {
    const int o = t >> 5, f = t & 31;
    if (f < gm.n_in) gm.w0[o * gm.n_in + f] += *m.omega[0] * a;
  } else if (t < GL::A(0)) {
    gm.b0[t - GL::b0] += *m.omega[0] * a;
  } else if (t < GL::om) {
    const int k = (t - GL::A(0)) / 2112, r = (t - GL::A(0)) % 2112;
    const float w = *m.omega[k + 1];
    const float other = tot[c ^ 1];          // A(k) and c(k) start at even offsets
    const bool re = (r & 1) == 0;
    const float tr = re ? a : other, ti = re ? other : a;
    float* gdst = r < 2048 ? gm.wh[k] + r : gm.bh[k] + (r - 2048);
    *gdst += re ? w * ti : -w * tr;
  } else if (t < GL::wo) {
    *gm.omega[t - GL::om] += a;
  } else if (t < GL::bo) {
    gm.wo[t - GL::wo] += a;
  } else {
    gm.bo[t - GL::bo] += a;
  }
}

static int validate_mlp(const PtyMlp* m, bool need_all) {
  PTY_REQUIRE(m != nullptr, PTY_EINVAL, "mlp descriptor is NULL");
  PTY_REQUIRE(m->n_in >= 1 && m->n_in <= 32, PTY_EUNSUP, "mlp: n_in %d outside 1..32", m->n_in);
  PTY_REQUIRE(m->n_hidden >= 1 && m->n_hidden <= 2, PTY_EUNSUP, "mlp: hidden_layers %d outside 1..2", m->n_hidden);
  PTY_REQUIRE(m->precision == PTY_PREC_FP32 || m->precision == PTY_PREC_BF16 || m->precision == PTY_PREC_FP8,
              PTY_EUNSUP, "mlp: precision %d not built", m->precision);
  PTY_REQUIRE(m->precision != PTY_PREC_FP8 || m->fp8_state, PTY_EINVAL, "mlp: FP8 needs fp8_state");
  if (need_all) {
    PTY_REQUIRE(m->w0 && m->b0 && m->wo && m->bo, PTY_EINVAL, "mlp: null weight pointer");
    for (int k = 0; k < m->n_hidden; ++k)
      PTY_REQUIRE(m->wh[k] && m->bh[k], PTY_EINVAL, "mlp: null hidden weight pointer");
    for (int k = 0; k <= m->n_hidden; ++k) PTY_REQUIRE(m->omega[k], PTY_EINVAL, "mlp: null omega pointer");
  }
  return PTY_OK;
}

static int validate_points(const PtyPoints* p) {
  PTY_REQUIRE(p != nullptr, PTY_EINVAL, "points descriptor is NULL");
  PTY_REQUIRE(p->n >= 0, PTY_EINVAL, "points: negative count");
  if (!p->coords) {
    PTY_REQUIRE(p->axis0 && p->axis1 && p->n_axis0 > 0 && p->n_axis1 > 0, PTY_EINVAL, "points: no coordinates");
    PTY_REQUIRE((int64_t)p->n_axis0 * p->n_axis1 == p->n, PTY_EINVAL, "points: n != n_axis0*n_axis1");
  }
  PTY_REQUIRE(((uintptr_t)p->features & 15u) == 0, PTY_EINVAL, "points: feature cache must be 16-byte aligned");
  return PTY_OK;
}

// One workgroup per CU (a backward workgroup fills the CU's LDS): every wave
// walks its tiles in one round, and the per-workgroup set-up and gradient
// flush are paid once per CU.  A workgroup covers 128 points per round in
// every variant (4 waves x 32 points, or 8 waves x 16 points in FP32).
static int bwd_blocks(int64_t n) {
  const int64_t tiles = (n + 31) / 32;
  int64_t b = (tiles + kWaves - 1) / kWaves;
  const int cus = device_cu_count();
  if (b > cus) b = cus;
  if (b < 1) b = 1;
  return (int)b;
}

// Forward grids: up to four 4-wave workgroups per CU (the forward kernels
// run three to four waves per SIMD), each wave walking several tiles, so the
// weights are staged in LDS once per workgroup rather than once per tile.
static int fwd_blocks(int64_t n) {
  const int64_t tiles = (n + 31) / 32;
  int64_t b = (tiles + kWaves - 1) / kWaves;
  const int64_t cap = 4 * (int64_t)device_cu_count();
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace pty

using namespace pty;

extern "C" size_t pty_points_feature_floats(int64_t n) {
  if (n <= 0) return 0;
  return 32 * (size_t)((n + 31) / 32) * 32;
}

extern "C" int pty_inr_fwd(const PtyGrid* g, const PtyMlp* mlp, const float* table, const PtyPoints* pts, float* out,
                           hipStream_t stream) {
  int rc = validate_grid(g);
  if (!rc) rc = validate_mlp(mlp, true);
  if (!rc) rc = validate_points(pts);
  if (rc) return rc;
  PTY_REQUIRE(mlp->n_in == g->n_levels * g->n_features, PTY_EINVAL, "mlp n_in %d != levels*features %d", mlp->n_in,
              g->n_levels * g->n_features);
  PTY_REQUIRE(table && out, PTY_EINVAL, "inr_fwd: null pointer");
  if (pts->n == 0) return PTY_OK;
  if (mlp->precision == PTY_PREC_FP8)
    launch_inr_fwd_lowp(*g, *mlp, (const float2*)table, *pts, (float2*)out, fwd_blocks(pts->n), stream);
  else if (mlp->precision == PTY_PREC_BF16)
    launch_inr_fwd_b16(*g, *mlp, (const float2*)table, *pts, (float2*)out, stream);
  else
    launch_inr_fwd_f32(*g, *mlp, (const float2*)table, *pts, (float2*)out, stream);
  return check_launch("inr_fwd");
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

static size_t grad_total(const PtyMlp* mlp) {
  return (mlp->n_hidden == 1) ? GradLayout<1>::total : GradLayout<2>::total;
}

// per-workgroup gradient partials
static size_t partial_bytes(const PtyMlp* mlp, const PtyPoints* pts) {
  return align256(sizeof(float) * grad_total(mlp) * (size_t)bwd_blocks(pts->n));
}

extern "C" size_t pty_inr_bwd_workspace(const PtyGrid* g, const PtyMlp* mlp, const PtyPoints* pts) {
  if (!g || !mlp || !pts) return 0;
  return partial_bytes(mlp, pts) + align256(sizeof(float2) * (size_t)g->n_levels * (size_t)pts->n) +
         table_bwd_workspace(*g, *pts);
}

static int check_bwd_args(const PtyGrid* g, const PtyMlp* mlp, const PtyPoints* pts, void* work,
                          size_t work_bytes) {
  int rc = validate_grid(g);
  if (!rc) rc = validate_mlp(mlp, true);
  if (!rc) rc = validate_points(pts);
  if (rc) return rc;
  PTY_REQUIRE(mlp->n_in == g->n_levels * g->n_features, PTY_EINVAL, "mlp n_in != levels*features");
  if (pts->n == 0) return PTY_OK;
  PTY_REQUIRE(work && work_bytes >= pty_inr_bwd_workspace(g, mlp, pts), PTY_EWORK, "inr_bwd: workspace too small");
  return PTY_OK;
}

// Workspace layout shared by the two stages: per-block MLP gradient partials,
// the feature gradient [n_levels][n] c64, then the table-gradient scratch.
struct BwdWork {
  float* partial;
  float2* gfeat;
  char* tab_work;
  int blocks, total;
};

static BwdWork bwd_work(const PtyGrid* g, const PtyMlp* mlp, const PtyPoints* pts, void* work) {
  BwdWork b;
  char* w = (char*)work;
  b.blocks = bwd_blocks(pts->n);
  b.total = (int)grad_total(mlp);
  b.partial = (float*)w;
  b.gfeat = (float2*)(w + partial_bytes(mlp, pts));
  b.tab_work = (char*)b.gfeat + align256(sizeof(float2) * (size_t)g->n_levels * (size_t)pts->n);
  return b;
}

extern "C" int pty_inr_bwd_mlp(const PtyGrid* g, const PtyMlp* mlp, const float* table, const PtyPoints* pts,
                               const float* dout, void* work, size_t work_bytes, hipStream_t stream) {
  int rc = check_bwd_args(g, mlp, pts, work, work_bytes);
  if (rc) return rc;
  PTY_REQUIRE(table && dout, PTY_EINVAL, "inr_bwd: null pointer");
  if (pts->n == 0) return PTY_OK;
  const BwdWork b = bwd_work(g, mlp, pts, work);
  if (mlp->precision == PTY_PREC_FP8)
    launch_inr_bwd_lowp(*g, *mlp, (const float2*)table, b.gfeat, *pts, (const float2*)dout, b.partial, b.blocks,
                        stream);
  else if (mlp->precision == PTY_PREC_BF16)
    launch_inr_bwd_b16(*g, *mlp, (const float2*)table, b.gfeat, *pts, (const float2*)dout, b.partial, b.blocks,
                       stream);
  else
    launch_inr_bwd_f32(*g, *mlp, (const float2*)table, b.gfeat, *pts, (const float2*)dout, b.partial, b.blocks,
                       stream);
  return check_launch("inr_bwd_mlp");
}

static int inr_bwd_finish(const PtyGrid* g, const PtyMlp* mlp, const PtyMlp* grad, float* dtable,
                          const PtyPoints* pts, void* work, size_t work_bytes, hipStream_t stream,
                          const TableAdam& ad) {
  int rc = check_bwd_args(g, mlp, pts, work, work_bytes);
  if (!rc) rc = validate_mlp(grad, true);
  if (rc) return rc;
  PTY_REQUIRE(grad->n_in == mlp->n_in && grad->n_hidden == mlp->n_hidden, PTY_EINVAL, "grad descriptor mismatch");
  PTY_REQUIRE(dtable || ad.p, PTY_EINVAL, "inr_bwd: null pointer");
  if (pts->n == 0) return PTY_OK;
  const BwdWork b = bwd_work(g, mlp, pts, work);
  const dim3 grid((b.total + kFinishCols - 1) / kFinishCols);
  if (mlp->n_hidden == 1)
    hipLaunchKernelGGL(k_inr_grad_finish<1>, grid, dim3(kFinishCols * kFinishGroups), 0, stream,
                       (const float*)b.partial, b.blocks, *mlp, *grad);
  else
    hipLaunchKernelGGL(k_inr_grad_finish<2>, grid, dim3(kFinishCols * kFinishGroups), 0, stream,
                       (const float*)b.partial, b.blocks, *mlp, *grad);
  rc = check_launch("inr_bwd_finish");
  if (rc) return rc;
  return launch_table_bwd(*g, *pts, b.gfeat, dtable, b.tab_work, stream, ad);
}

extern "C" int pty_inr_bwd_finish(const PtyGrid* g, const PtyMlp* mlp, const PtyMlp* grad, float* dtable,
                                  const PtyPoints* pts, void* work, size_t work_bytes, hipStream_t stream) {
  TableAdam none{};
  PTY_REQUIRE(dtable, PTY_EINVAL, "inr_bwd: null pointer");
  return inr_bwd_finish(g, mlp, grad, dtable, pts, work, work_bytes, stream, none);
}

extern "C" int pty_inr_bwd_finish_adam(const PtyGrid* g, const PtyMlp* mlp, const PtyMlp* grad, float* table,
                                       float* exp_avg, float* exp_avg_sq, const float* hyper, float beta1,
                                       float beta2, float eps, const PtyPoints* pts, void* work, size_t work_bytes,
                                       hipStream_t stream) {
  PTY_REQUIRE(table && exp_avg && exp_avg_sq && hyper, PTY_EINVAL, "inr_bwd_finish_adam: null pointer");
  PTY_REQUIRE(((((uintptr_t)table) | ((uintptr_t)exp_avg) | ((uintptr_t)exp_avg_sq)) & 7) == 0, PTY_EINVAL,
              "inr_bwd_finish_adam: table and moments must be 8-byte aligned");
  PTY_REQUIRE(pts && pts->n > 0 && !pts->coords, PTY_EINVAL,
              "inr_bwd_finish_adam: needs a non-empty grid of points (axis vectors)");
  const TableAdam ad{(float2*)table, (float2*)exp_avg, (float2*)exp_avg_sq, hyper, beta1, beta2, eps};
  return inr_bwd_finish(g, mlp, grad, nullptr, pts, work, work_bytes, stream, ad);
}

extern "C" int pty_inr_bwd(const PtyGrid* g, const PtyMlp* mlp, const PtyMlp* grad, const float* table, float* dtable,
                           const PtyPoints* pts, const float* dout, void* work, size_t work_bytes,
                           hipStream_t stream) {
  int rc = pty_inr_bwd_mlp(g, mlp, table, pts, dout, work, work_bytes, stream);
  if (rc) return rc;
  return pty_inr_bwd_finish(g, mlp, grad, dtable, pts, work, work_bytes, stream);
}

extern "C" int pty_mlp_fwd(const PtyMlp* mlp, const float* feats, int64_t n, float* out, hipStream_t stream) {
  int rc = validate_mlp(mlp, true);
  if (rc) return rc;
  PTY_REQUIRE(n >= 0 && (n == 0 || (feats && out)), PTY_EINVAL, "mlp_fwd: null pointer");
  if (n == 0) return PTY_OK;
  const int blocks = fwd_blocks(n);
  if (mlp->n_hidden == 1)
    hipLaunchKernelGGL(k_mlp_fwd<1>, dim3(blocks), dim3(kThreads), 0, stream, *mlp, feats, n, (float2*)out);
  else
    hipLaunchKernelGGL(k_mlp_fwd<2>, dim3(blocks), dim3(kThreads), 0, stream, *mlp, feats, n, (float2*)out);
  return check_launch("mlp_fwd");
}
