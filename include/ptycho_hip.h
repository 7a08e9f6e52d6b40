/*
 * ptycho_hip.h -- C ABI of the MI355X (gfx950) INR-ptychography hot path.
 *
 * Every entry point takes raw device pointers, plain sizes and a hipStream_t,
 * launches asynchronously on that stream, never allocates device memory and
 * never synchronises (caller-provided workspace, sized by the *_workspace
 * queries); the one exception is pty_gs_num_rendered, which reads a count
 * back to the host as the reference's rasterizer does.  Return value: 0 on success, a negative PTY_E* code on error,
 * with a message available from pty_last_error() (thread-local).
 *
 * Complex arrays are interleaved (re, im) float pairs, i.e. the memory of a
 * torch complex64 tensor.  2-D arrays are row-major.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   tcnn.Encoding "HashGrid" fwd/bwd   models/complex_inr.py:130-145, 279, 294
 *   ComplexMLP.net (Gabor layers)      models/complex_inr.py:6-82, 282, 295, 313
 *   fourier_upsample + its autograd    models/complex_inr.py:252-271, 287
 *   apply_residual                     models/complex_inr.py:316-334
 *   probe COM roll + energy norm       Recovery.py:262-285
 *   per-position far field + losses    Recovery.py:287-307 (utils/optics.py:102-112)
 *   haar_wavelet_sparsity_loss         utils/loss.py:60-83, Recovery.py:309-313
 *   torch.optim.Adam step              Recovery.py:238-241, 321
 *   Gaussian-field rasterizer (gscp)   gscp/cuda/_wrapper.py:18-166, gscp/csrc
 */
#ifndef PTYCHO_HIP_H
#define PTYCHO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* hipStream_t;

#define PTY_ABI_VERSION 7
#define PTY_MAX_LEVELS 16
#define PTY_MAX_HIDDEN 4
#define PTY_FP8_STATE_FLOATS 520

enum {
  PTY_OK = 0,
  PTY_EINVAL = -1,    /* bad argument / descriptor */
  PTY_EUNSUP = -2,    /* shape or option outside what the kernels implement */
  PTY_EWORK = -3,     /* workspace too small */
  PTY_ELAUNCH = -4,   /* HIP launch error */
};

enum { PTY_LOSS_SMOOTH_L1 = 0, PTY_LOSS_FD = 1, PTY_LOSS_POISSON = 2 };
enum { PTY_PREC_FP32 = 0, PTY_PREC_BF16 = 1, PTY_PREC_FP8 = 2 };
/* Hash-table storage read by the forward kernels. */
enum { PTY_TABLE_F32 = 0, PTY_TABLE_F16 = 1 };

/* Multiresolution hash-grid layout (2-D input, F = 2 features per level).
 * Built on the host from (n_levels, log2_hashmap_size, base_resolution,
 * per_level_scale): scale[l] = f32(exp2(l*log2(pls)))*base - 1 (f32 steps),
 * res[l] = ceil(scale[l]) + 1, size[l] = min(next_mult8(res^2), 2^log2T),
 * offset[l] = sum of earlier sizes, dense when res^2 <= size.
 * Table memory: fp32 [n_entries * 2], entry-major, features interleaved
 * (the layout of tcnn.Encoding.params). */
typedef struct PtyGrid {
  int32_t abi_version;
  int32_t n_levels;
  int32_t n_features;
  int32_t log2_hashmap_size;
  float scale[PTY_MAX_LEVELS];
  uint32_t res[PTY_MAX_LEVELS];
  uint32_t size[PTY_MAX_LEVELS];
  uint32_t offset[PTY_MAX_LEVELS];
  uint32_t dense_mask;
  uint32_t n_entries;
  /* Storage of the table that forward kernels read: PTY_TABLE_F32 (float
   * [n_entries*2]) or PTY_TABLE_F16 (half [n_entries*2], the FP16 copy that
   * pty_cast_f16 makes of the f32 master).  Gradients are always f32. */
  uint32_t table_precision;
} PtyGrid;

/* Complex-exponential MLP (ComplexMLP with hidden_features = 64, i.e. 32
 * complex units).  Device pointers; complex tensors interleaved.
 *   w0 [32, n_in] f32, b0 [32] f32            first (real) Gabor layer
 *   wh[k] [32, 32] c64, bh[k] [32] c64        hidden Gabor layers k < n_hidden
 *   omega[k] -> [1] f32                       omega_0 of layer k (0 = first)
 *   wo [1, 32] c64, bo [1] c64                final complex Linear(32, 1)
 * For gradients (PtyMlp used as the gradient descriptor) every pointer is an
 * accumulation target of the same shape (the kernels add into it). */
typedef struct PtyMlp {
  int32_t n_in;        /* = n_levels * n_features, <= 32 */
  int32_t n_hidden;    /* 0 .. PTY_MAX_HIDDEN */
  int32_t precision;   /* PTY_PREC_* for the GEMMs (accumulation is fp32) */
  int32_t pad_;
  float* w0;
  float* b0;
  float* wh[PTY_MAX_HIDDEN];
  float* bh[PTY_MAX_HIDDEN];
  float* omega[PTY_MAX_HIDDEN + 1];
  float* wo;
  float* bo;
  /* PTY_PREC_FP8 only (else ignored, may be NULL): PTY_FP8_STATE_FLOATS
   * floats of delayed per-tensor scaling state, [0, 8) the scales in use
   * (224 / amax of the previous step, one binade below the e4m3 maximum 448;
   * 1 before any), then the amax recorded
   * by this step's kernels (features, hidden inputs, and the backward's
   * activation-side gradients g_z conj(z), i.e. the pre-activation gradients
   * without their -i*omega factor; 64 sub-slots per tensor,
   * zero-initialised).  The forward and
   * backward of one step read the same scales; pty_fp8_roll moves the
   * recorded amax into the scales between steps. */
  float* fp8_state;
} PtyMlp;

/* Coordinate source of the INR evaluation: either explicit coords [n, 2]
 * (column 0 = first grid input) or, when coords == NULL, the regular grid of
 * ComplexINRModel2D.init_scale_grids (models/complex_inr.py:179-207): point
 * n = i*n_axis0 + j has input (axis0[j], axis1[i]), n = n_axis1 * n_axis0. */
typedef struct PtyPoints {
  const float* coords;
  const float* axis0;
  const float* axis1;
  int32_t n_axis0;
  int32_t n_axis1;
  int64_t n;
  /* Regular grid only: device plan built once by pty_points_plan_build for
   * this (grid, axes) pair (cell ranges, fractions, hash buckets); NULL makes
   * pty_inr_bwd build it in its workspace on every call. */
  const void* plan;
  /* Optional feature cache (no counterpart in the reference): a device
   * buffer of pty_points_feature_floats(n) floats.  pty_inr_fwd stores the
   * encoded features of every point there (in the kernels' tile layout) and
   * pty_inr_bwd / pty_inr_bwd_mlp read them instead of gathering the table
   * again; the caller must not change the table in between.  NULL: no cache. */
  float* features;
} PtyPoints;

/* Far-field batch descriptor (Recovery.py:287-307). */
typedef struct PtyFar {
  int32_t M, N;              /* pattern / probe size: 8..4096, prime factors <= 64 (any
                              * shape the LDS holds; powers of two take the fastest path) */
  int32_t canvas_h, canvas_w;
  int32_t n_batch;           /* positions in this batch */
  int32_t n_frames;          /* patterns in the measured stack */
  int32_t loss_type;         /* PTY_LOSS_* */
  int32_t pad_;
} PtyFar;

int pty_abi_version(void);
const char* pty_last_error(void);

/* ---- hash grid (replaces tcnn.Encoding fwd/bwd, complex_inr.py:138-145) ---- */
/* out [n, n_levels*2] f32 features. */
int pty_hashgrid_fwd(const PtyGrid* g, const float* coords, int64_t n, const float* table,
                     float* out, hipStream_t stream);
/* dtable[entry*2+f] += sum over points/corners of w_c * dout (atomics). */
int pty_hashgrid_bwd(const PtyGrid* g, const float* coords, int64_t n, const float* dout,
                     float* dtable, hipStream_t stream);
/* Corner entry indices [n, n_levels, 4] (u32, level offset included): the
 * integer part of the encoding, exposed for bit-exact checks. */
int pty_hashgrid_indices(const PtyGrid* g, const float* coords, int64_t n, uint32_t* idx,
                         hipStream_t stream);

/* ---- fused hash encode + complex MLP (complex_inr.py:279-282, 294-295) ---- */
/* out [n] c64 = ComplexMLP.net(HashGrid(points)). */
int pty_inr_fwd(const PtyGrid* g, const PtyMlp* mlp, const float* table, const PtyPoints* pts,
                float* out, hipStream_t stream);
size_t pty_inr_bwd_workspace(const PtyGrid* g, const PtyMlp* mlp, const PtyPoints* pts);
/* Given dout [n] c64, recomputes the activations and accumulates the
 * gradients of every MLP parameter into *grad and of the table into dtable. */
int pty_inr_bwd(const PtyGrid* g, const PtyMlp* mlp, const PtyMlp* grad, const float* table,
                float* dtable, const PtyPoints* pts, const float* dout, void* work,
                size_t work_bytes, hipStream_t stream);
/* The two stages of pty_inr_bwd, for callers that time or schedule them
 * apart (no separate counterpart in the reference; together they are the
 * same backward).  _mlp runs the MLP backward with recompute and leaves the
 * feature gradient and per-block MLP-gradient partials in `work`; _finish
 * reduces the partials into *grad and forms the table gradient.  Both take
 * the same workspace, and _finish must follow _mlp on the same stream. */
int pty_inr_bwd_mlp(const PtyGrid* g, const PtyMlp* mlp, const float* table, const PtyPoints* pts,
                    const float* dout, void* work, size_t work_bytes, hipStream_t stream);
int pty_inr_bwd_finish(const PtyGrid* g, const PtyMlp* mlp, const PtyMlp* grad, float* dtable,
                       const PtyPoints* pts, void* work, size_t work_bytes, hipStream_t stream);
/* pty_inr_bwd_finish with the table's Adam step (torch.optim.Adam.step,
 * Recovery.py:321, over the hash-table parameter) applied in place: each
 * table entry's gradient goes straight into its update of table / exp_avg /
 * exp_avg_sq ([n_entries, 2] f32 each, 8-byte aligned), with hyper =
 * [lr, step] as for pty_adam (advance step with pty_adam_tick first), so the
 * table gradient is never stored.  Same bits as pty_inr_bwd_finish followed by
 * pty_adam over the table.  Needs axis-generated points (a plan) with n > 0.
 * The MLP gradients still accumulate into *grad. */
int pty_inr_bwd_finish_adam(const PtyGrid* g, const PtyMlp* mlp, const PtyMlp* grad, float* table,
                            float* exp_avg, float* exp_avg_sq, const float* hyper, float beta1, float beta2,
                            float eps, const PtyPoints* pts, void* work, size_t work_bytes, hipStream_t stream);
/* Floats of the feature cache of n points (32 per point, n rounded up to a
 * multiple of 32). */
size_t pty_points_feature_floats(int64_t n);
/* Static plan of a regular grid for the table gradient (no counterpart in
 * the reference: it only depends on the hash-grid layout and the axes). */
size_t pty_points_plan_bytes(const PtyGrid* g, const PtyPoints* pts);
int pty_points_plan_build(const PtyGrid* g, const PtyPoints* pts, void* plan, size_t plan_bytes,
                          hipStream_t stream);
/* Delayed-scaling update of an FP8 MLP's state (no counterpart in the
 * reference, which has no FP8 path): scales = 224 / recorded amax, amax
 * cleared.  Call once per optimizer step, after the backward. */
int pty_fp8_roll(float* fp8_state, hipStream_t stream);
/* ComplexMLP.net over given features [n, n_in] (complex_inr.py:313). */
int pty_mlp_fwd(const PtyMlp* mlp, const float* feats, int64_t n, float* out, hipStream_t stream);

/* ---- single layers of ComplexMLP.net (complex_inr.py:38-40, 73-78, 308-314) ----
 * The fused kernels above evaluate the whole network; these run one layer,
 * for callers that use the modules layer by layer.
 *   activation 1: ComplexGaborLayer2D.forward, out = exp(1j*omega*(W x + b))
 *   activation 0: nn.Linear, out = W x + b
 *   complex_io 0: real x [n, n_in], weight [n_out, n_in], bias [n_out] (the
 *   first Gabor layer); 1: all complex (interleaved).  The output is complex
 *   [n, n_out] unless both flags are 0 (real). */
typedef struct PtyLayer {
  int32_t n_in, n_out;       /* 1 .. 4096 */
  int32_t complex_io;
  int32_t activation;
  const float* weight;
  const float* bias;         /* may be NULL (bias=False) */
  const float* omega;        /* [1]; activation only */
} PtyLayer;
int pty_layer_fwd(const PtyLayer* d, const float* x, int64_t n, float* out, hipStream_t stream);
size_t pty_layer_bwd_workspace(const PtyLayer* d, int64_t n);
/* Given gout (the output's gradient, PyTorch's complex convention), writes
 * gx (d x, may be NULL) and overwrites gweight / gbias / gomega (each may be
 * NULL).  Deterministic (fixed-order sums over point ranges). */
int pty_layer_bwd(const PtyLayer* d, const float* x, int64_t n, const float* gout, float* gx, float* gweight,
                  float* gbias, float* gomega, void* work, size_t work_bytes, hipStream_t stream);

/* ---- Fourier upsample (complex_inr.py:252-271) and its adjoint ---- */
size_t pty_upsample_workspace(int32_t n, int32_t ds);
/* big [ds*n, ds*n] c64 = ds^2 ifft2(ifftshift(pad(fftshift(fft2(small))))). */
int pty_upsample_fwd(const float* small, int32_t n, int32_t ds, float* big, void* work,
                     size_t work_bytes, hipStream_t stream);
/* gsmall [n, n] c64 = ifft2_n(crop(fft2_{ds n}(gbig))) (exact adjoint). */
int pty_upsample_adj(const float* gbig, int32_t n, int32_t ds, float* gsmall, void* work,
                     size_t work_bytes, hipStream_t stream);

/* ---- residual out = k * x + initial (complex_inr.py:316-334) ---- */
int pty_residual_fwd(const float* k, const float* x, const float* initial, int64_t n, float* out,
                     hipStream_t stream);
size_t pty_residual_workspace(int64_t n);
/* gx = k * gout; gk += Re sum(gout * conj(x)). */
int pty_residual_bwd(const float* k, const float* x, const float* gout, int64_t n, float* gx,
                     float* gk, void* work, size_t work_bytes, hipStream_t stream);

/* ---- probe COM roll + energy normalisation (Recovery.py:262-285) ---- */
size_t pty_probe_workspace(int32_t M, int32_t N);
/* stats (device, 8 floats): [cp0, cp1, scale, energy, applied_roll, applied_scale, 0, 0]. */
int pty_probe_normalize_fwd(const float* probe, int32_t M, int32_t N, float* probe_n, float* stats,
                            void* work, size_t work_bytes, hipStream_t stream);
int pty_probe_normalize_bwd(const float* probe, const float* gprobe_n, int32_t M, int32_t N,
                            const float* stats, float* gprobe, void* work, size_t work_bytes,
                            hipStream_t stream);

/* ---- far field, losses and their adjoint (Recovery.py:287-307) ---- */
/* Reorders the measured stack from the reference layout [M, N, frames] into
 * pattern-major [frames, M, N] with ifftshift applied (natural FFT order). */
int pty_stack_prepare(const float* stack_mnf, int32_t M, int32_t N, int32_t frames, float* out,
                      hipStream_t stream);
size_t pty_farfield_workspace(const PtyFar* d);
/* loss[b] = per-pattern loss (mean over pixels) of position batch_idx[b];
 * dobject (c64 canvas) = d(sum_b loss[b])/d object and dprobe (c64 [M,N]) =
 * d(sum_b loss[b])/d probe_n, both overwritten (the whole canvas is written,
 * zero outside the windows).  tl is [frames, 2] = (tlY, tlX).  The result is
 * deterministic (fixed-order sums, no atomics). */
int pty_farfield_fwd_bwd(const PtyFar* d, const float* object, const float* probe_n,
                         const float* stack, const int32_t* batch_idx, const int32_t* tl,
                         float* loss, float* dobject, float* dprobe, void* work,
                         size_t work_bytes, hipStream_t stream);
/* pty_farfield_fwd_bwd without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * workspace is a static device array of PTY_FAR_SCRATCH_BYTES per GPU (held
 * until the process ends); a batch whose pty_farfield_workspace exceeds it is
 * rejected with PTY_EWORK (16 MiB: e.g. batch 21 at 256 x 256, batch 85 at
 * 128 x 128, any loss).  The library allocates nothing and the kernels need no
 * cleared memory.  Every call on one GPU shares that GPU's scratch: calls
 * that overlap on one GPU (different streams or threads) race.  Same kernels
 * and same bits as pty_farfield_fwd_bwd. */
#define PTY_FAR_SCRATCH_BYTES (16u << 20)
int pty_farfield_fwd_bwd_noalloc(const PtyFar* d, const float* object, const float* probe_n,
                                 const float* stack, const int32_t* batch_idx, const int32_t* tl,
                                 float* loss, float* dobject, float* dprobe, hipStream_t stream);
/* Intensities [n_batch, M, N] = |fftshift(fft2(probe_n * patch))|^2 (forward only). */
int pty_farfield_intensity(const PtyFar* d, const float* object, const float* probe_n,
                           const int32_t* batch_idx, const int32_t* tl, float* intensity,
                           void* work, size_t work_bytes, hipStream_t stream);

/* ---- Haar wavelet sparsity (utils/loss.py:60-83; Recovery.py:309-313) ---- */
/* object: `rows` x W c64, rows and W even (a stack of images with even
 * heights may be passed as one image: no 2x2 block crosses an image).
 * loss[0] = scale * sum over the 2x2 blocks of the real and imaginary parts of
 * |LH| + |HL| + |HH| (scale 1: reduction='sum'; 1 / (rows/2 * W/2):
 * reduction='mean'), overwritten.  When dobject != NULL,
 * dobject += grad_weight * d loss[0] / d object (PyTorch's complex gradient
 * convention, sign(0) = 0), which is how Recovery.py:310 adds
 * sparsity_weight * loss to the step's objective.  Deterministic. */
size_t pty_haar_workspace(int64_t rows, int32_t W);
int pty_haar_sparsity(const float* object, int64_t rows, int32_t W, float scale, float grad_weight,
                      float* loss, float* dobject, void* work, size_t work_bytes, hipStream_t stream);

/* ---- Adam (torch.optim.Adam, amsgrad=False, weight_decay=0; Recovery.py:238, 321) ---- */
/* hyper (device, 2 floats): [lr, step]; bias corrections 1 - beta^step are
 * formed on the device.  pty_adam_tick adds 1 to step (call once per step
 * before the pty_adam calls of that step).  clear_grad != 0 zeroes grad after
 * use (saves the separate clear of the next step). */
int pty_adam(float* param, float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
             const float* hyper, float beta1, float beta2, float eps, int32_t clear_grad,
             hipStream_t stream);
int pty_adam_tick(float* hyper, hipStream_t stream);

/* ---- reconstruction loop (Recovery.py:243-248, 300-346) ---- */
/* The loop's batch slicing batch_idx = perm[batch_start : batch_start + batch]
 * and its per-batch loss bookkeeping, on the device, so that a captured step
 * needs no host work between replays.  order: the epoch's permutation
 * (n_order int32); cursor (1 int32): the batch number, 0 at the epoch start.
 * With c = *cursor and start = c * batch (start = 0 and c = 0 once start + batch
 * would pass n_order): idx[k] = order[start + k] for k < batch; when hist is
 * non-null, c >= 1 and c * batch <= n_hist, the previous batch's per-pattern
 * losses loss[0..batch) are copied to hist[(c - 1) * batch + k]; then
 * *cursor = c + 1, and with hyper non-null Adam's step count advances as in
 * pty_adam_tick (one launch instead of two).  One workgroup; batch > 0 and
 * n_order >= batch. */
int pty_batch_next(const int32_t* order, int32_t n_order, int32_t* cursor, int32_t batch, int32_t* idx,
                   float* hyper, const float* loss, float* hist, int64_t n_hist, hipStream_t stream);

/* ---- learned warm start: conv encoder-decoder (no counterpart in the reference) ----
 * Maps each diffraction pattern (log(1 + I) pooled to 64 x 64 plus the
 * window's offset, NHWC [B, 64, 64, 3]) to a 32 x 32 complex object patch
 * (NCHW [B, 2, 32, 32]); the patches seed ComplexINRModel2D's object_initial
 * (models/complex_inr.py:209-250).  Base width `base` in {16, 32, 48, 64}.
 * Parameters: one f32 vector of pty_wsnet_param_count(base) floats in the
 * order of the torch module's parameters, conv weights [Co][kh][kw][Ci],
 * transposed-conv weights [kh][kw][Co][Ci].  GEMMs in BF16 on MFMA with FP32
 * accumulation; deterministic.  Batches of 1 .. 1024. */
size_t pty_wsnet_param_count(int32_t base);
size_t pty_wsnet_workspace(int32_t base, int32_t batch);
/* x [P, 64, 64, 3] from fftshifted patterns intens [P, M, M] (M a multiple of
 * 64) and windows tl [P, 2] on a square canvas. */
int pty_wsnet_inputs(const float* intens, int32_t P, int32_t M, const int32_t* tl, int32_t canvas, float* x,
                     hipStream_t stream);
/* out = net(x); keeps the activations in `work` for pty_wsnet_backward. */
int pty_wsnet_forward(int32_t base, const float* params, const float* x, int32_t batch, float* out, void* work,
                      size_t work_bytes, hipStream_t stream);
/* loss[0] = mean (out - target)^2 over n values; dout (may be NULL) =
 * d loss / d out.  work: at least 4096 bytes. */
int pty_wsnet_mse(const float* out, const float* target, int64_t n, float* loss, float* dout, void* work,
                  size_t work_bytes, hipStream_t stream);
/* grads (overwritten, same layout as params) = d L / d params given dout =
 * d L / d out, using the activations of the last pty_wsnet_forward on the
 * same workspace, batch and parameters. */
int pty_wsnet_backward(int32_t base, const float* params, int32_t batch, const float* dout, float* grads, void* work,
                       size_t work_bytes, hipStream_t stream);

/* ---- 2-D Gaussian-field rasterizer (gscp/csrc: model_type gaussian_field) ----
 * Replaces the reference's CUDA rasterizer behind
 * gscp/cuda/_wrapper.py:18-166 (GaussianRasterize2D[Cholesky]) and its
 * bindings gscp/csrc/rasterize_gaussians.cpp:26-254, ext.cpp:69-74:
 * preprocess (gscp_rasterizer/forward.cu:18-170), tile binning
 * (rasterizer_impl.cu:45-106, 149-249), tile render (forward.cu:176-258),
 * render backward (backward.cu:21-133) and preprocess backward
 * (backward.cu:138-266).  16 x 16 pixel tiles, two channels (real, imag),
 * legacy circular AABB (USE_SNUGBOX 0).
 * Inputs: xy [n, 2] pixel positions (clamped to the canvas), weights [n, 2];
 * param PTY_GS_RS: p0 = scaling [n, 2] (log sigma), p1 = rotation [n];
 * param PTY_GS_CHOLESKY: p0 = log_L_diag [n, 2], p1 = L_offdiag [n].
 * out [2, height, width] f32.  Within a tile the Gaussians are summed in
 * ascending index (the reference's stable sort); the backward sums each
 * Gaussian's per-tile partials in a fixed order (no atomics): every result is
 * deterministic. */
enum { PTY_GS_RS = 0, PTY_GS_CHOLESKY = 1 };
typedef struct PtyGs {
  int32_t n;                /* Gaussians, 0 .. 2^24 */
  int32_t width, height;    /* canvas W, H: 1 .. 16384 */
  int32_t max_patch_radius; /* upper clamp of the activated scales (pixels) */
  float min_scale;          /* lower clamp of the activated scales (pixels) */
  int32_t param;            /* PTY_GS_RS or PTY_GS_CHOLESKY */
} PtyGs;
/* Per-Gaussian and per-tile state (the reference's geomBuffer + imgBuffer). */
size_t pty_gs_geom_bytes(const PtyGs* d);
/* radii [n] int32 (0: culled) and the geometry state; counts the
 * (Gaussian, tile) pairs per tile (forward.cu:18-170, rasterizer_impl.cu:186-192). */
int pty_gs_preprocess(const PtyGs* d, const float* xy, const float* p0, const float* p1, int32_t* radii,
                      void* geom, hipStream_t stream);
/* The number of (Gaussian, tile) pairs: copies it to the host and waits for
 * the stream (the one synchronising call, as rasterizer_impl.cu:194-197). */
int pty_gs_num_rendered(const PtyGs* d, const void* geom, int64_t* n_rendered, hipStream_t stream);
/* Per-pair state (the reference's binningBuffer): tile lists, and with
 * with_grad != 0 the backward's per-pair partial sums. */
size_t pty_gs_bin_bytes(const PtyGs* d, int64_t n_rendered, int32_t with_grad);
/* Builds the tile lists in `bin` (per-tile cursors in `geom`) and writes
 * every pixel of out. */
int pty_gs_render_fwd(const PtyGs* d, void* geom, int64_t n_rendered, const float* weights, void* bin,
                      float* out, hipStream_t stream);
/* dxy [n,2], dp0 [n,2], dp1 [n], dweight [n,2] (overwritten; zero where
 * radii <= 0) from dout [2, height, width], using the tile lists of the last
 * pty_gs_render_fwd on the same geom and bin (bin sized with with_grad). */
int pty_gs_render_bwd(const PtyGs* d, const float* xy, const float* p0, const float* p1, const float* weights,
                      const int32_t* radii, const void* geom, int64_t n_rendered, void* bin, const float* dout,
                      float* dxy, float* dp0, float* dp1, float* dweight, hipStream_t stream);

/* fill kernel (graph-safe replacement for hipMemsetAsync) */
int pty_fill(float* p, int64_t n, float value, hipStream_t stream);

/* dst [n] half = round-to-nearest-even(src [n] f32): the FP16 table copy the
 * forward kernels read with PTY_TABLE_F16 (tcnn's forward casts its f32
 * parameters to half the same way, complex_inr.py:138-145 uses its default
 * half precision). */
int pty_cast_f16(const float* src, int64_t n, void* dst, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* PTYCHO_HIP_H */
