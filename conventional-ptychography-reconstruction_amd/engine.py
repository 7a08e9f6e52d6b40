"""Fused reconstruction step: one optimizer step of Recovery.py:247-321 as a
fixed sequence of HIP launches, with no autograd graph and no host
synchronisation (every pointer is fixed at construction, so the step can be
replayed or captured).

The object and probe chains are independent between the start of the step
and the far field, and again between the far field and Adam: the probe chain
runs on a second stream (default priority) forked from and joined back into
the caller's stream, so its launches fill the gaps the object chain leaves
(small kernels, launch boundaries).  Each chain has its own buffers and
workspaces, so the result does not depend on how the two interleave.

Per step (batch of positions ``batch_idx`` on the device):
  object  = Upsample(INR_obj(grid))            [+ k_o * . + initial_o]
  probe   = INR_prb(grid)                      [+ k_p * . + initial_p]
  probe_n = COM-roll + energy-normalise(probe)
  loss_b, dobject, dprobe_n = far field + loss + adjoint over the batch
  dobject += sparsity_weight * d Haar(object)   (only when the weight is set)
  dprobe  = normalise^T(dprobe_n);  residual^T;  dsmall = Upsample^T(dobject)
  grads   = INR^T (recomputed activations; the table gradient gathered per
            entry from a static plan of the grid, no atomics)
  Adam over all trainable parameters: the hash tables inside their
  table-gradient kernels (the table gradient is never stored), the rest in
  one flat launch
Parameters of the model are rebound as views into one flat fp32 buffer so
that Adam is a single elementwise kernel with torch.optim.Adam semantics.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from .complex_inr import ComplexINRModel2D, mlp_descriptor
from .farfield import check_windows, far_descriptor, positions_tensor, prepare_stack


def _real_numel(p: torch.Tensor) -> int:
    return p.numel() * (2 if p.is_complex() else 1)


def _padded(n: int) -> int:
    return (n + 3) // 4 * 4


class FlatParams:
    """All kernel-visible parameters of the model packed into one fp32 buffer.

    Trainable parameters come first (Adam runs over that prefix); frozen
    ones (e.g. omega_0 with trainable_omega0=False) follow.  Each
    nn.Parameter is rebound to a view of the buffer; a grad view of the same
    shape is kept for the kernels to accumulate into.
    """

    def __init__(self, params: list[torch.nn.Parameter], device):
        train = [p for p in params if p.requires_grad]
        frozen = [p for p in params if not p.requires_grad]
        order = train + frozen
        # every parameter starts on a 16-byte boundary (complex views need an
        # even offset; padding entries keep zero gradients)
        total = sum(_padded(_real_numel(p)) for p in order)
        self.n_train = sum(_padded(_real_numel(p)) for p in train)
        self.data = torch.zeros(total, dtype=torch.float32, device=device)
        self.grad = torch.zeros(total, dtype=torch.float32, device=device)
        self.grad_views = {}
        self.offsets = {}
        off = 0
        for p in order:
            n = _real_numel(p)
            dv = self.data[off:off + n]
            gv = self.grad[off:off + n]
            if p.is_complex():
                dv.copy_(torch.view_as_real(p.detach().to(device)).reshape(-1))
                pview = torch.view_as_complex(dv.view(*p.shape, 2))
                gview = torch.view_as_complex(gv.view(*p.shape, 2))
            else:
                dv.copy_(p.detach().to(device).reshape(-1))
                pview = dv.view(p.shape)
                gview = gv.view(p.shape)
            p.data = pview
            self.grad_views[id(p)] = gview
            self.offsets[id(p)] = off
            off += _padded(n)

    def g(self, p):
        return self.grad_views[id(p)]

    def span(self, p):
        """(offset, floats) of p in the flat buffer."""
        return self.offsets[id(p)], _real_numel(p)


def _check_initial(name, t, shape):
    """A residual's initial field as the kernels read it: complex, of the
    field's shape; returned as a private contiguous complex64 copy."""
    if not torch.is_complex(t):
        raise ValueError(f"{name} must be a complex tensor")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, the field is {tuple(shape)}")
    return t.detach().to(torch.complex64).contiguous().clone()


def check_inputs(model, stack_mnf, tlY, tlX, batch_size):
    """Host-side shape checks of the fused step's inputs (the kernels index
    the stack, the scan windows and the initial fields without bounds
    checks).  Raises ValueError; returns (frames, tlY, tlX) as int64 arrays."""
    M, N = model.probe_width, model.probe_height
    if stack_mnf.dim() != 3 or tuple(stack_mnf.shape[:2]) != (M, N):
        raise ValueError(f"measured stack has shape {tuple(stack_mnf.shape)}; expected [{M}, {N}, frames] "
                         f"(the reference layout, Recovery.py:103)")
    frames = int(stack_mnf.shape[2])
    tlY = np.asarray(tlY).reshape(-1)
    tlX = np.asarray(tlX).reshape(-1)
    if frames < 1 or tlY.size != frames or tlX.size != frames:
        raise ValueError(f"{tlY.size} / {tlX.size} scan windows for {frames} measured patterns")
    if int(batch_size) < 1:
        raise ValueError("batch_size must be positive")
    check_windows(tlY, tlX, M, N, model.output_width, model.output_height)
    for which in ("object", "probe"):
        if not model.coords_are_grid(which):
            raise ValueError(f"the fused step evaluates the INR on the regular {which} grid of init_scale_grids; "
                             f"the model's {which} coordinates were changed (use model.forward() for other "
                             f"coordinates)")
    if model.has_object_residual:
        _check_initial("object_initial", model.object_initial, (model.output_width, model.output_height))
    if model.has_probe_residual:
        _check_initial("probe_initial", model.probe_initial, (M, N))
    return frames, tlY.astype(np.int64), tlX.astype(np.int64)


class FusedReconstruction:
    """Fused training step for a ComplexINRModel2D on a measured stack.

    The residual initial fields (``object_initial`` / ``probe_initial``) are
    buffers of the model that the step reads as complex64: they are copied at
    construction (a complex128 initial is rounded once), so later changes to
    the model's buffers do not reach a constructed step.
    """

    def __init__(self, model: ComplexINRModel2D, stack_mnf: torch.Tensor, tlY, tlX, loss_type="smooth_L1_loss",
                 lr=1e-3, batch_size=10, lr_decay_step=10000, lr_decay_gamma=0.1, betas=(0.9, 0.999),
                 eps=1e-8, sparsity_weight=0.0, fuse_table_adam=True):
        frames, tlY, tlX = check_inputs(model, stack_mnf, tlY, tlX, batch_size)
        dev = model.xy_hash_encoder.params.device
        if dev.type != "cuda":
            raise RuntimeError("FusedReconstruction needs the model on the GPU")
        if not model.update_probe:
            raise RuntimeError("update_probe must be True for conventional ptychography")
        self.model, self.device, self.loss_type = model, dev, loss_type
        self.batch_size = int(batch_size)
        self.lr0, self.gamma, self.decay_step = float(lr), float(lr_decay_gamma), int(lr_decay_step)
        self.beta1, self.beta2, self.eps = float(betas[0]), float(betas[1]), float(eps)
        self.M, self.N = model.probe_width, model.probe_height
        self.H, self.W = model.output_width, model.output_height
        self.n_small = model.downsampled_width
        self.ds = int(model.downsample_factor)
        if model.downsampled_width != model.downsampled_height or self.n_small * self.ds != self.H:
            raise ValueError("the fused step needs a square canvas equal to downsample_factor * INR grid")
        if self.M != self.N:
            raise ValueError("the fused step needs a square probe")
        check_windows(tlY, tlX, self.M, self.N, self.H, self.W)
        # Recovery.py:309-313: sparsity_weight * Haar sparsity of the object
        # joins the objective; at weight 0 the term is skipped (the reference
        # evaluates it and multiplies by 0)
        self.sparsity_weight = float(sparsity_weight)
        if self.sparsity_weight and (self.H % 2 or self.W % 2):
            raise ValueError("the Haar sparsity term needs an even canvas")
        self.frames = frames
        self.stack = prepare_stack(stack_mnf.to(dev))
        self.obj_init = (_check_initial("object_initial", model.object_initial, (self.H, self.W)).to(dev)
                         if model.has_object_residual else None)
        self.prb_init = (_check_initial("probe_initial", model.probe_initial, (self.M, self.N)).to(dev)
                         if model.has_probe_residual else None)
        self.tl = positions_tensor(tlY, tlX, dev)

        for mlp in (model.object_complex_mlp, model.probe_complex_mlp):
            mlp.check_supported()
        self.obj_ts = model.object_complex_mlp.tensors()
        self.prb_ts = model.probe_complex_mlp.tensors()
        params = [model.xy_hash_encoder.params, model.probe_hash_encoder.params, *self.obj_ts, *self.prb_ts]
        if model.has_object_residual:
            params.append(model.k_object)
        if model.has_probe_residual:
            params.append(model.k_probe)
        self.flat = FlatParams(params, dev)
        self.m = torch.zeros(self.flat.n_train, dtype=torch.float32, device=dev)
        self.v = torch.zeros(self.flat.n_train, dtype=torch.float32, device=dev)
        self.hyper = torch.tensor([self.lr0, 0.0], dtype=torch.float32, device=dev)
        # The hash tables (almost every parameter) take their Adam step inside
        # the table-gradient kernels, which then never store the table
        # gradient; the flat Adam launch covers the rest.  Needs both tables
        # trainable, first in the flat buffer.
        tabs = [model.xy_hash_encoder.params, model.probe_hash_encoder.params]
        self.fuse_table_adam = bool(fuse_table_adam) and all(t.requires_grad for t in tabs)
        self._adam_from = 0
        if self.fuse_table_adam:
            (o0, n0), (o1, n1) = self.flat.span(tabs[0]), self.flat.span(tabs[1])
            assert o0 == 0 and o1 == _padded(n0)
            self._adam_from = o1 + _padded(n1)
            self._tab_adam = {
                "obj": (tabs[0].data_ptr(), self.m[o0:o0 + n0].data_ptr(), self.v[o0:o0 + n0].data_ptr()),
                "prb": (tabs[1].data_ptr(), self.m[o1:o1 + n1].data_ptr(), self.v[o1:o1 + n1].data_ptr()),
            }
        self._lr_epoch = 0

        prec = model.mlp_precision
        nin = model.num_features
        self.obj_spec = model.xy_hash_encoder.spec
        self.prb_spec = model.probe_hash_encoder.spec
        enc_o, enc_p = model.xy_hash_encoder, model.probe_hash_encoder
        self.obj_grid = enc_o.descriptor()
        self.prb_grid = enc_p.descriptor()
        # tables the kernels read: the f32 master itself, or an FP16 copy
        # refreshed at the start of every forward
        self.obj_tab = (enc_o.params if enc_o.table_precision == "fp32" else
                        torch.empty(enc_o.params.numel(), dtype=torch.float16, device=dev))
        self.prb_tab = (enc_p.params if enc_p.table_precision == "fp32" else
                        torch.empty(enc_p.params.numel(), dtype=torch.float16, device=dev))
        self.obj_pts = model.grid_points("object")
        self.prb_pts = model.grid_points("probe")
        self.obj_pdesc = self.obj_pts.descriptor(self.obj_spec)
        self.prb_pdesc = self.prb_pts.descriptor(self.prb_spec)
        # feature caches: the forward of a step stores the encoded features,
        # its backward reads them instead of gathering the tables again (the
        # tables only change in Adam, after the backward)
        self.obj_feat = torch.empty(_lib.query("pty_points_feature_floats", self.obj_pts.n), dtype=torch.float32,
                                    device=dev)
        self.prb_feat = torch.empty(_lib.query("pty_points_feature_floats", self.prb_pts.n), dtype=torch.float32,
                                    device=dev)
        self.obj_pdesc.features = self.obj_feat.data_ptr()
        self.prb_pdesc.features = self.prb_feat.data_ptr()
        # FP8: per-MLP delayed-scaling state, rolled after every backward
        self.fp8_states = []
        st_o = st_p = None
        if prec == "fp8":
            st_o = model.object_complex_mlp.fp8_state(dev)
            st_p = model.probe_complex_mlp.fp8_state(dev)
            self.fp8_states = [st_o, st_p]
        self.obj_mlp = mlp_descriptor(self.obj_ts, nin, prec, st_o)
        self.prb_mlp = mlp_descriptor(self.prb_ts, nin, prec, st_p)
        self.obj_gmlp = mlp_descriptor([self.flat.g(t) for t in self.obj_ts], nin, prec, st_o)
        self.prb_gmlp = mlp_descriptor([self.flat.g(t) for t in self.prb_ts], nin, prec, st_p)
        self.obj_dtable = self.flat.g(model.xy_hash_encoder.params)
        self.prb_dtable = self.flat.g(model.probe_hash_encoder.params)

        c64 = torch.complex64
        n, M, N, H, W = self.n_small, self.M, self.N, self.H, self.W
        self.obj_small = torch.empty(n * n, dtype=c64, device=dev)
        self.obj_up = torch.empty(H, W, dtype=c64, device=dev)
        self.obj = torch.empty(H, W, dtype=c64, device=dev) if model.has_object_residual else self.obj_up
        self.prb_mlp_out = torch.empty(M * N, dtype=c64, device=dev)
        self.probe = torch.empty(M, N, dtype=c64, device=dev) if model.has_probe_residual else \
            self.prb_mlp_out.view(M, N)
        self.probe_n = torch.empty(M, N, dtype=c64, device=dev)
        self.stats = torch.zeros(8, dtype=torch.float32, device=dev)
        self.dobj = torch.empty(H, W, dtype=c64, device=dev)
        self.dobj_up = torch.empty(H, W, dtype=c64, device=dev) if model.has_object_residual else self.dobj
        self.dprobe_n = torch.empty(M, N, dtype=c64, device=dev)
        self.dprobe = torch.empty(M, N, dtype=c64, device=dev)
        self.dprb_mlp = torch.empty(M, N, dtype=c64, device=dev) if model.has_probe_residual else self.dprobe
        self.dsmall = torch.empty(n, n, dtype=c64, device=dev)
        self.loss = torch.zeros(self.batch_size, dtype=torch.float32, device=dev)
        # unweighted Haar sparsity of the object of the last step (0 when off)
        self.sparsity = torch.zeros(1, dtype=torch.float32, device=dev)

        q = _lib.query
        self.ws_up = _lib.workspace(q("pty_upsample_workspace", n, self.ds), dev)
        self.ws_probe = _lib.workspace(q("pty_probe_workspace", M, N), dev)
        self.ws_haar = _lib.workspace(q("pty_haar_workspace", H, W) if self.sparsity_weight else 1, dev)
        self.ws_res_o = _lib.workspace(q("pty_residual_workspace", H * W), dev)
        self.ws_res_p = _lib.workspace(q("pty_residual_workspace", M * N), dev)
        self.ws_obj = _lib.workspace(q("pty_inr_bwd_workspace", self.obj_grid, self.obj_mlp, self.obj_pdesc), dev)
        self.ws_prb = _lib.workspace(q("pty_inr_bwd_workspace", self.prb_grid, self.prb_mlp, self.prb_pdesc), dev)
        self.far = {}
        for b in {self.batch_size, self.frames % self.batch_size}:
            if b > 0:
                d = far_descriptor(M, N, H, W, b, self.frames, loss_type)
                self.far[b] = (d, _lib.workspace(q("pty_farfield_workspace", d), dev))
        self.epoch = 0
        self._timers = None
        # the probe chain's stream (default priority, like every stream here);
        # overlap = False runs both chains on the caller's stream
        self.side = torch.cuda.Stream(device=dev)
        self.overlap = True
        self.graph = None
        # epoch order on the device for graph replays: the captured step takes
        # its batch from order at cursor (pty_batch_next) and records the
        # previous batch's losses in loss_hist (one slot per position)
        self.order = torch.zeros(self.frames, dtype=torch.int32, device=dev)
        self.cursor = torch.zeros(1, dtype=torch.int32, device=dev)
        self.loss_hist = torch.zeros(self.frames, dtype=torch.float32, device=dev)
        self._from_order = False
        self._order_len = self._order_used = 0
        self._grad_clear = True   # FlatParams starts with zero gradients
        if self.fp8_states:
            # calibration: steps without the update record the amax of every
            # FP8 tensor, so the first real step has scales (their gradients
            # are cleared by the next step).  A layer's gradient amax is only
            # seen once the layer after it has a scale: with the initial
            # scale 1 a gradient below e4m3's smallest subnormal (2^-9), such
            # as the probe MLP's behind k_probe = 1e-8, converts to zero and
            # so does every gradient before it.  Each pass calibrates one more
            # layer of the backward; n_hidden + 2 passes cover all of them.
            calib = torch.arange(min(self.batch_size, self.frames), dtype=torch.int32, device=dev)
            n_pass = max(model.object_complex_mlp.n_hidden(), model.probe_complex_mlp.n_hidden()) + 2
            for _ in range(n_pass):
                self.step(calib, update=False)

    # ------------------------------------------------------------------ timing
    def enable_timing(self, names):
        """Record HIP events (on the launch stream) around the named launches
        ('obj_inr_fwd', 'obj_inr_bwd', 'prb_inr_bwd', 'farfield', 'obj_mlp_bwd' /
        'prb_mlp_bwd', which time the MLP-backward kernel alone, and
        'obj_finish_adam' / 'prb_finish_adam', the gradient finish with the
        tables' gradient and Adam step)."""
        self._timers = {n: [] for n in names}

    def _mark(self, name, begin, stream=None):
        if self._timers is None or name not in self._timers:
            return None
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(stream)
        if begin:
            self._timers[name].append([ev, None])
        else:
            self._timers[name][-1][1] = ev
        return ev

    def timings_ms(self):
        """{name: [ms per recorded launch]} (call after synchronising)."""
        return {n: [a.elapsed_time(b) for a, b in v] for n, v in (self._timers or {}).items()}

    # ------------------------------------------------------------------ step
    def set_lr(self, lr: float):
        _lib.call("pty_fill", self.hyper.data_ptr(), 1, float(lr), _lib.stream_ptr(self.device))

    def _refresh_tables(self, s):
        m = self.model
        for enc, tab in ((m.xy_hash_encoder, self.obj_tab), (m.probe_hash_encoder, self.prb_tab)):
            if tab.dtype == torch.float16:
                _lib.call("pty_cast_f16", enc.params.data_ptr(), enc.params.numel(), tab.data_ptr(), s)

    def _streams(self):
        main = torch.cuda.current_stream(self.device)
        return main, (self.side if self.overlap else main)

    def forward_fields(self):
        main, side = self._streams()
        C = _lib.call
        m = self.model
        self._refresh_tables(main.cuda_stream)
        if self.overlap:
            side.wait_stream(main)
        # object chain on the caller's stream
        s = main.cuda_stream
        self._mark("obj_inr_fwd", True, main)
        C("pty_inr_fwd", self.obj_grid, self.obj_mlp, self.obj_tab.data_ptr(), self.obj_pdesc,
          self.obj_small.data_ptr(), s)
        self._mark("obj_inr_fwd", False, main)
        # probe chain on the side stream (enqueued while the object encoder runs)
        p = side.cuda_stream
        C("pty_inr_fwd", self.prb_grid, self.prb_mlp, self.prb_tab.data_ptr(), self.prb_pdesc,
          self.prb_mlp_out.data_ptr(), p)
        if m.has_probe_residual:
            C("pty_residual_fwd", m.k_probe.data_ptr(), self.prb_mlp_out.data_ptr(), self.prb_init.data_ptr(),
              self.prb_mlp_out.numel(), self.probe.data_ptr(), p)
        C("pty_probe_normalize_fwd", self.probe.data_ptr(), self.M, self.N, self.probe_n.data_ptr(),
          self.stats.data_ptr(), self.ws_probe.data_ptr(), self.ws_probe.numel(), p)
        C("pty_upsample_fwd", self.obj_small.data_ptr(), self.n_small, self.ds, self.obj_up.data_ptr(),
          self.ws_up.data_ptr(), self.ws_up.numel(), s)
        if m.has_object_residual:
            C("pty_residual_fwd", m.k_object.data_ptr(), self.obj_up.data_ptr(), self.obj_init.data_ptr(),
              self.obj_up.numel(), self.obj.data_ptr(), s)
        if self.overlap:
            main.wait_stream(side)

    def _inr_bwd(self, which, grid, mlp, gmlp, table, dtable, pdesc, dout, ws, stream, adam=False):
        """INR backward on ``stream``; with timing on for '<which>_mlp_bwd' the
        MLP stage is bracketed by events on its own (the reduce and table
        gradient follow).  adam: the table takes its Adam step in place."""
        C = _lib.call
        s = stream.cuda_stream
        name = which + "_mlp_bwd"
        if adam:
            self._mark(name, True, stream)
            C("pty_inr_bwd_mlp", grid, mlp, table.data_ptr(), pdesc, dout.data_ptr(), ws.data_ptr(), ws.numel(), s)
            self._mark(name, False, stream)
            p, m, v = self._tab_adam[which]
            self._mark(which + "_finish_adam", True, stream)
            C("pty_inr_bwd_finish_adam", grid, mlp, gmlp, p, m, v, self.hyper.data_ptr(), self.beta1, self.beta2,
              self.eps, pdesc, ws.data_ptr(), ws.numel(), s)
            self._mark(which + "_finish_adam", False, stream)
            return
        if self._timers is None or name not in self._timers:
            self._mark(which + "_inr_bwd", True, stream)
            C("pty_inr_bwd", grid, mlp, gmlp, table.data_ptr(), dtable.data_ptr(), pdesc, dout.data_ptr(),
              ws.data_ptr(), ws.numel(), s)
            self._mark(which + "_inr_bwd", False, stream)
            return
        self._mark(name, True, stream)
        C("pty_inr_bwd_mlp", grid, mlp, table.data_ptr(), pdesc, dout.data_ptr(), ws.data_ptr(), ws.numel(), s)
        self._mark(name, False, stream)
        C("pty_inr_bwd_finish", grid, mlp, gmlp, dtable.data_ptr(), pdesc, ws.data_ptr(), ws.numel(), s)

    def step(self, batch_idx: torch.Tensor, update: bool = True):
        """One optimizer step on the positions ``batch_idx`` (int32, device)."""
        main, side = self._streams()
        s = main.cuda_stream
        C = _lib.call
        m = self.model
        b = batch_idx.numel()
        d, ws_far = self.far[b]
        fused = update and self.fuse_table_adam
        if not self._grad_clear:
            C("pty_fill", self.flat.grad.data_ptr(), self.flat.grad.numel(), 0.0, s)
        if self._from_order:
            # (capture only) the batch comes from the epoch order on the
            # device, and Adam's count advances in the same launch
            C("pty_batch_next", self.order.data_ptr(), self.order.numel(), self.cursor.data_ptr(), b,
              batch_idx.data_ptr(), self.hyper.data_ptr() if fused else None, self.loss.data_ptr(),
              self.loss_hist.data_ptr(), self.loss_hist.numel(), s)
        elif fused:
            # the table-gradient kernels apply Adam with this step's count
            C("pty_adam_tick", self.hyper.data_ptr(), s)
        self.forward_fields()
        self._mark("farfield", True, main)
        C("pty_farfield_fwd_bwd", d, self.obj.data_ptr(), self.probe_n.data_ptr(), self.stack.data_ptr(),
          batch_idx.data_ptr(), self.tl.data_ptr(), self.loss.data_ptr(), self.dobj.data_ptr(),
          self.dprobe_n.data_ptr(), ws_far.data_ptr(), ws_far.numel(), s)
        self._mark("farfield", False, main)
        if self.sparsity_weight:
            C("pty_haar_sparsity", self.obj.data_ptr(), self.H, self.W, 1.0, self.sparsity_weight,
              self.sparsity.data_ptr(), self.dobj.data_ptr(), self.ws_haar.data_ptr(), self.ws_haar.numel(), s)
        if self.overlap:
            side.wait_stream(main)
        # object adjoint of the upsample first (the object chain is the longer
        # one), then the probe chain on the side stream, then the object INR
        if m.has_object_residual:
            C("pty_residual_bwd", m.k_object.data_ptr(), self.obj_up.data_ptr(), self.dobj.data_ptr(),
              self.dobj.numel(), self.dobj_up.data_ptr(), self.flat.g(m.k_object).data_ptr(),
              self.ws_res_o.data_ptr(), self.ws_res_o.numel(), s)
        C("pty_upsample_adj", self.dobj_up.data_ptr(), self.n_small, self.ds, self.dsmall.data_ptr(),
          self.ws_up.data_ptr(), self.ws_up.numel(), s)
        p = side.cuda_stream
        C("pty_probe_normalize_bwd", self.probe.data_ptr(), self.dprobe_n.data_ptr(), self.M, self.N,
          self.stats.data_ptr(), self.dprobe.data_ptr(), self.ws_probe.data_ptr(), self.ws_probe.numel(), p)
        if m.has_probe_residual:
            C("pty_residual_bwd", m.k_probe.data_ptr(), self.prb_mlp_out.data_ptr(), self.dprobe.data_ptr(),
              self.dprobe.numel(), self.dprb_mlp.data_ptr(), self.flat.g(m.k_probe).data_ptr(),
              self.ws_res_p.data_ptr(), self.ws_res_p.numel(), p)
        self._inr_bwd("prb", self.prb_grid, self.prb_mlp, self.prb_gmlp, self.prb_tab,
                      self.prb_dtable, self.prb_pdesc, self.dprb_mlp, self.ws_prb, side, adam=fused)
        self._inr_bwd("obj", self.obj_grid, self.obj_mlp, self.obj_gmlp, self.obj_tab,
                      self.obj_dtable, self.obj_pdesc, self.dsmall, self.ws_obj, main, adam=fused)
        if self.fp8_states:
            st_o, st_p = self.fp8_states
            C("pty_fp8_roll", st_o.data_ptr(), s)
            C("pty_fp8_roll", st_p.data_ptr(), p)
        if self.overlap:
            main.wait_stream(side)
        if update:
            # Adam clears the trainable gradients it consumes; frozen ones are
            # cleared with them when they exist
            a0 = self._adam_from if fused else 0
            if not fused:
                C("pty_adam_tick", self.hyper.data_ptr(), s)
            C("pty_adam", self.flat.data.data_ptr() + 4 * a0, self.flat.grad.data_ptr() + 4 * a0,
              self.m.data_ptr() + 4 * a0, self.v.data_ptr() + 4 * a0, self.flat.n_train - a0,
              self.hyper.data_ptr(), self.beta1, self.beta2, self.eps, 1, s)
            n_frozen = self.flat.grad.numel() - self.flat.n_train
            if n_frozen:
                C("pty_fill", self.flat.grad.data_ptr() + 4 * self.flat.n_train, n_frozen, 0.0, s)
        self._grad_clear = update
        return self.loss[:b]

    # ------------------------------------------------------------------ graph
    def capture(self):
        """Capture one full training step (batch size ``batch_size``) in a HIP graph.

        Every launch of the step goes to the capture stream.  The captured step
        starts with ``pty_batch_next``: it reads its batch from the epoch order
        (``order``) at ``cursor`` and advances the cursor, so consecutive replays
        walk an epoch with no host work in between (``begin_order`` +
        ``graph_next``).  The graph holds no memset nodes (clears are fill
        kernels) and no host synchronisation.
        """
        self.batch_buf = torch.zeros(self.batch_size, dtype=torch.int32, device=self.device)
        # one eager pass without the update loads every kernel before capture;
        # it must not move the FP8 scaling state (its batch is position 0
        # repeated), so the state is put back afterwards
        saved = [st.clone() for st in self.fp8_states]
        self.step(self.batch_buf, update=False)
        for st, sv in zip(self.fp8_states, saved):
            st.copy_(sv)
        _lib.call("pty_fill", self.flat.grad.data_ptr(), self.flat.grad.numel(), 0.0, _lib.stream_ptr(self.device))
        self._grad_clear = True
        torch.cuda.synchronize(self.device)
        timers, self._timers = self._timers, None
        self.graph = torch.cuda.CUDAGraph()
        self._from_order = True
        try:
            with torch.cuda.graph(self.graph):
                self.step(self.batch_buf, update=True)
        finally:
            self._from_order = False
        self._timers = timers
        torch.cuda.synchronize(self.device)
        return self.graph

    def begin_order(self, order: torch.Tensor):
        """Load an epoch's position order (int32, device, length <= frames) for
        ``graph_next``; the next replay takes order[0:batch_size]."""
        n = order.numel()
        if n < self.batch_size or n > self.frames:
            raise ValueError(f"order of {n} positions for batch {self.batch_size} and {self.frames} frames")
        self.order[:n].copy_(order, non_blocking=True)
        self.cursor.zero_()
        # the captured step reads the order buffer of all ``frames`` entries;
        # replays stop at the end of what was loaded
        self._order_len = n
        self._order_used = 0

    def graph_next(self):
        """Replay the captured step on the next batch of the loaded order
        (``begin_order``); returns the loss buffer (valid until the next step)."""
        if self.graph is None:
            self.capture()
        if self._order_used + self.batch_size > self._order_len:
            raise RuntimeError(f"the loaded order ({self._order_len} positions) has no full batch left after "
                               f"{self._order_used}; load the next one with begin_order")
        self._order_used += self.batch_size
        if not self._grad_clear:
            # the captured step starts from cleared gradients; an eager step
            # without the update since the last one left some behind
            _lib.call("pty_fill", self.flat.grad.data_ptr(), self.flat.grad.numel(), 0.0,
                      _lib.stream_ptr(self.device))
            self._grad_clear = True
        self.graph.replay()
        return self.loss

    def graph_step(self, batch_idx: torch.Tensor):
        """Replay the captured step on ``batch_idx`` (int32, device, length batch_size)."""
        if batch_idx.numel() != self.batch_size:
            raise ValueError(f"graph_step takes full batches of {self.batch_size}")
        self.begin_order(batch_idx)
        return self.graph_next()

    def run_epoch(self, generator: torch.Generator | None = None, use_graph: bool = True):
        """One epoch: randperm over frames, ceil(frames/batch) steps (Recovery.py:243-350).

        Full batches replay the captured step (captured on first use), which
        takes each batch from the permutation on the device; the trailing
        partial batch, if any, runs eagerly.  Returns the sum of the epoch's
        per-pattern losses (a device scalar).  Both paths give the same bits
        (the graph holds the same launches, and both total the same per-position
        loss record)."""
        lr = self.lr0 * self.gamma ** (self.epoch // self.decay_step)
        if self.epoch == 0 or self.epoch % self.decay_step == 0:
            self.set_lr(lr)
        perm = torch.randperm(self.frames, device=self.device, generator=generator).to(torch.int32)
        B = self.batch_size
        n_full = self.frames // B
        if use_graph and n_full > 0:
            if self.graph is None:
                self.capture()
            self.begin_order(perm)
            for _ in range(n_full):
                self.graph_next()
            # each replay recorded the previous batch's losses; the last is left
            self.loss_hist[(n_full - 1) * B:n_full * B].copy_(self.loss[:B])
        else:
            for i in range(n_full):
                self.loss_hist[i * B:(i + 1) * B].copy_(self.step(perm[i * B:(i + 1) * B]))
        if n_full * B < self.frames:
            self.loss_hist[n_full * B:].copy_(self.step(perm[n_full * B:]))
        self.epoch += 1
        return self.loss_hist.sum()

    def current_fields(self):
        """(object [H,W], normalised probe [M,N]) of the current parameters."""
        self.forward_fields()
        return self.obj, self.probe_n
