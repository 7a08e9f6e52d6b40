"""Drop-in ``ComplexINRModel2D`` running on the gfx950 kernels.

Mirrors models/complex_inr.py of the reference:
  * ComplexGaborLayer2D (6-40) and ComplexMLP (42-82): same constructor,
    same parameters, same registration order and the same nn.Linear
    initialisation (so a seeded model starts from the reference's weights);
  * ComplexINRModel2D (85-334): same constructor signature, attributes
    (xy_coords, probe_coords, k_object, k_probe, has_*_residual, buffers
    object_initial / probe_initial), same ValueError for a non-complex
    initial, same ``forward() -> (object [W,H] c64, probe [Wp,Hp] c64)``.
The forward is one fused HIP launch per field (hash encode + MLP,
``pty_inr_fwd``), the batched-FFT Fourier upsample and the residual kernel;
the backward is the hand-written adjoint (``pty_inr_bwd``,
``pty_upsample_adj``, ``pty_residual_bwd``) behind torch.autograd.Function.
"""
from __future__ import annotations

import warnings

import torch
import torch.nn as nn

from . import _lib
from .hashgrid import HashGridEncoding, table_storage
from .levels import GridSpec

FP8_WARNING = ("mlp_precision='fp8' is experimental: single steps stay within their error bars, but "
               "reconstructions do not converge (e4m3 operands perturb the pre-activations that set "
               "every unit's magnitude); use 'bf16' or 'fp32' for real runs")


# ----------------------------------------------------------------------------
# modules with the reference's parameters
# ----------------------------------------------------------------------------
class LayerFn(torch.autograd.Function):
    """One layer of ComplexMLP.net on the device (``pty_layer_fwd/bwd``):
    exp(1j * omega * (W x + b)) with ``activation``, W x + b without."""

    @staticmethod
    def forward(ctx, x, weight, bias, omega, activation: bool):
        _lib.require_cuda(x, weight, bias, omega)
        cplx = weight.is_complex()
        n_out, n_in = weight.shape
        if x.is_complex() != cplx:
            raise RuntimeError(f"layer input is {x.dtype}, weight is {weight.dtype}")
        if x.shape[-1] != n_in:
            raise RuntimeError(f"layer input has {x.shape[-1]} features, the layer takes {n_in}")
        x2 = x.detach().resolve_conj().reshape(-1, n_in).to(torch.complex64 if cplx else torch.float32).contiguous()
        w = weight.detach().resolve_conj().contiguous()
        b = bias.detach().resolve_conj().contiguous() if bias is not None else None
        om = omega.detach().contiguous() if omega is not None else None
        n = x2.shape[0]
        out_dtype = torch.complex64 if (activation or cplx) else torch.float32
        out = torch.empty(n, n_out, dtype=out_dtype, device=x.device)
        d = _layer_descriptor(w, b, om, activation)
        _lib.call("pty_layer_fwd", d, x2.data_ptr(), n, out.data_ptr(), _lib.stream_ptr(x.device))
        ctx.save_for_backward(x2, w, b, om)
        ctx.activation, ctx.lead = activation, x.shape[:-1]
        return out.reshape(*x.shape[:-1], n_out)

    @staticmethod
    def backward(ctx, gout):
        x2, w, b, om = ctx.saved_tensors
        n, n_in = x2.shape
        n_out = w.shape[0]
        d = _layer_descriptor(w, b, om, ctx.activation)
        g = gout.resolve_conj().reshape(n, n_out).contiguous()
        need_x, need_w, need_b, need_om, _ = ctx.needs_input_grad
        gx = torch.empty_like(x2) if need_x else None
        gw = torch.empty_like(w) if need_w else None
        gb = torch.empty_like(b) if (need_b and b is not None) else None
        gom = torch.empty_like(om) if (need_om and om is not None) else None
        ws = _lib.workspace(_lib.query("pty_layer_bwd_workspace", d, n), x2.device)
        _lib.call("pty_layer_bwd", d, x2.data_ptr(), n, g.data_ptr(), _lib.ptr(gx), _lib.ptr(gw), _lib.ptr(gb),
                  _lib.ptr(gom), ws.data_ptr(), ws.numel(), _lib.stream_ptr(x2.device))
        if gx is not None:
            gx = gx.reshape(*ctx.lead, n_in)
        return gx, gw, gb, gom, None


def _layer_descriptor(w, b, om, activation) -> _lib.PtyLayer:
    d = _lib.PtyLayer()
    d.n_out, d.n_in = w.shape
    d.complex_io = int(w.is_complex())
    d.activation = int(bool(activation))
    d.weight = w.data_ptr()
    d.bias = b.data_ptr() if b is not None else None
    d.omega = om.data_ptr() if (activation and om is not None) else None
    return d


class ComplexGaborLayer2D(nn.Module):
    """exp(1j * omega_0 * Linear(x)); real Linear when is_first (complex_inr.py:6-40)."""

    def __init__(self, in_features, out_features, bias=True, is_first=False, omega0=100.0, trainable=False):
        super().__init__()
        self.is_first = is_first
        self.in_features = in_features
        dtype = torch.float if is_first else torch.cfloat
        self.omega_0 = nn.Parameter(omega0 * torch.ones(1), trainable)
        self.linear = nn.Linear(in_features, out_features, bias=bias, dtype=dtype)

    def forward(self, input):
        """exp(1j * omega_0 * linear(input)) (complex_inr.py:38-40), one HIP layer launch."""
        return LayerFn.apply(input, self.linear.weight, self.linear.bias, self.omega_0, True)


class ComplexLinear(nn.Linear):
    """The final ``nn.Linear(hidden, out, dtype=cfloat)`` of ComplexMLP.net
    (complex_inr.py:73-78): same parameters, forward on the HIP layer kernel."""

    def forward(self, input):
        return LayerFn.apply(input, self.weight, self.bias, None, False)


class ComplexMLP(nn.Module):
    """Complex-exponential MLP (complex_inr.py:42-82); hidden_features//2 complex units."""

    def __init__(self, in_features, hidden_features, hidden_layers, out_features, first_omega_0=10,
                 hidden_omega_0=10.0, trainable_omega0=False):
        super().__init__()
        hidden_features = int(hidden_features / 2)
        net = [ComplexGaborLayer2D(in_features, hidden_features, omega0=first_omega_0, is_first=True,
                                   trainable=trainable_omega0)]
        for _ in range(hidden_layers):
            net.append(ComplexGaborLayer2D(hidden_features, hidden_features, omega0=hidden_omega_0,
                                           trainable=trainable_omega0))
        net.append(ComplexLinear(hidden_features, out_features, dtype=torch.cfloat))
        self.net = nn.Sequential(*net)
        self.precision = "fp32"
        self._fp8_state = None

    def fp8_state(self, device) -> torch.Tensor:
        """This MLP's FP8 scaling state on ``device`` (created on first use)."""
        if self._fp8_state is None or self._fp8_state.device != torch.device(device):
            self._fp8_state = fp8_state_tensor(device)
        return self._fp8_state

    def n_hidden(self) -> int:
        """Number of hidden complex Gabor layers."""
        return len(self.net) - 2

    def tensors(self):
        """[w0, b0, omega0, (A_k, c_k, omega_k)*, wo, bo] in kernel order."""
        layers = list(self.net)
        ts = [layers[0].linear.weight, layers[0].linear.bias, layers[0].omega_0]
        for layer in layers[1:-1]:
            ts += [layer.linear.weight, layer.linear.bias, layer.omega_0]
        ts += [layers[-1].weight, layers[-1].bias]
        return ts

    def check_supported(self):
        layers = list(self.net)
        if layers[0].linear.out_features != 32 or layers[-1].out_features != 1:
            raise ValueError("the gfx950 MLP kernels implement hidden_features=64 (32 complex units) "
                             "and one complex output")
        if layers[0].linear.bias is None or any(l.linear.bias is None for l in layers[1:-1]):
            raise ValueError("the gfx950 MLP kernels need biased layers")
        if not 1 <= len(layers) - 2 <= 2:
            raise ValueError("the gfx950 MLP kernels implement hidden_layers in 1..2")

    def _fused_ok(self, coords) -> bool:
        try:
            self.check_supported()
        except ValueError:
            return False
        return coords.dim() == 2 and not coords.is_complex()

    def forward(self, coords):
        """ComplexMLP.forward: ``.real`` of the network output (complex_inr.py:80-82).

        Without autograd the whole network runs in one fused launch
        (``pty_mlp_fwd``); with autograd (or a network shape the fused kernel
        does not take) layer by layer (``pty_layer_fwd/bwd``).
        """
        return self.complex_forward(coords).real

    def complex_forward(self, coords):
        """The complex network output ``net(coords)`` (what _get_complex_output returns)."""
        grad = torch.is_grad_enabled() and (coords.requires_grad or any(p.requires_grad for p in self.parameters()))
        if not grad and self._fused_ok(coords):
            return mlp_forward(self, coords)
        return self.net(coords)


FP8_STATE_FLOATS = 520   # PTY_FP8_STATE_FLOATS: 8 scales + 8 x 64 amax sub-slots


def fp8_state_tensor(device) -> torch.Tensor:
    """Delayed-scaling state of an FP8 MLP: 8 scales (1 until a step has
    recorded amax values), then the zeroed amax sub-slots."""
    t = torch.zeros(FP8_STATE_FLOATS, dtype=torch.float32, device=device)
    t[:8] = 1.0
    return t


def mlp_descriptor(ts, n_in: int, precision: str = "fp32", fp8_state: torch.Tensor | None = None) -> _lib.PtyMlp:
    n_hidden = (len(ts) - 5) // 3
    d = _lib.PtyMlp()
    d.n_in = n_in
    d.n_hidden = n_hidden
    d.precision = _lib.PREC_CODES[precision]
    if precision == "fp8":
        if fp8_state is None:
            raise ValueError("an FP8 MLP needs its fp8_state tensor")
        d.fp8_state = fp8_state.data_ptr()
    d.w0, d.b0 = ts[0].data_ptr(), ts[1].data_ptr()
    d.omega[0] = ts[2].data_ptr()
    for k in range(n_hidden):
        d.wh[k] = ts[3 + 3 * k].data_ptr()
        d.bh[k] = ts[4 + 3 * k].data_ptr()
        d.omega[k + 1] = ts[5 + 3 * k].data_ptr()
    d.wo, d.bo = ts[-2].data_ptr(), ts[-1].data_ptr()
    return d


def mlp_forward(mlp: ComplexMLP, feats: torch.Tensor) -> torch.Tensor:
    mlp.check_supported()
    ts = [t.detach().contiguous() for t in mlp.tensors()]
    _lib.require_cuda(feats, *ts)
    feats = feats.detach().contiguous().float()
    n = feats.shape[0]
    out = torch.empty(n, 1, dtype=torch.complex64, device=feats.device)
    _lib.call("pty_mlp_fwd", mlp_descriptor(ts, feats.shape[1], mlp.precision), feats.data_ptr(), n,
              out.data_ptr(), _lib.stream_ptr(feats.device))
    return out


# ----------------------------------------------------------------------------
# point sets
# ----------------------------------------------------------------------------
class GridPoints:
    """Regular cell-centred grid of init_scale_grids (complex_inr.py:179-207).

    Point n = i*height + j has input (ys[j], xs[i]) with
    xs = linspace(0.5/width, 1-0.5/width, width) (f32), ys likewise.
    """

    def __init__(self, width: int, height: int, device):
        half_dx = 0.5 / width
        half_dy = 0.5 / height
        self.xs = torch.linspace(half_dx, 1 - half_dx, width).to(device)
        self.ys = torch.linspace(half_dy, 1 - half_dy, height).to(device)
        self.width, self.height = width, height
        self.n = width * height
        self._plans = {}

    def _raw(self) -> _lib.PtyPoints:
        d = _lib.PtyPoints()
        d.coords = None
        d.axis0 = self.ys.data_ptr()
        d.axis1 = self.xs.data_ptr()
        d.n_axis0 = self.height
        d.n_axis1 = self.width
        d.n = self.n
        d.plan = None
        return d

    def plan(self, spec: GridSpec) -> torch.Tensor:
        """Device plan of this grid for the table gradient (built once per level table)."""
        if spec not in self._plans:
            raw = self._raw()
            g = spec.descriptor()
            nbytes = _lib.query("pty_points_plan_bytes", g, raw)
            buf = _lib.workspace(nbytes, self.xs.device)
            _lib.call("pty_points_plan_build", g, raw, buf.data_ptr(), buf.numel(), _lib.stream_ptr(self.xs.device))
            self._plans[spec] = buf
        return self._plans[spec]

    def descriptor(self, spec: GridSpec | None = None) -> _lib.PtyPoints:
        d = self._raw()
        if spec is not None:
            d.plan = self.plan(spec).data_ptr()
        return d


class CoordPoints:
    """Explicit coordinate list [n, 2]."""

    def __init__(self, coords: torch.Tensor):
        self.coords = coords.detach().contiguous().float()
        self.n = self.coords.shape[0]

    def descriptor(self, spec: GridSpec | None = None) -> _lib.PtyPoints:
        d = _lib.PtyPoints()
        d.coords = self.coords.data_ptr()
        d.n = self.n
        d.plan = None
        return d


# ----------------------------------------------------------------------------
# autograd functions over the HIP kernels
# ----------------------------------------------------------------------------
class InrFieldFn(torch.autograd.Function):
    """Complex field [n] = ComplexMLP.net(HashGrid(points)), fused on the device."""

    @staticmethod
    def forward(ctx, spec: GridSpec, points, precision: str, table_precision: str, fp8_state, table, *ts):
        _lib.require_cuda(table, *ts)
        ts = tuple(t.contiguous() for t in ts)
        out = torch.empty(points.n, dtype=torch.complex64, device=table.device)
        storage = table_storage(table, table_precision)
        _lib.call("pty_inr_fwd", spec.descriptor(table_precision),
                  mlp_descriptor(ts, spec.n_output_dims, precision, fp8_state),
                  storage.data_ptr(), points.descriptor(spec), out.data_ptr(), _lib.stream_ptr(table.device))
        ctx.spec, ctx.points, ctx.precision, ctx.table_precision = spec, points, precision, table_precision
        ctx.fp8_state = fp8_state
        ctx.n_table = table.numel()
        ctx.save_for_backward(storage, *ts)
        return out

    @staticmethod
    def backward(ctx, gout):
        storage, *ts = ctx.saved_tensors
        spec, points, precision = ctx.spec, ctx.points, ctx.precision
        grid = spec.descriptor(ctx.table_precision)
        gout = gout.resolve_conj().contiguous()
        grads = [torch.zeros_like(t) for t in ts]
        dtable = torch.zeros(ctx.n_table, dtype=torch.float32, device=storage.device)
        desc = mlp_descriptor(ts, spec.n_output_dims, precision, ctx.fp8_state)
        gdesc = mlp_descriptor(grads, spec.n_output_dims, precision, ctx.fp8_state)
        pdesc = points.descriptor(spec)
        nbytes = _lib.query("pty_inr_bwd_workspace", grid, desc, pdesc)
        ws = _lib.workspace(nbytes, storage.device)
        stream = _lib.stream_ptr(storage.device)
        _lib.call("pty_inr_bwd", grid, desc, gdesc, storage.data_ptr(), dtable.data_ptr(), pdesc,
                  gout.data_ptr(), ws.data_ptr(), ws.numel(), stream)
        if precision == "fp8":
            # scales of the next step from the amax this step recorded
            _lib.call("pty_fp8_roll", ctx.fp8_state.data_ptr(), stream)
        return (None, None, None, None, None, dtable, *grads)


class UpsampleFn(torch.autograd.Function):
    """fourier_upsample (complex_inr.py:252-271) of a square [n, n] field."""

    @staticmethod
    def forward(ctx, small, ds: int):
        _lib.require_cuda(small)
        n = small.shape[-1]
        if small.shape[-2] != n:
            raise ValueError("the gfx950 Fourier upsample implements square grids")
        small = small.resolve_conj().contiguous()
        big = torch.empty(n * ds, n * ds, dtype=torch.complex64, device=small.device)
        ws = _lib.workspace(_lib.query("pty_upsample_workspace", n, ds), small.device)
        _lib.call("pty_upsample_fwd", small.data_ptr(), n, ds, big.data_ptr(), ws.data_ptr(), ws.numel(),
                  _lib.stream_ptr(small.device))
        ctx.n, ctx.ds = n, ds
        return big

    @staticmethod
    def backward(ctx, gbig):
        n, ds = ctx.n, ctx.ds
        gbig = gbig.resolve_conj().contiguous()
        gsmall = torch.empty(n, n, dtype=torch.complex64, device=gbig.device)
        ws = _lib.workspace(_lib.query("pty_upsample_workspace", n, ds), gbig.device)
        _lib.call("pty_upsample_adj", gbig.data_ptr(), n, ds, gsmall.data_ptr(), ws.data_ptr(), ws.numel(),
                  _lib.stream_ptr(gbig.device))
        return gsmall, None


def fourier_upsample(x: torch.Tensor, scale_factor: int) -> torch.Tensor:
    """Function form of the upsample for a [..., n, n] complex field (leading dims of size 1)."""
    lead = x.shape[:-2]
    y = UpsampleFn.apply(x.reshape(x.shape[-2], x.shape[-1]), int(scale_factor))
    return y.reshape(*lead, y.shape[-2], y.shape[-1])


class ResidualFn(torch.autograd.Function):
    """k * x + initial (complex_inr.py:324, 331); k real scalar parameter."""

    @staticmethod
    def forward(ctx, k, x, initial):
        _lib.require_cuda(k, x, initial)
        xc = x.resolve_conj().contiguous()
        init = initial.to(torch.complex64).contiguous()
        out = torch.empty_like(xc)
        _lib.call("pty_residual_fwd", k.data_ptr(), xc.data_ptr(), init.data_ptr(), xc.numel(), out.data_ptr(),
                  _lib.stream_ptr(x.device))
        ctx.save_for_backward(k, xc)
        return out

    @staticmethod
    def backward(ctx, g):
        k, x = ctx.saved_tensors
        g = g.resolve_conj().contiguous()
        gx = torch.empty_like(x)
        gk = torch.zeros(1, dtype=torch.float32, device=x.device)
        ws = _lib.workspace(_lib.query("pty_residual_workspace", x.numel()), x.device)
        _lib.call("pty_residual_bwd", k.data_ptr(), x.data_ptr(), g.data_ptr(), x.numel(), gx.data_ptr(),
                  gk.data_ptr(), ws.data_ptr(), ws.numel(), _lib.stream_ptr(x.device))
        return gk.reshape(k.shape), gx, None


# ----------------------------------------------------------------------------
# the model
# ----------------------------------------------------------------------------
class ComplexINRModel2D(nn.Module):
    """Drop-in for models/complex_inr.py:85-334 on the gfx950 kernels."""

    def __init__(self, output_width, output_height, downsample_factor, update_probe=True, probe_width=None,
                 probe_height=None, n_levels=16, n_features_per_level=2, log2_hashmap_size=19, base_resolution=16,
                 per_level_scale=1.5, first_omega_0=10.0, hidden_omega_0=10.0, hidden_features=64,
                 hidden_layers=2, trainable_omega0=False, use_residual=False, object_initial=None,
                 probe_initial=None, mlp_precision: str = "fp32", table_precision: str = "fp32"):
        super().__init__()
        self.update_probe = update_probe
        self.use_residual = use_residual
        self.output_width = output_width
        self.output_height = output_height
        self.probe_width = probe_width if probe_width is not None else output_width
        self.probe_height = probe_height if probe_height is not None else output_height
        self.num_features = n_levels * n_features_per_level
        encoding_config = {
            "otype": "HashGrid",
            "n_levels": n_levels,
            "n_features_per_level": n_features_per_level,
            "log2_hashmap_size": log2_hashmap_size,
            "base_resolution": base_resolution,
            "per_level_scale": per_level_scale,
        }
        if table_precision not in _lib.TABLE_CODES:
            raise ValueError(f"table_precision must be one of {tuple(_lib.TABLE_CODES)}")
        tdtype = torch.float16 if table_precision == "fp16" else None
        self.xy_hash_encoder = HashGridEncoding(n_input_dims=2, encoding_config=encoding_config, dtype=tdtype)
        self.probe_hash_encoder = (HashGridEncoding(n_input_dims=2, encoding_config=encoding_config.copy(),
                                                    dtype=tdtype)
                                   if update_probe else None)
        self.object_complex_mlp = ComplexMLP(self.num_features, hidden_features, hidden_layers, 1,
                                             first_omega_0, hidden_omega_0, trainable_omega0)
        if update_probe:
            self.probe_complex_mlp = ComplexMLP(self.num_features, hidden_features, hidden_layers, 1,
                                                first_omega_0, hidden_omega_0, trainable_omega0)
        else:
            self.probe_complex_mlp = None
        if mlp_precision not in _lib.PREC_CODES:
            raise ValueError(f"mlp_precision must be one of {tuple(_lib.PREC_CODES)}")
        if mlp_precision == "fp8":
            warnings.warn(FP8_WARNING, stacklevel=2)
        self.mlp_precision = mlp_precision
        self.init_scale_grids(downsample_factor)
        self.has_object_residual = False
        self.has_probe_residual = False
        if self.use_residual:
            self.init_residual_params(object_initial, probe_initial)
        self._grid_points = {}
        self._grid_checked = {}

    def init_scale_grids(self, downsample_factor):
        downsampled_width = self.output_width // downsample_factor
        downsampled_height = self.output_height // downsample_factor
        half_dx = 0.5 / downsampled_width
        half_dy = 0.5 / downsampled_height
        xs = torch.linspace(half_dx, 1 - half_dx, downsampled_width)
        ys = torch.linspace(half_dy, 1 - half_dy, downsampled_height)
        xv, yv = torch.meshgrid([xs, ys], indexing="ij")
        coords = torch.stack((yv.flatten(), xv.flatten())).t()
        self.downsampled_width = downsampled_width
        self.downsampled_height = downsampled_height
        self.xy_coords = nn.Parameter(coords, requires_grad=False)
        if self.update_probe:
            half_dx_probe = 0.5 / self.probe_width
            half_dy_probe = 0.5 / self.probe_height
            xs_probe = torch.linspace(half_dx_probe, 1 - half_dx_probe, self.probe_width)
            ys_probe = torch.linspace(half_dy_probe, 1 - half_dy_probe, self.probe_height)
            xv_probe, yv_probe = torch.meshgrid([xs_probe, ys_probe], indexing="ij")
            probe_coords = torch.stack((yv_probe.flatten(), xv_probe.flatten())).t()
            self.probe_coords = nn.Parameter(probe_coords, requires_grad=False)
        else:
            self.probe_coords = None
        self.downsample_factor = downsample_factor

    def init_residual_params(self, object_initial, probe_initial):
        if object_initial is not None:
            if not torch.is_complex(object_initial):
                raise ValueError("object_initial must be a complex tensor")
            self.register_buffer("object_initial", object_initial.clone())
            self.k_object = nn.Parameter(torch.tensor(1e-8, dtype=torch.float32))
            self.has_object_residual = True
        else:
            self.has_object_residual = False
        if self.update_probe:
            if probe_initial is not None:
                if not torch.is_complex(probe_initial):
                    raise ValueError("probe_initial must be a complex tensor")
                self.register_buffer("probe_initial", probe_initial.clone())
                self.k_probe = nn.Parameter(torch.tensor(1e-8, dtype=torch.float32))
                self.has_probe_residual = True
            else:
                self.has_probe_residual = False
        else:
            self.has_probe_residual = False

    # -------------------------------------------------------------- helpers
    def grid_points(self, which: str):
        """Device-resident regular grid of the object ('object') or probe ('probe')."""
        dev = self.xy_hash_encoder.params.device
        key = (which, str(dev))
        if key not in self._grid_points:
            if which == "object":
                self._grid_points[key] = GridPoints(self.downsampled_width, self.downsampled_height, dev)
            else:
                self._grid_points[key] = GridPoints(self.probe_width, self.probe_height, dev)
        return self._grid_points[key]

    def coords_are_grid(self, which: str) -> bool:
        """Whether ``xy_coords`` ('object') / ``probe_coords`` ('probe') still
        hold the regular grid of init_scale_grids.  Checked on the device when
        the tensor has changed (new storage or an in-place write) since the
        last check, cached otherwise."""
        c = self.xy_coords if which == "object" else self.probe_coords
        key = (which, c.data_ptr(), c._version, str(c.device))
        hit = self._grid_checked.get(which)
        if hit is not None and hit[0] == key:
            return hit[1]
        gp = self.grid_points(which)
        ok = tuple(c.shape) == (gp.n, 2)
        if ok:
            xs, ys = gp.xs.to(c.device), gp.ys.to(c.device)
            ok = (torch.equal(c[:, 0].reshape(gp.width, gp.height), ys.expand(gp.width, gp.height))
                  and torch.equal(c[:, 1].reshape(gp.width, gp.height), xs[:, None].expand(gp.width, gp.height)))
        self._grid_checked[which] = (key, ok)
        return ok

    def _points(self, which: str):
        """The regular grid (kernels generate its coordinates from two axis
        vectors) while the coordinates are the default ones; the coordinate
        list itself once a caller has changed them."""
        if self.coords_are_grid(which):
            return self.grid_points(which)
        c = self.xy_coords if which == "object" else self.probe_coords
        return CoordPoints(c.detach())

    def _field(self, encoder: HashGridEncoding, mlp: ComplexMLP, which: str):
        mlp.check_supported()
        state = mlp.fp8_state(encoder.params.device) if self.mlp_precision == "fp8" else None
        return InrFieldFn.apply(encoder.spec, self._points(which), self.mlp_precision, encoder.table_precision,
                                state, encoder.params, *mlp.tensors())

    def _get_complex_output(self, model, coords):
        """``model.net(coords)``: the complex network output, bypassing the
        ``.real`` of ComplexMLP.forward (complex_inr.py:308-314)."""
        return model.complex_forward(coords)

    def fourier_upsample(self, x, scale_factor):
        return fourier_upsample(x, scale_factor)

    def apply_residual(self, complex_tensor, is_object=True):
        if is_object:
            if self.has_object_residual:
                return ResidualFn.apply(self.k_object, complex_tensor, self.object_initial)
            return complex_tensor
        if self.has_probe_residual:
            return ResidualFn.apply(self.k_probe, complex_tensor, self.probe_initial)
        return complex_tensor

    def forward(self):
        obj = self._field(self.xy_hash_encoder, self.object_complex_mlp, "object")
        obj = obj.reshape(self.downsampled_width, self.downsampled_height)
        obj = UpsampleFn.apply(obj, int(self.downsample_factor))
        if self.update_probe:
            probe = self._field(self.probe_hash_encoder, self.probe_complex_mlp, "probe")
            probe = probe.reshape(self.probe_width, self.probe_height)
        else:
            probe = None
        if self.use_residual:
            obj = self.apply_residual(obj, is_object=True)
            if self.update_probe and probe is not None:
                probe = self.apply_residual(probe, is_object=False)
        return obj, probe
