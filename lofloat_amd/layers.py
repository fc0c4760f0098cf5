"""Autograd layers over the LoFloat kernels (reference LoFloat/layers.py:8-168, 221-447).

``STERound`` / ``STEMXRound`` / ``LoF_Quantize`` keep the reference's straight-through
semantics (layers.py:8-57, 60-104).  ``LoF_Linear`` keeps its constructor, ``mx`` and
``per_tensor`` scaling and forward numerics (layers.py:221-399); unlike the reference,
whose GEMM has no backward (test_mx_round_py.py:57-58), the GEMM here is differentiable:

    forward   out = alpha * (X_q @ W_q^T)                   emulated accumulation
    backward  G_q = virtual_round(grad_out, grad_format, grad_round_mode, bits)
              dX_q = alpha * (G_q @ W_q)                     emulated accumulation
              dW_q = alpha * (G_q^T @ X_q)                   emulated accumulation
    and the straight-through estimator carries dX_q / dW_q through the input rounding.

Both backward GEMMs read transposed operands in place (no transposed copies).
"""
from __future__ import annotations

import math

import torch
import torch.nn as nn

from . import ops
from ._custom_ops import register_format
from .core import FloatFormatDescriptor, RoundingMode
from .formats import create_e8m0_params, create_p3109_params, make_format


def _rne_if_none(mode):
    return RoundingMode.RoundToNearestEven if mode is None else mode


def _round_into(tensor, params, mode, scale=1.0):
    """scale * tensor rounded into ``params``: a registered format id (an int) goes through
    the traceable torch.ops.lofloat.virtual_round, a descriptor through ops.virtual_round."""
    if isinstance(params, int):
        return torch.ops.lofloat.virtual_round(tensor, params, int(mode), 0, float(scale))
    return ops.virtual_round(tensor, params, round_mode=mode, scale=float(scale))


def _mx_round_into(tensor, element, scale_format, block_size, mode):
    """MX rounding of contiguous blocks of the flattened tensor; registered ids take the
    traceable op (scale rounded to nearest even, no stochastic bits)."""
    if isinstance(element, int):
        return torch.ops.lofloat.virtual_mx_round(tensor, int(block_size), element,
                                                  scale_format, int(mode), 0, 0)
    return ops.virtual_mx_round(tensor, int(block_size), element, scale_format, round_mode=mode)


class STERound(torch.autograd.Function):
    """Fake quantisation y = round(scale * x); the straight-through gradient of the rounding
    is the identity, so dy/dx = scale (layers.py:8-31)."""

    @staticmethod
    def forward(ctx, tensor, params, rounding_mode=RoundingMode.RoundToNearestEven, scale=1.0):
        ctx.scale = float(scale)
        return _round_into(tensor, params, rounding_mode, scale)

    @staticmethod
    def backward(ctx, grad_output):
        grad = grad_output if ctx.scale == 1.0 else grad_output * ctx.scale
        return grad, None, None, None


class STEMXRound(torch.autograd.Function):
    """MX fake quantisation over blocks of the last axis; its output is in the input's own
    range, so the straight-through gradient is the identity (layers.py:34-57)."""

    @staticmethod
    def forward(ctx, tensor, element_params, scale_params, block_size,
                rounding_mode=RoundingMode.RoundToNearestEven):
        return _mx_round_into(tensor, element_params, scale_params, block_size, rounding_mode)

    @staticmethod
    def backward(ctx, grad_output):
        return (grad_output,) + (None,) * 4


_SCALINGS = ("per_tensor", "mx")


def _check_scaling(scaling):
    if scaling not in _SCALINGS:
        raise ValueError(f"unknown scaling {scaling!r}: use one of {_SCALINGS}")


def _check_blocks(length, block, what):
    if length % block:
        raise ValueError(f"MX blocks of {block} do not tile {what} of length {length}")


class LoF_Quantize(nn.Module):
    """A fake-quantisation module: per tensor, or MX blocks along the last axis with a
    shared ``scale_format`` scale (E8M0 by default) (layers.py:60-104)."""

    def __init__(self, params, rounding_mode=None, scaling="per_tensor", mx_block_size=32,
                 scale_format=None):
        super().__init__()
        _check_scaling(scaling)
        self.params = params
        self.rounding_mode = _rne_if_none(rounding_mode)
        self.scaling = scaling
        self.mx_block_size = mx_block_size
        self.scale_format = create_e8m0_params() if scale_format is None else scale_format

    def forward(self, x):
        if self.scaling == "per_tensor":
            return STERound.apply(x, self.params, self.rounding_mode)
        _check_blocks(x.shape[-1], self.mx_block_size, "the last axis")
        return STEMXRound.apply(x, self.params, self.scale_format, self.mx_block_size,
                                self.rounding_mode)

    def extra_repr(self):
        parts = [f"params={self.params}", f"rounding_mode={self.rounding_mode}",
                 f"scaling={self.scaling}"]
        if self.scaling == "mx":
            parts.append(f"mx_block_size={self.mx_block_size}")
        return ", ".join(parts)


def _p3109_fields(exp_bits, mant_bits):
    """The signed, saturating P3109 format with these exponent and mantissa field widths."""
    return create_p3109_params(exp_bits + mant_bits + 1, mant_bits + 1, True, True)


def _quantize(layer, tensor, params):
    """Fake-quantise ``tensor`` to ``params`` under ``layer.rounding_mode``; no format, no
    rounding (layers.py:106-109)."""
    return tensor if params is None else STERound.apply(tensor, params, layer.rounding_mode)


def mantissa_quantize(tensor, mantissa_bits):
    """Round to ``mantissa_bits`` with an 8-bit exponent field (layers.py:111-114)."""
    return tensor if mantissa_bits is None else STERound.apply(
        tensor, _p3109_fields(8, mantissa_bits))


def exp_mant_quantize(tensor, exp_bits, mantissa_bits):
    """Round to the given exponent / mantissa field widths (layers.py:116-119)."""
    if None in (exp_bits, mantissa_bits):
        return tensor
    return STERound.apply(tensor, _p3109_fields(exp_bits, mantissa_bits))


def _fwht(x: torch.Tensor, block_size=None) -> torch.Tensor:
    """Orthonormal Walsh-Hadamard transform along the last dim, optionally per block of a
    power-of-two size (layers.py:182-216)."""
    n = x.shape[-1]
    bs = n if block_size is None else block_size
    if bs <= 0 or (bs & (bs - 1)) != 0:
        raise ValueError(f"block_size must be a power of 2, got {bs}")
    if n % bs != 0:
        raise ValueError(f"block_size {bs} must divide last dim {n}")
    shape = x.shape
    y = x.reshape(-1, bs)
    h = 1
    while h < bs:
        y = y.reshape(y.shape[0], bs // (2 * h), 2, h)
        y = torch.stack((y[:, :, 0, :] + y[:, :, 1, :], y[:, :, 0, :] - y[:, :, 1, :]), dim=2)
        y = y.reshape(-1, bs)
        h *= 2
    return (y * (1.0 / math.sqrt(bs))).reshape(shape)


def _lof_gemm_2d(A, B, accum_mant_bits, gemm_round_mode, stochastic_rounding_bits,
                 scale_a=1.0, scale_b=1.0):
    """(..., M, K) x (..., K, N) -> (..., M, N) through lof_gemm (layers.py:121-150)."""
    if A.dim() == 2 and B.dim() == 2:
        return ops.lof_gemm(A, B, accum_mant_bits, gemm_round_mode, stochastic_rounding_bits,
                            scale_a, scale_b)
    batch = A.shape[:-2]
    M, K, N = A.shape[-2], A.shape[-1], B.shape[-1]
    Af = A.reshape(-1, M, K)
    Bf = B.reshape(-1, K, N)
    out = torch.stack([ops.lof_gemm(Af[i], Bf[i], accum_mant_bits, gemm_round_mode,
                                    stochastic_rounding_bits, scale_a, scale_b)
                       for i in range(Af.shape[0])])
    return out.reshape(*batch, M, N)


def _alpha(scale_a: float, scale_b: float) -> float:
    import numpy as np
    denom = np.float32(scale_a) * np.float32(scale_b)
    return 1.0 if denom == 0 else float(np.float32(1.0) / denom)


class LinearGradSpec:
    """How LoF_Linear's backward rounds grad_out before the dgrad / wgrad GEMMs."""

    def __init__(self, fmt: FloatFormatDescriptor | None = None,
                 round_mode=RoundingMode.StochasticRoundingA, stoch_len: int = 8,
                 seed: int | None = None):
        self.fmt = fmt if fmt is not None else make_format(8, 2, 15)   # e5m2
        self.round_mode = RoundingMode(int(round_mode))
        self.stoch_len = int(stoch_len)
        self.seed = seed   # None: a fresh seed per backward (ops.next_seed)


class LoFLinearFunction(torch.autograd.Function):
    """out = alpha * (x_q @ w_q^T) with emulated accumulation, hand-written backward."""

    @staticmethod
    def forward(ctx, x_q, w_q, accum: ops.AccumSpec, alpha: float, grad: LinearGradSpec):
        lead = x_q.shape[:-1]
        K = x_q.shape[-1]
        x2 = x_q.reshape(-1, K)
        out = ops.gemm(x2, w_q.t(), accum, alpha=alpha)          # w_q^T read in place
        ctx.save_for_backward(x2, w_q)
        ctx.accum, ctx.alpha, ctx.grad, ctx.lead = accum, alpha, grad, lead
        return out.reshape(*lead, w_q.shape[0])

    @staticmethod
    def backward(ctx, grad_out):
        x2, w_q = ctx.saved_tensors
        g = ctx.grad
        G = grad_out.reshape(-1, w_q.shape[0]).contiguous()
        seed = g.seed if g.seed is not None else ops.next_seed()
        G_q = ops.virtual_round(G, g.fmt, g.round_mode, g.stoch_len, seed=seed)
        dx = dw = None
        if ctx.needs_input_grad[0]:
            dx = ops.gemm(G_q, w_q, ctx.accum, alpha=ctx.alpha).reshape(*ctx.lead, x2.shape[1])
        if ctx.needs_input_grad[1]:
            dw = ops.gemm(G_q.t(), x2, ctx.accum, alpha=ctx.alpha)   # G_q^T read in place
        return dx, dw, None, None, None


# How each of the reference layers' format setters rebuilds a role's format from the old one
# (layers.py:402-432, 757-787): rule(exp_bits, mant_bits, old, value) gives the arguments
# (k, p, saturating, bias) of a signed P3109 format.  set_exponentbias and
# set_saturation_mode hand the old mantissa_bits to create_p3109_params as its precision p,
# so their rebuilt formats carry one mantissa bit fewer; that is kept as written.
_FORMAT_RULES = {
    "set_mantissa": lambda e, m, old, v: (1 + e + v, v + 1, True, None),
    "set_exponent": lambda e, m, old, v: (1 + v + m, m + 1, True, None),
    "set_exponentbias": lambda e, m, old, v: (old.total_bits, m, True, v),
    "set_saturation_mode": lambda e, m, old, v: (old.total_bits, m, v, None),
}
_ROLES = ("act", "weight", "bias")


def _role_format(params, exp_bits, mant_bits, role, required):
    """A role's format: ``params`` if given, else the P3109 format of its field widths."""
    if params is not None:
        return params
    if exp_bits is not None and mant_bits is not None:
        return _p3109_fields(exp_bits, mant_bits)
    if required:
        raise ValueError(f"the {role} format needs {role}_params or both {role}_exp and "
                         f"{role}_mant")
    return None


class _LoFLayer(nn.Module):
    """What LoF_Linear and LoF_Conv2d share: the activation / weight / bias formats, the
    rounding modes and scale factors, the accumulator and the scaling scheme, and the
    reference layers' format setters."""

    def _init_quantisation(self, *, formats, rounding_mode, scale_factors, accum_mant_bits,
                           gemm_round_mode, stochastic_rounding_bits, scaling, mx_block_size,
                           scale_format, hadamard, accum_format, accum_chunk):
        _check_scaling(scaling)
        self.rounding_mode = _rne_if_none(rounding_mode)
        self.gemm_round_mode = _rne_if_none(gemm_round_mode)
        self.act_scale_factor, self.w_scale_factor, self.b_scale_factor = scale_factors
        self.accum_mant_bits = accum_mant_bits
        self.stochastic_rounding_bits = stochastic_rounding_bits
        self.accum_format = accum_format
        self.accum_chunk = int(accum_chunk)
        self.scaling = scaling
        self.mx_block_size = mx_block_size
        self.scale_format = create_e8m0_params() if scale_format is None else scale_format
        self.hadamard_transform, self.hadamard_block_size = hadamard
        for role, (params, e, m) in zip(_ROLES, formats):
            setattr(self, f"{role}_params", _role_format(params, e, m, role, role != "bias"))
        self._refresh_traced_ids()

    def _refresh_traced_ids(self):
        # the traceable ops take formats as registry ids and modes as ints
        self._act_fid = register_format(self.act_params)
        self._weight_fid = register_format(self.weight_params)
        self._bias_fid = None if self.bias_params is None else register_format(self.bias_params)
        self._round_mode_int = int(self.rounding_mode)
        self._gemm_round_mode_int = int(self.gemm_round_mode)

    def _rebuild_formats(self, rule, values):
        derive = _FORMAT_RULES[rule]
        for role, value in zip(_ROLES, values):
            old = getattr(self, f"{role}_params")
            if old is None or value is None:
                continue
            k, p, saturating, bias = derive(old.total_bits - old.mantissa_bits - 1,
                                            old.mantissa_bits, old, value)
            setattr(self, f"{role}_params", create_p3109_params(k, p, True, saturating, bias))
        self._refresh_traced_ids()

    def set_mantissa(self, activ_mant, weight_mant, bias_mant):
        self._rebuild_formats("set_mantissa", (activ_mant, weight_mant, bias_mant))

    def set_exponent(self, activ_exp, weight_exp, bias_exp):
        self._rebuild_formats("set_exponent", (activ_exp, weight_exp, bias_exp))

    def set_exponentbias(self, activ_expbias, weight_expbias, bias_expbias):
        self._rebuild_formats("set_exponentbias", (activ_expbias, weight_expbias, bias_expbias))

    def set_saturation_mode(self, activ_sat, weight_sat, bias_sat):
        self._rebuild_formats("set_saturation_mode", (activ_sat, weight_sat, bias_sat))

    def set_accumulation_precision(self, accum_mant_bits):
        self.accum_mant_bits = accum_mant_bits

    def accum_spec(self) -> ops.AccumSpec:
        if self.accum_format is not None:
            return ops.AccumSpec(fmt=self.accum_format, chunk=self.accum_chunk)
        return ops.AccumSpec(mant_bits=int(self.accum_mant_bits), chunk=self.accum_chunk)

    def _make_parameters(self, weight_shape, n_out, bias):
        self.weight = nn.Parameter(torch.empty(*weight_shape))
        nn.init.kaiming_uniform_(self.weight)
        if bias:
            self.bias = nn.Parameter(torch.zeros(n_out))
        else:
            self.register_parameter("bias", None)

    def _adopt(self, source: nn.Module):
        """Take over a float layer's weight and bias (copies)."""
        self.weight = nn.Parameter(source.weight.detach().clone())
        if source.bias is not None:
            self.bias = nn.Parameter(source.bias.detach().clone())
        return self


class LoF_Linear(_LoFLayer):
    """Drop-in for the reference LoF_Linear (layers.py:221-447) with a differentiable GEMM.

    Extra keyword arguments (all optional): ``accum_format`` (round the accumulator to a
    full format instead of ``accum_mant_bits`` mantissa bits), ``accum_chunk`` (k values
    accumulated in FP32 between roundings; 1 = after every FMA as in the reference) and
    ``grad_spec`` (a LinearGradSpec: format / mode / stoch_len / seed for grad_out)."""

    def __init__(self, in_features, out_features, act_params=None, weight_params=None,
                 bias=True, bias_params=None, act_exp=None, act_mant=None, weight_exp=None,
                 weight_mant=None, bias_exp=None, bias_mant=None, rounding_mode=None,
                 act_scale_factor=1.0, w_scale_factor=1.0, b_scale_factor=1.0,
                 accum_mant_bits=23, gemm_round_mode=None, stochastic_rounding_bits=0,
                 hadamard_transform=False, hadamard_block_size=None, scaling="per_tensor",
                 mx_block_size=32, scale_format=None, *, accum_format=None, accum_chunk=1,
                 grad_spec: LinearGradSpec | None = None):
        super().__init__()
        self.in_features = in_features
        self.out_features = out_features
        self._init_quantisation(
            formats=((act_params, act_exp, act_mant), (weight_params, weight_exp, weight_mant),
                     (bias_params, bias_exp, bias_mant)),
            rounding_mode=rounding_mode,
            scale_factors=(act_scale_factor, w_scale_factor, b_scale_factor),
            accum_mant_bits=accum_mant_bits, gemm_round_mode=gemm_round_mode,
            stochastic_rounding_bits=stochastic_rounding_bits, scaling=scaling,
            mx_block_size=mx_block_size, scale_format=scale_format,
            hadamard=(hadamard_transform, hadamard_block_size), accum_format=accum_format,
            accum_chunk=accum_chunk)
        if scaling == "mx":   # blocks run along K, the last axis of x and of the weight
            _check_blocks(in_features, mx_block_size, "in_features")
        self.grad_spec = grad_spec if grad_spec is not None else LinearGradSpec()
        self._make_parameters((out_features, in_features), out_features, bias)

    @classmethod
    def from_linear(cls, linear: nn.Linear, act_params, weight_params, bias_params=None,
                    rounding_mode=None, accum_mant_bits=23, gemm_round_mode=None,
                    stochastic_rounding_bits=0, hadamard_transform=False,
                    hadamard_block_size=None, **kw):
        """A LoF_Linear holding copies of ``linear``'s weight and bias (layers.py:301-323)."""
        return cls(linear.in_features, linear.out_features, act_params=act_params,
                   weight_params=weight_params, bias=linear.bias is not None,
                   bias_params=bias_params, rounding_mode=rounding_mode,
                   accum_mant_bits=accum_mant_bits, gemm_round_mode=gemm_round_mode,
                   stochastic_rounding_bits=stochastic_rounding_bits,
                   hadamard_transform=hadamard_transform,
                   hadamard_block_size=hadamard_block_size, **kw)._adopt(linear)

    def quantized_operands(self, x):
        """(x_q, w_q, alpha) exactly as forward() feeds them to the GEMM."""
        if self.hadamard_transform:
            x = _fwht(x, block_size=self.hadamard_block_size)
        if self.scaling == "mx":
            x_q = STEMXRound.apply(x, self.act_params, self.scale_format, self.mx_block_size,
                                   self.rounding_mode)
            w_q = STEMXRound.apply(self.weight, self.weight_params, self.scale_format,
                                   self.mx_block_size, self.rounding_mode)
            return x_q, w_q, 1.0
        x_q = STERound.apply(x, self._act_fid, self._round_mode_int, self.act_scale_factor)
        w_q = STERound.apply(self.weight, self._weight_fid, self._round_mode_int,
                             self.w_scale_factor)
        return x_q, w_q, _alpha(self.act_scale_factor, self.w_scale_factor)

    def forward(self, x):
        x_q, w_q, alpha = self.quantized_operands(x)
        out = LoFLinearFunction.apply(x_q, w_q, self.accum_spec(), alpha, self.grad_spec)
        if self.bias is not None:
            if self._bias_fid is not None:
                b_q = STERound.apply(self.bias, self._bias_fid, self._round_mode_int,
                                     self.b_scale_factor)
                if self.b_scale_factor != 1.0:
                    b_q = b_q / self.b_scale_factor
            else:
                b_q = self.bias
            out = out + b_q
        return out

    def extra_repr(self):
        return (f"in_features={self.in_features}, out_features={self.out_features}, "
                f"bias={self.bias is not None}, scaling={self.scaling}, "
                f"accum_mant_bits={self.accum_mant_bits}, accum_format={self.accum_format}, "
                f"accum_chunk={self.accum_chunk}")


def _as_pair(v):
    return tuple(v) if isinstance(v, (tuple, list)) else (v, v)


def _hadamard_block(block_size, channels):
    """The Walsh-Hadamard block along the input channels of one group: ``block_size`` or,
    when None, all of them; a power of two dividing the channel count."""
    b = channels if block_size is None else block_size
    if b <= 0 or b & (b - 1):
        raise ValueError(f"the Hadamard block ({b}) must be a power of two")
    if channels % b:
        raise ValueError(f"the Hadamard block ({b}) must divide the {channels} input channels "
                         "of a group")
    return b


class LoF_Conv2d(_LoFLayer):
    """Drop-in for the reference LoF_Conv2d (layers.py:450-800): per-tensor or MX
    quantisation of the activation and weight, then the emulated-accumulation convolution,
    then the (quantised) bias.  Forward only, as in the reference.

    ``backend="im2col"`` (default) and ``"cutlass"`` keep their reference numerics:
    im2col = one lof_gemm(w_col, unfold(x)[b]) per image with alpha = 1 / (s_w * s_a) in
    float (layers.py:687-737); cutlass = lof_conv2d with the scales folded in double
    (layers.py:641-660).  Both run the implicit-GEMM kernel (no unfold is materialised);
    each output column of the per-image GEMMs is independent, so stacking them changes
    nothing.  ``scaling="mx"`` rounds the unfolded activation in blocks along its reduction
    axis, which needs the unfold materialised, then runs one GEMM over all images.
    ``groups > 1`` runs one convolution per group.  Extra keywords: ``accum_format``,
    ``accum_chunk`` (as in LoF_Linear)."""

    _BACKENDS = ("im2col", "cutlass")

    def __init__(self, in_channels, out_channels, kernel_size, stride=1, padding=0,
                 dilation=1, groups=1, act_params=None, weight_params=None, bias=True,
                 bias_params=None, act_exp=None, act_mant=None, weight_exp=None,
                 weight_mant=None, bias_exp=None, bias_mant=None, rounding_mode=None,
                 act_scale_factor=1.0, w_scale_factor=1.0, b_scale_factor=1.0,
                 accum_mant_bits=23, gemm_round_mode=None, stochastic_rounding_bits=0,
                 hadamard_transform=False, hadamard_block_size=None, scaling="per_tensor",
                 mx_block_size=32, scale_format=None, backend="im2col", *, accum_format=None,
                 accum_chunk=1):
        super().__init__()
        if backend not in self._BACKENDS:
            raise ValueError(f"unknown backend {backend!r}: use one of {self._BACKENDS}")
        self.backend = backend
        self.in_channels, self.out_channels, self.groups = in_channels, out_channels, groups
        self.kernel_size = _as_pair(kernel_size)
        self.stride = _as_pair(stride)
        self.padding = _as_pair(padding)
        self.dilation = _as_pair(dilation)
        self._init_quantisation(
            formats=((act_params, act_exp, act_mant), (weight_params, weight_exp, weight_mant),
                     (bias_params, bias_exp, bias_mant)),
            rounding_mode=rounding_mode,
            scale_factors=(act_scale_factor, w_scale_factor, b_scale_factor),
            accum_mant_bits=accum_mant_bits, gemm_round_mode=gemm_round_mode,
            stochastic_rounding_bits=stochastic_rounding_bits, scaling=scaling,
            mx_block_size=mx_block_size, scale_format=scale_format,
            hadamard=(hadamard_transform, hadamard_block_size), accum_format=accum_format,
            accum_chunk=accum_chunk)
        # what each path supports: the direct convolution takes one group and per-tensor
        # scales; MX blocks run along the unfolded reduction axis C/groups * kH * kW
        if (backend == "cutlass" or scaling == "mx") and groups != 1:
            raise NotImplementedError(f"{backend} / {scaling}: grouped convolution is not "
                                      "supported on this path (groups must be 1)")
        if backend == "cutlass" and scaling == "mx":
            raise NotImplementedError("the cutlass backend takes per-tensor scaling only")
        if scaling == "mx":
            kh, kw = self.kernel_size
            _check_blocks(in_channels // groups * kh * kw, mx_block_size,
                          "the unfolded reduction axis (C/groups * kH * kW)")
        if hadamard_transform:
            _hadamard_block(hadamard_block_size, in_channels // groups)
        self._make_parameters((out_channels, in_channels // groups, *self.kernel_size),
                              out_channels, bias)

    @classmethod
    def from_conv2d(cls, conv: nn.Conv2d, act_params, weight_params, bias_params=None,
                    rounding_mode=None, accum_mant_bits=23, gemm_round_mode=None,
                    stochastic_rounding_bits=0, hadamard_transform=False,
                    hadamard_block_size=None, **kw):
        """A LoF_Conv2d with ``conv``'s geometry holding copies of its weight and bias
        (layers.py:560-587)."""
        return cls(conv.in_channels, conv.out_channels, conv.kernel_size, stride=conv.stride,
                   padding=conv.padding, dilation=conv.dilation, groups=conv.groups,
                   act_params=act_params, weight_params=weight_params,
                   bias=conv.bias is not None, bias_params=bias_params,
                   rounding_mode=rounding_mode, accum_mant_bits=accum_mant_bits,
                   gemm_round_mode=gemm_round_mode,
                   stochastic_rounding_bits=stochastic_rounding_bits,
                   hadamard_transform=hadamard_transform,
                   hadamard_block_size=hadamard_block_size, **kw)._adopt(conv)

    def _bias(self, out, scaled=True):
        if self.bias is None:
            return out
        if self._bias_fid is None:
            b_q = self.bias
        elif scaled:
            b_q = STERound.apply(self.bias, self._bias_fid, self._round_mode_int,
                                 self.b_scale_factor)
            if self.b_scale_factor != 1.0:
                b_q = b_q / self.b_scale_factor
        else:
            b_q = STERound.apply(self.bias, self._bias_fid, self._round_mode_int)
        return out + b_q.view(1, -1, 1, 1)

    def forward(self, x):
        if self.hadamard_transform:   # along the channel axis, per group (layers.py:622-636)
            if self.groups == 1:
                x = _fwht(x.movedim(1, -1).contiguous(), self.hadamard_block_size)
                x = x.movedim(-1, 1).contiguous()
            else:
                B, C_in, H_in, W_in = x.shape
                x = x.view(B, self.groups, C_in // self.groups, H_in, W_in)
                x = _fwht(x.movedim(2, -1).contiguous(), self.hadamard_block_size)
                x = x.movedim(-1, 2).contiguous().view(B, C_in, H_in, W_in)
        spec = self.accum_spec()
        if self.scaling == "mx":
            return self._bias(self._forward_mx(x, spec), scaled=False)
        x_q = STERound.apply(x, self._act_fid, self._round_mode_int, self.act_scale_factor)
        w_q = STERound.apply(self.weight, self._weight_fid, self._round_mode_int,
                             self.w_scale_factor)
        if self.backend == "cutlass":   # scales folded in double (Lof_kernel.cu:208-211)
            import numpy as np
            denom = float(self.w_scale_factor) * float(self.act_scale_factor)
            alpha = 1.0 if denom == 0.0 else float(np.float32(1.0 / denom))
        else:                           # lof_gemm's run_scaled, in float
            alpha = _alpha(self.w_scale_factor, self.act_scale_factor)
        kw = dict(stride=self.stride, padding=self.padding, dilation=self.dilation, alpha=alpha)
        if self.groups == 1:
            out = ops.conv2d(x_q.detach(), w_q.detach(), spec, **kw)
        else:
            ci = x_q.shape[1] // self.groups
            co = self.out_channels // self.groups
            out = torch.cat([ops.conv2d(x_q[:, g * ci:(g + 1) * ci].detach(),
                                        w_q[g * co:(g + 1) * co].detach(), spec, **kw)
                             for g in range(self.groups)], dim=1)
        return self._bias(out)

    def _forward_mx(self, x, spec):
        """MX along the unfolded reduction axis (layers.py:663-685): round unfold(x) per
        column in blocks of mx_block_size, and the weight rows likewise, then one GEMM."""
        B, C_in, H_in, W_in = x.shape
        kH, kW = self.kernel_size
        P = ops.conv_out_size(H_in, self.padding[0], self.dilation[0], kH, self.stride[0])
        Q = ops.conv_out_size(W_in, self.padding[1], self.dilation[1], kW, self.stride[1])
        x_col = torch.nn.functional.unfold(x, kernel_size=self.kernel_size,
                                           dilation=self.dilation, padding=self.padding,
                                           stride=self.stride)                 # B x K x L
        x_col = STEMXRound.apply(x_col.transpose(1, 2).contiguous(), self.act_params,
                                 self.scale_format, self.mx_block_size, self.rounding_mode)
        w_col = STEMXRound.apply(self.weight.reshape(self.out_channels, -1), self.weight_params,
                                 self.scale_format, self.mx_block_size, self.rounding_mode)
        K = x_col.shape[2]
        cols = x_col.detach().reshape(B * P * Q, K)                            # (b, l) x K
        out = ops.gemm(w_col.detach(), cols.t(), spec)                         # C_out x B*L
        return out.view(self.out_channels, B, P, Q).permute(1, 0, 2, 3).contiguous()

    def extra_repr(self):
        bits = lambda p: (p.mantissa_bits, p.total_bits - p.mantissa_bits - 1)   # noqa: E731
        am, ae = bits(self.act_params)
        wm, we = bits(self.weight_params)
        s = (f"in_channels={self.in_channels}, out_channels={self.out_channels}, "
             f"kernel_size={self.kernel_size}, stride={self.stride}, padding={self.padding}, "
             f"dilation={self.dilation}, groups={self.groups}, bias={self.bias is not None}, "
             f"act_mantissa={am}, act_exponent={ae}, weight_mantissa={wm}, "
             f"weight_exponent={we}, ")
        if self.bias_params is not None:
            bm, be = bits(self.bias_params)
            s += f"bias_mantissa={bm}, bias_exponent={be}, "
        return s + (f"accum_mant_bits={self.accum_mant_bits}, scaling={self.scaling}, "
                    f"backend={self.backend}")


class _PWLSiLUApprox(torch.autograd.Function):
    """pwl_silu forward; backward with the analytic SiLU derivative
    sigmoid(x) * (1 + x * (1 - sigmoid(x))) as the surrogate (layers.py:1050-1069): the
    piecewise-linear forward's own derivative is the segment slope."""

    @staticmethod
    def forward(ctx, x, lut, R):
        ctx.save_for_backward(x)
        return ops.pwl_silu(x.detach(), lut, float(R))

    @staticmethod
    def backward(ctx, grad_output):
        (x,) = ctx.saved_tensors
        s = torch.sigmoid(x)
        return grad_output * (s * (1 + x * (1 - s))), None, None


def _silu(x):
    return x * torch.sigmoid(x)


# The table's knot values under each lut_method (layers.py:1098-1121), from the knots t_0..t_N
# and the segment midpoints c_0..c_{N-1}.  These formulas define the table the kernel reads.
def _knots_uniform(t, c):
    return _silu(t)


def _knots_midpoint(t, c):
    v = _silu(t)
    v[1:-1] = _silu(c[:-1])   # interior knot j takes SiLU at the midpoint of segment j - 1
    return v


def _knots_minimax(t, c):
    v = _silu(t)
    gap = _silu(c) - 0.5 * (v[:-1] + v[1:])   # SiLU above the chord at each midpoint
    # a knot rises by half the mean gap of the segments that meet at it (one at the ends)
    edge = gap[[0, -1]] / 2
    return v + torch.cat((edge[:1], (gap[:-1] + gap[1:]) / 4, edge[1:]))


_LUT_RULES = {"uniform": _knots_uniform, "midpoint": _knots_midpoint,
              "minimax": _knots_minimax}


class PWLSiLU(nn.Module):
    """SiLU as a piecewise-linear table of 2^lut_bits uniform segments on [-R, R], the
    asymptotes 0 / x outside (layers.py:1072-1137).  The table's knot values follow
    ``lut_method``: "uniform" (SiLU at the knots), "midpoint" (SiLU at the segment
    midpoints for the interior knots) or "minimax" (knot values shifted by the average of
    the adjacent segments' midpoint errors, halving the maximum error of the chord).
    ``use_ext=False`` runs the torch op chain the kernel replaces."""

    _MAX_SEGMENTS = 4096   # the largest table the kernel stages in LDS

    def __init__(self, R=8.0, lut_bits=4, lut_method="minimax", use_ext=True):
        super().__init__()
        self.R = float(R)
        self.lut_bits = lut_bits
        self.lut_size = 1 << lut_bits
        if self.lut_size < 2:
            raise ValueError("PWLSiLU needs an even number of segments (lut_bits >= 1), so "
                             "that x = 0 is a knot")
        self.use_ext = use_ext and self.lut_size <= self._MAX_SEGMENTS
        self.register_buffer("lut", torch.empty(self.lut_size + 1))
        self.init_lut(lut_method)

    def init_lut(self, method="minimax"):
        rule = _LUT_RULES.get(method)
        if rule is None:
            raise ValueError(f"lut_method must be one of {sorted(_LUT_RULES)} (got {method!r})")
        t = torch.linspace(-self.R, self.R, self.lut_size + 1, dtype=self.lut.dtype)
        with torch.no_grad():
            self.lut.copy_(rule(t, 0.5 * (t[:-1] + t[1:])))
        self.lut_method = method

    def forward(self, x):
        if self.use_ext:
            return _PWLSiLUApprox.apply(x, self.lut, self.R)
        return self._forward_ref(x)

    def _forward_ref(self, x):
        """The torch op chain the kernel replaces (layers.py:1131-1150): the clamped input's
        position in knot units, its segment, linear interpolation between the segment's two
        knots, and relu(x) beyond [-R, R]."""
        R, n = self.R, self.lut_size
        pos = (x.clamp(-R, R) + R) * (n / (2 * R))
        seg = pos.long().clamp_(max=n - 1)
        inside = torch.lerp(self.lut[seg], self.lut[seg + 1], pos - seg.to(x.dtype))
        return torch.where(x.abs() > R, torch.relu(x), inside)


def _not_in_scope(name):
    class _Missing(nn.Module):
        def __init__(self, *a, **k):
            raise NotImplementedError(
                f"lofloat_amd.{name}: outside the accelerated hot path; use the reference "
                "LoFloat package for it")
    _Missing.__name__ = name
    return _Missing


ExplicitMultiheadAttention = _not_in_scope("ExplicitMultiheadAttention")
L1BatchNorm = _not_in_scope("L1BatchNorm")
LinfBatchNorm = _not_in_scope("LinfBatchNorm")
FISRBatchNorm = _not_in_scope("FISRBatchNorm")
PWLBatchNorm = _not_in_scope("PWLBatchNorm")


def replace_mha_with_explicit(model: nn.Module) -> nn.Module:
    """The reference swaps every nn.MultiheadAttention for its explicit attention layer
    (layers.py:1499-1510), which this package does not accelerate: a model without one is
    returned unchanged, a model with one raises NotImplementedError."""
    for name, m in model.named_modules():
        if isinstance(m, nn.MultiheadAttention):
            raise NotImplementedError(f"replace_mha_with_explicit: '{name}' is an "
                                      "nn.MultiheadAttention, outside the accelerated hot path")
    return model
