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


class STERound(torch.autograd.Function):
    """round(scale * x) forward; backward multiplies by scale (layers.py:8-31)."""

    @staticmethod
    def forward(ctx, tensor, params, rounding_mode=RoundingMode.RoundToNearestEven, scale=1.0):
        ctx.scale = float(scale)
        if isinstance(params, int):
            return torch.ops.lofloat.virtual_round(tensor, params, int(rounding_mode), 0,
                                                   float(scale))
        return ops.virtual_round(tensor, params, round_mode=rounding_mode, stoch_len=0,
                                 scale=float(scale))

    @staticmethod
    def backward(ctx, grad_output):
        if ctx.scale == 1.0:
            return grad_output, None, None, None
        return grad_output * ctx.scale, None, None, None


class STEMXRound(torch.autograd.Function):
    """MX fake-quantisation along contiguous blocks of the last axis; identity backward
    (layers.py:34-57)."""

    @staticmethod
    def forward(ctx, tensor, element_params, scale_params, block_size,
                rounding_mode=RoundingMode.RoundToNearestEven):
        if isinstance(element_params, int):
            return torch.ops.lofloat.virtual_mx_round(
                tensor, int(block_size), element_params, scale_params, int(rounding_mode),
                int(RoundingMode.RoundToNearestEven), 0)
        return ops.virtual_mx_round(tensor, int(block_size), element_params, scale_params,
                                    round_mode=rounding_mode)

    @staticmethod
    def backward(ctx, grad_output):
        return grad_output, None, None, None, None


class LoF_Quantize(nn.Module):
    """Fake-quantise to a format, per tensor or MX along the last axis (layers.py:60-104)."""

    def __init__(self, params, rounding_mode=None, scaling="per_tensor", mx_block_size=32,
                 scale_format=None):
        super().__init__()
        if scaling not in ("per_tensor", "mx"):
            raise ValueError(f"scaling must be 'per_tensor' or 'mx', got {scaling!r}")
        self.params = params
        self.rounding_mode = (rounding_mode if rounding_mode is not None
                              else RoundingMode.RoundToNearestEven)
        self.scaling = scaling
        self.mx_block_size = mx_block_size
        self.scale_format = scale_format if scale_format is not None else create_e8m0_params()

    def forward(self, x):
        if self.scaling == "mx":
            if x.shape[-1] % self.mx_block_size != 0:
                raise ValueError(
                    f"mx scaling needs the last dim ({x.shape[-1]}) divisible by "
                    f"mx_block_size ({self.mx_block_size})")
            return STEMXRound.apply(x, self.params, self.scale_format, self.mx_block_size,
                                    self.rounding_mode)
        return STERound.apply(x, self.params, self.rounding_mode)

    def extra_repr(self):
        s = f"params={self.params}, rounding_mode={self.rounding_mode}, scaling={self.scaling}"
        if self.scaling == "mx":
            s += f", mx_block_size={self.mx_block_size}"
        return s


def _quantize(self, tensor, params):
    if params is None:
        return tensor
    return STERound.apply(tensor, params, self.rounding_mode)


def mantissa_quantize(tensor, mantissa_bits):
    if mantissa_bits is None:
        return tensor
    return STERound.apply(tensor, create_p3109_params(8 + mantissa_bits + 1, mantissa_bits + 1,
                                                      True, True))


def exp_mant_quantize(tensor, exp_bits, mantissa_bits):
    if exp_bits is None or mantissa_bits is None:
        return tensor
    return STERound.apply(tensor, create_p3109_params(exp_bits + mantissa_bits + 1,
                                                      mantissa_bits + 1, True, True))


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


class _FormatSetters:
    """set_mantissa / set_exponent / set_exponentbias / set_saturation_mode of the reference
    layers (layers.py:756-783): each rebuilds the three P3109 formats from the current ones,
    keeping the other fields."""

    def set_mantissa(self, activ_mant, weight_mant, bias_mant):
        act_exp = self.act_params.total_bits - self.act_params.mantissa_bits - 1
        weight_exp = self.weight_params.total_bits - self.weight_params.mantissa_bits - 1
        bias_exp = self.bias_params.total_bits - self.bias_params.mantissa_bits - 1
        self.act_params = create_p3109_params(1 + act_exp + activ_mant, activ_mant + 1, True, True)
        self.weight_params = create_p3109_params(1 + weight_exp + weight_mant, weight_mant + 1,
                                                 True, True)
        self.bias_params = create_p3109_params(1 + bias_exp + bias_mant, bias_mant + 1, True, True)
        self._refresh_traced_ids()

    def set_exponent(self, activ_exp, weight_exp, bias_exp):
        act_mant = self.act_params.mantissa_bits
        weight_mant = self.weight_params.mantissa_bits
        bias_mant = self.bias_params.mantissa_bits
        self.act_params = create_p3109_params(1 + activ_exp + act_mant, act_mant + 1, True, True)
        self.weight_params = create_p3109_params(1 + weight_exp + weight_mant, weight_mant + 1,
                                                 True, True)
        self.bias_params = create_p3109_params(1 + bias_exp + bias_mant, bias_mant + 1, True, True)
        self._refresh_traced_ids()

    def set_exponentbias(self, activ_expbias, weight_expbias, bias_expbias):
        self.act_params = create_p3109_params(self.act_params.total_bits,
                                              self.act_params.mantissa_bits, True, True,
                                              activ_expbias)
        self.weight_params = create_p3109_params(self.weight_params.total_bits,
                                                 self.weight_params.mantissa_bits, True, True,
                                                 weight_expbias)
        self.bias_params = create_p3109_params(self.bias_params.total_bits,
                                               self.bias_params.mantissa_bits, True, True,
                                               bias_expbias)
        self._refresh_traced_ids()

    def set_saturation_mode(self, activ_sat, weight_sat, bias_sat):
        self.act_params = create_p3109_params(self.act_params.total_bits,
                                              self.act_params.mantissa_bits, True, activ_sat)
        self.weight_params = create_p3109_params(self.weight_params.total_bits,
                                                 self.weight_params.mantissa_bits, True,
                                                 weight_sat)
        self.bias_params = create_p3109_params(self.bias_params.total_bits,
                                               self.bias_params.mantissa_bits, True, bias_sat)
        self._refresh_traced_ids()

    def set_accumulation_precision(self, accum_mant_bits):
        self.accum_mant_bits = accum_mant_bits


class LoF_Linear(_FormatSetters, nn.Module):
    """Drop-in for the reference LoF_Linear (layers.py:221-447) with a differentiable GEMM.

    Extra keyword arguments (all optional): ``accum_format`` (round the accumulator to a
    full format instead of ``accum_mant_bits`` mantissa bits), ``accum_chunk`` (k values
    accumulated in FP32 between roundings; 1 = after every FMA as in the reference) and
    ``grad_spec`` (a LinearGradSpec: format / mode / stoch_len / seed for grad_out)."""

    @staticmethod
    def _resolve_params(exp, mant, params):
        if params is not None:
            return params
        if exp is not None and mant is not None:
            return create_p3109_params(exp + mant + 1, mant)
        raise ValueError("Must provide either (exp, mant) or params for each tensor")

    def __init__(self, in_features, out_features, act_params=None, weight_params=None,
                 bias=True, bias_params=None, act_exp=None, act_mant=None, weight_exp=None,
                 weight_mant=None, bias_exp=None, bias_mant=None, rounding_mode=None,
                 act_scale_factor=1.0, w_scale_factor=1.0, b_scale_factor=1.0,
                 accum_mant_bits=23, gemm_round_mode=None, stochastic_rounding_bits=0,
                 hadamard_transform=False, hadamard_block_size=None, scaling="per_tensor",
                 mx_block_size=32, scale_format=None, *, accum_format=None, accum_chunk=1,
                 grad_spec: LinearGradSpec | None = None):
        super().__init__()
        if scaling not in ("per_tensor", "mx"):
            raise ValueError(f"scaling must be 'per_tensor' or 'mx', got {scaling!r}")
        self.in_features = in_features
        self.out_features = out_features
        self.rounding_mode = (rounding_mode if rounding_mode is not None
                              else RoundingMode.RoundToNearestEven)
        self.act_scale_factor = act_scale_factor
        self.w_scale_factor = w_scale_factor
        self.b_scale_factor = b_scale_factor
        self.scaling = scaling
        self.mx_block_size = mx_block_size
        self.scale_format = scale_format if scale_format is not None else create_e8m0_params()
        if scaling == "mx" and in_features % mx_block_size != 0:
            raise ValueError(f"mx scaling needs in_features ({in_features}) divisible by "
                             f"mx_block_size ({mx_block_size})")
        self.accum_mant_bits = accum_mant_bits
        self.gemm_round_mode = (gemm_round_mode if gemm_round_mode is not None
                                else RoundingMode.RoundToNearestEven)
        self.stochastic_rounding_bits = stochastic_rounding_bits
        self.hadamard_transform = hadamard_transform
        self.hadamard_block_size = hadamard_block_size
        self.accum_format = accum_format
        self.accum_chunk = int(accum_chunk)
        self.grad_spec = grad_spec if grad_spec is not None else LinearGradSpec()
        self.act_params = (act_params if act_params is not None else
                           create_p3109_params(act_exp + act_mant + 1, act_mant + 1, True, True))
        self.weight_params = (weight_params if weight_params is not None else
                              create_p3109_params(weight_exp + weight_mant + 1, weight_mant + 1,
                                                  True, True))
        if bias_params is not None:
            self.bias_params = bias_params
        elif bias_exp is not None and bias_mant is not None:
            self.bias_params = create_p3109_params(bias_exp + bias_mant + 1, bias_mant + 1, True,
                                                   True)
        else:
            self.bias_params = None
        self._refresh_traced_ids()
        self.weight = nn.Parameter(torch.empty(out_features, in_features))
        if bias:
            self.bias = nn.Parameter(torch.zeros(out_features))
        else:
            self.register_parameter("bias", None)
        nn.init.kaiming_uniform_(self.weight)

    def _refresh_traced_ids(self):
        self._act_fid = register_format(self.act_params)
        self._weight_fid = register_format(self.weight_params)
        self._bias_fid = register_format(self.bias_params) if self.bias_params is not None else None
        self._round_mode_int = int(self.rounding_mode)
        self._gemm_round_mode_int = int(self.gemm_round_mode)

    def accum_spec(self) -> ops.AccumSpec:
        if self.accum_format is not None:
            return ops.AccumSpec(fmt=self.accum_format, chunk=self.accum_chunk)
        return ops.AccumSpec(mant_bits=int(self.accum_mant_bits), chunk=self.accum_chunk)

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


class LoF_Conv2d(_FormatSetters, nn.Module):
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

    @staticmethod
    def _resolve_params(exp, mant, params):
        if params is not None:
            return params
        if exp is not None and mant is not None:
            return create_p3109_params(exp + mant + 1, mant + 1)
        raise ValueError("Must provide either (exp, mant) or params for each tensor")

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
        if scaling not in ("per_tensor", "mx"):
            raise ValueError(f"scaling must be 'per_tensor' or 'mx', got {scaling!r}")
        if backend not in ("im2col", "cutlass"):
            raise ValueError(f"backend must be 'im2col' or 'cutlass', got {backend!r}")
        if backend == "cutlass":
            if scaling == "mx":
                raise NotImplementedError("backend='cutlass' does not yet support scaling='mx'")
            if groups != 1:
                raise NotImplementedError("backend='cutlass' currently supports groups=1 only")
        pair = lambda v: v if isinstance(v, tuple) else (v, v)   # noqa: E731
        self.backend = backend
        self.in_channels = in_channels
        self.out_channels = out_channels
        self.kernel_size = pair(kernel_size)
        self.stride = pair(stride)
        self.padding = pair(padding)
        self.dilation = pair(dilation)
        self.groups = groups
        self.rounding_mode = (rounding_mode if rounding_mode is not None
                              else RoundingMode.RoundToNearestEven)
        self.act_scale_factor = act_scale_factor
        self.w_scale_factor = w_scale_factor
        self.b_scale_factor = b_scale_factor
        self.accum_mant_bits = accum_mant_bits
        self.gemm_round_mode = (gemm_round_mode if gemm_round_mode is not None
                                else RoundingMode.RoundToNearestEven)
        self.stochastic_rounding_bits = stochastic_rounding_bits
        self.accum_format = accum_format
        self.accum_chunk = int(accum_chunk)
        self.scaling = scaling
        self.mx_block_size = mx_block_size
        self.scale_format = scale_format if scale_format is not None else create_e8m0_params()
        if scaling == "mx":
            kH, kW = self.kernel_size
            K = (in_channels // groups) * kH * kW
            if groups != 1:
                raise NotImplementedError(
                    "scaling='mx' for LoF_Conv2d currently supports groups=1 only")
            if K % mx_block_size != 0:
                raise ValueError(
                    f"mx scaling needs the im2col K=(C_in/groups)*kH*kW={K} divisible by "
                    f"mx_block_size ({mx_block_size})")
        self.hadamard_transform = hadamard_transform
        self.hadamard_block_size = hadamard_block_size
        if hadamard_transform:
            c_per_group_in = in_channels // groups
            eff = hadamard_block_size if hadamard_block_size is not None else c_per_group_in
            if eff <= 0 or (eff & (eff - 1)) != 0:
                raise ValueError(
                    "hadamard_transform=True requires (hadamard_block_size or "
                    f"in_channels//groups) to be a power of 2, got {eff}")
            if c_per_group_in % eff != 0:
                raise ValueError(f"hadamard_block_size={eff} must divide in_channels//groups="
                                 f"{c_per_group_in}")
        self.act_params = (act_params if act_params is not None else
                           create_p3109_params(act_exp + act_mant + 1, act_mant + 1, True, True))
        self.weight_params = (weight_params if weight_params is not None else
                              create_p3109_params(weight_exp + weight_mant + 1, weight_mant + 1,
                                                  True, True))
        if bias_params is not None:
            self.bias_params = bias_params
        elif bias_exp is not None and bias_mant is not None:
            self.bias_params = create_p3109_params(bias_exp + bias_mant + 1, bias_mant + 1, True,
                                                   True)
        else:
            self.bias_params = None
        self._refresh_traced_ids()
        self.weight = nn.Parameter(torch.empty(out_channels, in_channels // groups,
                                               *self.kernel_size))
        if bias:
            self.bias = nn.Parameter(torch.zeros(out_channels))
        else:
            self.register_parameter("bias", None)
        nn.init.kaiming_uniform_(self.weight)

    @classmethod
    def from_conv2d(cls, conv: nn.Conv2d, act_params, weight_params, bias_params=None,
                    rounding_mode=None, accum_mant_bits=23, gemm_round_mode=None,
                    stochastic_rounding_bits=0, hadamard_transform=False,
                    hadamard_block_size=None):
        layer = cls(conv.in_channels, conv.out_channels, conv.kernel_size, stride=conv.stride,
                    padding=conv.padding, dilation=conv.dilation, groups=conv.groups,
                    act_params=act_params, weight_params=weight_params,
                    bias=conv.bias is not None, bias_params=bias_params,
                    rounding_mode=rounding_mode, accum_mant_bits=accum_mant_bits,
                    gemm_round_mode=gemm_round_mode,
                    stochastic_rounding_bits=stochastic_rounding_bits,
                    hadamard_transform=hadamard_transform,
                    hadamard_block_size=hadamard_block_size)
        layer.weight = nn.Parameter(conv.weight.detach().clone())
        if conv.bias is not None:
            layer.bias = nn.Parameter(conv.bias.detach().clone())
        return layer

    def _refresh_traced_ids(self):
        self._act_fid = register_format(self.act_params)
        self._weight_fid = register_format(self.weight_params)
        self._bias_fid = register_format(self.bias_params) if self.bias_params is not None else None
        self._round_mode_int = int(self.rounding_mode)

    def accum_spec(self) -> ops.AccumSpec:
        if self.accum_format is not None:
            return ops.AccumSpec(fmt=self.accum_format, chunk=self.accum_chunk)
        return ops.AccumSpec(mant_bits=int(self.accum_mant_bits), chunk=self.accum_chunk)

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


class PWLSiLU(nn.Module):
    """SiLU as a piecewise-linear table of 2^lut_bits uniform segments on [-R, R], the
    asymptotes 0 / x outside (layers.py:1072-1137).  The table's knot values follow
    ``lut_method``: "uniform" (SiLU at the knots), "midpoint" (SiLU at the segment
    midpoints for the interior knots) or "minimax" (knot values shifted by the average of
    the adjacent segments' midpoint errors, halving the maximum error of the chord).
    ``use_ext=False`` runs the torch op chain the kernel replaces."""

    def __init__(self, R=8.0, lut_bits=4, lut_method="minimax", use_ext=True):
        super().__init__()
        self.R = float(R)
        self.lut_bits = lut_bits
        self.lut_size = 1 << lut_bits
        if self.lut_size % 2 != 0:
            raise ValueError("lut_size must be even so x=0 is a knot.")
        self.use_ext = use_ext and self.lut_size <= 4096
        self.register_buffer("lut", torch.empty(self.lut_size + 1))
        self.init_lut(lut_method)

    @staticmethod
    def _silu(x):
        return x * torch.sigmoid(x)

    def init_lut(self, method="minimax"):
        N = self.lut_size
        knots = torch.linspace(-self.R, self.R, N + 1, dtype=self.lut.dtype)
        at_knots = self._silu(knots)
        mids = 0.5 * (knots[:-1] + knots[1:])
        if method == "uniform":
            lut = at_knots.clone()
        elif method == "midpoint":
            lut = at_knots.clone()
            lut[1:-1] = self._silu(mids)[:-1]
        elif method == "minimax":
            err = self._silu(mids) - 0.5 * (at_knots[:-1] + at_knots[1:])   # chord error
            shift = torch.empty_like(at_knots)
            shift[0] = err[0] / 2
            shift[1:-1] = (err[:-1] + err[1:]) / 4
            shift[-1] = err[-1] / 2
            lut = at_knots + shift
        else:
            raise ValueError(f"Unknown lut_method {method!r}")
        with torch.no_grad():
            self.lut.copy_(lut)
        self.lut_method = method

    def forward(self, x):
        if self.use_ext:
            return _PWLSiLUApprox.apply(x, self.lut, self.R)
        return self._forward_ref(x)

    def _forward_ref(self, x):
        N, R = self.lut_size, self.R
        u = (x.clamp(-R, R) + R) * (N / (2 * R))
        idx = u.long().clamp_(max=N - 1)
        frac = u - idx.to(x.dtype)
        out = torch.lerp(self.lut[idx], self.lut[idx + 1], frac)
        return torch.where(x.abs() > R, torch.relu(x), out)


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


def replace_mha_with_explicit(*args, **kwargs):
    raise NotImplementedError("replace_mha_with_explicit: outside the accelerated hot path")
