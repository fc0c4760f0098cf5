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


class LoF_Linear(nn.Module):
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

    def set_accumulation_precision(self, accum_mant_bits):
        self.accum_mant_bits = accum_mant_bits

    def extra_repr(self):
        return (f"in_features={self.in_features}, out_features={self.out_features}, "
                f"bias={self.bias is not None}, scaling={self.scaling}, "
                f"accum_mant_bits={self.accum_mant_bits}, accum_format={self.accum_format}, "
                f"accum_chunk={self.accum_chunk}")


def _not_in_scope(name):
    class _Missing(nn.Module):
        def __init__(self, *a, **k):
            raise NotImplementedError(
                f"lofloat_amd.{name}: outside the accelerated hot path; use the reference "
                "LoFloat package for it")
    _Missing.__name__ = name
    return _Missing


LoF_Conv2d = _not_in_scope("LoF_Conv2d")
ExplicitMultiheadAttention = _not_in_scope("ExplicitMultiheadAttention")
L1BatchNorm = _not_in_scope("L1BatchNorm")
LinfBatchNorm = _not_in_scope("LinfBatchNorm")
FISRBatchNorm = _not_in_scope("FISRBatchNorm")
PWLBatchNorm = _not_in_scope("PWLBatchNorm")
PWLSiLU = _not_in_scope("PWLSiLU")


def replace_mha_with_explicit(*args, **kwargs):
    raise NotImplementedError("replace_mha_with_explicit: outside the accelerated hot path")
