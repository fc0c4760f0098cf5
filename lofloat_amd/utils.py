"""Model-level helpers of the reference package (LoFloat/utils.py:1-353), over this package's
layers: converting a model's nn.Linear / nn.Conv2d layers to LoF_Linear / LoF_Conv2d,
setting per-layer formats from name-keyed dicts, and reporting formats, FLOP shares and
weight sparsity.

The conversions share one walker (``_convert``) and the per-layer setters one driver
(``_set_per_layer``); a layer name is the dotted path ``nn.Module.named_modules`` gives.
"""
from __future__ import annotations

import copy

import torch
import torch.nn as nn

from .core import RoundingMode
from .formats import create_half_params, create_p3109_params, create_single_params
from .layers import LoF_Conv2d, LoF_Linear, replace_mha_with_explicit

_LOF_LAYERS = (LoF_Linear, LoF_Conv2d)


def _convert(model: nn.Module, make, skip=frozenset(), prefix: str = "") -> nn.Module:
    """Replace, in place and recursively, every nn.Linear / nn.Conv2d child whose dotted
    name is not in ``skip`` by ``make(child, name)``; returns ``model``."""
    for name, child in list(model.named_children()):
        path = f"{prefix}.{name}" if prefix else name
        if isinstance(child, (nn.Linear, nn.Conv2d)) and not isinstance(child, _LOF_LAYERS):
            if path not in skip:
                setattr(model, name, make(child, path))
        else:
            _convert(child, make, skip, path)
    return model


def _same_formats(params, rounding_mode=None):
    """Layer factory giving activation, weight and bias one format."""
    def make(child, path):
        kw = dict(act_params=params, weight_params=params, bias_params=params)
        if rounding_mode is not None:
            kw["rounding_mode"] = rounding_mode
        if isinstance(child, nn.Linear):
            return LoF_Linear.from_linear(child, **kw)
        return LoF_Conv2d.from_conv2d(child, **kw)
    return make


def lofloatify(model: nn.Module, rounding_mode=RoundingMode.RoundToNearestEven, device="cuda",
               debug: bool = False, skip_layer_names: set | None = None,
               _prefix: str = "") -> nn.Module:
    """A copy of ``model`` with every nn.Linear / nn.Conv2d (except the dotted names in
    ``skip_layer_names``) replaced by its LoF layer in FP32 formats under
    ``rounding_mode`` (LoFloat/utils.py:6-98).  ``device`` is accepted for interface
    compatibility; the copy keeps its parameters where they are.  ``debug`` prints each
    converted layer's weight difference statistics."""
    del device
    model = copy.deepcopy(model)
    replace_mha_with_explicit(model)
    make = _same_formats(create_single_params(), rounding_mode)
    if debug:
        inner = make

        def make(child, path):
            layer = inner(child, path)
            d = layer.weight.detach().cpu().float() - child.weight.detach().cpu().float()
            kind = "Linear" if isinstance(child, nn.Linear) else "Conv2d"
            print(f"  orig dtype: {child.weight.dtype}\n  new dtype:  {layer.weight.dtype}")
            print(f"{path} ({kind})\n  diff mean: {d.mean():.10f}\n"
                  f"  diff max:  {d.abs().max():.10f}\n"
                  f"  diff rms:  {d.pow(2).mean().sqrt():.10f}\n"
                  f"  has NaN:   {torch.isnan(d).any()}")
            return layer
    return _convert(model, make, frozenset(skip_layer_names or ()), _prefix)


def set_all_to_half(model: nn.Module) -> nn.Module:
    """Convert (in place) every nn.Linear / nn.Conv2d to its LoF layer in binary16."""
    return _convert(model, _same_formats(create_half_params()))


def set_all_to_3109(model: nn.Module, k: int, p: int) -> nn.Module:
    """Convert (in place) every nn.Linear / nn.Conv2d to its LoF layer in the signed,
    saturating P3109 binary<k>p<p> format."""
    return _convert(model, _same_formats(create_p3109_params(k=k, p=p, is_signed=True,
                                                             saturating=True)))


def _set_per_layer(model, tables, setter: str, required: int, what: str) -> None:
    """For every LoF layer, look its name up in each dict of ``tables`` (activation, weight,
    bias) and call ``layer.<setter>(*values)``; a layer missing from one of the first
    ``required`` dicts is skipped with a message."""
    for name, layer in model.named_modules():
        if not isinstance(layer, _LOF_LAYERS):
            continue
        values = [t.get(name) for t in tables]
        if any(v is None for v in values[:required]):
            print(f"[{what}] skipping '{name}': missing entry in one or more dicts")
            continue
        getattr(layer, setter)(*values)


def set_mantissa_fields(model, activation_mantissa_bits: dict, weight_mantissa_bits: dict,
                        bias_mantissa_bits: dict) -> None:
    _set_per_layer(model, (activation_mantissa_bits, weight_mantissa_bits, bias_mantissa_bits),
                   "set_mantissa", 0, "set_mantissa_fields")


def set_exponent_fields(model, activation_exp_bits: dict, weight_exp_bits: dict,
                        bias_exp_bits: dict) -> None:
    _set_per_layer(model, (activation_exp_bits, weight_exp_bits, bias_exp_bits),
                   "set_exponent", 2, "set_exponent_fields")


def set_exponentbias_fields(model, activation_expbias: dict, weight_expbias: dict,
                            bias_expbias: dict) -> None:
    _set_per_layer(model, (activation_expbias, weight_expbias, bias_expbias),
                   "set_exponentbias", 2, "set_exponentbias_fields")


def set_saturation_modes(model, activation_sat: dict, weight_sat: dict, bias_sat: dict) -> None:
    _set_per_layer(model, (activation_sat, weight_sat, bias_sat), "set_saturation_mode", 2,
                   "set_saturation_modes")


def set_accumulation_precisions(model, accum_dict: dict) -> None:
    """Set each LoF layer's accumulator mantissa bits from ``accum_dict`` (1..23)."""
    for name, layer in model.named_modules():
        if not isinstance(layer, _LOF_LAYERS):
            continue
        bits = accum_dict.get(name)
        if bits is None or not 0 < bits <= 23:
            print(f"[set_accumulation_precisions] skipping '{name}': invalid mantissa bits "
                  f"{bits}")
            continue
        layer.set_accumulation_precision(bits)


def _exp_mant(params):
    return params.mantissa_bits, params.total_bits - params.mantissa_bits - 1


def print_exp_mant(model) -> None:
    for name, layer in model.named_modules():
        if not isinstance(layer, _LOF_LAYERS):
            continue
        print(f"Layer: {name}")
        for role, params in (("Activation", layer.act_params), ("Weight", layer.weight_params),
                             ("Bias", layer.bias_params)):
            if params is not None:
                m, e = _exp_mant(params)
                print(f"  {role}: mantissa bits={m}, exponent bits={e}")


def name_type(total_bits: int, mantissa_bits: int) -> str:
    return f"binary{total_bits}p{mantissa_bits}"


def _hw(module, attr="output_size"):
    return getattr(module, attr) if hasattr(module, attr) else (1, 1)


def _has(module, *attrs):
    return all(hasattr(module, a) for a in attrs)


def _pool_window(module):
    k = module.kernel_size
    return k * k if isinstance(k, int) else k[0] * k[1]


# FLOPs of the non-LoF layers that the reference counts as FP32 work, per output
# (LoFloat/utils.py:266-330): (layer types, flops(module)); spatial sizes and channel
# counts come from attributes a caller attached (output_size, num_channels, ...).
def _elementwise(per_output):
    def f(m):
        if not _has(m, "output_size", "num_channels"):
            return 0
        h, w = m.output_size
        return per_output * m.num_channels * h * w
    return f


def _batchnorm(m):
    h, w = _hw(m)
    return 4 * m.num_features * (h * w if isinstance(m, nn.BatchNorm2d) else 1)


def _pool(m):
    if not _has(m, "output_size", "num_channels"):
        return 0
    h, w = m.output_size
    return (_pool_window(m) - 1) * m.num_channels * h * w


def _adaptive_avg(m):
    if not _has(m, "output_size", "input_size", "num_channels"):
        return 0
    (ih, iw), (oh, ow) = m.input_size, m.output_size
    return ((ih // oh) * (iw // ow) - 1) * m.num_channels * oh * ow


def _upsample(m):
    return _elementwise(7)(m) if getattr(m, "mode", None) == "bilinear" else 0


_FP32_FLOPS = (
    ((nn.BatchNorm2d, nn.BatchNorm1d), _batchnorm),
    ((nn.LeakyReLU,), _elementwise(2)),
    ((nn.SiLU,), _elementwise(5)),
    ((nn.MaxPool2d, nn.AvgPool2d), _pool),
    ((nn.AdaptiveAvgPool2d,), _adaptive_avg),
    ((nn.Upsample,), _upsample),
    ((nn.Sigmoid,), _elementwise(4)),
    ((nn.Softmax,), lambda m: 3 * m.num_elements if hasattr(m, "num_elements") else 0),
)


def _lof_gemm_shape(layer, batch_size):
    if isinstance(layer, LoF_Linear):
        return batch_size, layer.out_features, layer.in_features
    h, w = _hw(layer)
    kh, kw = layer.kernel_size
    return (layer.out_channels, batch_size * h * w,
            (layer.in_channels // layer.groups) * kh * kw)


def record_formats(model, batch_size: int = 1):
    """(share of FLOPs per (activation, weight, bias, accumulator) format in percent,
    {(M, N, K): {format key: count}}) over the LoF layers' GEMMs and the FP32 layers of
    the table above (LoFloat/utils.py:220-339)."""
    flops: dict = {}
    shapes: dict = {}
    for layer in (m for _, m in model.named_modules()):
        if isinstance(layer, _LOF_LAYERS):
            key = tuple(name_type(p.total_bits, p.mantissa_bits) for p in
                        (layer.act_params, layer.weight_params, layer.bias_params)) + (
                name_type(9 + layer.accum_mant_bits, 1 + layer.accum_mant_bits),)
            mnk = _lof_gemm_shape(layer, batch_size)
            flops[key] = flops.get(key, 0) + 2 * mnk[0] * mnk[1] * mnk[2]
            per = shapes.setdefault(mnk, {})
            per[key] = per.get(key, 0) + 1
            continue
        for types, fn in _FP32_FLOPS:
            if isinstance(layer, types):
                f = fn(layer)
                if f > 0:
                    key = ("fp32",) * 4
                    flops[key] = flops.get(key, 0) + f
                break
    total = sum(flops.values())
    return ({k: (v / total) * 100 if total > 0 else 0 for k, v in flops.items()}, shapes)


def sparsity_in_weights(model) -> dict:
    """Fraction of exact zeros in every trainable parameter, by parameter name."""
    return {name: torch.sum(p.data == 0).item() / p.data.numel()
            for name, p in model.named_parameters() if p.requires_grad}
