"""The reference package's import path: every name LoFloat/__init__.py:1-5 and
LoFloat/layers.py:5-6 import resolves through ``LoFloat`` to this package's
implementation, and the model helpers of LoFloat/utils.py work on CPU-held models (no kernel
runs here: building and converting layers launches nothing)."""
import importlib

import pytest
import torch
import torch.nn as nn

# LoFloat/__init__.py:1-5, by source line
REFERENCE_NAMES = {
    "LoFloat.LoFloat": ["virtual_round", "virtual_mx_round", "RoundingMode", "InfBehavior",
                        "NaNBehavior", "Signedness", "FloatFormatDescriptor", "lof_gemm",
                        "lof_conv2d", "pwl_silu"],
    "LoFloat.formats": ["create_p3109_params", "create_half_params", "create_single_params",
                        "create_e8m0_params"],
    "LoFloat.layers": ["LoF_Quantize", "LoF_Linear", "LoF_Conv2d", "_quantize",
                       "mantissa_quantize", "exp_mant_quantize", "ExplicitMultiheadAttention",
                       "replace_mha_with_explicit", "L1BatchNorm", "LinfBatchNorm",
                       "FISRBatchNorm", "PWLBatchNorm", "PWLSiLU", "_fwht"],
    "LoFloat.utils": ["set_mantissa_fields", "set_exponent_fields", "set_exponentbias_fields",
                      "set_accumulation_precisions", "lofloatify", "set_all_to_half",
                      "set_all_to_3109", "print_exp_mant", "record_formats",
                      "sparsity_in_weights"],
    # LoFloat/layers.py:6 and LoFloat/_custom_ops.py's public functions
    "LoFloat._custom_ops": ["register_format", "get_format", "virtual_round_op",
                            "virtual_mx_round_op", "lof_gemm_op"],
}


@pytest.mark.parametrize("module", sorted(REFERENCE_NAMES))
def test_reference_names_resolve(module):
    import lofloat_amd
    import LoFloat
    mod = importlib.import_module(module)
    for name in REFERENCE_NAMES[module]:
        obj = getattr(mod, name)
        if module != "LoFloat._custom_ops" and not name.startswith("_"):
            assert getattr(LoFloat, name) is obj, name      # re-exported at the top level
        if module == "LoFloat.LoFloat":
            assert getattr(lofloat_amd, name) is obj, name  # the same implementation


def test_submodules_are_the_implementation():
    import LoFloat
    import lofloat_amd
    from LoFloat._custom_ops import register_format
    assert LoFloat.layers is lofloat_amd.layers and LoFloat.utils is lofloat_amd.utils
    assert LoFloat._custom_ops is lofloat_amd._custom_ops
    f = LoFloat.create_half_params()
    assert register_format(f) == lofloat_amd.register_format(f)
    assert hasattr(torch.ops.lofloat, "virtual_round")


def _fields(p):
    return (p.total_bits, p.mantissa_bits, p.bias, int(p.inf_behavior))


def test_format_setters_follow_the_reference():
    """The four setters rebuild P3109 formats as LoFloat/layers.py:402-432 writes them,
    including set_exponentbias / set_saturation_mode passing mantissa_bits as the
    precision (one mantissa bit fewer)."""
    import LoFloat as lof
    e4m3 = lof.create_p3109_params(8, 4)                     # (8, 3, 8), saturating
    lin = lof.LoF_Linear(16, 8, act_params=e4m3, weight_params=e4m3, bias_params=e4m3)
    lin.set_mantissa(2, 5, 1)
    assert _fields(lin.act_params) == (7, 2, 8, 1)           # 4 exponent bits kept
    assert _fields(lin.weight_params) == (10, 5, 8, 1)
    assert _fields(lin.bias_params) == (6, 1, 8, 1)
    lin.set_exponent(5, 4, 3)
    assert _fields(lin.act_params) == (8, 2, 16, 1)
    assert _fields(lin.bias_params) == (5, 1, 4, 1)
    lin.set_exponentbias(10, 3, 2)
    assert _fields(lin.act_params) == (8, 1, 10, 1)
    lin.set_saturation_mode(False, True, False)
    assert _fields(lin.act_params) == (8, 0, 64, 0)          # extended, default bias
    assert lin._act_fid == lof._custom_ops.register_format(lin.act_params)
    lin.set_accumulation_precision(7)
    assert lin.accum_mant_bits == 7 and lin.accum_spec().mant_bits == 7


class _Net(nn.Module):
    def __init__(self):
        super().__init__()
        self.body = nn.Sequential(nn.Conv2d(3, 8, 3), nn.ReLU(), nn.Conv2d(8, 4, 1, bias=False))
        self.head = nn.Linear(16, 5)
        self.keep = nn.Linear(5, 5)


def test_lofloatify_and_helpers():
    import LoFloat as lof
    torch.manual_seed(0)
    net = _Net()
    conv = lof.lofloatify(net, skip_layer_names={"keep"})
    assert isinstance(net.head, nn.Linear) and not isinstance(net.head, lof.LoF_Linear)
    assert isinstance(conv.head, lof.LoF_Linear) and type(conv.keep) is nn.Linear
    assert isinstance(conv.body[0], lof.LoF_Conv2d) and isinstance(conv.body[2], lof.LoF_Conv2d)
    assert conv.body[2].bias is None and conv.body[0].kernel_size == (3, 3)
    assert torch.equal(conv.head.weight, net.head.weight)
    assert conv.head.act_params.mantissa_bits == 23
    lof.set_mantissa_fields(conv, {"head": 4, "body.0": 2}, {"head": 3, "body.0": 2},
                            {"head": 3, "body.0": 2})
    assert conv.head.act_params.mantissa_bits == 4 and conv.body[0].weight_params.mantissa_bits == 2
    assert conv.body[2].act_params.mantissa_bits == 23          # not in the dicts: untouched
    lof.set_exponent_fields(conv, {"head": 5}, {}, {})          # missing weight entry: skipped
    assert conv.head.act_params.total_bits - conv.head.act_params.mantissa_bits - 1 == 8
    lof.set_accumulation_precisions(conv, {"head": 10, "body.0": 0, "body.2": 7})
    assert (conv.head.accum_mant_bits, conv.body[0].accum_mant_bits,
            conv.body[2].accum_mant_bits) == (10, 23, 7)
    shares, shapes = lof.record_formats(conv, batch_size=2)
    assert abs(sum(shares.values()) - 100.0) < 1e-9
    assert (2, 5, 16) in shapes                                  # head: M = batch, N, K
    assert (8, 2, 27) in shapes                                  # conv: C_out, batch * 1 * 1, C*kh*kw
    half = lof.set_all_to_half(_Net())
    assert half.head.act_params.mantissa_bits == 10 and isinstance(half.keep, lof.LoF_Linear)
    p = lof.set_all_to_3109(_Net(), k=8, p=4)
    assert _fields(p.body[0].weight_params) == (8, 3, 8, 1)
    sp = lof.sparsity_in_weights(nn.Linear(4, 4, bias=False).requires_grad_(True))
    assert set(sp) == {"weight"} and 0.0 <= sp["weight"] <= 1.0
    lof.print_exp_mant(conv)
    with pytest.raises(NotImplementedError):
        lof.lofloatify(nn.Sequential(nn.MultiheadAttention(8, 2)))


def test_record_formats_fp32_layers():
    import LoFloat as lof
    bn = nn.BatchNorm2d(6)
    bn.output_size = (4, 4)
    act = nn.SiLU()
    act.output_size, act.num_channels = (4, 4), 6
    pool = nn.MaxPool2d(2)
    pool.output_size, pool.num_channels = (2, 2), 6
    lin = lof.LoF_Linear(8, 4, act_params=lof.create_half_params(),
                         weight_params=lof.create_half_params(),
                         bias_params=lof.create_half_params())
    shares, _ = lof.record_formats(nn.Sequential(bn, act, pool, lin), batch_size=3)
    fp32 = 4 * 6 * 16 + 5 * 6 * 16 + 3 * 6 * 4
    gemm = 2 * 3 * 4 * 8
    assert abs(shares[("fp32",) * 4] - 100.0 * fp32 / (fp32 + gemm)) < 1e-9
    key = ("binary16p10",) * 3 + ("binary32p24",)
    assert abs(shares[key] - 100.0 * gemm / (fp32 + gemm)) < 1e-9
