"""``LoFloat``: the reference package's import path (LoFloat/__init__.py:1-5), served by
``lofloat_amd``'s gfx950 kernels.  Code written against the reference -- ``import LoFloat
as lof``, ``from LoFloat._custom_ops import register_format``, ``LoFloat.LoFloat`` (the
compiled extension's module), ``LoFloat.formats``, ``LoFloat.layers``, ``LoFloat.utils``
-- runs unchanged.  The submodules are ``lofloat_amd``'s own modules under these names (the
same objects), so formats registered through one path are seen through the other.
"""
from .LoFloat import (FloatFormatDescriptor, InfBehavior, NaNBehavior, RoundingMode,  # noqa: F401
                      Signedness, lof_conv2d, lof_gemm, pwl_silu, virtual_mx_round,
                      virtual_round)
from . import _custom_ops  # noqa: F401,E402  (torch.ops.lofloat.* are registered)
from .formats import (create_e8m0_params, create_half_params,  # noqa: F401,E402
                      create_p3109_params, create_single_params)
from .layers import (ExplicitMultiheadAttention, FISRBatchNorm, L1BatchNorm,  # noqa: F401,E402
                     LinfBatchNorm, LoF_Conv2d, LoF_Linear, LoF_Quantize, PWLBatchNorm,
                     PWLSiLU, _fwht, _quantize, exp_mant_quantize, mantissa_quantize,
                     replace_mha_with_explicit)
from .utils import (lofloatify, print_exp_mant, record_formats,  # noqa: F401,E402
                    set_accumulation_precisions, set_all_to_3109, set_all_to_half,
                    set_exponent_fields, set_exponentbias_fields, set_mantissa_fields,
                    sparsity_in_weights)
