"""lofloat_amd -- MI355X (gfx950) implementation of LoFloat's data-parallel hot path.

Drop-in names of the reference package ``LoFloat`` (LoFloat/__init__.py:1-5):
``virtual_round``, ``virtual_mx_round``, ``lof_gemm``, ``lof_conv2d``, ``pwl_silu``,
``RoundingMode``, ``InfBehavior``,
``NaNBehavior``, ``Signedness``, ``FloatFormatDescriptor``, the ``create_*_params``
factories, the ``torch.ops.lofloat.*`` operators, the layers and the model helpers of
LoFloat/utils.py; ``lofloat_amd.pack`` is
the device bit-pack codec of lo_pack.hpp.  The arithmetic runs in
hand-written HIP kernels in ``liblofloat_hip.so`` (C-ABI: include/lofloat.h); importing
this package fails if that library has not been built.
"""
from . import _lib  # noqa: F401  (loads liblofloat_hip.so or raises)
from .core import (FloatFormatDescriptor, InfBehavior, NaNBehavior, RoundingMode,  # noqa: F401
                   Signedness)
from .formats import (create_bfloat16_params, create_e8m0_params,  # noqa: F401
                      create_half_params, create_ocp_e2m1_params, create_ocp_e3m2_params,
                      create_ocp_e4m3fn_params, create_ocp_e5m2_params, create_p3109_params,
                      create_single_params, make_format)
from .ops import (AccumSpec, conv2d, fill_random_u32, gemm, gemm_autotune,  # noqa: F401
                  gemm_config_name, gemm_config_valid, gemm_num_configs, lof_conv2d,
                  lof_gemm, pwl_silu, set_seed, virtual_mx_round, virtual_round)
from . import _custom_ops  # noqa: F401,E402  (registers torch.ops.lofloat.*)
from ._custom_ops import get_format, register_format  # noqa: F401,E402
from .layers import (ExplicitMultiheadAttention, FISRBatchNorm, L1BatchNorm,  # noqa: F401,E402
                     LinearGradSpec, LinfBatchNorm, LoF_Conv2d, LoF_Linear, LoF_Quantize,
                     LoFLinearFunction, PWLBatchNorm, PWLSiLU, STEMXRound, STERound, _fwht,
                     _quantize, exp_mant_quantize, mantissa_quantize, replace_mha_with_explicit)
from . import utils  # noqa: F401,E402
from .utils import (lofloatify, print_exp_mant, record_formats,  # noqa: F401,E402
                    set_accumulation_precisions, set_all_to_3109, set_all_to_half,
                    set_exponent_fields, set_exponentbias_fields, set_mantissa_fields,
                    set_saturation_modes, sparsity_in_weights)
from . import grouped  # noqa: F401,E402
from . import pack  # noqa: F401,E402
from .grouped import GemmJob, GroupedPlan, MXJob, RoundJob  # noqa: F401,E402

__version__ = "0.1.0"

