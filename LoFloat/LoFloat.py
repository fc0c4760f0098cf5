"""``LoFloat.LoFloat``: the names of the reference's compiled extension module
(LoPy_bind.cpp:281-394) -- virtual_round, virtual_mx_round, lof_gemm, lof_conv2d, pwl_silu,
the enums and FloatFormatDescriptor -- bound to ``lofloat_amd``'s implementations, which
call the gfx950 kernels through liblofloat_hip.so."""
from lofloat_amd.core import (FloatFormatDescriptor, InfBehavior, NaNBehavior,  # noqa: F401
                              RoundingMode, Signedness)
from lofloat_amd.ops import (lof_conv2d, lof_gemm, pwl_silu, virtual_mx_round,  # noqa: F401
                             virtual_round)

__all__ = ["virtual_round", "virtual_mx_round", "lof_gemm", "lof_conv2d", "pwl_silu",
           "RoundingMode", "InfBehavior", "NaNBehavior", "Signedness", "FloatFormatDescriptor"]
