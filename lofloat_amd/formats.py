"""Format factories (reference LoFloat/formats.py:55-165) and Inf / NaN checkers.

virtual_round never consults the checkers.  The bit-pack codec (lofloat_amd.pack) does: each
checker class here describes itself to the device as data through ``lof_checker()``
(kind, codes, masks, patterns; see lof_checker in include/lofloat.h).
"""
from __future__ import annotations

from .core import FloatFormatDescriptor, InfBehavior, NaNBehavior, Signedness


class P3109_NaNChecker:
    """NaN encoding of a k-bit P3109 format: 1 << (k-1) when signed, all-ones when
    unsigned (formats.py:2-23)."""

    def __init__(self, k: int, is_signed: bool = True):
        self.k = k
        self.is_signed = is_signed

    def _pattern(self) -> int:
        return (1 << (self.k - 1)) if self.is_signed else ((1 << self.k) - 1)

    def __call__(self, bits: int) -> bool:
        return bits == self._pattern()

    def qNanBitPattern(self) -> int:
        return self._pattern()

    def sNanBitPattern(self) -> int:
        return self._pattern()

    def lof_checker(self):
        return (1, (self._pattern(),), 0, 0, self._pattern(), self._pattern())


class P3109_InfChecker:
    """Infinity encoding of a k-bit P3109 format; a saturating format has none
    (formats.py:26-52)."""

    def __init__(self, k: int, is_signed: bool = True, has_inf_saturating: bool = True):
        self.k = k
        self.is_signed = is_signed
        self.has_inf_saturating = has_inf_saturating

    def __call__(self, bits: int) -> bool:
        if self.has_inf_saturating:
            return False
        top = (1 << self.k) - 1 if self.is_signed else (1 << self.k) - 2
        return (bits | (1 << (self.k - 1))) == top

    def minNegInf(self) -> int:
        return (1 << self.k) - 1 if self.is_signed else 0

    def minPosInf(self) -> int:
        return (1 << (self.k - 1)) - 1 if self.is_signed else (1 << self.k) - 2

    def lof_checker(self):
        if self.has_inf_saturating:
            return (0, (), 0, 0, self.minPosInf(), self.minNegInf())
        top = (1 << self.k) - 1 if self.is_signed else (1 << self.k) - 2
        return (3, (top,), 1 << (self.k - 1), 0, self.minPosInf(), self.minNegInf())


def create_p3109_params(k: int, p: int, is_signed: bool = True, saturating: bool = True,
                        bias: int | None = None) -> FloatFormatDescriptor:
    """P3109 binary<k>p<p>: p-1 mantissa bits, default bias 2^(k-p-1) (formats.py:55-71)."""
    if bias is None:
        bias = 1 << (k - p - 1)
    return FloatFormatDescriptor(
        k, p - 1, bias,
        inf_behavior=InfBehavior.Saturating if saturating else InfBehavior.Extended,
        nan_behavior=NaNBehavior._3109,
        signedness=Signedness.Signed if is_signed else Signedness.Unsigned,
        is_inf_checker=P3109_InfChecker(k, is_signed, saturating),
        is_nan_checker=P3109_NaNChecker(k, is_signed))


def create_e8m0_params() -> FloatFormatDescriptor:
    """E8M0 MX shared-scale format: (8, 0, 127), unsigned, extended (formats.py:74-88)."""
    return FloatFormatDescriptor(
        8, 0, 127, inf_behavior=InfBehavior.Extended, nan_behavior=NaNBehavior._3109,
        signedness=Signedness.Unsigned, is_inf_checker=P3109_InfChecker(8, False, False),
        is_nan_checker=P3109_NaNChecker(8, False))


class HalfPrecisionInfChecker:
    def __call__(self, bits: int) -> bool:
        return (bits & 0x7C00) == 0x7C00 and (bits & 0x03FF) == 0

    def minNegInf(self) -> int:
        return 0xFC00

    def minPosInf(self) -> int:
        return 0x7C00

    def lof_checker(self):
        return (2, (), 0x7C00, 0x03FF, 0x7C00, 0xFC00)


class HalfPrecisionNaNChecker:
    def __call__(self, bits: int) -> bool:
        return (bits & 0x7C00) == 0x7C00 and (bits & 0x03FF) != 0

    def qNanBitPattern(self) -> int:
        return 0x7E00

    def sNanBitPattern(self) -> int:
        return 0x7D00

    def lof_checker(self):
        return (2, (), 0x7C00, 0x03FF, 0x7E00, 0x7D00)


def create_half_params() -> FloatFormatDescriptor:
    """IEEE half as (16, 10, 15), extended (formats.py:117-127)."""
    return FloatFormatDescriptor(
        16, 10, 15, inf_behavior=InfBehavior.Extended, nan_behavior=NaNBehavior._3109,
        signedness=Signedness.Signed, is_inf_checker=HalfPrecisionInfChecker(),
        is_nan_checker=HalfPrecisionNaNChecker())


class SinglePrecisionInfChecker:
    def __call__(self, bits: int) -> bool:
        return (bits & 0x7F800000) == 0x7F800000 and (bits & 0x007FFFFF) == 0

    def minNegInf(self) -> int:
        return 0xFF800000

    def minPosInf(self) -> int:
        return 0x7F800000


class SinglePrecisionNaNChecker:
    def __call__(self, bits: int) -> bool:
        return (bits & 0x7F800000) == 0x7F800000 and (bits & 0x007FFFFF) != 0

    def qNanBitPattern(self) -> int:
        return 0x7FC00000

    def sNanBitPattern(self) -> int:
        return 0x7FA00000


def create_single_params() -> FloatFormatDescriptor:
    """IEEE single as (32, 23, 127), extended (formats.py:155-165)."""
    return FloatFormatDescriptor(
        32, 23, 127, inf_behavior=InfBehavior.Extended, nan_behavior=NaNBehavior._3109,
        signedness=Signedness.Signed, is_inf_checker=SinglePrecisionInfChecker(),
        is_nan_checker=SinglePrecisionNaNChecker())


def make_format(total_bits: int, mantissa_bits: int, bias: int, saturating: bool = True,
                signed: bool = True) -> FloatFormatDescriptor:
    """Convenience: an IEEE-style (total_bits, mantissa_bits, bias) format, e.g.
    make_format(8, 3, 7) for OCP-style e4m3 or make_format(16, 10, 15) for 5e10m."""
    inf_b = InfBehavior.Saturating if saturating else InfBehavior.Extended
    return FloatFormatDescriptor(
        total_bits, mantissa_bits, bias, inf_behavior=inf_b, nan_behavior=NaNBehavior._3109,
        signedness=Signedness.Signed if signed else Signedness.Unsigned,
        is_inf_checker=P3109_InfChecker(total_bits, signed, saturating),
        is_nan_checker=P3109_NaNChecker(total_bits, signed))


# ------------------------------------------------------------------ more checkers
class FiniteNaNChecker:
    """NaN is the single all-ones byte 0xFF and nothing else: the OCP e4m3fn-family
    checker the reference also uses for e5m2, e3m2, e2m1 (lo_float.h:1448-1483, 1521-1666)
    and the E8M0 one (1777-1789)."""

    def __call__(self, bits: int) -> bool:
        return bits == 0xFF

    def qNanBitPattern(self) -> int:
        return 0xFF

    def sNanBitPattern(self) -> int:
        return 0xFF

    def lof_checker(self):
        return (1, (0xFF,), 0, 0, 0xFF, 0xFF)


class NoInfChecker:
    """No infinities (lo_float.h:1466-1483, 1429-1446)."""

    def __call__(self, bits: int) -> bool:
        return False

    def minNegInf(self) -> int:
        return 0

    def minPosInf(self) -> int:
        return 0

    def lof_checker(self):
        return (0, (), 0, 0, 0, 0)


class NoNaNChecker:
    """No NaN encodings (lo_float.h:1411-1427)."""

    def __call__(self, bits: int) -> bool:
        return False

    def qNanBitPattern(self) -> int:
        return 0

    def sNanBitPattern(self) -> int:
        return 0

    def lof_checker(self):
        return (0, (), 0, 0, 0, 0)


class IEEEF8NaNChecker:
    """NaN = 0x80 (signed) or 0xFF (unsigned) (lo_float.h:1485-1501)."""

    def __init__(self, is_signed: bool = True):
        self.code = 0x80 if is_signed else 0xFF

    def __call__(self, bits: int) -> bool:
        return bits == self.code

    def qNanBitPattern(self) -> int:
        return self.code

    def sNanBitPattern(self) -> int:
        return self.code

    def lof_checker(self):
        return (1, (self.code,), 0, 0, self.code, self.code)


class IEEEF8InfChecker:
    """Inf = 0x7F or 0xFF (lo_float.h:1503-1519)."""

    def __call__(self, bits: int) -> bool:
        return bits in (0x7F, 0xFF)

    def minNegInf(self) -> int:
        return 0xFF

    def minPosInf(self) -> int:
        return 0x7F

    def lof_checker(self):
        return (1, (0x7F, 0xFF), 0, 0, 0x7F, 0xFF)


class MaskedInfChecker:
    """IEEE-style: exponent field all ones, mantissa zero (DeviceInfChecker,
    lo_float.h:4386-4391; BF16InfChecker, fp_tools.hpp:372-382)."""

    def __init__(self, exp_mask: int, mant_mask: int, pos_inf: int, neg_inf: int):
        self.exp_mask, self.mant_mask, self.pos, self.neg = exp_mask, mant_mask, pos_inf, neg_inf

    def __call__(self, bits: int) -> bool:
        return (bits & self.exp_mask) == self.exp_mask and (bits & self.mant_mask) == 0

    def minNegInf(self) -> int:
        return self.neg

    def minPosInf(self) -> int:
        return self.pos

    def lof_checker(self):
        return (2, (), self.exp_mask, self.mant_mask, self.pos, self.neg)


class MaskedNaNChecker:
    """IEEE-style: exponent field all ones, mantissa nonzero (DeviceNaNChecker,
    lo_float.h:4393-4398; BF16NaNChecker, fp_tools.hpp:384-395)."""

    def __init__(self, exp_mask: int, mant_mask: int, qnan: int, snan: int):
        self.exp_mask, self.mant_mask, self.q, self.s = exp_mask, mant_mask, qnan, snan

    def __call__(self, bits: int) -> bool:
        return (bits & self.exp_mask) == self.exp_mask and (bits & self.mant_mask) != 0

    def qNanBitPattern(self) -> int:
        return self.q

    def sNanBitPattern(self) -> int:
        return self.s

    def lof_checker(self):
        return (2, (), self.exp_mask, self.mant_mask, self.q, self.s)


def _ocp(total, mant, bias) -> FloatFormatDescriptor:
    return FloatFormatDescriptor(total, mant, bias, inf_behavior=InfBehavior.Saturating,
                                 nan_behavior=NaNBehavior._3109, signedness=Signedness.Signed,
                                 is_inf_checker=NoInfChecker(), is_nan_checker=FiniteNaNChecker())


def create_ocp_e4m3fn_params() -> FloatFormatDescriptor:
    """param_float8_e4m3fn (lo_float.h:1521-1531)."""
    return _ocp(8, 3, 7)


def create_ocp_e5m2_params() -> FloatFormatDescriptor:
    """param_float8_e5m2 (lo_float.h:1578-1588; OCP_F8E4M3 checkers)."""
    return _ocp(8, 2, 15)


def create_ocp_e3m2_params() -> FloatFormatDescriptor:
    """param_float6_e3m2 (lo_float.h:1615-1625)."""
    return _ocp(6, 2, 3)


def create_ocp_e2m1_params() -> FloatFormatDescriptor:
    """param_float4_e2m1 (lo_float.h:1653-1663)."""
    return _ocp(4, 1, 1)


def create_bfloat16_params() -> FloatFormatDescriptor:
    """bfloatPrecisionParams (fp_tools.hpp:397-407)."""
    return FloatFormatDescriptor(
        16, 7, 127, inf_behavior=InfBehavior.Extended, nan_behavior=NaNBehavior._3109,
        signedness=Signedness.Signed, is_inf_checker=MaskedInfChecker(0x7F80, 0x7F, 0x7F80, 0xFF80),
        is_nan_checker=MaskedNaNChecker(0x7F80, 0x7F, 0x7FC0, 0x7F81))
