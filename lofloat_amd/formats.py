"""Format factories (reference LoFloat/formats.py:55-165).

The rounding never consults the Inf/NaN checkers; they exist so descriptors built here
are accepted wherever the reference's are (same constructor contract, same bit patterns
reported by the checker methods).
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


class HalfPrecisionNaNChecker:
    def __call__(self, bits: int) -> bool:
        return (bits & 0x7C00) == 0x7C00 and (bits & 0x03FF) != 0

    def qNanBitPattern(self) -> int:
        return 0x7E00

    def sNanBitPattern(self) -> int:
        return 0x7D00


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
