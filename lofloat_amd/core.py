"""Format descriptor and enums of the LoFloat API (pybind module ``LoFloat.LoFloat``).

Mirrors the reference binding (src/LoPy_bind.cpp:284-327): ``RoundingMode``,
``InfBehavior``, ``NaNBehavior``, ``Signedness`` with the reference's integer values
(fp_tools.hpp:13-50), and ``FloatFormatDescriptor`` with read-only fields and the same
checker validation (LoPy_bind.cpp:45-68).
"""
from __future__ import annotations

import enum

from . import _lib


class RoundingMode(enum.IntEnum):
    """Rounding_Mode (fp_tools.hpp:13-28).  All 15 values are named here; the reference's
    Python enum leaves out RoundToNearestOdd and RoundTiesToAway (LoPy_bind.cpp:294-308)
    but accepts them by integer."""
    RoundToNearestEven = 0
    RoundTowardsZero = 1
    RoundAwayFromZero = 2
    StochasticRoundingA = 3
    RoundToNearestOdd = 4
    RoundDown = 5
    RoundUp = 6
    RoundTiesToAway = 7
    StochasticRoundingB = 8
    StochasticRoundingC = 9
    True_StochasticRounding = 10
    ProbabilisticRounding = 11
    RoundToOdd = 12
    RoundTiesTowardsZero = 13
    StochasticRoundingD = 14


class InfBehavior(enum.IntEnum):
    """Inf_Behaviors (fp_tools.hpp:34-36)."""
    Extended = 0
    Saturating = 1


class NaNBehavior(enum.IntEnum):
    """NaN_Behaviors (fp_tools.hpp:44-46)."""
    _3109 = 0
    _754 = 1


class Signedness(enum.IntEnum):
    """Signedness (fp_tools.hpp:30-32)."""
    Signed = 0
    Unsigned = 1


STOCHASTIC_MODES = frozenset({
    RoundingMode.StochasticRoundingA, RoundingMode.StochasticRoundingB,
    RoundingMode.StochasticRoundingC, RoundingMode.StochasticRoundingD,
    RoundingMode.True_StochasticRounding, RoundingMode.ProbabilisticRounding})


class FloatFormatDescriptor:
    """Runtime float format: FloatingPointParamsPy (LoPy_bind.cpp:45-68, 315-327).

    ``is_inf_checker`` must be callable and have ``minNegInf()``/``minPosInf()``;
    ``is_nan_checker`` must be callable and have ``qNanBitPattern()`` or
    ``sNanBitPattern()``.  The checkers are validated and kept but, as in the reference,
    never consulted by the rounding (LoPy_bind.cpp:178-185).
    """

    __slots__ = ("_total_bits", "_mantissa_bits", "_bias", "_inf_behavior", "_nan_behavior",
                 "_signedness", "_is_inf_checker", "_is_nan_checker", "__weakref__")

    def __init__(self, total_bits: int, mantissa_bits: int, bias: int,
                 inf_behavior: InfBehavior, nan_behavior: NaNBehavior, signedness: Signedness,
                 is_inf_checker, is_nan_checker):
        if not callable(is_inf_checker):
            raise ValueError("inf_checker must be callable")
        if not hasattr(is_inf_checker, "minNegInf"):
            raise ValueError("inf_checker must have minNegInf()")
        if not hasattr(is_inf_checker, "minPosInf"):
            raise ValueError("inf_checker must have minPosInf()")
        if not callable(is_nan_checker):
            raise ValueError("nan_checker must be callable")
        if not (hasattr(is_nan_checker, "qNanBitPattern")
                or hasattr(is_nan_checker, "sNanBitPattern")):
            raise ValueError("nan_checker must have qNanBitPattern() or sNanBitPattern()")
        self._total_bits = int(total_bits)
        self._mantissa_bits = int(mantissa_bits)
        self._bias = int(bias)
        self._inf_behavior = InfBehavior(int(inf_behavior))
        self._nan_behavior = NaNBehavior(int(nan_behavior))
        self._signedness = Signedness(int(signedness))
        self._is_inf_checker = is_inf_checker
        self._is_nan_checker = is_nan_checker

    total_bits = property(lambda self: self._total_bits)
    mantissa_bits = property(lambda self: self._mantissa_bits)
    bias = property(lambda self: self._bias)
    inf_behavior = property(lambda self: self._inf_behavior)
    nan_behavior = property(lambda self: self._nan_behavior)
    signedness = property(lambda self: self._signedness)
    is_inf_checker = property(lambda self: self._is_inf_checker)
    is_nan_checker = property(lambda self: self._is_nan_checker)

    def c_format(self) -> _lib.Format:
        return _lib.Format(self._total_bits, self._mantissa_bits, self._bias,
                           int(self._inf_behavior), int(self._nan_behavior),
                           int(self._signedness))

    def as_tuple(self) -> tuple:
        return (self._total_bits, self._mantissa_bits, self._bias, int(self._inf_behavior),
                int(self._nan_behavior), int(self._signedness))

    def __repr__(self) -> str:
        return (f"FloatFormatDescriptor(total_bits={self._total_bits}, "
                f"mantissa_bits={self._mantissa_bits}, bias={self._bias}, "
                f"inf_behavior={self._inf_behavior.name}, nan_behavior={self._nan_behavior.name}, "
                f"signedness={self._signedness.name})")
