"""CPU oracle for the LoFloat hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to compare against (or
time beside) the HIP product path in ``lofloat_amd``.  The product never imports it.

It wraps ``liblof_oracle.so`` (built from ``lof_oracle.c`` by ``make -C oracle``), a
plain-C restatement of the reference's scalar semantics; see the file:line citations
in ``lof_oracle.c``.  Formats are plain tuples
``(bitwidth, mantissa_bits, bias, inf_behavior, nan_behavior, signedness)`` with the
reference's enum integer values (fp_tools.hpp:13-50).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liblof_oracle.so")

# Rounding_Mode values (fp_tools.hpp:13-28)
RNE, RTZ, RAFZ, SRA, RNO, RDOWN, RUP, RTA, SRB, SRC, TRUESR, PROB, RTO, RTTZ, SRD = range(15)
DETERMINISTIC_MODES = (RNE, RTZ, RAFZ, RNO, RDOWN, RUP, RTA, RTO, RTTZ)
STOCHASTIC_MODES = (SRA, SRB, SRC, TRUESR, PROB, SRD)


class _Fmt(ctypes.Structure):
    _fields_ = [("bitwidth", ctypes.c_int32), ("mantissa_bits", ctypes.c_int32),
                ("bias", ctypes.c_int32), ("inf_behavior", ctypes.c_int32),
                ("nan_behavior", ctypes.c_int32), ("signedness", ctypes.c_int32)]


class _Spec(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("saturation", ctypes.c_int32),
                ("stoch_len", ctypes.c_int32)]


class _Accum(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("mant_bits", ctypes.c_int32), ("fmt", _Fmt),
                ("spec", _Spec), ("chunk", ctypes.c_int32)]


def build() -> None:
    """Compile liblof_oracle.so (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load() -> ctypes.CDLL:
    if not os.path.exists(_SO):
        build()
    lib = ctypes.CDLL(_SO)
    P = ctypes.POINTER
    f32p, u32p, i64p = P(ctypes.c_float), P(ctypes.c_uint32), P(ctypes.c_int64)
    i64, u64, i32 = ctypes.c_int64, ctypes.c_uint64, ctypes.c_int
    lib.orc_rand_u32.restype = ctypes.c_uint32
    lib.orc_rand_u32.argtypes = [u64, u64]
    lib.orc_round_mantissa_bits.restype = ctypes.c_uint32
    lib.orc_round_mantissa_bits.argtypes = [ctypes.c_uint32, i32, i32, i32, i32, ctypes.c_uint32]
    lib.orc_virtual_round_fmt.restype = ctypes.c_float
    lib.orc_virtual_round_fmt.argtypes = [ctypes.c_float, P(_Fmt), _Spec, ctypes.c_uint32]
    lib.orc_virtual_round_mant.restype = ctypes.c_float
    lib.orc_virtual_round_mant.argtypes = [ctypes.c_float, i32, _Spec, ctypes.c_uint32]
    lib.orc_virtual_round_fmt_array.restype = None
    lib.orc_virtual_round_fmt_array.argtypes = [f32p, f32p, i64, P(_Fmt), _Spec, ctypes.c_float,
                                                u32p, u64, i32]
    lib.orc_virtual_round_mant_array.restype = None
    lib.orc_virtual_round_mant_array.argtypes = [f32p, f32p, i64, i32, _Spec, ctypes.c_float,
                                                 u32p, u64, i32]
    lib.orc_virtual_mx_round.restype = None
    lib.orc_virtual_mx_round.argtypes = [f32p, f32p, i64, i32, P(_Fmt), P(_Fmt), _Spec, i32,
                                         u32p, u64, i32]
    lib.orc_rto_fma.restype = ctypes.c_float
    lib.orc_rto_fma.argtypes = [ctypes.c_float] * 3
    lib.orc_rto_fma_dir.restype = ctypes.c_float
    lib.orc_rto_fma_dir.argtypes = [ctypes.c_float] * 3
    lib.orc_gemm_f32.restype = None
    lib.orc_gemm_f32.argtypes = [i32, i32, i32, f32p, i64, i32, f32p, i64, i32, f32p, i64,
                                 ctypes.c_float, P(_Accum), i64p, i64, i32, f32p, i64,
                                 ctypes.c_float]
    lib.orc_max_threads.restype = i32
    return lib


_lib = _load()


def _fmt(f) -> _Fmt:
    f = tuple(int(v) for v in f)
    if len(f) == 3:  # (bitwidth, mantissa, bias): Saturating, signed
        f = f + (1, 0, 0)
    return _Fmt(*f)


def _f32(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.float32)


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype)) if a is not None else None


def max_threads() -> int:
    return _lib.orc_max_threads()


def rand_u32(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Random words orc_rand_u32(seed, offset + i) for i < n (vectorised in numpy)."""
    idx = np.arange(offset, offset + n, dtype=np.uint64) + np.uint64(1)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return (z >> np.uint64(32)).astype(np.uint32)


def rand_u32_scalar(seed: int, index: int) -> int:
    return int(_lib.orc_rand_u32(seed, index))


def round_mantissa_bits(bits: int, r: int, mode: int, stoch_len: int = 0, pos: bool = True,
                        u: int = 0) -> int:
    return int(_lib.orc_round_mantissa_bits(bits, r, mode, stoch_len, int(pos), u))


def virtual_round_scalar(x: float, fmt, mode: int = RNE, stoch_len: int = 0, u: int = 0) -> float:
    return float(_lib.orc_virtual_round_fmt(float(x), ctypes.byref(_fmt(fmt)),
                                            _Spec(mode, 0, stoch_len), u))


def virtual_round_mant_scalar(x: float, mant_bits: int, mode: int = RNE, stoch_len: int = 0,
                              u: int = 0) -> float:
    return float(_lib.orc_virtual_round_mant(float(x), mant_bits, _Spec(mode, 0, stoch_len), u))


def virtual_round(x, fmt, mode: int = RNE, stoch_len: int = 0, scale: float = 1.0,
                  bits=None, seed: int = 0, nthreads: int = 0) -> np.ndarray:
    """Elementwise virtual_round(scale*x, fmt) (lo_float.h:2844-2939, Lof_kernel.cu:49-56)."""
    x = _f32(x)
    out = np.empty_like(x)
    b = None if bits is None else np.ascontiguousarray(bits, dtype=np.uint32)
    assert b is None or b.size == x.size
    _lib.orc_virtual_round_fmt_array(_ptr(x, ctypes.c_float), _ptr(out, ctypes.c_float), x.size,
                                     ctypes.byref(_fmt(fmt)), _Spec(mode, 0, stoch_len),
                                     float(np.float32(scale)), _ptr(b, ctypes.c_uint32), seed,
                                     nthreads)
    return out


def virtual_round_mant(x, mant_bits: int, mode: int = RNE, stoch_len: int = 0,
                       scale: float = 1.0, bits=None, seed: int = 0,
                       nthreads: int = 0) -> np.ndarray:
    """Elementwise virtual_round(scale*x, mant_bits) (lo_float.h:2716-2756)."""
    x = _f32(x)
    out = np.empty_like(x)
    b = None if bits is None else np.ascontiguousarray(bits, dtype=np.uint32)
    assert b is None or b.size == x.size
    _lib.orc_virtual_round_mant_array(_ptr(x, ctypes.c_float), _ptr(out, ctypes.c_float), x.size,
                                      mant_bits, _Spec(mode, 0, stoch_len),
                                      float(np.float32(scale)), _ptr(b, ctypes.c_uint32), seed,
                                      nthreads)
    return out


def virtual_mx_round(x, block_size: int, elem_fmt, scale_fmt, mode: int = RNE,
                     scale_mode: int = RNE, stoch_len: int = 0, bits=None, seed: int = 0,
                     nthreads: int = 0) -> np.ndarray:
    """MX block rounding over contiguous blocks of the flat tensor (LoPy_bind.cpp:198-214)."""
    x = _f32(x)
    out = np.empty_like(x)
    b = None if bits is None else np.ascontiguousarray(bits, dtype=np.uint32)
    assert b is None or b.size == x.size
    _lib.orc_virtual_mx_round(_ptr(x, ctypes.c_float), _ptr(out, ctypes.c_float), x.size,
                              block_size, ctypes.byref(_fmt(elem_fmt)),
                              ctypes.byref(_fmt(scale_fmt)), _Spec(mode, 0, stoch_len),
                              scale_mode, _ptr(b, ctypes.c_uint32), seed, nthreads)
    return out


def rto_fma(a: float, b: float, c: float) -> float:
    return float(_lib.orc_rto_fma(a, b, c))


def rto_fma_dir(a: float, b: float, c: float) -> float:
    return float(_lib.orc_rto_fma_dir(a, b, c))


def accum_mant(mant_bits: int, chunk: int = 1, mode: int = RNE) -> _Accum:
    return _Accum(0, mant_bits, _Fmt(0, 0, 0, 0, 0, 0), _Spec(mode, 0, 0), chunk)


def accum_fmt(fmt, chunk: int = 1, mode: int = RNE) -> _Accum:
    return _Accum(1, 0, _fmt(fmt), _Spec(mode, 0, 0), chunk)


def gemm(A, B, accum: _Accum, trans_a: bool = False, trans_b: bool = False,
         alpha: float = 1.0, samples=None, nthreads: int = 0, C=None,
         beta: float = 0.0) -> np.ndarray:
    """Emulated-accumulation GEMM oracle.  A is M x K (or K x M if trans_a), B is K x N
    (or N x K if trans_b), C (read when beta != 0) is M x N.  With samples (flat i*N+j
    indices) returns just those outputs."""
    A = _f32(A)
    B = _f32(B)
    Cc = _f32(C) if beta != 0.0 else None
    if trans_a:
        K, M = A.shape
    else:
        M, K = A.shape
    if trans_b:
        N, K2 = B.shape
    else:
        K2, N = B.shape
    assert K == K2, (A.shape, B.shape)
    lda, ldb = A.shape[1], B.shape[1]
    if samples is not None:
        s = np.ascontiguousarray(samples, dtype=np.int64)
        out = np.empty(s.size, dtype=np.float32)
        _lib.orc_gemm_f32(M, N, K, _ptr(A, ctypes.c_float), lda, int(trans_a),
                          _ptr(B, ctypes.c_float), ldb, int(trans_b), _ptr(out, ctypes.c_float),
                          0, float(alpha), ctypes.byref(accum), _ptr(s, ctypes.c_int64), s.size,
                          nthreads, _ptr(Cc, ctypes.c_float), N, float(beta))
        return out
    out = np.empty((M, N), dtype=np.float32)
    _lib.orc_gemm_f32(M, N, K, _ptr(A, ctypes.c_float), lda, int(trans_a),
                      _ptr(B, ctypes.c_float), ldb, int(trans_b), _ptr(out, ctypes.c_float), N,
                      float(alpha), ctypes.byref(accum), None, 0, nthreads,
                      _ptr(Cc, ctypes.c_float), N, float(beta))
    return out
