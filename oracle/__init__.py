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


class _Checker(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("ncodes", ctypes.c_int32),
                ("codes", ctypes.c_uint32 * 3), ("a", ctypes.c_uint32), ("b", ctypes.c_uint32),
                ("pat0", ctypes.c_uint32), ("pat1", ctypes.c_uint32)]


class _PackFmt(ctypes.Structure):
    _fields_ = [("fmt", _Fmt), ("unsigned_behavior", ctypes.c_int32),
                ("is_inf", _Checker), ("is_nan", _Checker)]


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
    u8p = P(ctypes.c_uint8)
    lib.orc_float_to_rep.restype = ctypes.c_uint32
    lib.orc_float_to_rep.argtypes = [ctypes.c_float, P(_PackFmt), _Spec, ctypes.c_uint32]
    lib.orc_rep_to_float.restype = ctypes.c_float
    lib.orc_rep_to_float.argtypes = [ctypes.c_uint32, P(_PackFmt)]
    lib.orc_packed_bytes.restype = i64
    lib.orc_packed_bytes.argtypes = [i32, i64]
    lib.orc_pack_encode.restype = i64
    lib.orc_pack_encode.argtypes = [u8p, f32p, i64, P(_PackFmt), _Spec, i32, u32p, u64]
    lib.orc_pack_decode.restype = i64
    lib.orc_pack_decode.argtypes = [f32p, u8p, i64, P(_PackFmt), i32]
    lib.orc_pack_reps.restype = i64
    lib.orc_pack_reps.argtypes = [u8p, u32p, i64, i32, i32]
    lib.orc_unpack_reps.restype = i64
    lib.orc_unpack_reps.argtypes = [u32p, u8p, i64, i32, i32]
    lib.orc_float_to_rep_array.restype = None
    lib.orc_float_to_rep_array.argtypes = [f32p, u32p, i64, P(_PackFmt), _Spec, u32p, u64]
    lib.orc_rep_to_float_array.restype = None
    lib.orc_rep_to_float_array.argtypes = [u32p, f32p, i64, P(_PackFmt)]
    lib.orc_max_threads.restype = i32
    lib.orc_pwl_silu_array.restype = None
    lib.orc_pwl_silu_array.argtypes = [f32p, f32p, i64, f32p, i32, ctypes.c_float]
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


# ------------------------------------------------------------------ piecewise-linear SiLU
def pwl_silu(x, lut, R: float) -> np.ndarray:
    """pwl_silu_kernel (Lof_kernel.cu:240-258) elementwise on float32 x with an
    (N+1)-entry float32 table over [-R, R]."""
    x = _f32(x)
    t = np.ascontiguousarray(lut, dtype=np.float32)
    out = np.empty_like(x)
    _lib.orc_pwl_silu_array(_ptr(x, ctypes.c_float), _ptr(out, ctypes.c_float), x.size,
                            _ptr(t, ctypes.c_float), t.size - 1, float(np.float32(R)))
    return out


# ------------------------------------------------------------------ convolution
def conv_out_size(n: int, pad: int, dil: int, k: int, stride: int) -> int:
    return (n + 2 * pad - dil * (k - 1) - 1) // stride + 1


def im2col(x, R: int, S: int, pad=(0, 0), stride=(1, 1), dil=(1, 1)) -> np.ndarray:
    """unfold of an N x C x H x W array (torch.nn.functional.unfold's order): a C*R*S x N*P*Q
    matrix, row (c, r, s) with s fastest, column (n, p, q) with q fastest, zero in the
    padding."""
    x = _f32(x)
    N, C, H, W = x.shape
    P = conv_out_size(H, pad[0], dil[0], R, stride[0])
    Q = conv_out_size(W, pad[1], dil[1], S, stride[1])
    xp = np.pad(x, ((0, 0), (0, 0), (pad[0], pad[0]), (pad[1], pad[1])))
    cols = np.empty((C, R, S, N, P, Q), np.float32)
    for r in range(R):
        for s in range(S):
            h0, w0 = r * dil[0], s * dil[1]
            win = xp[:, :, h0:h0 + stride[0] * (P - 1) + 1:stride[0],
                     w0:w0 + stride[1] * (Q - 1) + 1:stride[1]]
            cols[:, r, s] = win.transpose(1, 0, 2, 3)
    return cols.reshape(C * R * S, N * P * Q)


def conv2d(x, w, accum: _Accum, pad=(0, 0), stride=(1, 1), dil=(1, 1), alpha: float = 1.0,
           layout: int = 0, C=None, beta: float = 0.0, nthreads: int = 0) -> np.ndarray:
    """Conv2d forward as LoF_Conv2d's im2col + lof_gemm path (layers.py:687-737): the
    filter as a K x (reduction) matrix times the unfolded input, through the GEMM oracle.
    layout 0: x NCHW, w KCRS (OIHW), result N x K x P x Q, reduction (c, r, s);
    layout 1: x NHWC, w KRSC, result N x P x Q x K, reduction (r, s, c)."""
    x = _f32(x)
    w = _f32(w)
    if layout == 0:
        N = x.shape[0]
        K, Cin, R, S = w.shape
        cols = im2col(x, R, S, pad, stride, dil)
        A = w.reshape(K, Cin * R * S)
    else:
        N = x.shape[0]
        K, R, S, Cin = w.shape
        cols = im2col(x.transpose(0, 3, 1, 2), R, S, pad, stride, dil)
        cols = np.ascontiguousarray(cols.reshape(Cin, R, S, -1).transpose(1, 2, 0, 3)
                                    .reshape(R * S * Cin, -1))
        A = w.reshape(K, R * S * Cin)
    P = conv_out_size(x.shape[2] if layout == 0 else x.shape[1], pad[0], dil[0], R, stride[0])
    Q = conv_out_size(x.shape[3] if layout == 0 else x.shape[2], pad[1], dil[1], S, stride[1])
    Cm = None
    if beta != 0.0:
        Cc = _f32(C)
        Cm = (Cc.reshape(N, K, P * Q).transpose(1, 0, 2).reshape(K, N * P * Q) if layout == 0
              else Cc.reshape(N * P * Q, K).T)
    D = gemm(A, cols, accum, alpha=alpha, nthreads=nthreads, C=Cm, beta=beta)   # K x N*P*Q
    if layout == 0:
        return np.ascontiguousarray(D.reshape(K, N, P, Q).transpose(1, 0, 2, 3))
    return np.ascontiguousarray(D.T.reshape(N, P, Q, K))


# ------------------------------------------------------------------ bit-pack codec
# Checker presets: the Inf / NaN functors of the reference's format parameter sets, as data
# (kind, codes, a, b, pat0, pat1); see orc_checker in lof_oracle.h.
def _chk(kind=0, codes=(), a=0, b=0, pat0=0, pat1=0) -> _Checker:
    c = list(codes) + [0] * (3 - len(codes))
    return _Checker(kind, len(codes), (ctypes.c_uint32 * 3)(*c), a, b, pat0, pat1)


def checkers(preset, fmt=None):
    """(is_inf, is_nan) checker pair for a preset name or tuple:
    'ocp_fn' (OCP_F8E4M3_*: NaN == 0xFF, no Inf; lo_float.h:1448-1483), ('p3109', k, signed,
    has_inf) (lo_float.h:1678-1759), ('ieee_f8', signed) (1485-1519), 'fp16' / 'bf16'
    (fp_tools.hpp:331-396), 'e8m0' (lo_float.h:1777-1789), 'trivial' (1411-1446), 'device'
    (DeviceInf/NaNChecker, lo_float.h:4386-4398, masks from fmt; patterns default 0)."""
    if preset == "ocp_fn":
        return _chk(), _chk(1, (0xFF,), pat0=0xFF, pat1=0xFF)
    if preset == "e8m0":
        return _chk(), _chk(1, (0xFF,), pat0=0xFF, pat1=0xFF)
    if preset == "trivial":
        return _chk(), _chk()
    if preset == "fp16":
        return (_chk(2, (), 0x7C00, 0x3FF, 0x7C00, 0xFC00),
                _chk(2, (), 0x7C00, 0x3FF, 0x7E00, 0x7C01))
    if preset == "bf16":
        return (_chk(2, (), 0x7F80, 0x7F, 0x7F80, 0xFF80),
                _chk(2, (), 0x7F80, 0x7F, 0x7FC0, 0x7F81))
    if preset == "device":
        n, mb = fmt[0], fmt[1]
        em, mm = ((1 << (n - mb - 1)) - 1) << mb, (1 << mb) - 1
        return _chk(2, (), em, mm), _chk(2, (), em, mm)
    kind = preset[0]
    if kind == "p3109":
        _, k, signed, has_inf = preset
        nan = (1 << (k - 1)) if signed else (1 << k) - 1
        pos = (1 << (k - 1)) - 1 if signed else (1 << k) - 2
        neg = (1 << k) - 1 if signed else 0
        target = (1 << k) - 1 if signed else (1 << k) - 2
        inf = _chk(3, (target,), 1 << (k - 1), 0, pos, neg) if has_inf else _chk(0, (), 0, 0, pos, neg)
        return inf, _chk(1, (nan,), pat0=nan, pat1=nan)
    if kind == "ieee_f8":
        nan = 0x80 if preset[1] else 0xFF
        return _chk(1, (0x7F, 0xFF), pat0=0x7F, pat1=0xFF), _chk(1, (nan,), pat0=nan, pat1=nan)
    raise ValueError(f"unknown checker preset {preset!r}")


def pack_format(fmt, preset, unsigned_behavior: int = 0) -> _PackFmt:
    inf, nan = checkers(preset, fmt)
    return _PackFmt(_fmt(fmt), int(unsigned_behavior), inf, nan)


def packed_bytes(n: int, count: int) -> int:
    return int(_lib.orc_packed_bytes(n, count))


def float_to_rep(x, pf: _PackFmt, mode: int = RNE, saturation: int = 0, stoch_len: int = 0,
                 bits=None, seed: int = 0) -> np.ndarray:
    x = _f32(x)
    out = np.empty(x.size, np.uint32)
    b = None if bits is None else np.ascontiguousarray(bits, dtype=np.uint32)
    _lib.orc_float_to_rep_array(_ptr(x, ctypes.c_float), _ptr(out, ctypes.c_uint32), x.size,
                                ctypes.byref(pf), _Spec(mode, saturation, stoch_len),
                                _ptr(b, ctypes.c_uint32), seed)
    return out


def rep_to_float(reps, pf: _PackFmt) -> np.ndarray:
    r = np.ascontiguousarray(reps, dtype=np.uint32)
    out = np.empty(r.size, np.float32)
    _lib.orc_rep_to_float_array(_ptr(r, ctypes.c_uint32), _ptr(out, ctypes.c_float), r.size,
                                ctypes.byref(pf))
    return out


def pack_encode(x, pf: _PackFmt, mode: int = RNE, saturation: int = 0, stoch_len: int = 0,
                endian: int = 0, bits=None, seed: int = 0) -> np.ndarray:
    x = _f32(x)
    out = np.zeros(packed_bytes(pf.fmt.bitwidth, x.size), np.uint8)
    b = None if bits is None else np.ascontiguousarray(bits, dtype=np.uint32)
    _lib.orc_pack_encode(_ptr(out, ctypes.c_uint8), _ptr(x, ctypes.c_float), x.size,
                         ctypes.byref(pf), _Spec(mode, saturation, stoch_len), endian,
                         _ptr(b, ctypes.c_uint32), seed)
    return out


def pack_decode(buf, count: int, pf: _PackFmt, endian: int = 0) -> np.ndarray:
    b = np.ascontiguousarray(buf, dtype=np.uint8)
    out = np.empty(count, np.float32)
    _lib.orc_pack_decode(_ptr(out, ctypes.c_float), _ptr(b, ctypes.c_uint8), count,
                         ctypes.byref(pf), endian)
    return out


def pack_reps(reps, n: int, endian: int = 0) -> np.ndarray:
    r = np.ascontiguousarray(reps, dtype=np.uint32)
    out = np.zeros(packed_bytes(n, r.size), np.uint8)
    _lib.orc_pack_reps(_ptr(out, ctypes.c_uint8), _ptr(r, ctypes.c_uint32), r.size, n, endian)
    return out


def unpack_reps(buf, count: int, n: int, endian: int = 0) -> np.ndarray:
    b = np.ascontiguousarray(buf, dtype=np.uint8)
    out = np.empty(count, np.uint32)
    _lib.orc_unpack_reps(_ptr(out, ctypes.c_uint32), _ptr(b, ctypes.c_uint8), count, n, endian)
    return out
