"""ctypes binding of liblofloat_hip.so (the C-ABI declared in include/lofloat.h).

torch is imported first so that its HIP runtime is the one the library binds to (both
carry the soname libamdhip64.so.7).  There is no fallback: if the library is missing the
import fails with instructions to build it.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (load torch's HIP runtime before ours)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblofloat_hip.so")


class LofError(RuntimeError):
    """A C-ABI call returned an error status."""


class Format(ctypes.Structure):
    _fields_ = [("bitwidth", ctypes.c_int32), ("mantissa_bits", ctypes.c_int32),
                ("bias", ctypes.c_int32), ("inf_behavior", ctypes.c_int32),
                ("nan_behavior", ctypes.c_int32), ("signedness", ctypes.c_int32)]


class RoundSpec(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("saturation", ctypes.c_int32),
                ("stoch_len", ctypes.c_int32)]


class Accum(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("mant_bits", ctypes.c_int32), ("fmt", Format),
                ("spec", RoundSpec), ("chunk", ctypes.c_int32)]


class GemmDesc(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int32), ("N", ctypes.c_int32), ("K", ctypes.c_int32),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_int64), ("trans_a", ctypes.c_int32),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_int64), ("trans_b", ctypes.c_int32),
                ("D", ctypes.c_void_p), ("ldd", ctypes.c_int64), ("alpha", ctypes.c_float),
                ("accum", Accum), ("C", ctypes.c_void_p), ("ldc", ctypes.c_int64),
                ("beta", ctypes.c_float)]


class Conv2dDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("N", "H", "W", "C", "K", "R", "S", "pad_h", "pad_w", "stride_h", "stride_w",
                 "dil_h", "dil_w", "layout")] + [
        ("x", ctypes.c_void_p), ("w", ctypes.c_void_p), ("y", ctypes.c_void_p),
        ("alpha", ctypes.c_float), ("accum", Accum), ("c", ctypes.c_void_p),
        ("beta", ctypes.c_float)]


class GroupItem(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("in_", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("n", ctypes.c_int64), ("fmt", Format), ("mant_bits", ctypes.c_int32),
                ("spec", RoundSpec), ("scale", ctypes.c_float), ("rand_bits", ctypes.c_void_p),
                ("seed", ctypes.c_uint64), ("block_size", ctypes.c_int32),
                ("scale_fmt", Format), ("scale_mode", ctypes.c_int32), ("gemm", GemmDesc)]


class Checker(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("ncodes", ctypes.c_int32),
                ("codes", ctypes.c_uint32 * 3), ("a", ctypes.c_uint32), ("b", ctypes.c_uint32),
                ("pat0", ctypes.c_uint32), ("pat1", ctypes.c_uint32)]


class PackFormat(ctypes.Structure):
    _fields_ = [("fmt", Format), ("unsigned_behavior", ctypes.c_int32),
                ("is_inf", Checker), ("is_nan", Checker)]


# C-ABI symbols and their signatures (include/lofloat.h)
_P = ctypes.POINTER
_vp, _i32, _i64, _u64, _f32 = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64,
                              ctypes.c_float)
SIGNATURES = {
    "lof_last_error": (ctypes.c_char_p, []),
    "lof_version": (_i32, []),
    "lof_rand_u32": (ctypes.c_uint32, [_u64, _u64]),
    "lof_fill_random_u32": (_i32, [_vp, _i64, _u64, _u64, _vp]),
    "lof_virtual_round_f32": (_i32, [_vp, _vp, _i64, _P(Format), RoundSpec, _f32, _vp, _u64, _vp]),
    "lof_virtual_round_mant_f32": (_i32, [_vp, _vp, _i64, _i32, RoundSpec, _f32, _vp, _u64, _vp]),
    "lof_virtual_mx_round_f32": (_i32, [_vp, _vp, _i64, _i32, _P(Format), _P(Format), RoundSpec,
                                        _i32, _vp, _u64, _vp]),
    "lof_gemm_f32": (_i32, [_P(GemmDesc), _i32, _vp]),
    "lof_conv2d_f32": (_i32, [_P(Conv2dDesc), _i32, _vp]),
    "lof_gemm_num_configs": (_i32, []),
    "lof_gemm_config_valid": (_i32, [_i32, _i32]),
    "lof_gemm_config_name": (_i32, [_i32, ctypes.c_char_p, _i32]),
    "lof_gemm_autotune": (_i32, [_P(GemmDesc), _i32, _vp, _P(_i32), _P(_f32)]),
    "lof_grouped_create": (_i32, [_P(GroupItem), _i32, _i32, _P(_vp)]),
    "lof_packed_size": (_i64, [_i32, _i64]),
    "lof_pack_encode_f32": (_i32, [_vp, _vp, _i64, _P(PackFormat), RoundSpec, _i32, _vp, _u64,
                                   _vp]),
    "lof_pack_decode_f32": (_i32, [_vp, _vp, _i64, _P(PackFormat), _i32, _vp]),
    "lof_pack_reps": (_i32, [_vp, _vp, _i64, _i32, _i32, _vp]),
    "lof_unpack_reps": (_i32, [_vp, _vp, _i64, _i32, _i32, _vp]),
    "lof_rep_bytes": (_i32, [_i32]),
    "lof_pwl_silu": (_i32, [_vp, _vp, _i64, _i32, _vp, _i32, ctypes.c_float, _vp]),
    "lof_cast_to_rep_f32": (_i32, [_vp, _vp, _i64, _P(PackFormat), RoundSpec, _vp, _u64, _vp]),
    "lof_cast_from_rep_f32": (_i32, [_vp, _vp, _i64, _P(PackFormat), _vp]),
    "lof_grouped_run": (_i32, [_vp, _vp]),
    "lof_grouped_run_noalloc": (_i32, [_P(GroupItem), _i32, _i32, _vp]),
    "lof_grouped_run_separate": (_i32, [_vp, _vp]),
    "lof_grouped_destroy": (_i32, [_vp]),
    "lof_grouped_num_configs": (_i32, []),
    "lof_grouped_config_name": (_i32, [_i32, ctypes.c_char_p, _i32]),
    "lof_grouped_info": (_i32, [_vp, _P(_i64), _P(_i32)]),
    "lof_grouped_autotune": (_i32, [_P(GroupItem), _i32, _i32, _vp, _P(_i32), _P(_f32)]),
    "lof_graph_capture_grouped": (_i32, [_vp, _i32, _vp, _P(_vp)]),
    "lof_graph_launch": (_i32, [_vp, _vp]),
    "lof_graph_destroy": (_i32, [_vp]),
}


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"lofloat_amd: {LIB_PATH} is missing. Build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` or "
            "`python -m lofloat_amd._build` (needs hipcc).")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib.lof_last_error().decode(errors="replace")
        kinds = {1: ValueError, 2: NotImplementedError}
        exc = kinds.get(status)
        if exc is not None:
            raise exc(f"{what}: {msg}" if what else msg)
        raise LofError(f"{what}: status {status}: {msg}" if what else msg)


def missing_symbols() -> list[str]:
    return [n for n in SIGNATURES if getattr(lib, n, None) is None]
