"""Bit-pack codec on the device: quantise FP32 tensors to n-bit reps (1 <= n <= 16) and pack
them tightly, or unpack and widen them back (reference lo_pack.hpp: LoF_encode / LoF_decode /
LoF_packed_size / LoF_encode_cuda / LoF_decode_cuda, lo_pack.hpp:392-531).

A group of lcm(n, 8) bits (n <= 8) or lcm(n, 16) bits holds bits / n elements; fields are
LSB-first within bytes; ``endian="forward"`` puts element i of a group in slot i,
``"reverse"`` in slot group_count-1-i.  A partial final group is padded with encodings of
+0.0f.  Unlike virtual_round, the codec reads the format's Inf / NaN checkers; the checker
classes of lofloat_amd.formats describe themselves to the device (``lof_checker()``).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .core import FloatFormatDescriptor, RoundingMode
from .ops import _bits_ptr, _mode, _require_device_f32, _seed_for, _stream

__all__ = ["packed_size", "encode", "decode", "pack_reps", "unpack_reps", "pack_format",
           "rep_dtype", "to_reps", "from_reps"]

_ENDIAN = {"forward": 0, "reverse": 1, 0: 0, 1: 1}


def _checker(obj, what: str) -> _lib.Checker:
    desc = getattr(obj, "lof_checker", None)
    if desc is None:
        raise NotImplementedError(
            f"pack codec: the {what} checker {type(obj).__name__} has no device description "
            "(use the checker classes of lofloat_amd.formats)")
    kind, codes, a, b, p0, p1 = desc()
    c = list(codes) + [0] * (3 - len(codes))
    return _lib.Checker(int(kind), len(codes), (ctypes.c_uint32 * 3)(*c), int(a), int(b),
                        int(p0), int(p1))


def pack_format(params: FloatFormatDescriptor, unsigned_behavior: int = 0) -> _lib.PackFormat:
    """The C-ABI lof_pack_format of a descriptor (Unsigned_behavior: 0 NegtoZero,
    1 NegtoNaN)."""
    if not 1 <= params.total_bits <= 16:
        raise NotImplementedError(f"pack codec: bitwidth {params.total_bits} outside [1, 16]")
    return _lib.PackFormat(params.c_format(), int(unsigned_behavior),
                           _checker(params.is_inf_checker, "Inf"),
                           _checker(params.is_nan_checker, "NaN"))


def packed_size(bitwidth: int, count: int) -> int:
    """Bytes of a packed buffer holding ``count`` elements of ``bitwidth`` bits."""
    v = int(_lib.lib.lof_packed_size(int(bitwidth), int(count)))
    if v < 0:
        raise ValueError(f"packed_size: bitwidth {bitwidth} outside [1, 16] or count < 0")
    return v


def encode(input: torch.Tensor, params: FloatFormatDescriptor,
           round_mode=RoundingMode.RoundToNearestEven, stoch_len: int = 0,
           saturation: int = 0, endian="forward", *, unsigned_behavior: int = 0,
           rand_bits=None, seed=None, out: torch.Tensor | None = None) -> torch.Tensor:
    """Round ``input`` (flattened) to ``params`` and pack it; returns a uint8 tensor of
    packed_size(bitwidth, numel) bytes.  ``saturation``: Saturation_Mode (0 OvfInf,
    1 SatFinite, 2 SatPropagate)."""
    _require_device_f32(input, "pack.encode")
    x = input.contiguous().reshape(-1)
    pf = pack_format(params, unsigned_behavior)
    nbytes = packed_size(params.total_bits, x.numel())
    y = torch.empty(nbytes, dtype=torch.uint8, device=x.device) if out is None else out
    if y.dtype != torch.uint8 or y.numel() < nbytes or not y.is_contiguous() or y.device != x.device:
        raise ValueError("pack.encode: out must be a contiguous uint8 device tensor of "
                         f">= {nbytes} bytes")
    mode = _mode(round_mode)
    bptr, keep = _bits_ptr(rand_bits, x, "pack.encode")
    with torch.cuda.device(x.device):
        st = _lib.lib.lof_pack_encode_f32(
            ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()), x.numel(),
            ctypes.byref(pf), _lib.RoundSpec(mode, int(saturation), int(stoch_len)),
            _ENDIAN[endian], bptr, _seed_for(mode, seed), _stream(x))
    _lib.check(st, "pack.encode")
    del keep
    return y


def decode(buf: torch.Tensor, params: FloatFormatDescriptor, count: int, endian="forward", *,
           unsigned_behavior: int = 0, out: torch.Tensor | None = None) -> torch.Tensor:
    """Unpack ``count`` elements of ``params`` from the uint8 tensor ``buf`` and widen them to
    float32 (exactly: every n-bit value is an FP32 value)."""
    if not buf.is_cuda or buf.dtype != torch.uint8 or not buf.is_contiguous():
        raise RuntimeError("pack.decode: buf must be a contiguous uint8 device tensor")
    pf = pack_format(params, unsigned_behavior)
    if buf.numel() < packed_size(params.total_bits, count):
        raise ValueError("pack.decode: buf is shorter than packed_size(bitwidth, count)")
    y = torch.empty(int(count), dtype=torch.float32, device=buf.device) if out is None else out
    if y.device != buf.device:
        raise ValueError(f"pack.decode: out must be on {buf.device} (got {y.device})")
    if y.dtype != torch.float32 or y.numel() != count or not y.is_contiguous():
        raise ValueError("pack.decode: out must be a contiguous float32 tensor of count elements")
    with torch.cuda.device(buf.device):
        st = _lib.lib.lof_pack_decode_f32(ctypes.c_void_p(y.data_ptr()),
                                          ctypes.c_void_p(buf.data_ptr()), int(count),
                                          ctypes.byref(pf), _ENDIAN[endian], _stream(buf))
    _lib.check(st, "pack.decode")
    return y


def pack_reps(reps: torch.Tensor, bitwidth: int, endian="forward") -> torch.Tensor:
    """Pack the low ``bitwidth`` bits of each element of an int32 / uint32 tensor."""
    if not reps.is_cuda or reps.dtype not in (torch.int32, torch.uint32):
        raise RuntimeError("pack.pack_reps: reps must be an int32 / uint32 device tensor")
    r = reps.contiguous().reshape(-1)
    y = torch.empty(packed_size(bitwidth, r.numel()), dtype=torch.uint8, device=r.device)
    with torch.cuda.device(r.device):
        st = _lib.lib.lof_pack_reps(ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(r.data_ptr()),
                                    r.numel(), int(bitwidth), _ENDIAN[endian], _stream(r))
    _lib.check(st, "pack.pack_reps")
    return y


def unpack_reps(buf: torch.Tensor, bitwidth: int, count: int, endian="forward") -> torch.Tensor:
    """The ``count`` n-bit fields of ``buf`` as an int32 tensor."""
    if not buf.is_cuda or buf.dtype != torch.uint8 or not buf.is_contiguous():
        raise RuntimeError("pack.unpack_reps: buf must be a contiguous uint8 device tensor")
    if buf.numel() < packed_size(bitwidth, count):
        raise ValueError("pack.unpack_reps: buf is shorter than packed_size(bitwidth, count)")
    y = torch.empty(int(count), dtype=torch.int32, device=buf.device)
    with torch.cuda.device(buf.device):
        st = _lib.lib.lof_unpack_reps(ctypes.c_void_p(y.data_ptr()),
                                      ctypes.c_void_p(buf.data_ptr()), int(count),
                                      int(bitwidth), _ENDIAN[endian], _stream(buf))
    _lib.check(st, "pack.unpack_reps")
    return y


# ---------------------------------------------------------------------------- typed cast
def rep_dtype(bitwidth: int) -> torch.dtype:
    """Container of one n-bit rep: uint8 for n <= 8, int16 (the uint16 bits) for n <= 16."""
    b = int(_lib.lib.lof_rep_bytes(int(bitwidth)))
    if b < 0:
        raise NotImplementedError(f"typed cast: bitwidth {bitwidth} outside [1, 16]")
    return torch.uint8 if b == 1 else torch.int16


def to_reps(input: torch.Tensor, params: FloatFormatDescriptor,
            round_mode=RoundingMode.RoundToNearestEven, stoch_len: int = 0, saturation: int = 0,
            *, unsigned_behavior: int = 0, rand_bits=None, seed=None,
            out: torch.Tensor | None = None) -> torch.Tensor:
    """The typed cast lo_float::Project<Templated_Float<Fp>>(x, y, n, ps) (lo_float.h:
    4014-4022, with the scalar ConvertImpl semantics of lo_float.h:3057-3364): each element
    rounded to ``params`` and stored as its rep, one per uint8 (n <= 8) or int16 (n <= 16)
    element, same shape as the input.  ``saturation``: 0 OvfInf, 1 SatFinite,
    2 SatPropagate."""
    _require_device_f32(input, "to_reps")
    x = input.contiguous()
    pf = pack_format(params, unsigned_behavior)
    dt = rep_dtype(params.total_bits)
    y = torch.empty(x.shape, dtype=dt, device=x.device) if out is None else out
    if y.dtype != dt or y.shape != x.shape or not y.is_contiguous() or y.device != x.device:
        raise ValueError(f"to_reps: out must be a contiguous {dt} tensor of shape {tuple(x.shape)}")
    mode = _mode(round_mode)
    bptr, keep = _bits_ptr(rand_bits, x, "to_reps")
    with torch.cuda.device(x.device):
        st = _lib.lib.lof_cast_to_rep_f32(
            ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()), x.numel(),
            ctypes.byref(pf), _lib.RoundSpec(mode, int(saturation), int(stoch_len)), bptr,
            _seed_for(mode, seed), _stream(x))
    _lib.check(st, "to_reps")
    del keep
    return y


def from_reps(reps: torch.Tensor, params: FloatFormatDescriptor, *, unsigned_behavior: int = 0,
              out: torch.Tensor | None = None) -> torch.Tensor:
    """The FP32 value of each rep (ConvertImpl<Templated_Float<Fp>, float>::run); ``reps``
    is a uint8 (n <= 8) or int16 (n <= 16) tensor as to_reps returns."""
    dt = rep_dtype(params.total_bits)
    if not reps.is_cuda or reps.dtype != dt:
        raise RuntimeError(f"from_reps: reps must be a {dt} device tensor for a "
                           f"{params.total_bits}-bit format")
    r = reps.contiguous()
    pf = pack_format(params, unsigned_behavior)
    y = torch.empty(r.shape, dtype=torch.float32, device=r.device) if out is None else out
    if y.dtype != torch.float32 or y.shape != r.shape or not y.is_contiguous():
        raise ValueError("from_reps: out must be a contiguous float32 tensor of reps' shape")
    with torch.cuda.device(r.device):
        st = _lib.lib.lof_cast_from_rep_f32(ctypes.c_void_p(y.data_ptr()),
                                            ctypes.c_void_p(r.data_ptr()), r.numel(),
                                            ctypes.byref(pf), _stream(r))
    _lib.check(st, "from_reps")
    return y
