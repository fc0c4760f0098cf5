"""Tensor-level operations: the reference binding's functions (LoPy_bind.cpp:80-217,
Lof_kernel.cu:96-154) running on hand-written gfx950 kernels through the C-ABI.

All work is launched on torch's current HIP stream of the tensor's device; nothing here
synchronises the host.  Inputs must be float32 tensors on a HIP device: there is no CPU
path (a CPU tensor raises).

Stochastic modes draw one uint32 word per element.  Pass ``rand_bits`` (an int32/uint32
tensor with one word per element) for reproducible, oracle-checkable results; otherwise
the words come from the library's counter-based generator with a per-call seed taken
from a host-side sequence seeded by :func:`set_seed` (the analogue of lo_float::set_seed,
lo_float.h:58-62).
"""
from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .core import FloatFormatDescriptor, RoundingMode

_seed_lock = threading.Lock()
_seed_state = {"seed": 12345, "calls": 0}


def set_seed(seed: int) -> None:
    """Seed the per-call seed sequence used by stochastic modes without rand_bits."""
    with _seed_lock:
        _seed_state["seed"] = int(seed) & 0xFFFFFFFFFFFFFFFF
        _seed_state["calls"] = 0


def next_seed() -> int:
    with _seed_lock:
        base, k = _seed_state["seed"], _seed_state["calls"]
        _seed_state["calls"] = k + 1
    hi = _lib.lib.lof_rand_u32(base, 2 * k)
    lo = _lib.lib.lof_rand_u32(base, 2 * k + 1)
    return (int(hi) << 32) | int(lo)


def _stream(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _require_device_f32(t: torch.Tensor, what: str) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what}: expected a torch.Tensor, got {type(t).__name__}")
    if t.device.type != "cuda":
        raise RuntimeError(f"{what}: lofloat_amd runs on HIP device tensors only "
                           f"(got a {t.device.type} tensor); move it to the GPU first")
    if t.dtype != torch.float32:
        raise RuntimeError(f"{what}: float32 tensors only (got {t.dtype})")


def _require_same_device(t: torch.Tensor | None, ref: torch.Tensor, what: str) -> None:
    """A caller-supplied output or source tensor must live on the input's device: its
    pointer goes straight to a kernel on that device."""
    if t is not None and t.device != ref.device:
        raise ValueError(f"{what} must be on {ref.device} (got {t.device})")


def _bits_ptr(rand_bits, ref: torch.Tensor, what: str):
    if rand_bits is None:
        return None, None
    if not isinstance(rand_bits, torch.Tensor) or rand_bits.dtype not in (
            torch.int32, torch.uint32):
        raise TypeError(f"{what}: rand_bits must be an int32/uint32 tensor")
    if rand_bits.device != ref.device or rand_bits.numel() != ref.numel():
        raise ValueError(f"{what}: rand_bits needs one word per element on {ref.device}")
    rb = rand_bits.contiguous()
    return ctypes.c_void_p(rb.data_ptr()), rb


def _mode(m) -> int:
    return int(RoundingMode(int(m)))


def _seed_for(mode: int, seed) -> int:
    if seed is not None:
        return int(seed) & 0xFFFFFFFFFFFFFFFF
    if RoundingMode(mode) in (RoundingMode.StochasticRoundingA, RoundingMode.StochasticRoundingB,
                              RoundingMode.StochasticRoundingC, RoundingMode.StochasticRoundingD,
                              RoundingMode.True_StochasticRounding,
                              RoundingMode.ProbabilisticRounding):
        return next_seed()
    return 0


def virtual_round(input: torch.Tensor, params, round_mode=RoundingMode.RoundToNearestEven,
                  stoch_len: int = 0, scale: float = 1.0, *, rand_bits=None, seed=None,
                  out: torch.Tensor | None = None) -> torch.Tensor:
    """round(scale * input) into a format (LoPy_bind.cpp:122-157) or, when ``params`` is an
    int, to that many mantissa bits (LoPy_bind.cpp:80-120).  Returns a new tensor (or
    writes ``out``, which may alias ``input``)."""
    _require_device_f32(input, "virtual_round")
    x = input.contiguous()
    _require_same_device(out, x, "virtual_round: out")
    y = torch.empty_like(x) if out is None else out
    if y.shape != x.shape or y.dtype != torch.float32 or not y.is_contiguous():
        raise ValueError("virtual_round: out must be a contiguous float32 tensor of input's shape")
    mode = _mode(round_mode)
    bptr, keep = _bits_ptr(rand_bits, x, "virtual_round")
    spec = _lib.RoundSpec(mode, 0, int(stoch_len))
    s = _seed_for(mode, seed)
    sc = float(np.float32(scale))  # static_cast<float>(scale); the kernel multiplies in FP32
    with torch.cuda.device(x.device):
        if isinstance(params, FloatFormatDescriptor):
            fmt = params.c_format()
            st = _lib.lib.lof_virtual_round_f32(
                ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), x.numel(),
                ctypes.byref(fmt), spec, sc, bptr, s, _stream(x))
        elif isinstance(params, int) and not isinstance(params, bool):
            st = _lib.lib.lof_virtual_round_mant_f32(
                ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), x.numel(),
                int(params), spec, sc, bptr, s, _stream(x))
        else:
            raise TypeError("virtual_round: params must be a FloatFormatDescriptor or an int "
                            "number of mantissa bits")
    _lib.check(st, "virtual_round")
    del keep
    return y


def virtual_mx_round(input: torch.Tensor, block_size: int, element_format: FloatFormatDescriptor,
                     scale_format: FloatFormatDescriptor,
                     round_mode=RoundingMode.RoundToNearestEven,
                     scale_round_mode=RoundingMode.RoundToNearestEven, stoch_len: int = 0, *,
                     rand_bits=None, seed=None, out: torch.Tensor | None = None) -> torch.Tensor:
    """MX block fake-quantisation over contiguous blocks of ``block_size`` elements of the
    flattened tensor (LoPy_bind.cpp:165-217).  Returns a new tensor of input's shape."""
    _require_device_f32(input, "virtual_mx_round")
    if int(block_size) <= 0:
        raise RuntimeError("virtual_mx_round: block_size must be positive")
    x = input.contiguous()
    _require_same_device(out, x, "virtual_mx_round: out")
    y = torch.empty_like(x) if out is None else out
    if y.shape != x.shape or y.dtype != torch.float32 or not y.is_contiguous():
        raise ValueError("virtual_mx_round: out must be a contiguous float32 tensor")
    mode = _mode(round_mode)
    bptr, keep = _bits_ptr(rand_bits, x, "virtual_mx_round")
    ef, sf = element_format.c_format(), scale_format.c_format()
    with torch.cuda.device(x.device):
        st = _lib.lib.lof_virtual_mx_round_f32(
            ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), x.numel(),
            int(block_size), ctypes.byref(ef), ctypes.byref(sf),
            _lib.RoundSpec(mode, 0, int(stoch_len)), _mode(scale_round_mode), bptr,
            _seed_for(mode, seed), _stream(x))
    _lib.check(st, "virtual_mx_round")
    del keep
    return y.view_as(input)


def fill_random_u32(n: int, seed: int, device="cuda", offset: int = 0) -> torch.Tensor:
    """A uint32 tensor of n counter-based random words lof_rand_u32(seed, offset + i)."""
    t = torch.empty(int(n), dtype=torch.uint32, device=device)
    with torch.cuda.device(t.device):
        st = _lib.lib.lof_fill_random_u32(ctypes.c_void_p(t.data_ptr()), int(n),
                                          int(seed) & 0xFFFFFFFFFFFFFFFF, int(offset), _stream(t))
    _lib.check(st, "fill_random_u32")
    return t


# ------------------------------------------------------------------------------------
# emulated-accumulation GEMM
# ------------------------------------------------------------------------------------
@dataclass(frozen=True)
class AccumSpec:
    """How the FP32 accumulator is re-rounded.  ``fmt=None``: mantissa-only to
    ``mant_bits`` (what the reference's lof_gemm does); else full format ``fmt``.
    ``chunk``: k values accumulated in plain FP32 between roundings (1 = every FMA)."""
    mant_bits: int = 23
    fmt: FloatFormatDescriptor | None = None
    chunk: int = 1
    mode: int = int(RoundingMode.RoundToNearestEven)

    def c_accum(self) -> _lib.Accum:
        if self.fmt is None:
            return _lib.Accum(0, int(self.mant_bits), _lib.Format(0, 0, 0, 0, 0, 0),
                              _lib.RoundSpec(int(self.mode), 0, 0), int(self.chunk))
        return _lib.Accum(1, 0, self.fmt.c_format(), _lib.RoundSpec(int(self.mode), 0, 0),
                          int(self.chunk))


def _mat_view(t: torch.Tensor, what: str):
    """(tensor, ld, transposed) for a 2-D float32 tensor that is row- or column-major."""
    if t.dim() != 2:
        raise RuntimeError(f"{what}: must be 2-D")
    r, c = t.shape
    if t.stride(1) == 1 and (t.stride(0) >= max(c, 1) or r <= 1):
        return t, max(t.stride(0), c, 1), False
    if t.stride(0) == 1 and (t.stride(1) >= max(r, 1) or c <= 1):
        return t, max(t.stride(1), r, 1), True
    tc = t.contiguous()
    return tc, max(c, 1), False


def gemm(A: torch.Tensor, B: torch.Tensor, accum: AccumSpec, alpha: float = 1.0,
         out: torch.Tensor | None = None, config: int = -1, C: torch.Tensor | None = None,
         beta: float = 0.0) -> torch.Tensor:
    """D = alpha * A @ B (+ beta * C, as fma(alpha, acc, beta * C)) with the emulated
    accumulator.  A (M x K) and B (K x N) may be row-major or transposed views (no copy is
    made for either layout)."""
    _require_device_f32(A, "lof_gemm")
    _require_device_f32(B, "lof_gemm")
    if A.dim() != 2 or B.dim() != 2:
        raise RuntimeError("LoF_gemm: A and B must be 2-D")
    if A.shape[1] != B.shape[0]:
        raise RuntimeError(f"LoF_gemm: A.cols ({A.shape[1]}) != B.rows ({B.shape[0]})")
    if A.device != B.device:
        raise RuntimeError("LoF_gemm: A and B must be on the same device")
    M, K = A.shape
    N = B.shape[1]
    Av, lda, ta = _mat_view(A, "lof_gemm A")
    Bv, ldb, tb = _mat_view(B, "lof_gemm B")
    _require_same_device(out, A, "lof_gemm: out")
    _require_same_device(C if beta != 0.0 else None, A, "lof_gemm: C")
    D = torch.empty((M, N), dtype=torch.float32, device=A.device) if out is None else out
    if D.shape != (M, N) or not D.is_contiguous() or D.dtype != torch.float32:
        raise ValueError("lof_gemm: out must be a contiguous float32 (M, N) tensor")
    if M == 0 or N == 0:
        return D
    Cc = None
    if beta != 0.0:
        if C is None or C.shape != (M, N):
            raise ValueError("lof_gemm: beta != 0 needs C of shape (M, N)")
        _require_device_f32(C, "lof_gemm C")
        Cc = C.contiguous()
    desc = _lib.GemmDesc(M, N, K, Av.data_ptr(), lda, int(ta), Bv.data_ptr(), ldb, int(tb),
                         D.data_ptr(), max(N, 1), float(alpha), accum.c_accum(),
                         Cc.data_ptr() if Cc is not None else None, max(N, 1), float(beta))
    with torch.cuda.device(A.device):
        st = _lib.lib.lof_gemm_f32(ctypes.byref(desc), int(config), _stream(A))
    _lib.check(st, "lof_gemm")
    return D


def lof_gemm(A: torch.Tensor, B: torch.Tensor, accum_mant_bits: int,
             round_mode=RoundingMode.RoundToNearestEven, stochastic_rounding_bits: int = 0,
             scale_a: float = 1.0, scale_b: float = 1.0) -> torch.Tensor:
    """The reference's lof_gemm (Lof_kernel.cu:96-154): D = (A @ B) / (scale_a * scale_b)
    with the accumulator re-rounded to ``accum_mant_bits`` after every FMA.  As in the
    reference kernel, the accumulator always rounds to nearest even: ``round_mode`` and
    ``stochastic_rounding_bits`` are accepted and ignored (cutlass_gemms.cuh:122)."""
    RoundingMode(int(round_mode))
    # run_scaled (cutlass_gemms.cuh:1777-1787), all in float
    denom = np.float32(scale_a) * np.float32(scale_b)
    alpha = 1.0 if denom == 0 else float(np.float32(1.0) / denom)
    return gemm(A, B, AccumSpec(mant_bits=int(accum_mant_bits), chunk=1), alpha=alpha)


def _pair(v, what: str):
    if isinstance(v, int):
        return int(v), int(v)
    t = tuple(int(e) for e in v)
    if len(t) != 2:
        raise ValueError(f"{what}: expected an int or a pair")
    return t


def conv_out_size(n: int, pad: int, dil: int, k: int, stride: int) -> int:
    return (n + 2 * pad - dil * (k - 1) - 1) // stride + 1


def conv2d(x: torch.Tensor, w: torch.Tensor, accum: AccumSpec, stride=1, padding=0,
           dilation=1, alpha: float = 1.0, layout: str = "NCHW",
           C: torch.Tensor | None = None, beta: float = 0.0,
           out: torch.Tensor | None = None, config: int = -1) -> torch.Tensor:
    """Conv2d forward (cross-correlation, groups = 1) as an implicit GEMM with the emulated
    accumulator: y = alpha * conv(x, w) (+ beta * C).  ``layout="NCHW"``: x N x C x H x W,
    w K x C x R x S, y N x K x P x Q, reduction over (c, r, s) with s fastest -- the
    order of the filter-as-matrix times unfold(x) GEMM; ``"NHWC"``: x N x H x W x C,
    w K x R x S x C, y N x P x Q x K, reduction over (r, s, c).  Each output is the
    ascending-k leaf chain of ``gemm`` over its reduction; the unfolded input is never
    materialised."""
    _require_device_f32(x, "conv2d")
    _require_device_f32(w, "conv2d")
    if x.dim() != 4 or w.dim() != 4:
        raise RuntimeError("conv2d: input and weight must be 4-D")
    if x.device != w.device:
        raise RuntimeError("conv2d: input and weight must be on the same device")
    nhwc = {"NCHW": False, "NHWC": True}.get(layout)
    if nhwc is None:
        raise ValueError(f"conv2d: layout must be 'NCHW' or 'NHWC' (got {layout!r})")
    (sh, sw), (ph, pw), (dh, dw) = (_pair(stride, "stride"), _pair(padding, "padding"),
                                    _pair(dilation, "dilation"))
    if nhwc:
        N, H, W, Cin = x.shape
        K, R, S, Cw = w.shape
    else:
        N, Cin, H, W = x.shape
        K, Cw, R, S = w.shape
    if Cw != Cin:
        raise RuntimeError(f"conv2d: input channels ({Cin}) != weight channels ({Cw})")
    P, Q = conv_out_size(H, ph, dh, R, sh), conv_out_size(W, pw, dw, S, sw)
    if P < 1 or Q < 1:
        raise RuntimeError(f"conv2d: non-positive output spatial size (P={P}, Q={Q})")
    shape = (N, P, Q, K) if nhwc else (N, K, P, Q)
    _require_same_device(out, x, "conv2d: out")
    _require_same_device(C if beta != 0.0 else None, x, "conv2d: C")
    y = torch.empty(shape, dtype=torch.float32, device=x.device) if out is None else out
    if tuple(y.shape) != shape or not y.is_contiguous() or y.dtype != torch.float32:
        raise ValueError(f"conv2d: out must be a contiguous float32 tensor of shape {shape}")
    xc, wc = x.contiguous(), w.contiguous()
    Cc = None
    if beta != 0.0:
        if C is None or tuple(C.shape) != shape:
            raise ValueError(f"conv2d: beta != 0 needs C of shape {shape}")
        _require_device_f32(C, "conv2d C")
        Cc = C.contiguous()
    desc = _lib.Conv2dDesc(N, H, W, Cin, K, R, S, ph, pw, sh, sw, dh, dw, int(nhwc),
                           xc.data_ptr(), wc.data_ptr(), y.data_ptr(), float(alpha),
                           accum.c_accum(), Cc.data_ptr() if Cc is not None else None,
                           float(beta))
    with torch.cuda.device(x.device):
        st = _lib.lib.lof_conv2d_f32(ctypes.byref(desc), int(config), _stream(x))
    _lib.check(st, "conv2d")
    return y


def lof_conv2d(input: torch.Tensor, weight: torch.Tensor, pad_h: int, pad_w: int,
               stride_h: int, stride_w: int, dilation_h: int, dilation_w: int,
               accum_mant_bits: int, round_mode=RoundingMode.RoundToNearestEven,
               stochastic_rounding_bits: int = 0, weight_scale: float = 1.0,
               input_scale: float = 1.0) -> torch.Tensor:
    """The reference's lof_conv2d (Lof_kernel.cu:156-228; LoPy_bind.cpp:374-392): NCHW
    input, OIHW weight, NCHW output divided by (weight_scale * input_scale) (no division
    when that product is 0), accumulator re-rounded to ``accum_mant_bits`` after every
    multiply-add.
    The conv runs the GEMM's leaf, which always rounds to nearest even, so ``round_mode``
    and ``stochastic_rounding_bits`` are accepted and ignored as in lof_gemm."""
    RoundingMode(int(round_mode))
    if input.dim() != 4 or weight.dim() != 4:
        raise RuntimeError("LoF_conv2d: input and weight must be 4-D (NCHW / OIHW)")
    if input.size(1) != weight.size(1):
        raise RuntimeError(f"LoF_conv2d: input C_in ({input.size(1)}) != weight C_in "
                           f"({weight.size(1)})")
    denom = float(weight_scale) * float(input_scale)          # in double, as the binding
    alpha = 1.0 if denom == 0.0 else float(np.float32(1.0 / denom))
    return conv2d(input, weight, AccumSpec(mant_bits=int(accum_mant_bits), chunk=1),
                  stride=(stride_h, stride_w), padding=(pad_h, pad_w),
                  dilation=(dilation_h, dilation_w), alpha=alpha)


_SILU_DTYPES = {torch.float32: 0, torch.float16: 1, torch.float64: 2}


def pwl_silu(input: torch.Tensor, lut: torch.Tensor, R: float) -> torch.Tensor:
    """The reference's pwl_silu (LoPy_bind.cpp:228-277, Lof_kernel.cu:240-258): SiLU by
    linear interpolation of the (N+1)-knot table ``lut`` over [-R, R] (N a power of two in
    [2, 4096]), relu(x) outside, computed in float for float32 / float16 / float64 inputs;
    returns a new tensor of the input's shape and dtype."""
    if not isinstance(input, torch.Tensor) or input.device.type != "cuda":
        raise RuntimeError("pwl_silu: lofloat_amd runs on HIP device tensors only")
    if lut.dim() != 1 or lut.numel() < 2:
        raise RuntimeError("pwl_silu: lut must be 1-D with >= 2 entries")
    if not R > 0.0:
        raise RuntimeError("pwl_silu: R must be positive")
    dt = _SILU_DTYPES.get(input.dtype)
    if dt is None:
        raise RuntimeError(f"pwl_silu: dtype {input.dtype} not supported")
    x = input.contiguous()
    t = lut.to(device=x.device, dtype=torch.float32).contiguous()
    y = torch.empty_like(x)
    with torch.cuda.device(x.device):
        st = _lib.lib.lof_pwl_silu(ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                   x.numel(), dt, ctypes.c_void_p(t.data_ptr()), t.numel() - 1,
                                   float(R), _stream(x))
    _lib.check(st, "pwl_silu")
    return y


def gemm_num_configs() -> int:
    return int(_lib.lib.lof_gemm_num_configs())


def gemm_config_valid(i: int, chunk: int) -> bool:
    """Whether tile config i can run a GEMM with k-chunk ``chunk``."""
    return bool(_lib.lib.lof_gemm_config_valid(int(i), int(chunk)))


def gemm_config_name(i: int) -> str:
    buf = ctypes.create_string_buffer(64)
    _lib.check(_lib.lib.lof_gemm_config_name(int(i), buf, 64), "gemm_config_name")
    return buf.value.decode()


def gemm_autotune(A: torch.Tensor, B: torch.Tensor, accum: AccumSpec, iters: int = 3):
    """Time every GEMM tile config on this problem; returns (best_config, best_ms)."""
    M, K = A.shape
    N = B.shape[1]
    Av, lda, ta = _mat_view(A, "autotune A")
    Bv, ldb, tb = _mat_view(B, "autotune B")
    D = torch.empty((M, N), dtype=torch.float32, device=A.device)
    desc = _lib.GemmDesc(M, N, K, Av.data_ptr(), lda, int(ta), Bv.data_ptr(), ldb, int(tb),
                         D.data_ptr(), N, 1.0, accum.c_accum())
    best, ms = ctypes.c_int32(-1), ctypes.c_float(0.0)
    with torch.cuda.device(A.device):
        st = _lib.lib.lof_gemm_autotune(ctypes.byref(desc), int(iters), _stream(A),
                                        ctypes.byref(best), ctypes.byref(ms))
    _lib.check(st, "gemm_autotune")
    return int(best.value), float(ms.value)
