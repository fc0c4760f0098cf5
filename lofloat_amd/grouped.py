"""Grouped persistent launch: many rounding / MX / GEMM jobs, each with its own format and
mode, executed by ONE kernel launch that walks a device-side work queue (lof_grouped_*).

    jobs = [RoundJob(x, fmt, mode=...), MXJob(w, 32, e4m3, e8m0), GemmJob(A, B, accum), ...]
    plan = GroupedPlan(jobs)           # validates and uploads the work list (synchronous)
    plan.run()                         # one launch on torch's current stream
    g = plan.capture_graph()           # HIP graph of that launch; g.replay()

``plan.run_separate()`` launches every job with its own per-job kernel (the baseline the
grouped launch is measured against); ``autotune(jobs)`` picks the tile configuration.
Outputs are allocated when the job is created (``job.out``) unless given.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import torch

from . import _lib
from .core import FloatFormatDescriptor, RoundingMode
from .ops import AccumSpec, _mat_view, _require_device_f32, next_seed

ITEM_ROUND, ITEM_ROUND_MANT, ITEM_MX, ITEM_GEMM = 0, 1, 2, 3
_STOCH = {int(m) for m in (RoundingMode.StochasticRoundingA, RoundingMode.StochasticRoundingB,
                           RoundingMode.StochasticRoundingC, RoundingMode.StochasticRoundingD,
                           RoundingMode.True_StochasticRounding,
                           RoundingMode.ProbabilisticRounding)}


def _bits(rand_bits, ref):
    if rand_bits is None:
        return None
    if rand_bits.dtype not in (torch.int32, torch.uint32) or rand_bits.numel() != ref.numel():
        raise ValueError("rand_bits needs one int32/uint32 word per element")
    return rand_bits.contiguous()


@dataclass
class RoundJob:
    """virtual_round(scale * input) into ``params`` (format or int mantissa bits)."""
    input: torch.Tensor
    params: object
    mode: int = 0
    stoch_len: int = 0
    scale: float = 1.0
    rand_bits: torch.Tensor | None = None
    seed: int | None = None
    out: torch.Tensor | None = None

    def item(self) -> _lib.GroupItem:
        _require_device_f32(self.input, "RoundJob")
        if not self.input.is_contiguous():
            raise ValueError("RoundJob: input must be contiguous")
        if self.out is None:
            self.out = torch.empty_like(self.input)
        self.rand_bits = _bits(self.rand_bits, self.input)
        it = _lib.GroupItem()
        mant = not isinstance(self.params, FloatFormatDescriptor)
        it.kind = ITEM_ROUND_MANT if mant else ITEM_ROUND
        it.in_ = self.input.data_ptr()
        it.out = self.out.data_ptr()
        it.n = self.input.numel()
        if mant:
            it.mant_bits = int(self.params)
        else:
            it.fmt = self.params.c_format()
        it.spec = _lib.RoundSpec(int(self.mode), 0, int(self.stoch_len))
        it.scale = float(self.scale)
        it.rand_bits = self.rand_bits.data_ptr() if self.rand_bits is not None else None
        if self.seed is None and int(self.mode) in _STOCH and self.rand_bits is None:
            self.seed = next_seed()
        it.seed = int(self.seed or 0)
        return it


@dataclass
class MXJob:
    """virtual_mx_round(input, block_size, element_format, scale_format, ...)."""
    input: torch.Tensor
    block_size: int
    element_format: FloatFormatDescriptor
    scale_format: FloatFormatDescriptor
    mode: int = 0
    scale_mode: int = 0
    stoch_len: int = 0
    rand_bits: torch.Tensor | None = None
    seed: int | None = None
    out: torch.Tensor | None = None

    def item(self) -> _lib.GroupItem:
        _require_device_f32(self.input, "MXJob")
        if not self.input.is_contiguous():
            raise ValueError("MXJob: input must be contiguous")
        if self.out is None:
            self.out = torch.empty_like(self.input)
        self.rand_bits = _bits(self.rand_bits, self.input)
        it = _lib.GroupItem()
        it.kind = ITEM_MX
        it.in_ = self.input.data_ptr()
        it.out = self.out.data_ptr()
        it.n = self.input.numel()
        it.fmt = self.element_format.c_format()
        it.scale_fmt = self.scale_format.c_format()
        it.block_size = int(self.block_size)
        it.spec = _lib.RoundSpec(int(self.mode), 0, int(self.stoch_len))
        it.scale_mode = int(self.scale_mode)
        it.rand_bits = self.rand_bits.data_ptr() if self.rand_bits is not None else None
        if self.seed is None and int(self.mode) in _STOCH and self.rand_bits is None:
            self.seed = next_seed()
        it.seed = int(self.seed or 0)
        it.scale = 1.0
        return it


@dataclass
class GemmJob:
    """D = alpha * A @ B with the emulated accumulator (A, B may be transposed views)."""
    A: torch.Tensor
    B: torch.Tensor
    accum: AccumSpec = field(default_factory=AccumSpec)
    alpha: float = 1.0
    out: torch.Tensor | None = None

    def item(self) -> _lib.GroupItem:
        _require_device_f32(self.A, "GemmJob")
        _require_device_f32(self.B, "GemmJob")
        M, K = self.A.shape
        K2, N = self.B.shape
        if K != K2:
            raise RuntimeError(f"GemmJob: A.cols ({K}) != B.rows ({K2})")
        self.A, lda, ta = _mat_view(self.A, "GemmJob A")
        self.B, ldb, tb = _mat_view(self.B, "GemmJob B")
        if self.out is None:
            self.out = torch.empty((M, N), dtype=torch.float32, device=self.A.device)
        it = _lib.GroupItem()
        it.kind = ITEM_GEMM
        it.gemm = _lib.GemmDesc(M, N, K, self.A.data_ptr(), lda, int(ta), self.B.data_ptr(), ldb,
                                int(tb), self.out.data_ptr(), max(N, 1), float(self.alpha),
                                self.accum.c_accum())
        return it


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class GraphHandle:
    """An instantiated HIP graph of a plan's launch(es)."""

    def __init__(self, ptr, plan, stream):
        self._ptr = ptr
        self._plan = plan      # keeps the plan (and its tensors) alive
        self._stream = stream

    def replay(self, stream: torch.cuda.Stream | None = None):
        s = stream or torch.cuda.current_stream()
        _lib.check(_lib.lib.lof_graph_launch(self._ptr, ctypes.c_void_p(s.cuda_stream)),
                   "graph replay")

    def __del__(self):
        # at interpreter exit the library handle may already be gone: nothing to free then
        lib = getattr(_lib, "lib", None)
        if getattr(self, "_ptr", None) and lib is not None:
            lib.lof_graph_destroy(self._ptr)
            self._ptr = None


class GroupedPlan:
    def __init__(self, jobs, config: int = -1):
        self.jobs = list(jobs)
        arr = (_lib.GroupItem * max(len(self.jobs), 1))()
        for i, j in enumerate(self.jobs):
            arr[i] = j.item()
        self._items = arr
        ptr = ctypes.c_void_p()
        _lib.check(_lib.lib.lof_grouped_create(arr, len(self.jobs), int(config),
                                               ctypes.byref(ptr)), "lof_grouped_create")
        self._ptr = ptr

    @property
    def outputs(self):
        return [j.out for j in self.jobs]

    def info(self) -> tuple[int, int]:
        tiles, grid = ctypes.c_int64(), ctypes.c_int32()
        _lib.check(_lib.lib.lof_grouped_info(self._ptr, ctypes.byref(tiles), ctypes.byref(grid)),
                   "lof_grouped_info")
        return tiles.value, grid.value

    def run(self):
        _lib.check(_lib.lib.lof_grouped_run(self._ptr, _stream()), "lof_grouped_run")

    def run_separate(self):
        _lib.check(_lib.lib.lof_grouped_run_separate(self._ptr, _stream()),
                   "lof_grouped_run_separate")

    def capture_graph(self, separate: bool = False) -> GraphHandle:
        """Capture the grouped launch (or the per-job launches) into a HIP graph on a side
        stream; replay it with GraphHandle.replay()."""
        torch.cuda.synchronize()
        side = torch.cuda.Stream()
        g = ctypes.c_void_p()
        _lib.check(_lib.lib.lof_graph_capture_grouped(self._ptr, int(bool(separate)),
                                                      ctypes.c_void_p(side.cuda_stream),
                                                      ctypes.byref(g)), "graph capture")
        torch.cuda.synchronize()
        return GraphHandle(g, self, side)

    def __del__(self):
        lib = getattr(_lib, "lib", None)
        if getattr(self, "_ptr", None) and lib is not None:
            lib.lof_grouped_destroy(self._ptr)
            self._ptr = None


def num_configs() -> int:
    return int(_lib.lib.lof_grouped_num_configs())


def config_name(i: int) -> str:
    buf = ctypes.create_string_buffer(64)
    _lib.check(_lib.lib.lof_grouped_config_name(int(i), buf, 64), "grouped config name")
    return buf.value.decode()


def autotune(jobs, iters: int = 3) -> tuple[int, float]:
    """Time each grouped tile configuration on these jobs (outputs are overwritten); the
    winner becomes the default for later plans.  Returns (config, ms)."""
    jobs = list(jobs)
    arr = (_lib.GroupItem * max(len(jobs), 1))()
    for i, j in enumerate(jobs):
        arr[i] = j.item()
    best, ms = ctypes.c_int32(-1), ctypes.c_float(0)
    _lib.check(_lib.lib.lof_grouped_autotune(arr, len(jobs), int(iters), _stream(),
                                             ctypes.byref(best), ctypes.byref(ms)),
               "lof_grouped_autotune")
    return best.value, ms.value


def run_noalloc(jobs, config: int = -1) -> None:
    """Run the jobs as one grouped launch without allocating device memory
    (lof_grouped_run_noalloc): the work list goes into the library's fixed device
    workspace, so at most 256 jobs; outputs must already exist (``job.out``, allocated here
    when missing, as the other entry points do).  Calls that overlap on one GPU share that
    workspace and must be ordered by the caller."""
    jobs = list(jobs)
    arr = (_lib.GroupItem * max(len(jobs), 1))()
    for i, j in enumerate(jobs):
        arr[i] = j.item()
    _lib.check(_lib.lib.lof_grouped_run_noalloc(arr, len(jobs), int(config), _stream()),
               "lof_grouped_run_noalloc")
