"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every symbol
include/lofloat.h declares, and rejects bad arguments before touching a device."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(REPO, "include", "lofloat.h")).read()
    return sorted(set(re.findall(r"LOF_API\s+[\w\s\*]+?\b(lof_\w+)\s*\(", src)))


def test_header_symbols_exported():
    from lofloat_amd import _lib
    names = _declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if getattr(_lib.lib, n, None) is None]
    assert not missing, f"declared in include/lofloat.h but not exported: {missing}"
    # and the Python binding knows every declared symbol's signature
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_version_and_rand():
    from lofloat_amd import _lib
    import oracle
    assert _lib.lib.lof_version() >= 1
    for seed, idx in [(0, 0), (12345, 7), (2**63 + 5, 2**40)]:
        assert _lib.lib.lof_rand_u32(seed, idx) == oracle.rand_u32_scalar(seed, idx)


def test_validation_without_device():
    """Argument errors are reported before any device work (n = 0 launches nothing)."""
    from lofloat_amd import _lib
    L = _lib.lib
    bad = _lib.Format(30, 24, 7, 1, 0, 0)
    st = L.lof_virtual_round_f32(None, None, 0, ctypes.byref(bad), _lib.RoundSpec(0, 0, 0),
                                 1.0, None, 0, None)
    assert st == 2 and b"mantissa_bits" in L.lof_last_error()
    ok = _lib.Format(8, 3, 7, 1, 0, 0)
    st = L.lof_virtual_round_f32(None, None, 0, ctypes.byref(ok), _lib.RoundSpec(3, 0, 21),
                                 1.0, None, 0, None)
    assert st == 1 and b"stoch_len" in L.lof_last_error()
    st = L.lof_virtual_round_f32(None, None, 0, ctypes.byref(ok), _lib.RoundSpec(0, 0, 0),
                                 1.0, None, 0, None)
    assert st == 0
    st = L.lof_virtual_round_f32(None, None, 5, ctypes.byref(ok), _lib.RoundSpec(0, 0, 0),
                                 1.0, None, 0, None)
    assert st == 1  # NULL data with n > 0
    st = L.lof_virtual_mx_round_f32(None, None, 0, 48, ctypes.byref(ok), ctypes.byref(ok),
                                    _lib.RoundSpec(0, 0, 0), 0, None, 0, None)
    assert st == 2 and b"block_size" in L.lof_last_error()
    st = L.lof_virtual_round_mant_f32(None, None, 0, 24, _lib.RoundSpec(0, 0, 0), 1.0, None, 0,
                                      None)
    assert st == 2
    # the allocation-free grouped launch has a fixed workspace of 256 items
    items = (_lib.GroupItem * 257)()
    st = L.lof_grouped_run_noalloc(items, 257, -1, None)
    assert st == 2 and b"256" in L.lof_last_error()
    st = L.lof_grouped_run_noalloc(None, 3, -1, None)
    assert st == 1


def test_format_descriptor_contract():
    import lofloat_amd as lof
    f = lof.create_p3109_params(8, 4)
    assert (f.total_bits, f.mantissa_bits, f.bias) == (8, 3, 8)
    assert f.inf_behavior == lof.InfBehavior.Saturating and f.signedness == lof.Signedness.Signed
    with pytest.raises(AttributeError):
        f.bias = 3
    e = lof.create_e8m0_params()
    assert (e.total_bits, e.mantissa_bits, e.bias, int(e.signedness)) == (8, 0, 127, 1)
    with pytest.raises(ValueError):
        lof.FloatFormatDescriptor(8, 3, 7, 0, 0, 0, object(), lof.formats.P3109_NaNChecker(8))
    with pytest.raises(ValueError):
        lof.FloatFormatDescriptor(8, 3, 7, 0, 0, 0, lof.formats.P3109_InfChecker(8), object())
    assert int(lof.RoundingMode.RoundToOdd) == 12 and lof.RoundingMode(4).name == "RoundToNearestOdd"


def test_custom_ops_registered():
    import torch
    import lofloat_amd as lof
    for name in ("virtual_round", "virtual_mx_round", "lof_gemm"):
        assert hasattr(torch.ops.lofloat, name)
    f = lof.create_half_params()
    assert lof.register_format(f) == lof.register_format(f)
    assert lof.get_format(lof.register_format(f)) is f
    # fake (meta) implementations give the reference's output shapes
    x = torch.empty(3, 5, device="meta")
    fid = lof.register_format(f)
    assert torch.ops.lofloat.virtual_round(x, fid, 0, 0, 1.0).shape == (3, 5)
    a, b = torch.empty(4, 6, device="meta"), torch.empty(6, 7, device="meta")
    assert torch.ops.lofloat.lof_gemm(a, b, 10, 0, 0, 1.0, 1.0).shape == (4, 7)


def test_cpu_tensors_rejected():
    import torch
    import lofloat_amd as lof
    with pytest.raises(RuntimeError, match="HIP device"):
        lof.virtual_round(torch.ones(8), lof.create_half_params())
    with pytest.raises(RuntimeError, match="HIP device"):
        lof.virtual_mx_round(torch.ones(64), 32, lof.create_p3109_params(8, 4),
                             lof.create_e8m0_params())
