"""GPU parity of the typed cast (lof_cast_to_rep_f32 / lof_cast_from_rep_f32 through
lofloat_amd.pack.to_reps / from_reps) against the oracle's float_to_rep / rep_to_float, bit
for bit: the codec's formats, every rounding mode (stochastic ones with explicit and
generated words), the three saturation modes, the test_sat_mode.cpp sweep, every rep's
value, partial tiles, unaligned containers and errors.

A container holds the rep's low n bits, as the codec stores it: a checker pattern wider
than the format (e2m1 takes e4m3's checkers, lo_float.h:1521-1530, so its qNaN is 0xFF) is
cut to n bits, so the oracle's reps are compared masked."""
import numpy as np
import pytest
import torch

import oracle
from conftest import assert_bits
from test_pack_gpu import NAMES, _descs, _inputs, _opf

pytestmark = pytest.mark.gpu


def _rep(pf, x, *a, **k):
    return oracle.float_to_rep(x, pf, *a, **k) & np.uint32((1 << pf.fmt.bitwidth) - 1)


def _to(lof, x, desc, dev, mode=0, L=0, sat=0, bits=None, seed=None, ub=0):
    xt = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    bt = None if bits is None else torch.from_numpy(bits.view(np.int32)).to(dev)
    r = lof.pack.to_reps(xt, desc, mode, L, sat, unsigned_behavior=ub, rand_bits=bt, seed=seed)
    return r.cpu().numpy().view(np.uint8 if desc.total_bits <= 8 else np.uint16).astype(np.uint32)


@pytest.mark.parametrize("name", NAMES)
def test_encode_and_every_rep(dev, name):
    import lofloat_amd as lof
    desc = _descs()[name]
    pf = _opf(desc)
    x = _inputs(pf, seed=len(name))
    got = _to(lof, x, desc, dev)
    assert np.array_equal(got, _rep(pf, x)), name
    reps = np.arange(1 << desc.total_bits, dtype=np.uint32)
    rt = torch.from_numpy(reps.astype(np.uint8 if desc.total_bits <= 8 else np.uint16)
                          .view(np.uint8 if desc.total_bits <= 8 else np.int16)).to(dev)
    vals = lof.pack.from_reps(rt, desc).cpu().numpy()
    assert_bits(vals, oracle.rep_to_float(reps, pf), f"from_reps {name}")


@pytest.mark.parametrize("mode", range(15))
@pytest.mark.parametrize("name", ["e4m3fn", "p3109<11,5>", "half"])
def test_every_mode(dev, mode, name):
    import lofloat_amd as lof
    desc = _descs()[name]
    pf = _opf(desc)
    x = _inputs(pf, n_random=20000, seed=100 + mode)
    if mode in oracle.STOCHASTIC_MODES:
        L = 0 if mode in (oracle.TRUESR, oracle.PROB) else 3
        w = np.random.default_rng(mode).integers(0, 2**32, x.size, dtype=np.uint64).astype(np.uint32)
        assert np.array_equal(_to(lof, x, desc, dev, mode, L, bits=w),
                              _rep(pf, x, mode, 0, L, bits=w))
        assert np.array_equal(_to(lof, x, desc, dev, mode, L, seed=77),
                              _rep(pf, x, mode, 0, L, seed=77))
    else:
        assert np.array_equal(_to(lof, x, desc, dev, mode), _rep(pf, x, mode))


@pytest.mark.parametrize("l", range(2, 9))
def test_saturation_sweep(dev, l):
    """test_sat_mode.cpp's sweep on the device: every P3109<l, p>, both infinity
    behaviours, over-range and infinite inputs of both signs and 1.0, in all three
    saturation modes, against the oracle (itself pinned to the sweep's rules)."""
    import lofloat_amd as lof
    x = np.array([np.inf, -np.inf, 1.0, -1.0, 3e38, -3e38, 1e6, -1e6, 4096.0, -37.0, 0.5],
                 np.float32)
    for p in range(1, l):
        for ext in (False, True):
            desc = lof.create_p3109_params(l, p, True, not ext)
            pf = _opf(desc)
            for sat in (0, 1, 2):
                assert np.array_equal(_to(lof, x, desc, dev, sat=sat),
                                      _rep(pf, x, saturation=sat)), (l, p, ext, sat)


def test_shapes_alignment_and_unsigned(dev):
    import lofloat_amd as lof
    g = np.random.default_rng(3)
    for name in ("e4m3fn", "p3109<9,3>"):
        desc = _descs()[name]
        pf = _opf(desc)
        for count in (1, 5, 4095, 4096, 4097, 3 * 4096 + 13):
            x = (g.standard_normal(count) * 50).astype(np.float32)
            assert np.array_equal(_to(lof, x, desc, dev), _rep(pf, x)), (name, count)
        # containers that are not 4 / 8-byte aligned, float buffers not 16-byte aligned
        x = (g.standard_normal(9000) * 50).astype(np.float32)
        dt = lof.pack.rep_dtype(desc.total_bits)
        big = torch.zeros(x.size + 1, dtype=dt, device=dev)
        xs = torch.zeros(x.size + 1, device=dev)
        xs[1:] = torch.from_numpy(x).to(dev)
        lof.pack.to_reps(xs[1:], desc, out=big[1:])
        ref = lof.pack.to_reps(torch.from_numpy(x).to(dev), desc)
        assert torch.equal(big[1:], ref) and big[0].item() == 0
        ys = torch.zeros(x.size + 1, device=dev)
        lof.pack.from_reps(big[1:], desc, out=ys[1:])
        assert torch.equal(ys[1:], lof.pack.from_reps(ref, desc))
    desc = _descs()["u_p3109<6,3>"]
    for ub in (0, 1):
        x = np.array([-1.5, -0.0, 0.0, 1.5, -np.inf, np.nan, 2.0, 7.0], np.float32)
        assert np.array_equal(_to(lof, x, desc, dev, ub=ub), _rep(_opf(desc, ub), x))


def test_large_e4m3fn_roundtrip(dev):
    """2^26 + 7 values: encode against the oracle; decode(encode(x)) against the pinned
    virtual_round oracle inside the finite range."""
    import lofloat_amd as lof
    desc = _descs()["e4m3fn"]
    pf = _opf(desc)
    g = np.random.default_rng(26)
    x = (g.standard_normal((1 << 26) + 7) * 16).astype(np.float32)
    xt = torch.from_numpy(x).to(dev)
    r = lof.pack.to_reps(xt, desc)
    assert np.array_equal(r.cpu().numpy().astype(np.uint32), _rep(pf, x))
    back = lof.pack.from_reps(r, desc).cpu().numpy()
    keep = np.abs(x) <= 448
    ref = oracle.virtual_round(x[keep], (8, 3, 7, 1, 0, 0))
    assert_bits(back[keep], ref, "roundtrip")


def test_errors(dev):
    import lofloat_amd as lof
    x = torch.ones(8, device=dev)
    with pytest.raises(NotImplementedError):
        lof.pack.to_reps(x, lof.create_single_params())
    with pytest.raises(RuntimeError):
        lof.pack.from_reps(torch.zeros(8, dtype=torch.int16, device=dev),
                           lof.create_ocp_e4m3fn_params())   # 8-bit reps are uint8
    with pytest.raises(ValueError):
        lof.pack.to_reps(x, lof.create_ocp_e4m3fn_params(), 3, 21)
