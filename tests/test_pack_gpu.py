"""GPU parity of the bit-pack codec (lof_pack_encode_f32 / lof_pack_decode_f32 /
lof_pack_reps / lof_unpack_reps) against the oracle's restatement of lo_pack.hpp, bit-exact
in bytes and in decoded floats (NaN by class).

Covers the formats of the reference's test_pack.cpp / test_pack_cuda.cu (P3109 <3,2> <5,2>
<7,3> <9,3> <11,5> <12,8>, e2m1, e3m2, e5m2, e4m3fn, half) plus bf16, an Extended P3109,
an unsigned format and the IEEE-F8 checkers; both group orders; every rounding mode (the
stochastic ones with explicit and generated words); the three saturation modes; negatives
into unsigned formats; partial groups, empty and odd sizes; unaligned buffers; raw rep
packing for n = 1..16.
"""
import numpy as np
import pytest
import torch

import oracle
from conftest import assert_bits

pytestmark = pytest.mark.gpu


def _descs():
    import lofloat_amd as lof
    F = lof.formats
    I, N, S = lof.InfBehavior, lof.NaNBehavior, lof.Signedness

    def p3109(k, p, signed=True, has_inf=False):
        return lof.FloatFormatDescriptor(
            k, p - 1, 1 << (k - p - 1), I.Extended if has_inf else I.Saturating, N._3109,
            S.Signed if signed else S.Unsigned, F.P3109_InfChecker(k, signed, not has_inf),
            F.P3109_NaNChecker(k, signed))
    return {
        "p3109<3,2>": p3109(3, 2), "e2m1": lof.create_ocp_e2m1_params(),
        "p3109<5,2>": p3109(5, 2), "e3m2": lof.create_ocp_e3m2_params(),
        "p3109<7,3>": p3109(7, 3), "e5m2": lof.create_ocp_e5m2_params(),
        "e4m3fn": lof.create_ocp_e4m3fn_params(), "p3109<9,3>": p3109(9, 3),
        "p3109<11,5>": p3109(11, 5), "p3109<12,8>": p3109(12, 8),
        "half": lof.create_half_params(), "bf16": lof.create_bfloat16_params(),
        "p3109<8,4>ext": p3109(8, 4, True, True), "u_p3109<6,3>": p3109(6, 3, False),
        "ieee_f8": lof.FloatFormatDescriptor(8, 3, 7, I.Extended, N._3109, S.Signed,
                                             F.IEEEF8InfChecker(), F.IEEEF8NaNChecker(True)),
        # 10 exponent bits, bias 150: values from 2^-154 to 2^873, beyond FP32 both ways
        "wide16": lof.FloatFormatDescriptor(
            16, 5, 150, I.Extended, N._3109, S.Signed,
            F.MaskedInfChecker(0x7FE0, 0x1F, 0x7FE0, 0xFFE0),
            F.MaskedNaNChecker(0x7FE0, 0x1F, 0x7FF0, 0x7FE1)),
    }


NAMES = ["p3109<3,2>", "e2m1", "p3109<5,2>", "e3m2", "p3109<7,3>", "e5m2", "e4m3fn",
         "p3109<9,3>", "p3109<11,5>", "p3109<12,8>", "half", "bf16", "p3109<8,4>ext",
         "u_p3109<6,3>", "ieee_f8", "wide16"]


def _opf(desc, ub=0):
    """The oracle's pack format built from the same checker data the device gets."""
    return oracle._PackFmt(oracle._fmt(desc.as_tuple()), ub,
                           oracle._chk(*desc.is_inf_checker.lof_checker()),
                           oracle._chk(*desc.is_nan_checker.lof_checker()))


def _inputs(pf, n_random=40000, seed=0):
    g = np.random.default_rng(seed)
    reps = np.arange(1 << pf.fmt.bitwidth, dtype=np.uint32)
    vals = oracle.rep_to_float(reps, pf)
    fin = np.unique(vals[np.isfinite(vals)])
    mids = ((fin[1:].astype(np.float64) + fin[:-1]) / 2).astype(np.float32)
    rnd = (g.standard_normal(n_random) * np.exp2(g.uniform(-30, 30, n_random))).astype(np.float32)
    specials = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-40, -1e-40, 1e-45, 3e38, -3e38],
                        np.float32)
    return np.concatenate([fin, np.nextafter(fin, np.float32(np.inf)),
                           np.nextafter(fin, np.float32(-np.inf)), mids, rnd, specials])


def _gpu_encode(lof, x, desc, dev, mode=0, L=0, sat=0, endian=0, bits=None, seed=None, ub=0):
    xt = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    bt = None if bits is None else torch.from_numpy(bits.view(np.int32)).to(dev)
    return lof.pack.encode(xt, desc, mode, L, sat, endian, unsigned_behavior=ub, rand_bits=bt,
                           seed=seed).cpu().numpy()


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("endian", [0, 1])
def test_encode_decode_match_oracle(dev, name, endian):
    import lofloat_amd as lof
    desc = _descs()[name]
    pf = _opf(desc)
    x = _inputs(pf, seed=len(name) + endian)
    got = _gpu_encode(lof, x, desc, dev, endian=endian)
    want = oracle.pack_encode(x, pf, endian=endian)
    assert np.array_equal(got, want), (name, np.flatnonzero(got != want)[:8])
    dec = lof.pack.decode(torch.from_numpy(got).to(dev), desc, x.size, endian).cpu().numpy()
    assert_bits(dec, oracle.pack_decode(want, x.size, pf, endian), f"decode {name}")


@pytest.mark.parametrize("mode", range(15))
@pytest.mark.parametrize("name", ["e4m3fn", "p3109<9,3>", "half", "bf16"])
def test_every_mode(dev, mode, name):
    """Deterministic modes, and stochastic ones with explicit words and with the
    counter-based generator (len 4; 0 for TrueSR / Prob which read none)."""
    import lofloat_amd as lof
    desc = _descs()[name]
    pf = _opf(desc)
    x = _inputs(pf, n_random=20000, seed=mode)
    if mode in oracle.STOCHASTIC_MODES:
        L = 0 if mode in (oracle.TRUESR, oracle.PROB) else 4
        w = np.random.default_rng(mode).integers(0, 2**32, x.size, dtype=np.uint64).astype(np.uint32)
        got = _gpu_encode(lof, x, desc, dev, mode, L, bits=w)
        assert np.array_equal(got, oracle.pack_encode(x, pf, mode, 0, L, bits=w)), "explicit words"
        got = _gpu_encode(lof, x, desc, dev, mode, L, seed=1234 + mode)
        assert np.array_equal(got, oracle.pack_encode(x, pf, mode, 0, L, seed=1234 + mode)), \
            "generated words"
    else:
        got = _gpu_encode(lof, x, desc, dev, mode)
        assert np.array_equal(got, oracle.pack_encode(x, pf, mode))


@pytest.mark.parametrize("sat", [0, 1, 2])
@pytest.mark.parametrize("name", ["p3109<8,4>ext", "half", "ieee_f8", "e4m3fn"])
def test_saturation_modes(dev, sat, name):
    import lofloat_amd as lof
    desc = _descs()[name]
    pf = _opf(desc)
    x = np.array([np.inf, -np.inf, 1e30, -1e30, 3e38, 70000.0, -70000.0, 500.0, -500.0, 1.0],
                 np.float32)
    got = _gpu_encode(lof, x, desc, dev, sat=sat)
    assert np.array_equal(got, oracle.pack_encode(x, pf, saturation=sat))


@pytest.mark.parametrize("ub", [0, 1])
def test_unsigned_behaviour(dev, ub):
    import lofloat_amd as lof
    desc = _descs()["u_p3109<6,3>"]
    pf = _opf(desc, ub)
    x = np.array([-1.5, -0.0, 0.0, 1.5, -np.inf, np.nan, -1e-30, 2.0], np.float32)
    got = _gpu_encode(lof, x, desc, dev, ub=ub)
    assert np.array_equal(got, oracle.pack_encode(x, pf))


@pytest.mark.parametrize("n", range(1, 17))
@pytest.mark.parametrize("endian", [0, 1])
def test_pack_reps_roundtrip(dev, n, endian):
    import lofloat_amd as lof
    g = np.random.default_rng(n * 3 + endian)
    for count in (1, 7, 33, 1000, 4099, 3 * 8192 + 5):
        reps = g.integers(0, 1 << 31, count, dtype=np.int64).astype(np.int32)
        rt = torch.from_numpy(reps).to(dev)
        buf = lof.pack.pack_reps(rt, n, endian)
        want = oracle.pack_reps(reps.view(np.uint32), n, endian)
        assert np.array_equal(buf.cpu().numpy(), want), (n, endian, count)
        back = lof.pack.unpack_reps(buf, n, count, endian).cpu().numpy()
        assert np.array_equal(back.view(np.uint32), reps.view(np.uint32) & ((1 << n) - 1))


def test_sizes_and_unaligned(dev):
    import lofloat_amd as lof
    desc = _descs()["p3109<11,5>"]            # 16 elements in 22 bytes
    pf = _opf(desc)
    g = np.random.default_rng(5)
    for count in (0, 1, 15, 16, 17, 33, (1 << 20) + 3):
        x = g.standard_normal(count).astype(np.float32)
        got = _gpu_encode(lof, x, desc, dev)
        assert got.size == lof.pack.packed_size(11, count)
        assert np.array_equal(got, oracle.pack_encode(x, pf)), count
    # unaligned packed buffers (byte offset 1) give the same bytes and values, and nothing
    # is written past the packed length; so do float buffers that are not 16-byte aligned
    for count in (1001, 5 * 8192 + 77):
        x = g.standard_normal(count).astype(np.float32)
        nb = lof.pack.packed_size(11, x.size)
        big = torch.full((nb + 5,), 0xA5, dtype=torch.uint8, device=dev)
        xt = torch.from_numpy(x).to(dev)
        lof.pack.encode(xt, desc, out=big[1:1 + nb])
        aligned = lof.pack.encode(xt, desc)
        assert torch.equal(big[1:1 + nb], aligned), count
        assert big[0].item() == 0xA5 and bool((big[1 + nb:] == 0xA5).all()), count
        assert np.array_equal(aligned.cpu().numpy(), oracle.pack_encode(x, pf)), count
        d1 = lof.pack.decode(big[1:1 + nb], desc, x.size)
        d0 = lof.pack.decode(aligned, desc, x.size)
        assert torch.equal(d1, d0), count
        xs = torch.zeros(count + 1, device=dev)
        xs[1:] = xt                                  # 4-byte but not 16-byte aligned
        assert torch.equal(lof.pack.encode(xs[1:], desc), aligned), count
        ys = torch.zeros(count + 1, device=dev)
        lof.pack.decode(aligned, desc, count, out=ys[1:])
        assert torch.equal(ys[1:], d0), count


def test_large_e4m3fn(dev):
    """2^24 + 5 elements (the round workload's size) against the oracle."""
    import lofloat_amd as lof
    desc = _descs()["e4m3fn"]
    pf = _opf(desc)
    g = np.random.default_rng(24)
    x = (g.standard_normal((1 << 24) + 5) * np.exp2(g.uniform(-12, 12, (1 << 24) + 5))).astype(np.float32)
    got = _gpu_encode(lof, x, desc, dev)
    assert np.array_equal(got, oracle.pack_encode(x, pf))
    dec = lof.pack.decode(torch.from_numpy(got).to(dev), desc, x.size).cpu().numpy()
    assert_bits(dec, oracle.pack_decode(got, x.size, pf), "large decode")


def test_errors(dev):
    import lofloat_amd as lof
    x = torch.ones(16, device=dev)
    with pytest.raises(NotImplementedError):
        lof.pack.encode(x, lof.create_single_params())            # 32 bits
    with pytest.raises(ValueError):
        lof.pack.encode(x, lof.create_ocp_e4m3fn_params(), 3, 21)  # stoch_len > 20
    with pytest.raises(RuntimeError):
        lof.pack.encode(x.cpu(), lof.create_ocp_e4m3fn_params())

    class Opaque:
        def __call__(self, bits):
            return False

        def minNegInf(self):
            return 0

        def minPosInf(self):
            return 0
    odd = lof.FloatFormatDescriptor(8, 3, 7, lof.InfBehavior.Saturating, lof.NaNBehavior._3109,
                                    lof.Signedness.Signed, Opaque(),
                                    lof.formats.FiniteNaNChecker())
    with pytest.raises(NotImplementedError):
        lof.pack.encode(x, odd)
