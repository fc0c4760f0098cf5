"""GPU parity: elementwise virtual_round (both overloads) against the oracle, bit-exact.

Covers every rounding mode (stochastic ones with explicit random words and with the
library's counter-based words), 13 formats (OCP/IEEE-style, P3109, extended vs
saturating, unsigned, E8M0, identity), the fused scale, +-0 / +-inf / NaN / FP32
subnormals, ragged sizes, in-place calls, and an exhaustive sweep of all 2^32 inputs for
the e4m3 workload format.
"""
import numpy as np
import pytest
import torch

import oracle
from conftest import assert_bits, c1_values, random_words

pytestmark = pytest.mark.gpu

FORMATS = {
    "e4m3_sat": (8, 3, 7, 1, 0, 0),
    "e4m3_ext": (8, 3, 7, 0, 0, 0),
    "e2m1": (4, 1, 1, 1, 0, 0),
    "e5m2": (8, 2, 15, 1, 0, 0),
    "e5m2_ext": (8, 2, 15, 0, 0, 0),
    "p3109_8p4": (8, 3, 8, 1, 0, 0),
    "e3m2": (6, 2, 3, 1, 0, 0),
    "bf16_like": (16, 7, 127, 1, 0, 0),
    "fp16_like": (16, 10, 15, 0, 0, 0),
    "e8m10": (19, 10, 127, 1, 0, 0),
    "single": (32, 23, 127, 0, 0, 0),
    "e8m0": (8, 0, 127, 0, 0, 1),
    "u_e5m2": (7, 2, 15, 1, 0, 1),
}
DET = oracle.DETERMINISTIC_MODES
STOCH = oracle.STOCHASTIC_MODES


def _fmt(t):
    import lofloat_amd as lof
    return lof.FloatFormatDescriptor(t[0], t[1], t[2], t[3], t[4], t[5],
                                     lof.formats.P3109_InfChecker(t[0]),
                                     lof.formats.P3109_NaNChecker(t[0]))


@pytest.fixture(scope="module")
def inputs():
    x = c1_values(1 << 16, seed=7)
    # strided sweep over the whole 32-bit pattern space
    pats = (np.arange(65536, dtype=np.uint64) * 65537 + 12345).astype(np.uint32)
    # values sitting on and next to rounding boundaries of narrow formats
    edge = np.array([0.5 * 2.0**-9, 1.5 * 2.0**-9, 2.0**-7 * 1.0625, 240.0, 248.0, 464.0,
                     480.0, 496.0, 512.0, 57344.0, 61440.0, 65504.0, 65520.0, 131008.0,
                     1.5 * 2**-149, 2.0**-126, 3.4e38], dtype=np.float32)
    edge = np.concatenate([edge, -edge])
    return np.concatenate([x, pats.view(np.float32), edge])


def _gpu_round(lof, x, dev, params, mode, stoch_len=0, scale=1.0, bits=None, seed=None):
    xt = torch.from_numpy(x).to(dev)
    bt = None if bits is None else torch.from_numpy(bits.view(np.int32)).to(dev)
    y = lof.virtual_round(xt, params, mode, stoch_len, scale, rand_bits=bt, seed=seed)
    return y.cpu().numpy()


@pytest.mark.parametrize("fname", list(FORMATS))
def test_format_deterministic_modes(dev, inputs, fname):
    import lofloat_amd as lof
    ft = FORMATS[fname]
    fd = _fmt(ft)
    for mode in DET:
        y = _gpu_round(lof, inputs, dev, fd, mode)
        ref = oracle.virtual_round(inputs, ft, mode)
        assert_bits(y, ref, f"{fname} mode {mode}")


@pytest.mark.parametrize("fname", ["e4m3_sat", "e5m2", "e2m1", "bf16_like", "fp16_like",
                                   "single", "u_e5m2"])
def test_format_stochastic_with_bits(dev, inputs, fname):
    import lofloat_amd as lof
    ft = FORMATS[fname]
    fd = _fmt(ft)
    r = 23 - ft[1]
    bits = random_words(inputs.size, seed=11)
    for mode in STOCH:
        for L in sorted({0, min(1, r), r // 2, r}):
            y = _gpu_round(lof, inputs, dev, fd, mode, L, bits=bits)
            ref = oracle.virtual_round(inputs, ft, mode, L, bits=bits)
            assert_bits(y, ref, f"{fname} mode {mode} len {L}")


@pytest.mark.parametrize("mode", STOCH)
def test_stochastic_internal_generator(dev, inputs, mode):
    """Without rand_bits the words are lof_rand_u32(seed, i); the oracle restates that
    generator, so these are bit-exact too."""
    import lofloat_amd as lof
    ft = FORMATS["e5m2"]
    seed = 0x1234_5678_9ABC_DEF0
    y = _gpu_round(lof, inputs, dev, _fmt(ft), mode, 8, seed=seed)
    ref = oracle.virtual_round(inputs, ft, mode, 8, seed=seed)
    assert_bits(y, ref, f"internal generator mode {mode}")


def _ideal_p_up(mode, j, L):
    """Probability that value 1 + j*2^-11 (tail j/2^L of one e4m3 ulp) rounds up, from the
    reference's ideal probabilities (test_prob_round.cpp:64-67, lo_float.h:815-819) with a
    uniform random integer R in [0, 2^L) and a fair coin."""
    t = j / 2.0 ** L
    if mode in (oracle.SRA, oracle.TRUESR):
        return t                                  # unbiased: E[result] = x
    if mode == oracle.PROB:
        return 0.5                                # up with probability 1/2 for any tail
    if mode == oracle.SRB:
        return (j + 1) / 2.0 ** L                 # increment (R + 1): biased by 2^-L
    if mode == oracle.SRC:                        # increment ceil(R / 2) * 2 units of 2^-L
        return sum(1 for R in range(2 ** L) if 2 * ((R >> 1) + (R & 1)) + j >= 2 ** L) / 2.0 ** L
    if mode == oracle.SRD:                        # increment R + coin
        return sum(1 for R in range(2 ** L) for c in (0, 1) if R + c + j >= 2 ** L) / 2.0 ** (L + 1)
    raise AssertionError(mode)


@pytest.mark.parametrize("mode", STOCH)
def test_stochastic_expected_values(dev, mode):
    """Generated words give each stochastic mode its ideal probability of rounding up.
    1 + j*2^-11 sits between e4m3 neighbours 1 and 1.125 with a tail of j/256 ulp; over
    2^20 copies the observed frequency of 1.125 must lie within 6 sigma of the ideal."""
    import lofloat_amd as lof
    L, copies = 8, 1 << 20
    js = [1, 37, 128, 200, 255]
    x = np.repeat(np.array([1 + j * 2.0 ** -11 for j in js], np.float32), copies)
    xt = torch.from_numpy(x).to(dev)
    y = lof.virtual_round(xt, _fmt(FORMATS["e4m3_sat"]), mode, L, seed=0xC0FFEE + mode)
    y = y.cpu().numpy().reshape(len(js), copies)
    assert np.all((y == 1.0) | (y == 1.125)), "results outside the two neighbours"
    for row, j in zip(y, js):
        p = _ideal_p_up(mode, j, L)
        got = float(np.mean(row == 1.125))
        tol = 6 * np.sqrt(max(p * (1 - p), 1e-12) / copies) + 1e-9
        assert abs(got - p) <= tol, (mode, j, got, p)
        if mode in (oracle.SRA, oracle.TRUESR):   # unbiased in value as well
            assert abs(float(np.mean(row.astype(np.float64))) - (1 + j * 2.0 ** -11)) <= 0.125 * tol


@pytest.mark.parametrize("scale", [2.0, 0.375, 1e-3, 3.0e4, -1.5])
def test_fused_scale(dev, inputs, scale):
    import lofloat_amd as lof
    for ft in (FORMATS["e4m3_sat"], FORMATS["e5m2_ext"]):
        y = _gpu_round(lof, inputs, dev, _fmt(ft), 0, scale=scale)
        ref = oracle.virtual_round(inputs, ft, 0, scale=scale)
        assert_bits(y, ref, f"scale {scale}")
        ym = _gpu_round(lof, inputs, dev, 10, 0, scale=scale)
        refm = oracle.virtual_round_mant(inputs, 10, 0, scale=scale)
        assert_bits(ym, refm, f"mantissa scale {scale}")


@pytest.mark.parametrize("m", [0, 1, 3, 7, 10, 15, 22, 23])
def test_mantissa_overload(dev, inputs, m):
    import lofloat_amd as lof
    bits = random_words(inputs.size, seed=m)
    for mode in DET:
        y = _gpu_round(lof, inputs, dev, m, mode)
        assert_bits(y, oracle.virtual_round_mant(inputs, m, mode), f"mant {m} mode {mode}")
    r = 23 - m
    for mode in STOCH:
        L = min(r, 5)
        y = _gpu_round(lof, inputs, dev, m, mode, L, bits=bits)
        ref = oracle.virtual_round_mant(inputs, m, mode, L, bits=bits)
        assert_bits(y, ref, f"mant {m} stochastic mode {mode}")


@pytest.mark.parametrize("n", [0, 1, 3, 4, 5, 1023, 4097, 262147])
def test_ragged_sizes_and_inplace(dev, n):
    import lofloat_amd as lof
    x = c1_values(max(n, 1), seed=n)[:n]
    ft = FORMATS["e4m3_sat"]
    xt = torch.from_numpy(x.copy()).to(dev)
    y = lof.virtual_round(xt, _fmt(ft)).cpu().numpy()
    assert_bits(y, oracle.virtual_round(x, ft), f"n={n}")
    lof.virtual_round(xt, _fmt(ft), out=xt)  # in place
    assert_bits(xt.cpu().numpy(), oracle.virtual_round(x, ft), f"in-place n={n}")


def test_unaligned_views(dev):
    """Offset views (not 16-byte aligned) take the scalar-load path."""
    import lofloat_amd as lof
    x = c1_values(10001, seed=3)
    base = torch.from_numpy(x).to(dev)
    ft = FORMATS["e5m2"]
    for off in (1, 2, 3):
        v = base[off:]
        y = lof.virtual_round(v, _fmt(ft), 3, 4, rand_bits=torch.from_numpy(
            random_words(v.numel(), 5).view(np.int32)).to(dev)).cpu().numpy()
        ref = oracle.virtual_round(x[off:], ft, 3, 4, bits=random_words(v.numel(), 5))
        assert_bits(y, ref, f"offset {off}")


def test_noncontiguous_input(dev):
    import lofloat_amd as lof
    x = c1_values(64 * 48, seed=9).reshape(64, 48)
    xt = torch.from_numpy(x).to(dev).t()  # transposed view
    ft = FORMATS["e4m3_sat"]
    y = lof.virtual_round(xt, _fmt(ft)).cpu().numpy()
    ref = oracle.virtual_round(np.ascontiguousarray(x.T), ft).reshape(48, 64)
    assert_bits(y, ref, "transposed view")


@pytest.mark.slow
@pytest.mark.timeout(600)
@pytest.mark.parametrize("ft,mode,mant", [((8, 3, 7, 1, 0, 0), 0, None),
                                          (None, 0, 10)])
def test_exhaustive_all_fp32_patterns(dev, ft, mode, mant):
    """All 2^32 input bit patterns, in 2^26-element chunks."""
    import lofloat_amd as lof
    chunk = 1 << 26
    fd = _fmt(ft) if ft is not None else None
    for c in range((1 << 32) // chunk):
        pat = torch.arange(c * chunk, (c + 1) * chunk, dtype=torch.int64, device=dev)
        xt = pat.to(torch.int32).view(torch.float32)  # wraps to the 32-bit pattern
        y = lof.virtual_round(xt, fd if fd is not None else mant, mode)
        x = xt.cpu().numpy()
        ref = (oracle.virtual_round(x, ft, mode) if ft is not None
               else oracle.virtual_round_mant(x, mant, mode))
        assert_bits(y.cpu().numpy(), ref, f"exhaustive chunk {c}")


def test_argument_errors(dev):
    import lofloat_amd as lof
    x = torch.ones(16, device=dev)
    with pytest.raises(NotImplementedError):
        lof.virtual_round(x, _fmt((30, 24, 7, 1, 0, 0)))       # mantissa > 23
    with pytest.raises(ValueError):
        lof.virtual_round(x, _fmt(FORMATS["e4m3_sat"]), 3, 21)  # stoch_len > dropped bits
    with pytest.raises(ValueError):
        lof.virtual_round(x, _fmt(FORMATS["e4m3_sat"]), 15)
    with pytest.raises(RuntimeError):
        lof.virtual_round(torch.ones(4), _fmt(FORMATS["e4m3_sat"]))  # CPU tensor
    with pytest.raises(RuntimeError):
        lof.virtual_round(x.double(), _fmt(FORMATS["e4m3_sat"]))


def test_foreign_device_outputs_rejected(dev):
    """A host tensor passed as ``out`` (or as a GEMM / conv source ``C``) is refused before
    any launch: its pointer is not device memory."""
    import lofloat_amd as lof
    x = torch.ones(64, device=dev)
    host = torch.empty(64)
    with pytest.raises(ValueError, match="out must be on"):
        lof.virtual_round(x, lof.create_half_params(), out=host)
    with pytest.raises(ValueError, match="out must be on"):
        lof.virtual_round(x, 10, out=host)
    with pytest.raises(ValueError, match="out must be on"):
        lof.virtual_mx_round(x, 32, lof.create_p3109_params(8, 4), lof.create_e8m0_params(),
                             out=host)
    A = torch.ones(8, 8, device=dev)
    acc = lof.AccumSpec(mant_bits=10)
    with pytest.raises(ValueError, match="out must be on"):
        lof.gemm(A, A, acc, out=torch.empty(8, 8))
    with pytest.raises(ValueError, match="C must be on"):
        lof.gemm(A, A, acc, C=torch.empty(8, 8), beta=1.0)
    xc = torch.ones(1, 2, 5, 5, device=dev)
    wc = torch.ones(3, 2, 3, 3, device=dev)
    with pytest.raises(ValueError, match="out must be on"):
        lof.conv2d(xc, wc, acc, out=torch.empty(1, 3, 3, 3))
    with pytest.raises(ValueError, match="C must be on"):
        lof.conv2d(xc, wc, acc, C=torch.empty(1, 3, 3, 3), beta=1.0)
    fmt = lof.create_p3109_params(8, 4)
    buf = lof.pack.encode(x, fmt)
    with pytest.raises(ValueError, match="out must be on"):
        lof.pack.decode(buf, fmt, 64, out=torch.empty(64))
