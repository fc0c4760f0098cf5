"""CPU tests that pin the oracle (oracle/lof_oracle.c) independently of its own code.

Each oracle function is checked against something that fixes it without sharing its
method:

* virtual_round (full format, every mode): an arithmetic closed form in float64 -- the
  value's grid spacing q = 2^(max(e, ToMin_exp) - mb), y = |x| / q, the integer k chosen
  from floor(y), frac(y) and the mode (stochastic modes from the documented random-word
  mapping), result k * q with the reference's overflow / NaN / underflow rules
  (lo_float.h:2844-2939) and its FP32-subnormal hidden-bit quirk -- plus library
  conversions: torch.float8_e4m3fn, torch.float8_e5m2, torch.float16, torch.bfloat16.
* virtual_round (mantissa only): the same closed form with FP32's exponent range, and
  torch.bfloat16 for 7 bits.
* RTO FMA: exact rational arithmetic (fractions.Fraction) and the FPU-rounding-mode
  restatement of cutlass_gemms.cuh:74-80.
* GEMM leaf: a Fraction-based per-step emulation, the gamma_K error bound and
  exact-input mode independence of test_blis_lof_accum.cpp:120-245, and the property
  RN_m(RTO(x)) == RN_m(x) behind round_to_odd.cu:30-46.
* MX: blockwise recomputation with the closed form and IEEE float32 division, the
  scale-is-the-nearest-E8M0-neighbour property (test_virtual_mx_round.cpp:151-193), the
  exact block of test_mx_round_py.py:33-38 and the all-zero block.
* Golden fixtures in tests/golden/ (made by tests/golden/make_fixtures.py) are reproduced.
"""
import os
from fractions import Fraction

import numpy as np
import pytest
import torch

import oracle
from conftest import assert_bits, c1_values, random_words

FORMATS = {
    "e4m3_sat": (8, 3, 7, 1, 0, 0), "e4m3_ext": (8, 3, 7, 0, 0, 0), "e2m1": (4, 1, 1, 1, 0, 0),
    "e5m2": (8, 2, 15, 1, 0, 0), "e5m2_ext": (8, 2, 15, 0, 0, 0), "p3109_8p4": (8, 3, 8, 1, 0, 0),
    "e3m2": (6, 2, 3, 1, 0, 0), "bf16_like": (16, 7, 127, 1, 0, 0),
    "fp16_like": (16, 10, 15, 0, 0, 0), "e8m10": (19, 10, 127, 1, 0, 0),
    "single": (32, 23, 127, 0, 0, 0), "e8m0": (8, 0, 127, 0, 0, 1), "u_e5m2": (7, 2, 15, 1, 0, 1),
}
ALL_MODES = list(range(15))


# --------------------------------------------------------------------------- closed form
def _choose(lo, frac, mode, pos, L, u, odd=None):
    """Integer k in {lo, lo+1} for one rounding mode; frac in [0, 1) exact.  `odd` is the
    last kept bit of the bit pattern (defaults to lo's parity)."""
    hi = lo + 1
    if odd is None:
        odd = (lo % 2) == 1
    if mode == oracle.RNE:
        return np.where(frac > 0.5, hi, np.where(frac < 0.5, lo, np.where(odd, hi, lo)))
    if mode == oracle.RNO:
        return np.where(frac > 0.5, hi, np.where(frac < 0.5, lo, np.where(odd, lo, hi)))
    if mode == oracle.RTA:
        return np.where(frac >= 0.5, hi, lo)
    if mode == oracle.RTTZ:
        return np.where(frac > 0.5, hi, lo)
    if mode == oracle.RTZ:
        return lo
    if mode == oracle.RAFZ:
        return np.where(frac > 0, hi, lo)
    if mode == oracle.RUP:
        return np.where((frac > 0) & pos, hi, lo)
    if mode == oracle.RDOWN:
        return np.where((frac > 0) & ~pos, hi, lo)
    if mode == oracle.RTO:
        return np.where(frac > 0, np.where(odd, lo, hi), lo)
    u = u.astype(np.uint64)
    mask = np.uint64((1 << L) - 1)
    R = (u & mask).astype(np.float64)
    sc = 2.0 ** -L
    if mode == oracle.SRA:
        return lo if L <= 0 else np.where(frac + R * sc >= 1, hi, lo)
    if mode == oracle.SRB:
        return np.where((frac > 0) & (frac + (R + 1) * sc >= 1), hi, lo)
    if mode == oracle.SRC:
        Ri = (u & mask)
        add = ((Ri >> np.uint64(1)) + (Ri & np.uint64(1))).astype(np.float64) * 2.0 ** (1 - L)
        return lo if L == 0 else np.where((frac > 0) & (frac + add >= 1), hi, lo)
    if mode == oracle.SRD:
        coin = ((u >> np.uint64(L)) & np.uint64(1)).astype(np.float64)
        return lo if L <= 0 else np.where(frac + (R + coin) * sc >= 1, hi, lo)
    if mode == oracle.TRUESR:
        return np.where(u.astype(np.float64) * 2.0 ** -32 < frac, hi, lo)
    if mode == oracle.PROB:
        return np.where((frac > 0) & ((u & np.uint64(1)) == 1), hi, lo)
    raise ValueError(mode)


def cf_round(x, fmt, mode, L=0, u=None):
    """Closed-form virtual_round(x, FloatingPointParams) (see module docstring)."""
    bw, mb, bias, inf_b, _, sgn = fmt
    to_max = ((1 << (bw - mb - 1)) if sgn == 0 else (1 << (bw - mb))) - 1 - bias
    emin = 1 - bias
    q_min = float(np.float32(2.0 ** (emin - mb)))
    x = np.asarray(x, dtype=np.float32)
    u = np.zeros(x.shape, np.uint32) if u is None else u
    with np.errstate(invalid="ignore"):
        ax = np.abs(x).astype(np.float64)
    pos = ~np.signbit(x)
    finite = np.isfinite(x)
    # FP32-subnormal inputs reaching the target-subnormal branch keep a set hidden bit
    quirk = finite & (ax < 2.0 ** -126) & (ax >= q_min) & (emin > -127)
    axe = np.where(quirk, 2.0 ** -127 + ax / 2, ax)
    safe = np.where(finite & (axe > 0), axe, 1.0)
    e = np.frexp(safe)[1] - 1
    q = np.exp2((np.maximum(e, emin) - mb).astype(np.float64))
    y = safe / q
    lo = np.floor(y)
    frac = y - lo
    k = _choose(lo, frac, mode, pos, L, u, _kept_lsb(lo, e, emin, mb))
    val = np.where(finite & (axe > 0), k * q, 0.0)
    if mode in oracle.STOCHASTIC_MODES:                       # below To_min_val: flush
        val = np.where(ax < q_min, 0.0, val)
    if mode in (oracle.RNO, oracle.RTA) and np.float32(q_min) * np.float32(0.5) == 0:
        # the reference compares |x| with To_min_val * 0.5f in float; when that underflows
        # to 0, a zero input is an exact "tie" and RNO / RTA return +-To_min_val
        val = np.where(finite & (ax == 0), q_min, val)
    ovf = val >= 2.0 ** (to_max + 1)
    with np.errstate(over="ignore"):
        out = np.copysign(val, np.where(pos, 1.0, -1.0)).astype(np.float32)
    sat = x if inf_b == 1 else np.float32(np.inf)
    out = np.where(ovf & finite, sat, out)
    inf_out = x if (to_max >= 128 or inf_b == 1) else np.float32(np.inf)
    if mode == oracle.SRD and L > 0:
        # SR-D adds R + coin units of 2^-len; R + coin == 2^len adds a whole unit even to a
        # zero tail: +-inf's magnitude 0x7F800000 becomes 0x7F800000 + 2^(23-mb), a NaN
        # pattern (or, with mb = 0, exponent 256), then the overflow rule applies
        bumped = 0x7F800000 + (1 << (23 - mb))
        if (bumped >> 23) - 127 > to_max:
            bump_out = x if inf_b == 1 else np.float32(np.inf)
        else:
            bump_out = np.float32(np.nan)
        inf_out = np.where(_srd_full_unit(u, L), bump_out, inf_out)
    out = np.where(np.isinf(x), inf_out, out)
    return np.where(np.isnan(x), np.float32(np.nan), out).astype(np.float32)


def _srd_full_unit(u, L):
    u = u.astype(np.uint64)
    R = u & np.uint64((1 << L) - 1)
    coin = (u >> np.uint64(L)) & np.uint64(1)
    return (R + coin) == np.uint64(1 << L)


def _kept_lsb(lo, e, emin, mb):
    """Last kept bit of the bit-level rounders.  With mb >= 1 it is the significand's
    LSB (lo's parity); with no mantissa bits the kept field ends in the exponent, so in
    the normal range it is the biased exponent's LSB."""
    if mb > 0:
        return (lo % 2) == 1
    return np.where(e >= emin, ((e + 127) & 1) == 1, (lo % 2) == 1)


def cf_round_mant(x, m, mode, L=0, u=None):
    """Closed-form virtual_round(x, m): FP32 exponent range, no NaN/overflow handling."""
    x = np.asarray(x, dtype=np.float32)
    u = np.zeros(x.shape, np.uint32) if u is None else u
    ax = np.abs(x).astype(np.float64)
    pos = ~np.signbit(x)
    finite = np.isfinite(x)
    safe = np.where(finite, ax, 0.0)
    # zeros sit on the FP32-subnormal grid like any other tiny value
    e = np.where(safe > 0, np.frexp(np.where(safe > 0, safe, 1.0))[1] - 1, -127)
    q = np.exp2((np.maximum(e, -126) - m).astype(np.float64))
    y = safe / q
    lo = np.floor(y)
    k = _choose(lo, y - lo, mode, pos, L, u, _kept_lsb(lo, e, -126, m))
    val = np.where(finite, k * q, 0.0)
    with np.errstate(over="ignore"):
        out = np.copysign(val, np.where(pos, 1.0, -1.0)).astype(np.float32)  # 2^128 -> inf
    inf_out = x
    if mode == oracle.SRD and L > 0 and m < 23:
        # see cf_round: a whole unit added to inf's magnitude bits; with m = 0 the carry
        # reaches bit 31 (no overflow check here), giving -0 for either sign
        r = 23 - m
        mag = np.uint32(((0x7F800000 + (1 << r)) & ~((1 << r) - 1)) & 0xFFFFFFFF)
        bumped = (x.view(np.uint32) & np.uint32(0x80000000)) | mag
        inf_out = np.where(_srd_full_unit(u, L), bumped.view(np.float32), x)
    return np.where(np.isinf(x), inf_out, out)


@pytest.fixture(scope="module")
def inputs():
    x = c1_values(1 << 16, seed=21)
    pats = (np.arange(1 << 17, dtype=np.uint64) * 32771 + 99).astype(np.uint32).view(np.float32)
    return np.concatenate([x, pats])


@pytest.mark.parametrize("fname", list(FORMATS))
def test_virtual_round_vs_closed_form(inputs, fname):
    ft = FORMATS[fname]
    r = 23 - ft[1]
    u = random_words(inputs.size, seed=5)
    for mode in ALL_MODES:
        for L in ([0] if mode in oracle.DETERMINISTIC_MODES else sorted({0, min(1, r), r})):
            got = oracle.virtual_round(inputs, ft, mode, L, bits=u)
            assert_bits(got, cf_round(inputs, ft, mode, L, u), f"{fname} mode {mode} len {L}")


@pytest.mark.parametrize("m", [0, 1, 3, 7, 10, 19, 22, 23])
def test_mantissa_round_vs_closed_form(inputs, m):
    x = inputs[~np.isnan(inputs)]
    u = random_words(x.size, seed=m)
    for mode in ALL_MODES:
        L = 0 if mode in oracle.DETERMINISTIC_MODES else min(23 - m, 6)
        got = oracle.virtual_round_mant(x, m, mode, L, bits=u)
        assert_bits(got, cf_round_mant(x, m, mode, L, u), f"mant {m} mode {mode}")


def test_mantissa_round_nan_bits():
    """NaNs go through the same bit arithmetic (no NaN check in lo_float.h:2716-2756):
    a small-payload NaN rounds to inf, 0x7FFFFFFF carries into the sign position."""
    x = np.array([0x7F800001, 0x7FFFFFFF, 0x7FC00000, 0xFFC00000], np.uint32).view(np.float32)
    got = oracle.virtual_round_mant(x, 10, oracle.RNE).view(np.uint32)
    assert list(got) == [0x7F800000, 0x80000000, 0x7FC00000, 0xFFC00000]


def test_torch_dtype_cross_checks(inputs):
    x = inputs[np.isfinite(inputs)]
    xt = torch.from_numpy(x)
    cases = [((8, 3, 7, 1, 0, 0), torch.float8_e4m3fn, 448.0),
             ((8, 2, 15, 1, 0, 0), torch.float8_e5m2, 57344.0),
             ((16, 10, 15, 1, 0, 0), torch.float16, 65504.0)]
    for ft, dt, mx in cases:
        sel = np.abs(x) <= mx
        ref = xt[torch.from_numpy(sel)].to(dt).float().numpy()
        assert_bits(oracle.virtual_round(x[sel], ft), ref, f"{dt}")
    normal = (np.abs(x) >= 2.0 ** -126) & (np.abs(x) < 3.38e38)
    ref = xt[torch.from_numpy(normal)].to(torch.bfloat16).float().numpy()
    assert_bits(oracle.virtual_round(x[normal], (16, 7, 127, 1, 0, 0)), ref, "bfloat16 format")
    ref = xt.to(torch.bfloat16).float().numpy()
    assert_bits(oracle.virtual_round_mant(x, 7), ref, "bfloat16 mantissa")


def test_reference_known_answers():
    """test_virtual_round.cpp:134-184 ties (mantissa overload) and test_fp_params_round
    zero / NaN handling (356-370)."""
    f = oracle.virtual_round_mant_scalar
    one_half_ulp = 1.0 + 2.0 ** -11          # tie between 1 and 1 + 2^-10 at 10 bits
    assert f(one_half_ulp, 10, oracle.RNE) == 1.0
    assert f(one_half_ulp, 10, oracle.RNO) == 1.0 + 2.0 ** -10
    assert f(one_half_ulp, 10, oracle.RTA) == 1.0 + 2.0 ** -10
    assert f(one_half_ulp, 10, oracle.RTTZ) == 1.0
    three_half = 1.0 + 3 * 2.0 ** -11         # tie between 1 + 2^-10 and 1 + 2^-9
    assert f(three_half, 10, oracle.RNE) == 1.0 + 2.0 ** -9
    assert f(-three_half, 10, oracle.RDOWN) == -(1.0 + 2.0 ** -9)
    assert f(-three_half, 10, oracle.RUP) == -(1.0 + 2.0 ** -10)
    e4m3 = (8, 3, 7, 1, 0, 0)
    for mode in oracle.DETERMINISTIC_MODES:
        assert np.signbit(oracle.virtual_round_scalar(-0.0, e4m3, mode))
        assert oracle.virtual_round_scalar(0.0, e4m3, mode) == 0.0
        assert np.isnan(oracle.virtual_round_scalar(float("nan"), e4m3, mode))
    # directed modes below denorm_min reach +-denorm_min (2^-9), nearest ties go to 0
    assert oracle.virtual_round_scalar(2.0 ** -12, e4m3, oracle.RAFZ) == 2.0 ** -9
    assert oracle.virtual_round_scalar(2.0 ** -10, e4m3, oracle.RNE) == 0.0
    assert oracle.virtual_round_scalar(2.0 ** -10, e4m3, oracle.RTA) == 2.0 ** -9
    # saturating overflow returns the input, extended overflow returns +inf
    assert oracle.virtual_round_scalar(500.0, e4m3) == 500.0
    assert oracle.virtual_round_scalar(-1000.0, (8, 3, 7, 0, 0, 0)) == float("inf")


def test_exact_values_round_trip():
    """Representable values are fixed points of every deterministic mode."""
    g = np.random.default_rng(0)
    # bounds stay below each format's largest finite value: a saturating overflow returns
    # the unrounded input (lo_float.h:2936), which is no fixed point
    for key, bound in (("e4m3_sat", 240.0), ("e2m1", 4.0), ("fp16_like", 60000.0)):
        ft = FORMATS[key]
        x = oracle.virtual_round(g.standard_normal(20000).astype(np.float32) * 8, ft)
        x = x[np.abs(x) <= bound]
        for mode in oracle.DETERMINISTIC_MODES:
            assert_bits(oracle.virtual_round(x, ft, mode), x, f"{ft} mode {mode}")


# --------------------------------------------------------------------------- RTO FMA
FLT_MAX = Fraction(float(np.finfo(np.float32).max))


def _f32_floor(v: Fraction) -> np.float32:
    f = np.float32(float(v))
    while np.isfinite(f) and Fraction(float(f)) > v:
        f = np.nextafter(f, np.float32(-np.inf))
    while np.isfinite(np.nextafter(f, np.float32(np.inf))) and \
            Fraction(float(np.nextafter(f, np.float32(np.inf)))) <= v:
        f = np.nextafter(f, np.float32(np.inf))
    return f


def rto_exact(a, b, c) -> np.float32:
    a, b, c = np.float32(a), np.float32(b), np.float32(c)
    v = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    if v == 0:
        prod_pos_zero = (a == 0 or b == 0) and not (np.signbit(a) ^ np.signbit(b))
        return np.float32(0.0) if (prod_pos_zero and c == 0 and not np.signbit(c)) \
            else np.float32(-0.0)
    if v > FLT_MAX:
        return np.finfo(np.float32).max
    if v < -FLT_MAX:
        return -np.finfo(np.float32).max
    rd = _f32_floor(v)
    if Fraction(float(rd)) == v:
        return rd
    ru = np.nextafter(rd, np.float32(np.inf))
    rz = rd if rd >= 0 else ru
    bits = np.array([rz], np.float32).view(np.uint32) | np.uint32(1)
    return bits.view(np.float32)[0]


def fma_rn_exact(a, b, c) -> np.float32:
    """fmaf(a, b, c): one round-to-nearest-even rounding of the exact value."""
    a, b, c = np.float32(a), np.float32(b), np.float32(c)
    v = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    if v == 0:
        prod_neg = (a == 0 or b == 0) and bool(np.signbit(a) ^ np.signbit(b))
        return np.float32(-0.0) if (prod_neg and np.signbit(c)) else np.float32(0.0)
    rd = _f32_floor(v)
    if Fraction(float(rd)) == v:
        return rd
    ru = np.nextafter(rd, np.float32(np.inf))
    dlo, dhi = v - Fraction(float(rd)), Fraction(float(ru)) - v
    if dlo != dhi:
        return rd if dlo < dhi else ru
    return rd if (np.array([rd], np.float32).view(np.uint32)[0] & 1) == 0 else ru


def _triples(n, seed):
    g = np.random.default_rng(seed)
    a = (g.standard_normal(n) * np.exp2(g.uniform(-30, 30, n))).astype(np.float32)
    b = (g.standard_normal(n) * np.exp2(g.uniform(-30, 30, n))).astype(np.float32)
    c = (g.standard_normal(n) * np.exp2(g.uniform(-40, 40, n))).astype(np.float32)
    c[: n // 8] = -(a[: n // 8] * b[: n // 8])            # cancellations
    a[n // 8: n // 6] = 0.0
    b[n // 6: n // 5] = -0.0
    c[n // 5: n // 4] = 0.0
    c[n // 4: n // 4 + 20] = -0.0
    a[-30:] = 3e38
    b[-30:] = 7.0                                            # overflow
    a[-60:-30] = 1e-30
    b[-60:-30] = 1e-20                                       # underflow
    return a, b, c


def test_rto_fma_exact():
    a, b, c = _triples(3000, 1)
    for i in range(a.size):
        got = np.float32(oracle.rto_fma(a[i], b[i], c[i]))
        want = rto_exact(a[i], b[i], c[i])
        assert got.view(np.uint32) == want.view(np.uint32), (a[i], b[i], c[i], got, want)


def test_rto_fma_matches_fpu_mode_restatement():
    a, b, c = _triples(20000, 2)
    for i in range(a.size):
        x = np.float32(oracle.rto_fma(a[i], b[i], c[i])).view(np.uint32)
        y = np.float32(oracle.rto_fma_dir(a[i], b[i], c[i])).view(np.uint32)
        assert x == y, (a[i], b[i], c[i])


def test_round_to_odd_double_rounding_property():
    """RN_m(RTO24(x)) equals RN_m(x) for m <= 21 (round_to_odd.cu:30-46 analogue)."""
    a, b, c = _triples(3000, 3)
    for i in range(0, a.size, 3):
        v = Fraction(float(a[i])) * Fraction(float(b[i])) + Fraction(float(c[i]))
        if v == 0 or abs(v) > FLT_MAX / 2 or abs(v) < Fraction(2) ** -120:
            continue
        r = np.float32(oracle.rto_fma(a[i], b[i], c[i]))
        for m in (5, 10, 21):
            e = int(np.frexp(float(abs(v)))[1]) - 1
            q = Fraction(2) ** (e - m)
            k = abs(v) / q
            lo = k.numerator // k.denominator
            fr = k - lo
            kk = lo + 1 if (fr > Fraction(1, 2) or (fr == Fraction(1, 2) and lo % 2)) else lo
            exact_rn = float(kk * q) * (1 if v > 0 else -1)
            assert oracle.virtual_round_mant_scalar(float(r), m) == np.float32(exact_rn), (v, m)


# --------------------------------------------------------------------------- GEMM leaf
def _leaf_emulation(A, B, m, chunk=1):
    M, K = A.shape
    N = B.shape[1]
    out = np.empty((M, N), np.float32)
    for i in range(M):
        for j in range(N):
            acc = np.float32(0.0)
            for k0 in range(0, K, chunk):
                if chunk == 1:
                    r = rto_exact(A[i, k0], B[k0, j], acc)
                else:
                    p = np.float32(0.0)
                    for k in range(k0, min(k0 + chunk, K)):
                        p = fma_rn_exact(A[i, k], B[k, j], p)
                    r = rto_exact(p, np.float32(1.0), acc)
                acc = cf_round_mant(np.array([r], np.float32), m, oracle.RNE)[0]
            out[i, j] = acc
    return out


@pytest.mark.parametrize("chunk", [1, 3])
def test_gemm_leaf_vs_exact_emulation(chunk):
    g = np.random.default_rng(chunk)
    A = (g.standard_normal((3, 17)) * np.exp2(g.uniform(-4, 4, (3, 17)))).astype(np.float32)
    B = (g.standard_normal((17, 4)) * np.exp2(g.uniform(-4, 4, (17, 4)))).astype(np.float32)
    A[0, 3] = 0.0
    B[5, 1] = -0.0
    for m in (3, 10, 23):
        got = oracle.gemm(A, B, oracle.accum_mant(m, chunk))
        assert_bits(got, _leaf_emulation(A, B, m, chunk), f"m={m} chunk {chunk}")


def test_gemm_gamma_k_bound_and_modes():
    """test_blis_lof_accum.cpp:120-177: |D - AB| <= gamma_K |A||B| for a 10-bit accumulator
    in every deterministic mode; a 23-bit accumulator is far more accurate."""
    g = np.random.default_rng(13)
    M, N, K = 32, 32, 64
    A = (g.random((M, K)) - 0.5).astype(np.float32)
    B = (g.random((K, N)) - 0.5).astype(np.float32)
    exact = A.astype(np.float64) @ B.astype(np.float64)
    absab = np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)
    keps = K * 2.0 ** -10
    gamma = keps / (1 - keps)
    errs = {}
    for mode in (oracle.RNE, oracle.RTZ, oracle.RAFZ, oracle.RNO, oracle.RDOWN, oracle.RUP,
                 oracle.RTA):
        D = oracle.gemm(A, B, oracle.accum_mant(10, 1, mode))
        err = np.abs(D - exact)
        assert (err <= gamma * absab + 1e-12).all(), mode
        errs[mode] = err.max()
    D23 = oracle.gemm(A, B, oracle.accum_mant(23, 1))
    assert np.abs(D23 - exact).max() * 100 < errs[oracle.RNE]


def test_gemm_exact_inputs_mode_independent():
    """test_blis_lof_accum.cpp:200-245: exactly representable partial sums give the exact
    product in every accumulator mode."""
    g = np.random.default_rng(7)
    A = (g.integers(-8, 8, (16, 24)) / 8).astype(np.float32)
    B = (g.integers(-8, 8, (24, 12)) / 4).astype(np.float32)
    exact = (A.astype(np.float64) @ B.astype(np.float64)).astype(np.float32)
    for mode in oracle.DETERMINISTIC_MODES:
        for chunk in (1, 8):
            assert_bits(oracle.gemm(A, B, oracle.accum_mant(10, chunk, mode)), exact,
                        f"mode {mode} chunk {chunk}")


# --------------------------------------------------------------------------- MX
def _mx_closed_form(x, bs, ef, sf, mode=0, smode=0, L=0, u=None):
    x = np.asarray(x, np.float32)
    u = np.zeros(x.size, np.uint32) if u is None else u
    bw, mb, bias, _, _, sgn = ef
    me = ((1 << (bw - mb - 1)) if sgn == 0 else (1 << (bw - mb))) - 1 - bias
    pmn = np.float32(np.float32(2.0 ** me) * np.float32(2.0 - 2.0 ** -mb))
    out = np.empty_like(x)
    for b0 in range(0, x.size, bs):
        blk = x[b0:b0 + bs]
        a = np.abs(blk)
        amax = np.float32(a[~np.isnan(a)].max()) if (~np.isnan(a)).any() else np.float32(0)
        amax = max(amax, np.float32(0))
        with np.errstate(all="ignore"):
            t = np.float32(amax) / pmn
        scale = cf_round(np.array([t], np.float32), sf, smode)[0]
        if not scale > 0:
            out[b0:b0 + bs] = 0.0
            continue
        with np.errstate(all="ignore"):
            q = (blk / scale).astype(np.float32)
        out[b0:b0 + bs] = cf_round(q, ef, mode, L, u[b0:b0 + bs]) * scale
    return out


@pytest.mark.parametrize("bs", [4, 16, 32, 33, 256])
def test_mx_vs_closed_form(bs):
    x = c1_values(4099, seed=bs)
    E4M3, E2M1, E8M0 = (8, 3, 7, 1, 0, 0), (4, 1, 1, 1, 0, 0), (8, 0, 127, 0, 0, 1)
    u = random_words(x.size, seed=1)
    for ef, sf in ((E4M3, E8M0), (E2M1, E8M0), (E4M3, (8, 2, 15, 1, 0, 0))):
        for mode, L in ((0, 0), (5, 0), (12, 0), (3, 6), (10, 0)):
            got = oracle.virtual_mx_round(x, bs, ef, sf, mode, 0, L, bits=u)
            assert_bits(got, _mx_closed_form(x, bs, ef, sf, mode, 0, L, u), f"bs {bs} {mode}")


def test_mx_scale_is_nearest_e8m0_neighbour():
    """test_virtual_mx_round.cpp:151-193: the shared scale equals the RoundDown or RoundUp
    E8M0 projection of amax / pmn, whichever is nearer."""
    E8M0 = (8, 0, 127, 0, 0, 1)
    g = np.random.default_rng(3)
    t = (g.random(20000) * np.exp2(g.integers(-40, 40, 20000))).astype(np.float32)
    s = oracle.virtual_round(t, E8M0, oracle.RNE)
    dn = oracle.virtual_round(t, E8M0, oracle.RDOWN)
    up = oracle.virtual_round(t, E8M0, oracle.RUP)
    assert ((s == dn) | (s == up)).all()
    d = np.abs(s.astype(np.float64) - t)
    assert (d <= np.minimum(np.abs(dn - t.astype(np.float64)),
                            np.abs(up - t.astype(np.float64))) * (1 + 1e-12)).all()


def test_mx_known_blocks():
    m = 448.0
    exact = np.array([m, m / 2, -m / 4, 1.0, 0.5, -2.0, 8.0, 0.0] * 4, np.float32)
    assert_bits(oracle.virtual_mx_round(exact, 32, (8, 3, 8, 1, 0, 0), (8, 0, 127, 0, 0, 1)),
                exact, "exact block")
    z = np.zeros(64, np.float32)
    z[32:] = -0.0
    y = oracle.virtual_mx_round(z, 32, (8, 3, 7, 1, 0, 0), (8, 0, 127, 0, 0, 1))
    assert (y == 0).all() and not np.signbit(y).any()


# --------------------------------------------------------------------------- fixtures
def test_golden_fixtures_reproduce():
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    import sys
    sys.path.insert(0, here)
    import make_fixtures
    fx = np.load(os.path.join(here, "round_fixtures.npz"))
    for key, (kind, args) in make_fixtures.cases().items():
        assert_bits(make_fixtures.compute(fx, kind, args), fx[key], key)
