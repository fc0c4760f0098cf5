"""Pinning the bit-pack codec oracle (oracle/lof_oracle.c, orc_pack_* / orc_float_to_rep /
orc_rep_to_float, restating lo_pack.hpp) by checks independent of its code:

* group geometry against an independent lcm formula (test_pack.cpp:30-39's oracle);
* decode, exhaustively over every rep of 11 formats, against a closed form computed here
  with exact rationals (sign / exponent / mantissa fields, checker-defined Inf and NaN);
* encode against torch's float8_e4m3fn, float8_e5m2, float16 and bfloat16 conversions on
  the ranges where those types and the LoFloat presets agree;
* encode as the inverse of decode on every non-NaN rep (the reference's test_encode_codec
  grid values), in every deterministic mode;
* decode(encode(x)) against the pinned virtual_round oracle for in-range values;
* pack / unpack as a bijection on reps, both group orders, partial groups, byte counts;
* special inputs (NaN, +-inf under each saturation mode, -0, negatives into unsigned).
"""
import math
from fractions import Fraction

import numpy as np
import pytest
import torch

import oracle

# (name, format tuple, checker preset) -- the formats of test_pack.cpp:224-238
# (param_float_p_3109<k,p> has bias 2^(k-p-1); e2m1 / e3m2 / e5m2 / e4m3fn use the
# OCP_F8E4M3 checkers, lo_float.h:1521-1666) plus half and bf16 (fp_tools.hpp:361-407).
FORMATS = [
    ("p3109<3,2>", (3, 1, 1, 1, 0, 0), ("p3109", 3, True, False)),
    ("e2m1", (4, 1, 1, 1, 0, 0), "ocp_fn"),
    ("p3109<5,2>", (5, 1, 4, 1, 0, 0), ("p3109", 5, True, False)),
    ("e3m2", (6, 2, 3, 1, 0, 0), "ocp_fn"),
    ("p3109<7,3>", (7, 2, 8, 1, 0, 0), ("p3109", 7, True, False)),
    ("e5m2", (8, 2, 15, 1, 0, 0), "ocp_fn"),
    ("e4m3fn", (8, 3, 7, 1, 0, 0), "ocp_fn"),
    ("p3109<9,3>", (9, 2, 32, 1, 0, 0), ("p3109", 9, True, False)),
    ("p3109<11,5>", (11, 4, 32, 1, 0, 0), ("p3109", 11, True, False)),
    ("p3109<12,8>", (12, 7, 8, 1, 0, 0), ("p3109", 12, True, False)),
    ("half", (16, 10, 15, 0, 0, 0), "fp16"),
    ("bf16", (16, 7, 127, 0, 0, 0), "bf16"),
    ("p3109<8,4>ext", (8, 3, 8, 0, 0, 0), ("p3109", 8, True, True)),
    ("u_p3109<6,3>", (6, 2, 4, 1, 0, 1), ("p3109", 6, False, False)),
    ("ieee_f8", (8, 3, 7, 0, 0, 0), ("ieee_f8", True)),
]
DET = oracle.DETERMINISTIC_MODES


def _pf(fmt, preset, ub=0):
    return oracle.pack_format(fmt, preset, ub)


def _bits_eq(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    same = a.view(np.uint32) == b.view(np.uint32)
    return same | (np.isnan(a) & np.isnan(b))


# ------------------------------------------------------------------------ geometry
def _lcm(a, b):
    return a * b // math.gcd(a, b)


@pytest.mark.parametrize("n", range(1, 17))
def test_group_geometry(n):
    align = 8 if n <= 8 else 16
    gbits = _lcm(n, align)
    gc, gb = gbits // n, gbits // 8
    for count in (0, 1, gc - 1, gc, gc + 1, 3 * gc + 2, 1000):
        if count < 0:
            continue
        assert oracle.packed_bytes(n, count) == -(-count // gc) * gb


# ------------------------------------------------------------------------ decode
def _matches(chk, rep, nan):
    if chk.kind == 1:
        return rep in list(chk.codes)[:chk.ncodes]
    if chk.kind == 2:
        return (rep & chk.a) == chk.a and ((rep & chk.b) != 0 if nan else (rep & chk.b) == 0)
    if chk.kind == 3:
        return (rep | chk.a) == chk.codes[0]
    return False


def _decode_closed_form(rep, fmt, pf):
    n, mb, bias, inf_b, _, signedness = fmt
    signed = signedness == 0
    rep &= (1 << n) - 1
    neg = signed and (rep >> (n - 1)) & 1
    mag = rep & ((1 << (n - 1)) - 1) if signed else rep
    if inf_b != 1 and _matches(pf.is_inf, rep, False):
        return -math.inf if neg else math.inf
    if _matches(pf.is_nan, rep, True):
        return math.nan
    e, m = mag >> mb, mag & ((1 << mb) - 1)
    v = Fraction(m, 1 << mb) * Fraction(2) ** (1 - bias) if e == 0 else \
        (1 + Fraction(m, 1 << mb)) * Fraction(2) ** (e - bias)
    f = float(np.float32(float(v)))   # values here are exact in FP32
    return -f if neg else f


@pytest.mark.parametrize("name,fmt,preset", FORMATS, ids=[f[0] for f in FORMATS])
def test_decode_exhaustive_closed_form(name, fmt, preset):
    pf = _pf(fmt, preset)
    reps = np.arange(1 << fmt[0], dtype=np.uint32)
    got = oracle.rep_to_float(reps, pf)
    want = np.array([_decode_closed_form(int(r), fmt, pf) for r in reps], np.float32)
    ok = _bits_eq(got, want)
    assert ok.all(), (name, reps[~ok][:5], got[~ok][:5], want[~ok][:5])


# ------------------------------------------------------------------------ encode
def _probe_values(fmt, pf):
    """Every representable value, its FP32 neighbours and the midpoints between adjacent
    values (test_pack.cpp:69-113), plus extras."""
    reps = np.arange(1 << fmt[0], dtype=np.uint32)
    vals = oracle.rep_to_float(reps, pf)
    fin = np.unique(vals[np.isfinite(vals)])
    up = np.nextafter(fin, np.float32(np.inf))
    dn = np.nextafter(fin, np.float32(-np.inf))
    mids = ((fin[1:].astype(np.float64) + fin[:-1]) / 2).astype(np.float32)
    extra = np.array([0.0, -0.0, 1e30, -1e30, np.inf, -np.inf, 1e-40, -1e-40, 1e-45],
                     np.float32)
    return np.concatenate([fin, up, dn, mids, extra]).astype(np.float32)


@pytest.mark.parametrize("name,fmt,preset", FORMATS, ids=[f[0] for f in FORMATS])
def test_encode_inverts_decode(name, fmt, preset):
    """Every non-NaN, non-Inf rep r: encode(decode(r)) == r in every deterministic mode
    (representable values are fixed points), except the negative zero pattern of formats
    whose -0 code is NaN."""
    pf = _pf(fmt, preset)
    reps = np.arange(1 << fmt[0], dtype=np.uint32)
    vals = oracle.rep_to_float(reps, pf)
    keep = np.isfinite(vals)
    for mode in DET:
        back = oracle.float_to_rep(vals[keep], pf, mode)
        r = reps[keep]
        # several reps can decode to the same value (+-0); compare values then
        ok = (back == r) | _bits_eq(oracle.rep_to_float(back, pf), vals[keep])
        assert ok.all(), (name, mode, r[~ok][:5], back[~ok][:5])


TORCH_CASES = [
    # (format, preset, torch dtype, |x| limit where the two agree)
    ((8, 3, 7, 1, 0, 0), "ocp_fn", torch.float8_e4m3fn, 448.0),
    ((8, 2, 15, 1, 0, 0), "ocp_fn", torch.float8_e5m2, 57344.0),
    ((16, 10, 15, 0, 0, 0), "fp16", torch.float16, 65504.0),
    ((16, 7, 127, 0, 0, 0), "bf16", torch.bfloat16, 3.3e38),
]


@pytest.mark.parametrize("fmt,preset,dtype,lim", TORCH_CASES,
                         ids=["e4m3fn", "e5m2", "fp16", "bf16"])
def test_encode_matches_torch(fmt, preset, dtype, lim):
    pf = _pf(fmt, preset)
    g = np.random.default_rng(fmt[0] * 31 + fmt[1])
    x = (g.standard_normal(200000) * np.exp2(g.uniform(-40, 40, 200000))).astype(np.float32)
    x = np.concatenate([x, _probe_values(fmt, pf)])
    x = x[np.isfinite(x) & (np.abs(x) <= lim)]
    r = oracle.float_to_rep(x, pf)
    t = torch.from_numpy(x).to(dtype).view(torch.uint8 if fmt[0] == 8 else torch.int16)
    t = t.numpy().astype(np.uint16 if fmt[0] == 16 else np.uint8).astype(np.uint32)
    assert np.array_equal(r, t), (x[r != t][:5], r[r != t][:5], t[r != t][:5])


@pytest.mark.parametrize("name,fmt,preset", FORMATS[:12], ids=[f[0] for f in FORMATS[:12]])
def test_encode_decode_equals_virtual_round(name, fmt, preset):
    """Inside the format's finite range, decode(encode(x)) is x rounded onto the format's
    grid: the pinned virtual_round oracle, in the sign-symmetric modes."""
    pf = _pf(fmt, preset)
    reps = np.arange(1 << fmt[0], dtype=np.uint32)
    vals = oracle.rep_to_float(reps, pf)
    top = float(np.max(np.abs(vals[np.isfinite(vals)])))
    g = np.random.default_rng(fmt[0])
    x = (g.uniform(-1, 1, 60000) * top).astype(np.float32)
    x = np.concatenate([x, _probe_values(fmt, pf)])
    x = x[np.isfinite(x) & (np.abs(x) <= top) & (np.abs(x) >= np.float32(2.0 ** -126))]
    for mode in (oracle.RNE, oracle.RTZ, oracle.RAFZ, oracle.RNO, oracle.RTA, oracle.RTO,
                 oracle.RTTZ):
        dec = oracle.rep_to_float(oracle.float_to_rep(x, pf, mode), pf)
        vr = oracle.virtual_round(x, fmt, mode)
        # a value that lands on a NaN code decodes as NaN (e.g. -max of e4m3fn is 0xFF)
        ok = _bits_eq(dec, vr) | np.isnan(dec) | ((dec == 0) & (vr == 0))
        assert ok.all(), (name, mode, x[~ok][:5], dec[~ok][:5], vr[~ok][:5])


def test_special_inputs():
    e4 = _pf((8, 3, 8, 0, 0, 0), ("p3109", 8, True, True))        # has Inf: 0x7F / 0xFF
    assert oracle.float_to_rep([np.nan], e4)[0] == 0x80           # qNaN pattern
    assert oracle.float_to_rep([np.inf], e4, saturation=0)[0] == 0x7F
    assert oracle.float_to_rep([-np.inf], e4, saturation=0)[0] == 0xFF
    assert oracle.float_to_rep([np.inf], e4, saturation=1)[0] == 0x7E   # SatFinite: max finite
    assert oracle.float_to_rep([1e30], e4, saturation=0)[0] == 0x7F     # overflow -> Inf
    assert oracle.float_to_rep([1e30], e4, saturation=1)[0] == 0x7E
    # -0 into a format whose sign-only code is NaN stays +0 (lo_pack.hpp:151-155)
    assert oracle.float_to_rep([-0.0], e4)[0] == 0x00
    half = _pf((16, 10, 15, 0, 0, 0), "fp16")
    assert oracle.float_to_rep([-0.0], half)[0] == 0x8000
    # negatives into an unsigned format: NegtoZero -> 0, NegtoNaN -> the NaN pattern
    u = (6, 2, 4, 1, 0, 1)
    assert oracle.float_to_rep([-1.5], _pf(u, ("p3109", 6, False, False), 0))[0] == 0
    assert oracle.float_to_rep([-1.5], _pf(u, ("p3109", 6, False, False), 1))[0] == 63


# ------------------------------------------------------------------------ packing
@pytest.mark.parametrize("n", range(1, 17))
@pytest.mark.parametrize("endian", [0, 1])
def test_pack_reps_bijection(n, endian):
    g = np.random.default_rng(n * 7 + endian)
    align = 8 if n <= 8 else 16
    gc = _lcm(n, align) // n
    for count in (1, gc, 2 * gc + 1, 257):
        reps = g.integers(0, 1 << n, count, dtype=np.uint64).astype(np.uint32)
        buf = oracle.pack_reps(reps, n, endian)
        assert buf.size == oracle.packed_bytes(n, count)
        assert np.array_equal(oracle.unpack_reps(buf, count, n, endian), reps)
        # forward order is the plain LSB-first bit stream of the n-bit fields
        if endian == 0:
            bits = np.unpackbits(buf, bitorder="little")
            fields = [int("".join(str(b) for b in bits[i * n:(i + 1) * n][::-1]), 2)
                      for i in range(count)]
            assert fields == reps.tolist()


@pytest.mark.parametrize("n", [3, 5, 9, 12])
def test_endianness_reverses_slots(n):
    align = 8 if n <= 8 else 16
    gc = _lcm(n, align) // n
    reps = np.arange(1, 2 * gc + 1, dtype=np.uint32) & ((1 << n) - 1)
    fwd = oracle.pack_reps(reps, n, 0)
    rev = oracle.pack_reps(reps, n, 1)
    assert not np.array_equal(fwd, rev)
    # reversing each group's elements makes the reverse buffer from the forward order
    perm = np.concatenate([np.arange(gc)[::-1] + k * gc for k in range(2)])
    assert np.array_equal(oracle.pack_reps(reps[perm], n, 0), rev)


@pytest.mark.parametrize("name,fmt,preset", FORMATS, ids=[f[0] for f in FORMATS])
def test_float_encode_decode_arrays(name, fmt, preset):
    """The array encoder packs float_to_rep of each element (a partial group is padded
    with encodings of 0.0f) and the decoder returns rep_to_float of each field."""
    pf = _pf(fmt, preset)
    x = np.concatenate([np.arange(-8, 8.125, 0.125, dtype=np.float32),
                        np.array([0.0, -0.0, 1e30, -1e30], np.float32)])
    n = fmt[0]
    for endian in (0, 1):
        buf = oracle.pack_encode(x, pf, endian=endian)
        reps = oracle.float_to_rep(x, pf)
        assert np.array_equal(buf, oracle.pack_reps(reps, n, endian))
        dec = oracle.pack_decode(buf, x.size, pf, endian)
        assert _bits_eq(dec, oracle.rep_to_float(reps, pf)).all()


def test_stochastic_encode_with_words():
    """Stochastic modes consume element i's word: explicit words and the counter-based
    generator give the same bytes, and the probability of rounding up matches the tail."""
    pf = _pf((8, 3, 7, 1, 0, 0), "ocp_fn")
    x = np.full(1 << 16, 1.0 + 3 * 2.0 ** -6, np.float32)        # 3/8 of an ulp above 1
    w = oracle.rand_u32(99, x.size)
    a = oracle.pack_encode(x, pf, oracle.SRA, stoch_len=8, bits=w)
    b = oracle.pack_encode(x, pf, oracle.SRA, stoch_len=8, seed=99)
    assert np.array_equal(a, b)
    up = np.mean(oracle.float_to_rep(x, pf, oracle.SRA, stoch_len=8, bits=w) == 0x39)
    assert abs(up - 0.375) < 0.01


@pytest.mark.parametrize("name,fmt,preset", FORMATS[:12], ids=[f[0] for f in FORMATS[:12]])
def test_stochastic_encode_equals_virtual_round(name, fmt, preset):
    """In the format's normal range the codec's stochastic rounding is the pinned
    virtual_round oracle's, word for word, in all six stochastic modes.  (Below the normal
    range the two reference functions align the significand differently; that range is
    pinned by test_stochastic_encode_closed_form.)"""
    pf = _pf(fmt, preset)
    reps = np.arange(1 << fmt[0], dtype=np.uint32)
    fin = oracle.rep_to_float(reps, pf)
    fin = fin[np.isfinite(fin)]
    top = float(np.max(np.abs(fin)))
    g = np.random.default_rng(fmt[0] + 100)
    x = (g.uniform(-1, 1, 60000) * top).astype(np.float32)
    x = np.concatenate([x, _probe_values(fmt, pf)])
    x = x[np.isfinite(x) & (np.abs(x) <= top) & (np.abs(x) >= max(2.0 ** (1 - fmt[2]), 2.0 ** -126))]
    for mode in oracle.STOCHASTIC_MODES:
        L = 0 if mode in (oracle.TRUESR, oracle.PROB) else min(5, 23 - fmt[1])
        w = oracle.rand_u32(7 + mode, x.size)
        dec = oracle.rep_to_float(oracle.float_to_rep(x, pf, mode, stoch_len=L, bits=w), pf)
        vr = oracle.virtual_round(x, fmt, mode, L, bits=w)
        on_grid = np.isin(np.abs(vr), np.abs(fin))   # overflow passes through unrounded
        ok = _bits_eq(dec, vr) | np.isnan(dec) | ~on_grid
        assert ok.all(), (name, mode, x[~ok][:5], dec[~ok][:5], vr[~ok][:5])
        assert on_grid.mean() > 0.9


def _stochastic_up(mode, f, L, u):
    """Whether a stochastic mode rounds a magnitude with fractional part f (0 <= f < 1, in
    units of the target's last place) up, given the element's 32-bit word u: the mode
    definitions (R = u mod 2^L; SR-D adds bit L of u; TrueSR compares u with f * 2^32;
    Probabilistic takes u's low bit)."""
    R = u & ((1 << L) - 1) if L > 0 else 0
    if mode == oracle.SRA:
        return L > 0 and math.floor(f * 2 ** L) + R >= 2 ** L
    if mode == oracle.SRB:
        return f > 0 and math.floor(f * 2 ** L) + R + 1 >= 2 ** L
    if mode == oracle.SRC:
        return L > 0 and f > 0 and math.floor(f * 2 ** (L - 1)) + (R >> 1) + (R & 1) >= 2 ** (L - 1)
    if mode == oracle.SRD:
        return L > 0 and math.floor(f * 2 ** L) + R + ((u >> L) & 1) >= 2 ** L
    if mode == oracle.TRUESR:
        return u < f * 2 ** 32
    if mode == oracle.PROB:
        return f > 0 and (u & 1) == 1
    raise ValueError(mode)


@pytest.mark.parametrize("name,fmt,preset", [FORMATS[i] for i in (2, 5, 6, 8, 10)],
                         ids=[FORMATS[i][0] for i in (2, 5, 6, 8, 10)])
def test_stochastic_encode_closed_form(name, fmt, preset):
    """Below the format's normal range (bias < 126, so these are the reference's
    'fewer exponents near zero' branch), including FP32-subnormal inputs and values far
    below the smallest subnormal: the encoded value is sign * (floor(q) + up) * 2^(1-bias-mb)
    with q = |x| / 2^(1-bias-mb) in exact rationals and `up` from the mode's definition
    (SR-D's coin can add a whole unit to an exact multiple, as in virtual_round)."""
    pf = _pf(fmt, preset)
    n, mb, bias = fmt[0], fmt[1], fmt[2]
    unit = Fraction(1, 2 ** (bias - 1 + mb))
    g = np.random.default_rng(n * 13 + mb)
    mags = np.exp2(g.uniform(-152, 1 - bias, 1500)).astype(np.float32)
    mags = np.concatenate([mags, np.float32(unit) * np.array([0.5, 0.25, 1.5, 3.0], np.float32),
                           np.array([1e-45, 3e-44, 1e-40], np.float32)])
    x = (mags * np.where(g.random(mags.size) < 0.5, -1, 1)).astype(np.float32)
    for mode in oracle.STOCHASTIC_MODES:
        for L in ((0,) if mode in (oracle.TRUESR, oracle.PROB) else (1, 5, 23 - mb)):
            w = oracle.rand_u32(1000 + 31 * mode + L, x.size)
            dec = oracle.rep_to_float(oracle.float_to_rep(x, pf, mode, stoch_len=L, bits=w), pf)
            for xi, di, ui in zip(x.tolist(), dec.tolist(), w.tolist()):
                q = Fraction(abs(xi)) / unit
                fl = math.floor(q)
                # an exact zero encodes as zero before any rounding (lo_pack.hpp:176)
                want = 0 if xi == 0 else (fl + _stochastic_up(mode, q - fl, L, int(ui))) * unit
                assert Fraction(abs(di)) == want and (di == 0 or (di < 0) == (xi < 0)), \
                    (name, mode, L, xi, di, float(want))


def test_stochastic_encode_unbiased_below_normal():
    """TrueSR is unbiased: the mean of decode(encode(x)) over many words is x, also for
    values below e4m3's smallest subnormal."""
    pf = _pf((8, 3, 7, 1, 0, 0), "ocp_fn")
    for v in (2.0 ** -10, 3 * 2.0 ** -12, 0.7 * 2.0 ** -9, 5.3 * 2.0 ** -9):
        x = np.full(1 << 16, v, np.float32)
        dec = oracle.rep_to_float(oracle.float_to_rep(x, pf, oracle.TRUESR, seed=5), pf)
        sd = 2.0 ** -9 / 2 / math.sqrt(x.size)
        assert abs(dec.astype(np.float64).mean() - float(np.float32(v))) < 5 * sd, v


def _p3109_max(l, p, ext):
    """Largest finite value of signed P3109<l, p> from its layout: the top code below the
    sign bit, minus one when it is Inf (Extended); bias 2^(l-p-1), p-1 mantissa bits."""
    code = (1 << (l - 1)) - 1 - (1 if ext else 0)
    mb, bias = p - 1, 1 << (l - p - 1)
    E, M = code >> mb, code & ((1 << mb) - 1)
    if E == 0:
        return math.ldexp(M, 1 - bias - mb)
    return math.ldexp((1 << mb) + M, E - bias - mb)


@pytest.mark.parametrize("l", range(2, 9))
def test_saturation_modes_sweep(l):
    """The reference's test_sat_mode.cpp sweep (test_sat_mode.cpp:38-150) on the typed cast:
    every P3109<l, p>, p = 1..l-1, Saturating and Extended.  Finite overflow: SatFinite and
    SatPropagate give +-max, OvfInf +-inf where the format has a real infinity (else +-max);
    +-inf input: SatFinite +-max, OvfInf / SatPropagate +-inf (else +-max); 1.0 is the same
    under all three modes."""
    for p in range(1, l):
        for ext in (False, True):
            pf = _pf((l, p - 1, 1 << (l - p - 1), 0 if ext else 1, 0, 0), ("p3109", l, True, ext))
            maxf = _p3109_max(l, p, ext)
            if not maxf > 0.0:
                continue   # degenerate (e.g. <2,1> Extended): no finite nonzero value
            inf_code = (1 << (l - 1)) - 1
            real_inf = ext and math.isinf(float(oracle.rep_to_float([inf_code], pf)[0]))
            dec = lambda x, sat: float(oracle.rep_to_float(      # noqa: E731
                oracle.float_to_rep(np.array([x], np.float32), pf, oracle.RNE, sat), pf)[0])
            for m in (2.0, 8.0, 137.0, 1e6):
                for sign in (1.0, -1.0):
                    x = sign * maxf * m
                    if not math.isfinite(np.float32(x)):
                        continue
                    assert dec(x, 1) == sign * maxf and dec(x, 2) == sign * maxf, (l, p, ext, x)
                    assert dec(x, 0) == (sign * math.inf if real_inf else sign * maxf), (l, p, ext, x)
            for sign in (1.0, -1.0):
                assert dec(sign * math.inf, 1) == sign * maxf
                want = sign * math.inf if real_inf else sign * maxf
                assert dec(sign * math.inf, 0) == want and dec(sign * math.inf, 2) == want
            r = [dec(1.0, s) for s in (0, 1, 2)]
            assert r[0] == r[1] == r[2] and math.isfinite(r[0])
