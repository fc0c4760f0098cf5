"""CPU checks of the arithmetic behind the MFMA tile's two short leaves, against the oracle's
reference leaf acc <- round(RTO_fma(p, 1, acc)) (cutlass_gemms.cuh:73-88, 139-165):

- exact-sum leaves: the per-slice, per-lane test (slice statistics of the products, the
  lane's accumulators, the growth bound) is re-derived here in numpy.  In every slice that
  passes, every sum acc + p must be exact in FP32, and the 3-op integer round-to-nearest-even
  of -(acc + p) must give the reference leaf, negated.
- RTO leaves: from t = RN24(nacc - p) and its TwoSum residual, the rebuilt round-to-odd value,
  rounded by the integer RNE (mantissa-only accumulators) or by binary16 conversion
  (binary16-shaped ones), must equal the reference leaf for any in-range sum.

Random formats, scales, start accumulators and chunks are drawn from a fixed seed, and the
sizes are kept small so the file runs in seconds."""
import math

import numpy as np

import oracle

F32 = np.float32
E4M3 = (8, 3, 7, 1, 0, 0)


def _bits(x):
    return int(np.array([x], np.float32).view(np.uint32)[0])


def _float(b):
    return np.array([b & 0xFFFFFFFF], np.uint32).view(np.float32)[0]


def _rne(b, rf):
    """Round-to-nearest-even of a float pattern at bit rf, sign kept."""
    return (b + (1 << (rf - 1)) - 1 + ((b >> rf) & 1)) & ~((1 << rf) - 1) & 0xFFFFFFFF


def _ref_leaf(p, acc, fmt, m):
    x = oracle.rto_fma(float(p), 1.0, float(acc))
    return F32(oracle.virtual_round_mant_scalar(x, m) if fmt is None
               else oracle.virtual_round_scalar(x, fmt))


def _slice_stats(v, key):
    """amax exponent, smallest-nonzero exponent check, widest significand and lowest-bit
    bound of a slice operand, as the tile's staging gathers them."""
    b = np.asarray(v, np.float32).ravel().view(np.uint32).astype(np.int64)
    mag = b & 0x7FFFFFFF
    if mag.max() == 0:
        return None
    mn = int(mag[mag != 0].min())
    orb = int(np.bitwise_or.reduce(b))
    o = (orb & 0x7FFFFF) | 0x800000
    ctz = (o & -o).bit_length() - 1
    lsb = (mn >> 23) - 150 + ctz
    if key:   # per value: E + min(ctz, 24), E from (|v| << 1) - 1
        nz = ((mag << 1) - 1) & 0xFFFFFFFF
        bb = b | 0x80000000
        ctzs = np.array([min((int(x) & -int(x)).bit_length() - 1, 24) for x in bb])
        lsb = max(lsb, int(((nz >> 24) + ctzs).min()) - 150)
    return dict(emax=(int(mag.max()) >> 23) - 127, normal=(mn >> 23) != 0, width=24 - ctz,
                lsb=lsb)


def _growth(m, chunk):
    n = 1 if chunk >= 32 else 32 // chunk
    gr = (1 + 2.0 ** -(m + 1)) ** n
    gf = F32(gr)
    if float(gf) < gr:
        gf = np.nextafter(gf, F32(1e30))
    return gf, n


def _bound_exponent(ea, sTop, m, chunk):
    gf, n = _growth(m, chunk)
    pa = F32(2.0 ** max(ea, -126)) if ea > -1000 else F32(0)
    pp = F32(2.0 ** max(sTop, -126)) if sTop > -1000 else F32(0)
    with np.errstate(over="ignore"):
        inner = F32(np.float64(n) * np.float64(pp) + np.float64(pa))
        bnd = F32(np.float64(gf) * np.float64(inner) + np.float64(pp))
        bnd = F32(bnd * F32(1.0000152587890625))
    return (_bits(bnd) >> 23) - 126


def test_exact_sum_leaves():
    g = np.random.default_rng(2024)
    passed = checked = 0
    for _ in range(60):
        if g.random() < 0.3:
            fmt, m, q, ehi = None, int(g.integers(1, 20)), -149, 127
        else:
            m, ew = int(g.integers(1, 20)), int(g.integers(2, 9))
            bias = int(g.integers(-2, 2 ** (ew - 1) + 20))
            fmt = (1 + ew + m, m, bias, 1, 0, 0)
            q = max(1 - bias - m, -149)
            ehi = max(min((1 << ew) - 1 - bias, 127), max(1 - bias, -126))
        chunk = int(g.choice([1, 2, 4, 8, 16, 32]))
        lanes, K = 8, 64
        sc = 2.0 ** g.integers(-12, 6)
        A = (g.standard_normal((lanes, K)) * sc * 2.0 ** g.integers(-3, 4, (lanes, 1))).astype(F32)
        B = (g.standard_normal(K) * sc * np.where(g.random(K) < 0.1, 2.0 ** g.integers(-10, 6, K), 1)).astype(F32)
        A = oracle.virtual_round(A.ravel(), E4M3).reshape(lanes, K)
        B = oracle.virtual_round(B, E4M3)
        A[g.random((lanes, K)) < 0.1] = 0
        acc0 = (g.standard_normal(lanes) * sc * sc * 2.0 ** g.integers(-8, 10, lanes)).astype(F32)
        acc0[g.random(lanes) < 0.2] = 0
        acc = (oracle.virtual_round_mant(acc0, m) if fmt is None else oracle.virtual_round(acc0, fmt)).astype(F32)
        if not np.all(np.isfinite(acc)):
            continue
        rf = 23 - m
        for k0 in range(0, K, 32):
            sa, sb = _slice_stats(A[:, k0:k0 + 32], chunk <= 4), _slice_stats(B[k0:k0 + 32], chunk <= 4)
            if sa is None or sb is None:
                exact, sL, sTop = True, 1000, -1000
            else:
                L = sa["lsb"] + sb["lsb"]
                if chunk == 1:
                    sTop = sa["emax"] + sb["emax"] + 2
                    exact = (sa["normal"] and sb["normal"] and sa["width"] + sb["width"] <= 24
                             and sTop <= 127 and L >= -149)
                else:
                    sTop = int(math.log2(chunk)) + sa["emax"] + sb["emax"] + 2
                    exact = sa["normal"] and sb["normal"] and L >= -126 and sTop <= L + 24 and sTop <= 127
                sL = L
            ok = False
            if exact and sL >= max(q, -126):
                mag = acc.view(np.uint32).astype(np.int64) & 0x7FFFFFFF
                ea = (int(mag.max()) >> 23) - 126 if mag.max() else -1000
                nz = mag[mag != 0]
                lsb_acc = max(q, (int(nz.min()) >> 23) - 127 - m) if nz.size else 1000
                Lf = min(sL, lsb_acc)
                S = _bound_exponent(ea, sTop, m, chunk)
                ok = S <= min(sL + 24, ehi) and sTop <= Lf + 23 and Lf >= -126
            passed += ok
            for j in range(0, 32, chunk):
                for i in range(lanes):
                    p = F32(0.0)
                    for k in range(k0 + j, k0 + j + chunk):   # the FP32 fmaf chain
                        p = F32(np.float64(A[i, k]) * np.float64(B[k]) + np.float64(p))
                    ref = _ref_leaf(p, acc[i], fmt, m)
                    if ok:
                        x = float(acc[i]) + float(p)
                        assert float(F32(x)) == x, ("inexact sum in a passing slice", acc[i], p)
                        t = F32(F32(-acc[i]) - p)
                        got = -_float(_rne(_bits(t), rf))
                        assert _bits(got) == _bits(ref), (fmt, m, chunk, acc[i], p, got, ref)
                        checked += 1
                    acc[i] = ref
    assert passed > 10 and checked > 1000, (passed, checked)


def test_rto_leaves():
    g = np.random.default_rng(77)
    checked = 0
    for it in range(12):
        f16 = it % 2 == 0
        m = 10 if f16 else int(g.integers(1, 20))
        fmt = (16, 10, 15, it % 4 // 2, 0, 0) if f16 else None
        ehi = 15 if f16 else 127
        rf = 23 - m
        n = 600
        sc = 2.0 ** g.uniform(-40 if f16 else -140, ehi - 3, n)
        acc = (g.standard_normal(n) * sc).astype(F32)
        acc = oracle.virtual_round_mant(acc, m) if fmt is None else oracle.virtual_round(acc, fmt)
        a = oracle.virtual_round_mant(g.standard_normal(n).astype(F32), 7)
        b = oracle.virtual_round_mant((g.standard_normal(n) * sc * 2.0 ** g.uniform(-40, 2, n)).astype(F32), 7)
        p = (a.astype(np.float64) * b.astype(np.float64)).astype(F32)   # exact 16-bit products
        p[g.random(n) < 0.05] = 0
        acc[g.random(n) < 0.05] = 0
        acc[(acc == 0) & (g.random(n) < 0.5)] = F32(-0.0)
        for i in range(n):
            ac, pv = F32(acc[i]), F32(p[i])
            if abs(float(ac) + float(pv)) >= 2.0 ** ehi * 0.9:
                continue
            na = F32(-ac)
            t = F32(na - pv)
            bb = F32(t - na)
            e = F32(F32(na - F32(t - bb)) + F32(F32(-pv) - bb))   # TwoSum residual
            tb = _bits(t)
            if e == 0:
                rto = tb
            elif (_bits(e) ^ tb) >> 31 == 0:
                rto = tb | 1
            else:
                rto = (tb - 1) | 1
            r = F32(np.float16(_float(rto))) if f16 else _float(_rne(rto, rf))
            ref = _ref_leaf(pv, ac, fmt, m)
            assert _bits(-r) == _bits(ref), (fmt, m, ac, pv, t, e, -r, ref)
            checked += 1
    assert checked > 5000


def test_binary16_held_leaves():
    """Chunk 1's binary16-held leaves: a lane that passes the exact-sum test, and for a 10-bit
    mantissa-only accumulator also S - Lf <= 39 with every scaled B value FP32-normal, holds
    h = -acc 2^hs as binary16 (hs = 15 - S, or 0 for the binary16-shaped format) and takes
    h <- binary16(-a (b 2^hs) + h), the sum exact, so binary16's one rounding of it (the
    packed form, a binary16 FMA, needs every A and scaled B value to be a binary16 value:
    checked here wherever its slice test passes).  Every
    leaf -h 2^-hs must equal the reference leaf, and the held values must convert exactly
    both ways.  Row scales spread the accumulators over wide exponent ranges (and into
    binary16's subnormal range for the format), so lanes pass and fail both tests."""
    g = np.random.default_rng(4242)
    passed = checked = packed = 0
    for it in range(200):
        fmt = None if it % 2 == 0 or it >= 160 else (16, 10, 15, it % 4 // 2, 0, 0)
        m = 10
        q = -149 if fmt is None else -24
        ehi = 127 if fmt is None else 15
        lanes = 8
        if it < 160:
            sc = 2.0 ** (g.integers(-30, 30) if fmt is None else g.integers(-9, 1))
            K = 32 * int(g.integers(2, 6))
            # per-slice row scales: lanes move between passing and failing slices
            w = 24 if fmt is None else 6
            rs = 2.0 ** g.integers(-w, w + 1, (lanes, K // 32))
            rs[:, g.random(K // 32) < 0.4] = 1.0
            rows = np.repeat(rs, 32, axis=1)
            A = (g.standard_normal((lanes, K)) * sc * rows).astype(F32)
            B = (g.standard_normal(K) / 8).astype(F32)
            A = oracle.virtual_round(A.ravel(), E4M3).reshape(lanes, K) if fmt is not None else \
                oracle.virtual_round_mant(A.ravel(), 3).reshape(lanes, K)
            B = oracle.virtual_round(B, E4M3)
            A[g.random((lanes, K)) < 0.1] = 0
        else:
            # slice 0 leaves row 0 with a large accumulator and row 1 with a tiny one that
            # has bits down to 2^(E - 10); slice 1's products lie between, so the
            # exact-sum test passes with S - Lf near 40 (then binary16 could not hold row 1)
            K = 64
            A = np.zeros((lanes, K), F32)
            B = np.zeros(K, F32)
            B[:32] = oracle.virtual_round((g.standard_normal(32) / 64).astype(F32), E4M3)
            B[0] = 0.125
            B[32:] = g.choice([0.125, 0.25, 0.375, -0.125, -0.25, -0.375], 32)
            A[0, 0] = 2.0 ** int(g.integers(10, 14))
            A[1, 1:32] = oracle.virtual_round_mant(
                (g.standard_normal(31) * 2.0 ** int(g.integers(-20, -14))).astype(F32), 3)
            A[:2, 32:] = g.choice([1.0, 1.5, -1.0, -1.5], (2, 32)) * 2.0 ** int(g.integers(-9, -5))
        acc = np.zeros(lanes, F32)
        for k0 in range(0, K, 32):
            sa, sb = _slice_stats(A[:, k0:k0 + 32], True), _slice_stats(B[k0:k0 + 32], True)
            ok = False
            if sa is not None and sb is not None:
                sL = sa["lsb"] + sb["lsb"]
                sTop = sa["emax"] + sb["emax"] + 2
                exact = (sa["normal"] and sb["normal"] and sa["width"] + sb["width"] <= 24
                         and sTop <= 127 and sL >= -149)
                if exact and sL >= max(q, -126):
                    mag = acc.view(np.uint32).astype(np.int64) & 0x7FFFFFFF
                    ea = (int(mag.max()) >> 23) - 126 if mag.max() else -1000
                    nz = mag[mag != 0]
                    lsb_acc = max(q, (int(nz.min()) >> 23) - 127 - m) if nz.size else 1000
                    Lf = min(sL, lsb_acc)
                    S = _bound_exponent(ea, sTop, m, 1)
                    ok = S <= min(sL + 24, ehi) and sTop <= Lf + 23 and Lf >= -126
                    if ok and fmt is None:
                        hs = 15 - S
                        ok = S - Lf <= 39 and sb["lsb"] + hs >= -126 and sb["emax"] + 1 + hs <= 127
                    elif ok:
                        hs = 0
            # packed form: every A value and every scaled B value a binary16 value
            pk = ok and (sa["width"] <= 11 and sa["lsb"] >= -24 and sa["emax"] <= 15 and
                         sb["width"] <= 11 and sb["lsb"] + hs >= -24 and sb["emax"] + hs <= 15)
            passed += ok
            packed += pk
            if pk:
                sl = A[:, k0:k0 + 32].ravel()
                assert np.array_equal(sl.astype(np.float16).astype(F32), sl)
                bsl = np.array([math.ldexp(float(v), hs) for v in B[k0:k0 + 32]])
                assert np.array_equal(bsl.astype(np.float16).astype(np.float64), bsl)
            if ok:
                h = np.array([np.float16(math.ldexp(-float(v), hs)) for v in acc])
                assert all(float(h[i]) == math.ldexp(-float(acc[i]), hs) for i in range(lanes))
            for k in range(k0, k0 + 32):
                for i in range(lanes):
                    p = F32(np.float64(A[i, k]) * np.float64(B[k]))
                    ref = _ref_leaf(p, acc[i], fmt, m)
                    if ok:
                        bs = F32(math.ldexp(float(B[k]), hs))
                        assert bs == 0 or abs(float(bs)) >= 2.0 ** -126
                        x = -float(A[i, k]) * float(bs) + float(h[i])   # exact in double
                        assert float(F32(x)) == x, ("inexact scaled sum", acc[i], p, hs)
                        h[i] = np.float16(x)
                        got = F32(math.ldexp(-float(h[i]), -hs))
                        assert _bits(got) == _bits(ref), (fmt, hs, acc[i], p, got, ref)
                        checked += 1
                    acc[i] = ref
    assert passed > 10 and checked > 1000 and 0 < packed < passed, (passed, checked, packed)


def _round_binary16(x):
    """Round-to-nearest-even of an exact rational to IEEE binary16 (subnormals kept;
    |x| < 2^15 here, so no overflow)."""
    from fractions import Fraction
    if x == 0:
        return 0.0
    mag = abs(x)
    e = math.floor(math.log2(mag))
    while Fraction(2) ** e > mag:
        e -= 1
    while Fraction(2) ** (e + 1) <= mag:
        e += 1
    ulp = Fraction(2) ** (max(e, -14) - 10)
    q, r = divmod(mag, ulp)
    if r * 2 > ulp or (r * 2 == ulp and q % 2 == 1):
        q += 1
    v = float(q * ulp)
    return v if x > 0 else -v


def test_packed_binary16_rto_leaves():
    """The packed binary16 FMA on RTO lanes of a binary16-shaped accumulator: with a, b and
    acc binary16 values and |acc + a b| below 2^15, binary16's single rounding of the exact
    -(a b + acc) (computed here in exact rational arithmetic), negated, must equal the
    reference leaf round(RTO_fma(a b, 1, acc)), inexact FP32 sums included."""
    from fractions import Fraction
    g = np.random.default_rng(515)
    checked = inexact = 0
    for it in range(1600):
        fmt = (16, 10, 15, it % 2, 0, 0)
        # 8e7m-like operands that binary16 holds: 8 significant bits, |v| in [2^-16, 2^7]
        a = F32(np.float16(oracle.virtual_round_mant(
            np.array([g.standard_normal() * 2.0 ** g.integers(-16, 4)], F32), 7)[0]))
        b = F32(np.float16(oracle.virtual_round_mant(
            np.array([g.standard_normal() * 2.0 ** g.integers(-16, 4)], F32), 7)[0]))
        acc = F32(oracle.virtual_round(np.array([g.standard_normal() * 2.0 ** g.integers(-24, 10)],
                                                F32), fmt)[0])
        r = g.random()
        if r < 0.15:
            acc = F32(-0.0) if g.random() < 0.5 else F32(0.0)
        elif r < 0.2:
            acc = F32(np.float16(-np.float64(a) * np.float64(b)))   # exact cancellations
        if g.random() < 0.1:
            a = F32(-0.0) if g.random() < 0.5 else F32(0.0)
        p = F32(np.float64(a) * np.float64(b))
        assert float(p) == float(a) * float(b)          # 16-bit products are exact
        exact = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(acc))
        if abs(exact) >= 2 ** 15:
            continue
        if exact == 0:
            # the FMA's t = (-a) b + nacc, nacc = -acc: round-to-nearest gives +0 unless both
            # terms are -0 (nonzero terms that cancel give +0)
            neg = lambda x: math.copysign(1.0, float(x)) < 0
            t_neg = float(p) == 0 and float(acc) == 0 and neg(a) == neg(b) and not neg(acc)
            got = F32(0.0) if t_neg else F32(-0.0)
        else:
            inexact += float(F32(float(exact))) != exact
            got = F32(-_round_binary16(-exact))
        ref = _ref_leaf(p, acc, fmt, 10)
        assert _bits(got) == _bits(ref), (a, b, acc, got, ref)
        checked += 1
    assert checked > 1200 and inexact > 120, (checked, inexact)


def _directed(x, down, zero_neg):
    """x (an exact rational) rounded to FP32 toward -inf (down) or +inf, subnormals and
    overflow included; zero_neg: the sign an exact zero takes under that rounding."""
    from fractions import Fraction
    fmax = Fraction(float(np.finfo(np.float32).max))
    if x == 0:
        return F32(-0.0) if zero_neg else F32(0.0)
    if x > fmax:
        return F32(fmax) if down else F32(np.inf)
    if x < -fmax:
        return F32(-np.inf) if down else F32(-fmax)
    c = F32(float(x))
    inf = F32(np.inf)
    if down:
        while Fraction(float(c)) > x:
            c = np.nextafter(c, -inf)
        while Fraction(float(np.nextafter(c, inf))) <= x:
            c = np.nextafter(c, inf)
        if c == 0:                       # a tiny x > 0 rounds down to +0
            c = F32(0.0)
    else:
        while Fraction(float(c)) < x:
            c = np.nextafter(c, inf)
        while Fraction(float(np.nextafter(c, -inf))) >= x:
            c = np.nextafter(c, -inf)
        if c == 0:                       # a tiny x < 0 rounds up to -0
            c = F32(-0.0)
    return F32(c)


def _zero_sign(t1, t2, down):
    """IEEE sign of an exact zero t1 + t2 under a directed rounding: two zeros of one sign keep
    it, otherwise -0 rounding down and +0 rounding up."""
    neg = lambda v: math.copysign(1.0, float(v)) < 0
    if float(t1) == 0 and float(t2) == 0 and neg(t1) == neg(t2):
        return neg(t1)
    return down


def _dr_round(t, m, fmt):
    """The leaf's round of the RTO pattern t: integer RNE at bit 23 - m with the kernel's
    constants (mantissa-only), or the binary16 conversion (binary16-shaped)."""
    if fmt is not None:
        return F32(np.float16(_float(t)))
    rf = 23 - m
    half = (1 << (rf - 1)) - 1 if rf > 0 else 0
    lsb = (t >> rf) & 1 if rf > 0 else 0
    return _float((t + half + lsb) & ~((1 << rf) - 1) & 0xFFFFFFFF)


def test_directed_rounding_leaves():
    """The directed-rounding leaves: from d = RD(x) and u = RU(x) of the exact sum (computed
    here in rational arithmetic), RTO24(x) is d when d == u, else min(d, u) | 1 on the bit
    patterns; it must equal the oracle's nextafter-form round-to-odd FMA (pinned by exact
    rationals), and its round must equal the reference leaf.  Chunk 1 (x = a b + acc, the
    plain form) and chunk sums (y = nacc - p = -(p + acc), the negated form, which takes u for
    an exact zero) over FP32 uniform and wide operands, overflow, FP32-subnormal products,
    exact cancellations and signed zeros, mantissa-only accumulators of 0-23 bits and the
    binary16-shaped format below 2^15."""
    from fractions import Fraction
    g = np.random.default_rng(919)
    fmax = float(np.finfo(np.float32).max)
    checked = odd = 0
    for it in range(3000):
        kind = it % 6
        m = int(g.integers(0, 24))
        fmt = None
        if kind == 0:
            a, b = (F32(g.uniform(-1, 1)) for _ in range(2))
            acc = F32(g.uniform(-4, 4))
        elif kind == 1:
            a, b = (F32(g.standard_normal() * 2.0 ** g.integers(-60, 60)) for _ in range(2))
            acc = F32(g.standard_normal() * 2.0 ** g.integers(-100, 100))
        elif kind == 2:                  # sums near and past FLT_MAX
            a = F32(g.uniform(1, 2) * 2.0 ** 63 * g.choice([-1, 1]))
            b = F32(g.uniform(1, 2) * 2.0 ** 64)
            acc = F32(g.choice([-1, 1]) * fmax * g.uniform(0.5, 1))
        elif kind == 3:                  # products in the FP32 subnormal range
            a = F32(g.standard_normal() * 2.0 ** -70)
            b = F32(g.standard_normal() * 2.0 ** -62)
            acc = F32(g.standard_normal() * 2.0 ** -140)
        elif kind == 4:                  # exact cancellations and signed zeros
            a, b = (F32(g.uniform(-1, 1)) for _ in range(2))
            acc = F32(-(np.float64(a) * np.float64(b))) if g.random() < 0.5 else F32(0.0)
            if g.random() < 0.5:
                a = F32(0.0)
            a, acc = (F32(-v) if g.random() < 0.5 else v for v in (a, acc))
        else:                            # binary16-shaped accumulator, sums below 2^15
            fmt = (16, 10, 15, int(g.integers(0, 2)), 0, 0)
            a, b = (F32(g.standard_normal() * 2.0 ** g.integers(-12, 4)) for _ in range(2))
            acc = F32(oracle.virtual_round(np.array([g.standard_normal() * 2.0 ** 8], F32),
                                           fmt)[0])
        if fmt is None:
            acc = F32(oracle.virtual_round_mant_scalar(float(acc), m))
        if not np.isfinite(acc):         # a start value near FLT_MAX rounded up to inf
            continue
        pm = Fraction(float(a)) * Fraction(float(b))
        x = pm + Fraction(float(acc))
        if fmt is not None and abs(x) >= 2 ** 15:
            continue
        neg_p = (math.copysign(1.0, float(a)) < 0) != (math.copysign(1.0, float(b)) < 0)
        p0 = F32(-0.0) if neg_p else F32(0.0)    # the product's zero, for the zero sign
        # chunk 1, plain form
        d = _directed(x, True, _zero_sign(p0 if pm == 0 else 1, acc, True))
        u = _directed(x, False, _zero_sign(p0 if pm == 0 else 1, acc, False))
        t = _bits(d) if d == u else (min(_bits(d), _bits(u)) | 1)
        rto = oracle.rto_fma(float(a), float(b), float(acc))
        assert t == _bits(rto), (a, b, acc, _float(t), rto)
        odd += d != u
        ref = F32(oracle.virtual_round_mant_scalar(rto, m) if fmt is None
                  else oracle.virtual_round_scalar(rto, fmt))
        assert _bits(_dr_round(t, m, fmt)) == _bits(ref), (a, b, acc, m, fmt)
        # chunk sum, negated form: p is an FP32 chunk sum, y = (-acc) - p
        if fmt is None and abs(float(a) * float(b)) <= fmax / 2:
            p = F32(np.float64(a) * np.float64(b))
            nacc = F32(-acc)
            y = Fraction(float(nacc)) - Fraction(float(p))
            dn = _directed(y, True, _zero_sign(nacc, F32(-p), True))
            un = _directed(y, False, _zero_sign(nacc, F32(-p), False))
            tn = _bits(un) if dn == un else (min(_bits(dn), _bits(un)) | 1)
            ref_n = _ref_leaf(p, acc, None, m)
            assert _bits(F32(-_dr_round(tn, m, None))) == _bits(ref_n), (p, acc, m)
        checked += 1
    assert checked > 2800 and odd > 1000, (checked, odd)


def test_checked_binary16_held_leaves_10bit():
    """The checked binary16-held chunk-1 leaves of the 10-bit mantissa-only accumulator: with
    h = -acc 2^hs (a binary16 value) and any FP32 a, b, the leaf computes binary16's single
    rounding of the exact h - a b 2^hs (the mixed-precision FMA).  Wherever that result is
    above 2^-14 in magnitude, scaling it back must give the reference leaf
    round(RTO24(a b + acc)) negated; results at most 2^-14 (which the kernel flags, sending
    the slice to the directed-rounding leaves) are where the two may differ, and the data
    must reach that case too."""
    from fractions import Fraction
    g = np.random.default_rng(1010)
    checked = flagged = differ_flagged = 0
    for it in range(4000):
        hs = int(g.integers(-20, 40))
        # a 10-bit accumulator value whose scaled image is a binary16-normal value
        acc = F32(oracle.virtual_round_mant_scalar(
            float(g.standard_normal() * 2.0 ** (g.integers(-14, 14) - hs)), 10))
        if acc == 0 or not (2.0 ** -14 <= abs(float(acc)) * 2.0 ** hs < 2.0 ** 15):
            continue
        a = F32(g.uniform(-1, 1) * 2.0 ** g.integers(-8, 8))
        if it % 3 == 0:   # near-cancellation: a b close to -acc
            b = F32(-float(acc) / float(a) * (1 + g.standard_normal() * 2.0 ** -g.integers(8, 30)))
        else:
            b = F32(g.uniform(-1, 1) * 2.0 ** (g.integers(-8, 8)))
        x = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(acc))
        if abs(x) * Fraction(2) ** hs >= 2 ** 15 or abs(x) >= Fraction(2) ** 127:
            continue
        h_old = -Fraction(float(acc)) * Fraction(2) ** hs
        hx = h_old - Fraction(float(a)) * Fraction(float(b)) * Fraction(2) ** hs
        h_new = _round_binary16(hx) if hx != 0 else 0.0
        rto = oracle.rto_fma(float(a), float(b), float(acc))
        ref = F32(oracle.virtual_round_mant_scalar(rto, 10))
        got = F32(-h_new * 2.0 ** -hs) if h_new != 0 else None
        if abs(h_new) <= 2.0 ** -14:
            flagged += 1
            if got is None or _bits(got) != _bits(ref):
                differ_flagged += 1
            continue
        assert _bits(got) == _bits(ref), (a, b, acc, hs, got, ref)
        checked += 1
    assert checked > 1500 and flagged > 50 and differ_flagged > 0, (checked, flagged, differ_flagged)
