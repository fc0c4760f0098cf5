"""Pinning the piecewise-linear SiLU oracle (oracle.pwl_silu, restating pwl_silu_kernel,
Lof_kernel.cu:240-258) by what the mathematics fixes -- the reference ships no test of it:

* knots are fixed points: at x = -R + j * 2R/N (exact in float here) the result is lut[j]
  when the scale N / 2R is exact in float, within one segment's rounding otherwise;
* outside [-R, R] the result is relu(x) exactly, NaN maps to lut[0] (fmaxf), +-inf to relu;
* inside a segment the result lies between the segment's two knot values and, against
  the exact (float64) chord through them, within the rounding budget of the float index
  computation, the rise and the fma;
* against the reference layer's torch op chain (PWLSiLU._forward_ref: clamp, index,
  torch.lerp, relu where |x| > R): within twice the chord's rounding budget inside
  [-R, R], equal outside;
* the tables PWLSiLU builds approximate SiLU to the chord bound of each method;
* bit for bit, the kernel's float operation sequence (Lof_kernel.cu:248-254, 261) evaluated
  in exact rational arithmetic with one round-to-nearest-even to float per operation:
  inv_step = N / (2R), u = (xc + R) * inv_step, frac = u - i, y = fma(frac, y1 - y0, y0)
  (one rounding); the same data shows that mul-then-add, x inv_step + R inv_step, and
  double-precision intermediates each give different bits, so none of them passes.
"""
import numpy as np
import pytest
import torch

import oracle


def _ulp(v):
    return np.spacing(np.abs(v).astype(np.float32)).astype(np.float64)


@pytest.mark.parametrize("N,R", [(2, 1.0), (16, 8.0), (64, 6.0), (1024, 8.0), (4096, 5.0)])
def test_knots_asymptotes_and_bracket(N, R):
    g = np.random.default_rng(N)
    lut = g.standard_normal(N + 1).astype(np.float32)
    step = 2.0 * R / N
    knots = np.array([-R + j * step for j in range(N + 1)], np.float64)
    exact = np.float32(knots) == knots
    y = oracle.pwl_silu(knots.astype(np.float32), lut, R)
    inv = np.float32(N) / np.float32(2 * R)
    if float(inv) == N / (2 * R) and (float(inv) * 2 ** 24).is_integer():
        # an exact scale N / 2R: every representable knot lands on its index
        assert np.array_equal(y[exact], lut[exact]) and exact.sum() > N // 2
    else:   # a rounded scale can put a knot a hair below its index: one segment's rounding
        d = np.abs(np.diff(lut.astype(np.float64)))
        near = 2 * _ulp(np.float32(N)) * np.concatenate([[0.0], d]) + 2 * _ulp(lut)
        assert (np.abs(y[exact].astype(np.float64) - lut[exact]) <= near[exact]).all()
    out = np.array([-1e30, -R * 1.5, np.nextafter(np.float32(R), np.float32(np.inf)), R * 3, 1e30,
                    np.inf, -np.inf], np.float32)
    assert np.array_equal(oracle.pwl_silu(out, lut, R), np.maximum(out, 0))
    assert oracle.pwl_silu(np.array([np.nan], np.float32), lut, R)[0] == lut[0]
    x = g.uniform(-R, R, 200000).astype(np.float32)
    y = oracle.pwl_silu(x, lut, R).astype(np.float64)
    u = (x.astype(np.float64) + R) / step
    i = np.minimum(np.floor(u).astype(np.int64), N - 1)
    y0, y1 = lut[i].astype(np.float64), lut[i + 1].astype(np.float64)
    lo, hi = np.minimum(y0, y1), np.maximum(y0, y1)
    assert ((y >= lo) & (y <= hi)).all()
    chord = y0 + (u - i) * (y1 - y0)
    # u = (x + R) * (N / 2R) is formed in float (two roundings, <= 2 ulp(u)), then the rise
    # y1 - y0 (one rounding) and the fma (one): the error budget against the exact chord
    d = y1 - y0
    tol = 2 * _ulp(u) * np.abs(d) + _ulp(d) + 2 * _ulp(np.maximum(np.abs(y0), np.abs(y1)))
    assert (np.abs(y - chord) <= tol).all()


@pytest.mark.parametrize("bits,method", [(4, "minimax"), (4, "uniform"), (8, "midpoint"),
                                         (10, "minimax")])
def test_matches_layer_reference_chain(bits, method):
    import lofloat_amd as lof
    m = lof.PWLSiLU(R=8.0, lut_bits=bits, lut_method=method, use_ext=False)
    g = np.random.default_rng(bits)
    x = np.concatenate([g.uniform(-10, 10, 100000), np.linspace(-8, 8, 2 * m.lut_size + 1)]
                       ).astype(np.float32)
    ref = m._forward_ref(torch.from_numpy(x)).numpy()
    lut = m.lut.numpy()
    y = oracle.pwl_silu(x, lut, 8.0)
    # both are the segment's chord under a few roundings: compare within twice the budget
    # of each against the exact chord (on the segment's magnitude, not the result's)
    N, step = m.lut_size, 16.0 / m.lut_size
    u = (np.clip(x.astype(np.float64), -8, 8) + 8) / step
    i = np.minimum(np.floor(u).astype(np.int64), N - 1)
    y0, y1 = lut[i].astype(np.float64), lut[i + 1].astype(np.float64)
    d = y1 - y0
    tol = 2 * (2 * _ulp(u) * np.abs(d) + _ulp(d) + 2 * _ulp(np.maximum(np.abs(y0), np.abs(y1))))
    inside = np.abs(x) <= 8
    assert (np.abs(y.astype(np.float64) - ref)[inside] <= tol[inside]).all()
    assert np.array_equal(y[~inside], ref[~inside])


# uniform: the chord error of SiLU over unit segments (~0.054); minimax halves it (~0.031);
# midpoint gives interior knot j SiLU at the midpoint of segment j-1 (layers.py:1108-1111),
# half a segment early, so its error is about half a segment's rise (~0.55).
@pytest.mark.parametrize("method,bound", [("uniform", 0.06), ("midpoint", 0.6), ("minimax", 0.035)])
def test_tables_approximate_silu(method, bound):
    import lofloat_amd as lof
    m = lof.PWLSiLU(R=8.0, lut_bits=4, lut_method=method)
    x = np.linspace(-12, 12, 200001).astype(np.float32)
    y = oracle.pwl_silu(x, m.lut.numpy(), 8.0).astype(np.float64)
    silu = x / (1 + np.exp(-x.astype(np.float64)))
    err = np.abs(y - silu).max()
    assert err < bound, err


def _rn32(x):
    """Round-to-nearest-even of an exact rational to float32 (no overflow in this use)."""
    from fractions import Fraction
    if x == 0:
        return np.float32(0.0)
    c = np.float32(float(x))
    inf = np.float32(np.inf)
    while Fraction(float(c)) > x:
        c = np.nextafter(c, -inf)
    while Fraction(float(np.nextafter(c, inf))) <= x:
        c = np.nextafter(c, inf)
    lo, hi = c, np.nextafter(c, inf)          # lo <= x < hi
    dl, dh = x - Fraction(float(lo)), Fraction(float(hi)) - x
    if dl < dh or (dl == dh and int(np.array([lo]).view(np.uint32)[0]) % 2 == 0):
        return np.float32(lo)
    return np.float32(hi)


@pytest.mark.parametrize("N,R", [(16, 6.0), (64, 3.7), (256, 8.0), (1024, 5.3)])
def test_op_sequence_exact_rationals(N, R):
    from fractions import Fraction as Q
    g = np.random.default_rng(N + 7)
    lut = (g.standard_normal(N + 1) * 3).astype(np.float32)
    Rf = np.float32(R)
    x = np.concatenate([g.uniform(-R, R, 3000), g.uniform(-R, R, 64) * 1e-6]).astype(np.float32)
    got = oracle.pwl_silu(x, lut, R)
    inv = _rn32(Q(N) / Q(float(_rn32(2 * Q(float(Rf))))))
    alt = {"mul_add": 0, "split_scale": 0, "double": 0}
    for k, xv in enumerate(x):
        xc = min(max(float(xv), -float(Rf)), float(Rf))
        u = _rn32(Q(float(_rn32(Q(xc) + Q(float(Rf))))) * Q(float(inv)))
        i = min(int(u), N - 1)
        frac = _rn32(Q(float(u)) - i)
        y0, y1 = Q(float(lut[i])), Q(float(lut[i + 1]))
        d = _rn32(y1 - y0)
        y = _rn32(Q(float(frac)) * Q(float(d)) + y0)
        assert got[k].view(np.uint32) == y.view(np.uint32), (xv, got[k], y)
        # the plausible slips, on the same element
        alt["mul_add"] += _rn32(Q(float(_rn32(Q(float(frac)) * Q(float(d))))) + y0) != y
        u2 = _rn32(Q(float(_rn32(Q(xc) * Q(float(inv))))) +
                   Q(float(_rn32(Q(float(Rf)) * Q(float(inv))))))
        i2 = min(int(u2), N - 1)
        if i2 == i:
            alt["split_scale"] += _rn32(Q(float(u2)) - i) != frac
        yd = (float(frac) * (float(lut[i + 1]) - float(lut[i]))) + float(lut[i])
        alt["double"] += np.float32(yd) != y
    # with a power-of-two inv_step both scalings are exact, so they cannot differ
    if np.frexp(inv)[0] == 0.5:
        alt.pop("split_scale")
    assert all(v > 0 for v in alt.values()), alt
