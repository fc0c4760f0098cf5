"""Pinning the convolution oracle (oracle.conv2d / oracle.im2col) by checks independent of it:

* im2col against torch.nn.functional.unfold (a separate implementation of the same gather);
* exactness: with small-integer inputs every partial sum is exact at 23 accumulator bits, so
  the oracle must equal torch's float64 conv2d exactly, in both layouts (a transposed
  operand, a wrong padding / stride / dilation or a swapped tap order fails this);
* the reference layer's own construction: one lof_gemm per image over the unfolded input
  (layers.py:718-727), stacked, is the oracle's single GEMM over all columns;
* the reference's test_conv2d_cuda.cu cases (test_conv2d_cuda.cu:197-215) under its own
  bound |y - y_ref| <= 4 K eps conv(|x|, |w|) + 1e-6 against a double-precision conv.
"""
import numpy as np
import pytest
import torch

import oracle

CASES = [   # (N, C, H, W, K, R, S, pad, stride, dil)
    (2, 3, 7, 9, 4, 3, 2, (1, 2), (2, 1), (2, 3)),
    (1, 5, 8, 8, 3, 3, 3, (1, 1), (1, 1), (1, 1)),
    (3, 2, 6, 11, 5, 1, 1, (0, 0), (1, 2), (1, 1)),
    (2, 4, 9, 9, 2, 5, 5, (2, 2), (3, 3), (1, 1)),
    (1, 1, 5, 4, 1, 2, 3, (0, 1), (1, 1), (2, 1)),
]


def _torch_conv64(x, w, pad, stride, dil):
    return torch.nn.functional.conv2d(torch.from_numpy(x).double(), torch.from_numpy(w).double(),
                                      padding=pad, stride=stride, dilation=dil).numpy()


@pytest.mark.parametrize("case", CASES)
def test_im2col_matches_unfold(case):
    N, C, H, W, K, R, S, pad, stride, dil = case
    x = np.random.default_rng(sum(case[:7])).standard_normal((N, C, H, W)).astype(np.float32)
    cols = oracle.im2col(x, R, S, pad, stride, dil)
    t = torch.nn.functional.unfold(torch.from_numpy(x), (R, S), dilation=dil, padding=pad,
                                   stride=stride)                       # N x CRS x L
    assert np.array_equal(cols, t.permute(1, 0, 2).reshape(C * R * S, -1).numpy())


@pytest.mark.parametrize("case", CASES)
def test_exact_inputs_equal_float64_conv(case):
    N, C, H, W, K, R, S, pad, stride, dil = case
    g = np.random.default_rng(7 + sum(case[:7]))
    x = g.integers(-8, 9, (N, C, H, W)).astype(np.float32)
    w = g.integers(-8, 9, (K, C, R, S)).astype(np.float32) * np.float32(0.25)
    ref = _torch_conv64(x, w, pad, stride, dil)
    acc = oracle.accum_mant(23, 1)
    y = oracle.conv2d(x, w, acc, pad, stride, dil)
    assert np.array_equal(y.astype(np.float64), ref)
    yh = oracle.conv2d(x.transpose(0, 2, 3, 1), w.transpose(0, 2, 3, 1), acc, pad, stride, dil,
                       layout=1)
    assert np.array_equal(yh.astype(np.float64), ref.transpose(0, 2, 3, 1))
    # alpha and beta * C as in the GEMM epilogue: exact here too
    c = g.integers(-4, 5, ref.shape).astype(np.float32)
    y2 = oracle.conv2d(x, w, acc, pad, stride, dil, alpha=0.5, C=c, beta=2.0)
    assert np.array_equal(y2.astype(np.float64), 0.5 * ref + 2.0 * c)


@pytest.mark.parametrize("case", CASES[:3])
def test_equals_per_image_gemm_loop(case):
    """LoF_Conv2d's im2col path runs lof_gemm(w_col, x_col[b]) per image and stacks."""
    N, C, H, W, K, R, S, pad, stride, dil = case
    g = np.random.default_rng(11 + sum(case[:7]))
    x = g.standard_normal((N, C, H, W)).astype(np.float32)
    w = g.standard_normal((K, C, R, S)).astype(np.float32)
    acc = oracle.accum_mant(7, 1)
    y = oracle.conv2d(x, w, acc, pad, stride, dil)
    per = np.stack([oracle.gemm(w.reshape(K, -1), oracle.im2col(x[b:b + 1], R, S, pad, stride, dil),
                                acc).reshape(K, y.shape[2], y.shape[3]) for b in range(N)])
    assert np.array_equal(y, per)


# test_conv2d_cuda.cu:197-215: (N, H, W, C, K, R, S, pad, stride, dilation), accum bits
REF_CASES = [
    ((2, 10, 10, 8, 12, 3, 3), (1, 1), (1, 1), (1, 1), 10),
    ((1, 9, 11, 6, 10, 3, 3), (1, 1), (2, 2), (1, 1), 10),
    ((1, 12, 12, 4, 8, 3, 3), (2, 2), (1, 1), (2, 2), 8),
    ((2, 7, 7, 16, 20, 1, 1), (0, 0), (1, 1), (1, 1), 12),
    ((1, 8, 8, 4, 6, 3, 3), (1, 1), (1, 1), (1, 1), 23),
]


@pytest.mark.parametrize("dims,pad,stride,dil,bits", REF_CASES)
def test_reference_error_bound(dims, pad, stride, dil, bits):
    N, H, W, C, K, R, S = dims
    g = np.random.default_rng(1000 + bits)
    x = g.uniform(-1, 1, (N, C, H, W)).astype(np.float32)
    w = g.uniform(-1, 1, (K, C, R, S)).astype(np.float32)
    y = oracle.conv2d(x, w, oracle.accum_mant(bits, 1), pad, stride, dil).astype(np.float64)
    ref = _torch_conv64(x, w, pad, stride, dil)
    absref = _torch_conv64(np.abs(x), np.abs(w), pad, stride, dil)
    bound = 4.0 * C * R * S * max(2.0 ** -bits, 1e-7) * absref + 1e-6
    assert (np.abs(y - ref) <= bound).all()
    if bits < 23:   # and the rounding is really there: not the exact result everywhere
        assert (y != ref).any()
