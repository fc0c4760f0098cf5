"""GPU parity of the implicit-GEMM Conv2d (lof_conv2d_f32 through lofloat_amd.conv2d /
lof_conv2d) against the oracle (unfold + the GEMM oracle), bit for bit.

Covers the reference's test_conv2d_cuda.cu cases through the reference signature, a sweep of
shapes (padding, stride, dilation, 1x1 to 7x7 filters, odd channel counts, partial tiles),
both wave tiles, chunks 1 / 8 / 32, both layouts, N(0, 1) inputs (FP32 chain path) and
e4m3-rounded inputs (MFMA exact path), a full-format accumulator, alpha / beta * C, the
forced general leaf and VALU chain, and argument errors.
"""
import zlib

import numpy as np
import pytest
import torch

import oracle
from conftest import assert_bits

pytestmark = pytest.mark.gpu

CFG_64, CFG_128 = 7, 6   # wave tiles 64 x 64 (mfma11) and 128 x 64 (mfma21)


def _inputs(kind, xshape, wshape, seed):
    g = np.random.default_rng(seed)
    x = g.standard_normal(xshape).astype(np.float32)
    w = g.standard_normal(wshape).astype(np.float32)
    if kind == "e4m3":   # values whose chunk sums are exact in FP32 (MFMA path)
        x = oracle.virtual_round(x / 8, (8, 3, 7, 1, 0, 0)).reshape(xshape)
        w = oracle.virtual_round(w / 8, (8, 3, 7, 1, 0, 0)).reshape(wshape)
    return x, w


def _run(dev, x, w, spec, acc, pad, stride, dil, layout="NCHW", config=-1, alpha=1.0, C=None,
         beta=0.0):
    import lofloat_amd as lof
    Ct = None if C is None else torch.from_numpy(C).to(dev)
    y = lof.conv2d(torch.from_numpy(x).to(dev), torch.from_numpy(w).to(dev), spec, stride=stride,
                   padding=pad, dilation=dil, alpha=alpha, layout=layout, C=Ct, beta=beta,
                   config=config).cpu().numpy()
    ref = oracle.conv2d(x, w, acc, pad, stride, dil, alpha=alpha, layout=0 if layout == "NCHW" else 1,
                        C=C, beta=beta)
    return y, ref


# test_conv2d_cuda.cu:197-215: (N, H, W, C, K, R, S), pad, stride, dilation, accum bits, mode
REF_CASES = [
    ((2, 10, 10, 8, 12, 3, 3), (1, 1), (1, 1), (1, 1), 10, 0),
    ((1, 9, 11, 6, 10, 3, 3), (1, 1), (2, 2), (1, 1), 10, 0),
    ((1, 12, 12, 4, 8, 3, 3), (2, 2), (1, 1), (2, 2), 8, 1),
    ((2, 7, 7, 16, 20, 1, 1), (0, 0), (1, 1), (1, 1), 12, 4),
    ((1, 8, 8, 4, 6, 3, 3), (1, 1), (1, 1), (1, 1), 23, 0),
]


@pytest.mark.parametrize("dims,pad,stride,dil,bits,mode", REF_CASES)
def test_reference_cases_lof_conv2d(dev, dims, pad, stride, dil, bits, mode):
    import lofloat_amd as lof
    N, H, W, C, K, R, S = dims
    g = np.random.default_rng(bits * 7 + mode)
    x = g.uniform(-1, 1, (N, C, H, W)).astype(np.float32)
    w = g.uniform(-1, 1, (K, C, R, S)).astype(np.float32)
    y = lof.lof_conv2d(torch.from_numpy(x).to(dev), torch.from_numpy(w).to(dev), pad[0], pad[1],
                       stride[0], stride[1], dil[0], dil[1], bits, lof.RoundingMode(mode), 0,
                       2.0, 0.5).cpu().numpy()
    ref = oracle.conv2d(x, w, oracle.accum_mant(bits, 1), pad, stride, dil, alpha=1.0)
    assert_bits(y, ref, f"lof_conv2d {dims}")
    # alpha = 1 / (weight_scale * input_scale), and 1 when that product is 0
    # (Lof_kernel.cu:210-211)
    for ws, xs, alpha in ((4.0, 1.0, 0.25), (4.0, 0.0, 1.0), (0.0, 3.0, 1.0)):
        y4 = lof.lof_conv2d(torch.from_numpy(x).to(dev), torch.from_numpy(w).to(dev), pad[0],
                            pad[1], stride[0], stride[1], dil[0], dil[1], bits,
                            weight_scale=ws, input_scale=xs).cpu().numpy()
        assert_bits(y4, oracle.conv2d(x, w, oracle.accum_mant(bits, 1), pad, stride, dil,
                                      alpha=alpha), f"scales {ws} {xs}")


SHAPES = [   # (N, C, H, W, K, R, S, pad, stride, dil)
    (2, 3, 16, 16, 16, 3, 3, (1, 1), (1, 1), (1, 1)),
    (1, 5, 13, 9, 7, 5, 3, (2, 1), (2, 1), (1, 2)),
    (3, 8, 9, 10, 70, 1, 1, (0, 0), (1, 1), (1, 1)),
    (1, 3, 30, 30, 24, 7, 7, (3, 3), (2, 2), (1, 1)),
    (2, 16, 11, 7, 33, 3, 3, (0, 2), (1, 3), (2, 1)),
    (4, 1, 6, 6, 1, 2, 2, (0, 0), (2, 2), (1, 1)),
]


@pytest.mark.parametrize("shape", SHAPES, ids=[f"s{i}" for i in range(len(SHAPES))])
@pytest.mark.parametrize("chunk", [1, 8, 32])
@pytest.mark.parametrize("kind", ["gauss", "e4m3"])
@pytest.mark.parametrize("layout", ["NCHW", "NHWC"])
def test_shapes(dev, shape, chunk, kind, layout):
    import lofloat_amd as lof
    N, C, H, W, K, R, S, pad, stride, dil = shape
    x, w = _inputs(kind, (N, C, H, W), (K, C, R, S), seed=zlib.crc32(repr((shape, chunk, kind)).encode()))
    if layout == "NHWC":
        x, w = np.ascontiguousarray(x.transpose(0, 2, 3, 1)), np.ascontiguousarray(w.transpose(0, 2, 3, 1))
    spec = lof.AccumSpec(mant_bits=10, chunk=chunk)
    acc = oracle.accum_mant(10, chunk)
    for cfg in (CFG_64, CFG_128):
        y, ref = _run(dev, x, w, spec, acc, pad, stride, dil, layout, config=cfg)
        assert_bits(y, ref, f"conv {shape} chunk {chunk} {kind} {layout} cfg {cfg}")


def test_large_mfma_path(dev):
    """64 -> 128 channels, 3x3 on 4 x 20 x 20, e4m3 inputs: every slice on MFMA at chunk 32,
    128 x 64 tiles (the default for long chunks and many tiles is also checked)."""
    import lofloat_amd as lof
    x, w = _inputs("e4m3", (4, 64, 20, 20), (128, 64, 3, 3), seed=5)
    for chunk in (8, 32):
        fmt = lof.make_format(16, 10, 15)
        spec = lof.AccumSpec(fmt=fmt, chunk=chunk)
        acc = oracle.accum_fmt(fmt.as_tuple(), chunk)
        for cfg in (-1, CFG_64, CFG_128):
            y, ref = _run(dev, x, w, spec, acc, (1, 1), (1, 1), (1, 1), config=cfg)
            assert_bits(y, ref, f"large chunk {chunk} cfg {cfg}")


def test_alpha_beta_and_forced_paths(dev):
    import lofloat_amd as lof
    x, w = _inputs("gauss", (2, 6, 12, 12), (20, 6, 3, 3), seed=9)
    P = Q = 12
    C = np.random.default_rng(3).standard_normal((2, 20, P, Q)).astype(np.float32)
    spec = lof.AccumSpec(mant_bits=7, chunk=8)
    acc = oracle.accum_mant(7, 8)
    y, ref = _run(dev, x, w, spec, acc, (1, 1), (1, 1), (1, 1), alpha=0.75, C=C, beta=-1.5)
    assert_bits(y, ref, "alpha/beta")
    for bits in (0x100, 0x200, 0x300):
        y2, _ = _run(dev, x, w, spec, acc, (1, 1), (1, 1), (1, 1), config=CFG_64 | bits,
                     alpha=0.75, C=C, beta=-1.5)
        assert_bits(y2, ref, f"forced path {bits:#x}")


@pytest.mark.parametrize("kind", ["bf16", "gauss"])
def test_chunk1_binary16_accumulators(dev, kind):
    """Chunk 1 with binary16-shaped accumulators (saturating and extended 5e10m): bf16-like
    8e7m inputs that binary16 holds (the packed binary16 leaves), one tiny input value whose
    slices take the round-to-odd leaves, and N(0, 1) inputs whose products FP32 cannot hold
    (the directed-rounding leaves); bit for bit against the oracle with every leaf kind on,
    and with the exact-sum / RTO / binary16-held / packed / directed-rounding leaves switched
    off (0x800 .. 0x8000)."""
    import lofloat_amd as lof
    x, w = _inputs("gauss", (2, 8, 11, 11), (24, 8, 3, 3), seed=77)
    if kind == "bf16":
        x = oracle.virtual_round(x / 8, (16, 7, 127, 1, 0, 0)).reshape(x.shape)
        w = oracle.virtual_round(w / 8, (16, 7, 127, 1, 0, 0)).reshape(w.shape)
        x[1, 3, 5, 5] = np.float32(2.0 ** -30)
    for fmt in ((16, 10, 15, 1, 0, 0), (16, 10, 15, 0, 0, 0)):
        spec = lof.AccumSpec(fmt=lof.make_format(16, 10, 15, saturating=bool(fmt[3])), chunk=1)
        acc = oracle.accum_fmt(fmt, 1)
        y, ref = _run(dev, x, w, spec, acc, (1, 1), (1, 1), (1, 1), config=CFG_64)
        assert_bits(y, ref, f"{kind} {fmt} chunk 1")
        for bits in (0x800, 0x1000, 0x2000, 0x4000, 0x8000, 0xF800):
            for cfg in (CFG_64, CFG_128):
                y2, _ = _run(dev, x, w, spec, acc, (1, 1), (1, 1), (1, 1), config=cfg | bits)
                assert_bits(y2, ref, f"{kind} {fmt} cfg {cfg} bits {bits:#x}")


def test_empty_and_errors(dev):
    import lofloat_amd as lof
    x = torch.randn(0, 3, 8, 8, device=dev)
    w = torch.randn(4, 3, 3, 3, device=dev)
    y = lof.conv2d(x, w, lof.AccumSpec(mant_bits=10))
    assert tuple(y.shape) == (0, 4, 6, 6)
    x = torch.randn(1, 3, 8, 8, device=dev)
    with pytest.raises(NotImplementedError):            # chunk 3: not a wave-tile chunk
        lof.conv2d(x, w, lof.AccumSpec(mant_bits=10, chunk=3))
    with pytest.raises(NotImplementedError):            # a VALU tile config
        lof.conv2d(x, w, lof.AccumSpec(mant_bits=10), config=0)
    with pytest.raises(RuntimeError):                   # channel mismatch
        lof.conv2d(x, torch.randn(4, 2, 3, 3, device=dev), lof.AccumSpec(mant_bits=10))
    with pytest.raises(RuntimeError):                   # empty output
        lof.conv2d(x, torch.randn(4, 3, 9, 9, device=dev), lof.AccumSpec(mant_bits=10))
    with pytest.raises(NotImplementedError):            # stochastic accumulator
        lof.conv2d(x, w, lof.AccumSpec(mant_bits=10, mode=3))


# ----------------------------------------------------------------------- LoF_Conv2d layer
def _layer_ref(layer, x, fmt_t):
    """The reference layer's forward composed from oracle pieces (layers.py:620-750)."""
    w = layer.weight.detach().cpu().numpy()
    pad, stride, dil = layer.padding, layer.stride, layer.dilation
    acc = oracle.accum_mant(layer.accum_mant_bits, 1)
    if layer.scaling == "mx":
        B = x.shape[0]
        cols = oracle.im2col(x, *layer.kernel_size, pad, stride, dil)          # K x B*L
        K = cols.shape[0]
        rows = np.ascontiguousarray(cols.T)                                   # (b, l) x K
        sf = fmt_t(layer.scale_format)
        rows = oracle.virtual_mx_round(rows.ravel(), layer.mx_block_size, fmt_t(layer.act_params),
                                       sf).reshape(rows.shape)
        wq = oracle.virtual_mx_round(w.reshape(w.shape[0], -1).ravel(), layer.mx_block_size,
                                     fmt_t(layer.weight_params), sf).reshape(w.shape[0], K)
        D = oracle.gemm(wq, np.ascontiguousarray(rows.T), acc)
        P = oracle.conv_out_size(x.shape[2], pad[0], dil[0], layer.kernel_size[0], stride[0])
        Q = oracle.conv_out_size(x.shape[3], pad[1], dil[1], layer.kernel_size[1], stride[1])
        out = D.reshape(w.shape[0], B, P, Q).transpose(1, 0, 2, 3)
        bq = oracle.virtual_round(layer.bias.detach().cpu().numpy(), fmt_t(layer.bias_params))
        return out + bq.reshape(1, -1, 1, 1)
    xq = oracle.virtual_round(x.ravel(), fmt_t(layer.act_params),
                              scale=layer.act_scale_factor).reshape(x.shape)
    wq = oracle.virtual_round(w.ravel(), fmt_t(layer.weight_params),
                              scale=layer.w_scale_factor).reshape(w.shape)
    if layer.backend == "cutlass":
        alpha = float(np.float32(1.0 / (layer.w_scale_factor * layer.act_scale_factor)))
    else:
        alpha = float(np.float32(1.0) / (np.float32(layer.w_scale_factor) *
                                          np.float32(layer.act_scale_factor)))
    G = layer.groups
    ci, co = x.shape[1] // G, w.shape[0] // G
    out = np.concatenate([oracle.conv2d(np.ascontiguousarray(xq[:, g * ci:(g + 1) * ci]),
                                        wq[g * co:(g + 1) * co], acc, pad, stride, dil,
                                        alpha=alpha) for g in range(G)], axis=1)
    bq = oracle.virtual_round(layer.bias.detach().cpu().numpy(), fmt_t(layer.bias_params),
                              scale=layer.b_scale_factor) / np.float32(layer.b_scale_factor)
    return out + bq.astype(np.float32).reshape(1, -1, 1, 1)


@pytest.mark.parametrize("variant", ["im2col", "cutlass", "groups", "mx", "hadamard"])
def test_lof_conv2d_layer(dev, variant):
    import lofloat_amd as lof
    e4m3 = lof.make_format(8, 3, 7)
    kw = dict(act_params=e4m3, weight_params=e4m3, bias_params=lof.make_format(16, 10, 15),
              accum_mant_bits=10, padding=1)
    if variant == "cutlass":
        kw.update(backend="cutlass", act_scale_factor=3.0, w_scale_factor=7.0)
    elif variant == "im2col":
        kw.update(act_scale_factor=3.0, w_scale_factor=7.0, b_scale_factor=2.0, stride=2)
    elif variant == "groups":
        kw.update(groups=2, dilation=2)
    elif variant == "mx":
        kw.update(scaling="mx", mx_block_size=16)
    else:
        kw.update(hadamard_transform=True)
    torch.manual_seed(0)
    layer = lof.LoF_Conv2d(16, 12, 3, **kw).to(dev)
    with torch.no_grad():
        layer.bias.copy_(torch.randn(12))
    x = torch.randn(2, 16, 9, 10)
    y = layer(x.to(dev)).detach().cpu().numpy()
    xin = x.numpy()
    if variant == "hadamard":
        xin = lof._fwht(x.movedim(1, -1).contiguous()).movedim(-1, 1).contiguous().numpy()
    ref = _layer_ref(layer, xin, lambda p: p.as_tuple())
    assert_bits(y, ref.astype(np.float32), f"LoF_Conv2d {variant}")
