"""GPU checks of the differentiable LoF_Linear (forward + hand-written backward).

Bit-exact against the oracle composition (the reference has no GEMM backward, so its
"autograd gradient" is defined as this composition):
    X_q = MX(X), W_q = MX(W)                      (virtual_mx_round, e4m3 / e8m0, bs 32)
    out = GEMM(X_q, W_q^T) (+ b)                  emulated accumulation
    G_q = virtual_round(G, e5m2, SR-A(8), words lof_rand_u32(seed, i))
    dX  = GEMM(G_q, W_q),  dW = GEMM(G_q^T, X_q)  emulated accumulation
and, with every rounding set to the identity format, against finite differences and
torch's float32 autograd within stated tolerances.
"""
import numpy as np
import pytest
import torch

import oracle
from conftest import assert_bits

pytestmark = pytest.mark.gpu

E4M3 = (8, 3, 7, 1, 0, 0)
E8M0 = (8, 0, 127, 0, 0, 1)
E5M2 = (8, 2, 15, 1, 0, 0)
FP16 = (16, 10, 15, 1, 0, 0)


def _fmt(t):
    import lofloat_amd as lof
    return lof.FloatFormatDescriptor(*t, lof.formats.P3109_InfChecker(t[0]),
                                     lof.formats.P3109_NaNChecker(t[0]))


@pytest.mark.parametrize("accum", ["mant10_c1", "fp16_c32", "fp16_c1"])
@pytest.mark.parametrize("shape", [(64, 96, 48), (130, 128, 70), (8, 4096, 256)])
def test_linear_mx_forward_backward_bitexact(dev, accum, shape):
    import lofloat_amd as lof
    M, K, N = shape
    g = np.random.default_rng(M + K + N)
    X = (g.standard_normal((M, K)) * 0.02).astype(np.float32)
    W = (g.standard_normal((N, K)) * 0.02).astype(np.float32)
    G = g.standard_normal((M, N)).astype(np.float32)
    seed = 0xC0FFEE + M
    kw = dict(accum_mant_bits=10, accum_chunk=1)
    if accum == "fp16_c32":
        kw = dict(accum_format=_fmt(FP16), accum_chunk=32)
    elif accum == "fp16_c1":
        kw = dict(accum_format=_fmt(FP16), accum_chunk=1)
    lin = lof.LoF_Linear(K, N, act_params=_fmt(E4M3), weight_params=_fmt(E4M3), bias=False,
                         scaling="mx", mx_block_size=32,
                         grad_spec=lof.LinearGradSpec(_fmt(E5M2), 3, 8, seed=seed), **kw).to(dev)
    with torch.no_grad():
        lin.weight.copy_(torch.from_numpy(W))
    x = torch.from_numpy(X).to(dev).requires_grad_(True)
    out = lin(x)
    out.backward(torch.from_numpy(G).to(dev))

    if accum == "mant10_c1":
        oacc = oracle.accum_mant(10, 1)
    else:
        oacc = oracle.accum_fmt(FP16, 32 if accum == "fp16_c32" else 1)
    Xq = oracle.virtual_mx_round(X, 32, E4M3, E8M0).reshape(M, K)
    Wq = oracle.virtual_mx_round(W, 32, E4M3, E8M0).reshape(N, K)
    ref_out = oracle.gemm(Xq, Wq, oacc, trans_b=True)
    assert_bits(out.detach().cpu().numpy(), ref_out, "forward")
    Gq = oracle.virtual_round(G, E5M2, oracle.SRA, 8, seed=seed).reshape(M, N)
    assert_bits(x.grad.cpu().numpy(), oracle.gemm(Gq, Wq, oacc), "dgrad")
    assert_bits(lin.weight.grad.cpu().numpy(), oracle.gemm(Gq, Xq, oacc, trans_a=True), "wgrad")


def test_linear_per_tensor_scales_and_bias(dev):
    """per_tensor scaling: X_q = round(sa*X), W_q = round(sw*W), out = GEMM / (sa*sw) + b_q;
    STE backward multiplies by the scales (layers.py:26-31, 379-398)."""
    import lofloat_amd as lof
    M, K, N = 48, 64, 40
    g = np.random.default_rng(5)
    X = g.standard_normal((M, K)).astype(np.float32)
    W = (g.standard_normal((N, K)) * 0.1).astype(np.float32)
    b = g.standard_normal(N).astype(np.float32)
    G = g.standard_normal((M, N)).astype(np.float32)
    sa, sw = 4.0, 8.0
    lin = lof.LoF_Linear(K, N, act_params=_fmt(E4M3), weight_params=_fmt(E4M3),
                         bias_params=_fmt(FP16), act_scale_factor=sa, w_scale_factor=sw,
                         accum_mant_bits=10,
                         grad_spec=lof.LinearGradSpec(_fmt(E5M2), 3, 8, seed=99)).to(dev)
    with torch.no_grad():
        lin.weight.copy_(torch.from_numpy(W))
        lin.bias.copy_(torch.from_numpy(b))
    x = torch.from_numpy(X).to(dev).requires_grad_(True)
    out = lin(x)
    out.backward(torch.from_numpy(G).to(dev))
    oacc = oracle.accum_mant(10, 1)
    alpha = np.float32(1) / (np.float32(sa) * np.float32(sw))
    Xq = oracle.virtual_round(X, E4M3, scale=sa).reshape(M, K)
    Wq = oracle.virtual_round(W, E4M3, scale=sw).reshape(N, K)
    bq = oracle.virtual_round(b, FP16)
    ref = oracle.gemm(Xq, Wq, oacc, trans_b=True, alpha=alpha) + bq[None, :]
    assert_bits(out.detach().cpu().numpy(), ref.astype(np.float32), "forward")
    Gq = oracle.virtual_round(G, E5M2, oracle.SRA, 8, seed=99).reshape(M, N)
    dXq = oracle.gemm(Gq, Wq, oacc, alpha=alpha)
    dWq = oracle.gemm(Gq, Xq, oacc, trans_a=True, alpha=alpha)
    assert_bits(x.grad.cpu().numpy(), (dXq * np.float32(sa)).astype(np.float32), "dgrad")
    assert_bits(lin.weight.grad.cpu().numpy(), (dWq * np.float32(sw)).astype(np.float32), "wgrad")
    np.testing.assert_allclose(lin.bias.grad.cpu().numpy(), G.sum(0), rtol=1e-5, atol=1e-5)


def test_linear_identity_format_vs_fd_and_torch(dev):
    """All roundings set to the identity (32,23,127): gradients match torch float32 autograd
    (rtol 2e-5: the 23-bit accumulator rounds to odd, torch to nearest) and central finite
    differences of the loss (rtol 2e-2, float32 forward)."""
    import lofloat_amd as lof
    ident = lof.create_single_params()
    M, K, N = 16, 64, 24
    g = np.random.default_rng(3)
    X = g.standard_normal((M, K)).astype(np.float32)
    W = g.standard_normal((N, K)).astype(np.float32) * 0.1
    G = g.standard_normal((M, N)).astype(np.float32)
    lin = lof.LoF_Linear(K, N, act_params=ident, weight_params=ident, bias=False,
                         accum_mant_bits=23,
                         grad_spec=lof.LinearGradSpec(ident, lof.RoundingMode.RoundToNearestEven,
                                                      0)).to(dev)
    with torch.no_grad():
        lin.weight.copy_(torch.from_numpy(W))
    x = torch.from_numpy(X).to(dev).requires_grad_(True)
    Gt = torch.from_numpy(G).to(dev)
    (lin(x) * Gt).sum().backward()
    xr = torch.from_numpy(X).double().requires_grad_(True)
    wr = torch.from_numpy(W).double().requires_grad_(True)
    (torch.nn.functional.linear(xr, wr) * torch.from_numpy(G).double()).sum().backward()
    np.testing.assert_allclose(x.grad.cpu().numpy(), xr.grad.numpy(), rtol=2e-5, atol=2e-5)
    np.testing.assert_allclose(lin.weight.grad.cpu().numpy(), wr.grad.numpy(), rtol=2e-5,
                               atol=2e-5)

    def loss(xv):
        with torch.no_grad():
            return float((lin(xv) * Gt).double().sum())
    h = 1e-2
    for (i, j) in [(0, 0), (3, 17), (15, 63), (7, 40)]:
        e = torch.zeros_like(x)
        e[i, j] = h
        fd = (loss(x.detach() + e) - loss(x.detach() - e)) / (2 * h)
        assert abs(fd - float(x.grad[i, j])) <= 2e-2 * max(1.0, abs(fd)), (i, j, fd)


def test_quantize_module_and_ste(dev):
    """test_mx_round_py.py:40-46, 83-88: LoF_Quantize(mx) forward shape, identity STE grad,
    non-divisible last dim raises."""
    import lofloat_amd as lof
    e4m3 = lof.create_p3109_params(8, 4, True, True)
    q = lof.LoF_Quantize(e4m3, scaling="mx", mx_block_size=32)
    x = ((torch.rand(4, 256, device=dev) - 0.5) * 10.0).requires_grad_(True)
    y = q(x)
    y.sum().backward()
    assert y.shape == x.shape and torch.isfinite(y).all() and (y != x).any()
    assert torch.equal(x.grad, torch.ones_like(x))
    with pytest.raises(ValueError):
        q(torch.rand(3, 30, device=dev))
    p = lof.LoF_Quantize(e4m3)
    xs = torch.randn(100, device=dev, requires_grad=True)
    p(xs).sum().backward()
    assert torch.equal(xs.grad, torch.ones_like(xs))
