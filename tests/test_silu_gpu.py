"""GPU parity of the piecewise-linear SiLU (lof_pwl_silu through lofloat_amd.pwl_silu and
PWLSiLU) against the oracle's restatement of pwl_silu_kernel, bit for bit: every supported
table length, float32 / float16 / float64 elements (computed in float as the kernel does),
special values, partial tiles, unaligned buffers, the layer and its surrogate gradient."""
import numpy as np
import pytest
import torch

import oracle
from conftest import assert_bits

pytestmark = pytest.mark.gpu


def _x(n, R, seed):
    g = np.random.default_rng(seed)
    x = np.concatenate([g.uniform(-1.5 * R, 1.5 * R, n),
                        np.array([0.0, -0.0, R, -R, np.inf, -np.inf, np.nan, 1e-40, -3e38],
                                 np.float32)])
    return x.astype(np.float32)


@pytest.mark.parametrize("bits", range(1, 13))
def test_every_table_length(dev, bits):
    import lofloat_amd as lof
    N = 1 << bits
    R = 8.0 if bits % 2 else 5.5
    g = np.random.default_rng(bits)
    lut = g.standard_normal(N + 1).astype(np.float32)
    x = _x(50000, R, bits)
    y = lof.pwl_silu(torch.from_numpy(x).to(dev), torch.from_numpy(lut), R).cpu().numpy()
    assert_bits(y, oracle.pwl_silu(x, lut, R), f"N={N}")


@pytest.mark.parametrize("dtype", [np.float16, np.float64])
def test_other_dtypes(dev, dtype):
    """float16 / float64 elements: converted to float, computed in float, converted back
    with round to nearest even (static_cast in pwl_silu_kernel)."""
    import lofloat_amd as lof
    lut = np.random.default_rng(7).standard_normal(65).astype(np.float32)
    x = _x(30000, 6.0, 3).astype(dtype)
    y = lof.pwl_silu(torch.from_numpy(x).to(dev), torch.from_numpy(lut), 6.0).cpu().numpy()
    ref = oracle.pwl_silu(x.astype(np.float32), lut, 6.0).astype(dtype)
    assert y.dtype == dtype
    same = (y == ref) | (np.isnan(y) & np.isnan(ref))
    assert same.all(), (x[~same][:5], y[~same][:5], ref[~same][:5])


def test_sizes_and_alignment(dev):
    import lofloat_amd as lof
    lut = np.random.default_rng(9).standard_normal(17).astype(np.float32)
    for n in (1, 7, 4095, 4096, 4097, 3 * 4096 + 5):
        x = _x(n, 8.0, n)[:n]
        y = lof.pwl_silu(torch.from_numpy(x).to(dev), torch.from_numpy(lut), 8.0).cpu().numpy()
        assert_bits(y, oracle.pwl_silu(x, lut, 8.0), f"n={n}")
    x = _x(9000, 8.0, 1)
    xs = torch.zeros(x.size + 1, device=dev)
    xs[1:] = torch.from_numpy(x).to(dev)                       # not 16-byte aligned
    y = lof.pwl_silu(xs[1:], torch.from_numpy(lut), 8.0).cpu().numpy()
    assert_bits(y, oracle.pwl_silu(x, lut, 8.0), "unaligned")
    x2 = torch.from_numpy(_x(4096, 8.0, 2)[:4096].reshape(16, 256)).to(dev).t()   # strided view
    y2 = lof.pwl_silu(x2, torch.from_numpy(lut), 8.0)
    assert y2.shape == x2.shape
    assert_bits(y2.cpu().numpy(), oracle.pwl_silu(x2.cpu().numpy(), lut, 8.0).reshape(y2.shape),
                "strided")


@pytest.mark.parametrize("method", ["minimax", "uniform", "midpoint"])
def test_layer_and_gradient(dev, method):
    import lofloat_amd as lof
    m = lof.PWLSiLU(R=8.0, lut_bits=6, lut_method=method).to(dev)
    x = torch.from_numpy(_x(20000, 8.0, 5)[:20000]).to(dev).requires_grad_(True)
    y = m(x)
    assert_bits(y.detach().cpu().numpy(),
                oracle.pwl_silu(x.detach().cpu().numpy(), m.lut.cpu().numpy(), 8.0), method)
    g = torch.randn_like(y)
    y.backward(g)
    xc = x.detach().double().cpu()
    s = torch.sigmoid(xc)
    want = (g.double().cpu() * (s * (1 + xc * (1 - s)))).float()
    torch.testing.assert_close(x.grad.cpu(), want, rtol=1e-5, atol=1e-6)


def test_errors(dev):
    import lofloat_amd as lof
    x = torch.ones(8, device=dev)
    with pytest.raises(ValueError):
        lof.pwl_silu(x, torch.zeros(6), 8.0)                   # 5 segments: not a power of 2
    with pytest.raises(ValueError):
        lof.pwl_silu(x, torch.zeros(8194), 8.0)                # 8193 segments
    with pytest.raises(RuntimeError):
        lof.pwl_silu(x, torch.zeros(17), 0.0)
    with pytest.raises(RuntimeError):
        lof.pwl_silu(x.cpu(), torch.zeros(17), 8.0)
    with pytest.raises(RuntimeError):
        lof.pwl_silu(x.to(torch.int32), torch.zeros(17), 8.0)
