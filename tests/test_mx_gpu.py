"""GPU parity: MX block rounding against the oracle, bit-exact.

Mirrors the reference's test_virtual_mx_round_cuda.cu (lengths {bs, 3bs, 3bs+1, 1000,
4096, 5003}, formats e4m3 / e5m2 / e8m0 / e5m2-as-scale, modes RNE/RD/RU/RTZ, the
all-zero block) and test_mx_round_py.py (shape, exact-representable block round-trip),
plus stochastic element rounding with explicit and generated random words, every
supported block size, in-place and unaligned inputs, and the 8192 x 8192 workload rows.
"""
import numpy as np
import pytest
import torch

import oracle
from conftest import assert_bits, c1_values, random_words

pytestmark = pytest.mark.gpu

E4M3 = (8, 3, 7, 1, 0, 0)
E5M2 = (8, 2, 15, 1, 0, 0)
E2M1 = (4, 1, 1, 1, 0, 0)
E8M0 = (8, 0, 127, 0, 0, 1)
P3109_E4M3 = (8, 3, 8, 1, 0, 0)


def _fmt(t):
    import lofloat_amd as lof
    return lof.FloatFormatDescriptor(t[0], t[1], t[2], t[3], t[4], t[5],
                                     lof.formats.P3109_InfChecker(t[0]),
                                     lof.formats.P3109_NaNChecker(t[0]))


def _mx(lof, x, dev, bs, ef, sf, mode=0, smode=0, L=0, bits=None, seed=None):
    xt = torch.from_numpy(x).to(dev)
    bt = None if bits is None else torch.from_numpy(bits.view(np.int32)).to(dev)
    return lof.virtual_mx_round(xt, bs, _fmt(ef), _fmt(sf), mode, smode, L, rand_bits=bt,
                                seed=seed).cpu().numpy()


def _rows(n_rows, n_cols, seed):
    g = np.random.default_rng(seed)
    x = g.standard_normal((n_rows, n_cols)) * np.exp2(g.uniform(-8, 8, (n_rows, 1)))
    return x.astype(np.float32).ravel()


@pytest.mark.parametrize("bs", [4, 8, 16, 32, 64, 128, 256])
def test_block_sizes_and_lengths(dev, bs):
    import lofloat_amd as lof
    for n in sorted({bs, 3 * bs, 3 * bs + 1, 1000, 4096, 5003, 65536 + 7}):
        x = c1_values(n, seed=n + bs)
        for ef, sf in ((E4M3, E8M0), (E5M2, E8M0), (E2M1, E8M0), (E4M3, E5M2),
                       (P3109_E4M3, E8M0)):
            for mode in (0, 5, 6, 1):
                y = _mx(lof, x, dev, bs, ef, sf, mode)
                ref = oracle.virtual_mx_round(x, bs, ef, sf, mode)
                assert_bits(y, ref, f"bs {bs} n {n} {ef}/{sf} mode {mode}")


@pytest.mark.parametrize("smode", [0, 1, 2, 4, 5, 6, 7, 12, 13, 3])
def test_scale_modes(dev, smode):
    import lofloat_amd as lof
    x = _rows(64, 256, seed=smode)
    y = _mx(lof, x, dev, 32, E4M3, E8M0, 0, smode)
    assert_bits(y, oracle.virtual_mx_round(x, 32, E4M3, E8M0, 0, smode), f"scale mode {smode}")


@pytest.mark.parametrize("mode", oracle.STOCHASTIC_MODES)
def test_stochastic_elements(dev, mode):
    import lofloat_amd as lof
    x = _rows(128, 512, seed=mode)
    for ef in (E2M1, E4M3):
        L = min(8, 23 - ef[1])
        bits = random_words(x.size, seed=100 + mode)
        y = _mx(lof, x, dev, 32, ef, E8M0, mode, 0, L, bits=bits)
        assert_bits(y, oracle.virtual_mx_round(x, 32, ef, E8M0, mode, 0, L, bits=bits),
                    f"mode {mode} with bits")
        y = _mx(lof, x, dev, 32, ef, E8M0, mode, 0, L, seed=77)
        assert_bits(y, oracle.virtual_mx_round(x, 32, ef, E8M0, mode, 0, L, seed=77),
                    f"mode {mode} generated words")


def test_all_zero_and_special_blocks(dev):
    import lofloat_amd as lof
    x = np.zeros(32 * 8, dtype=np.float32)
    x[32:64] = -0.0
    x[64] = np.inf
    x[96] = np.nan
    x[97:128] = 1.0
    x[128] = 1e-45
    x[160:192] = np.float32(3e38)
    x[192] = -np.inf
    x[193] = np.nan
    y = _mx(lof, x, dev, 32, E4M3, E8M0)
    ref = oracle.virtual_mx_round(x, 32, E4M3, E8M0)
    assert_bits(y, ref, "special blocks")
    assert (y[:64] == 0).all() and not np.signbit(y[:64]).any()  # zero block -> +0


def test_exact_block_round_trips(dev):
    """test_mx_round_py.py:33-38: a block of e4m3-representable values comes back
    unchanged (P3109 e4m3 has bias 8, max normal 240: amax 448 gets the E8M0 scale 2 and
    every element / 2 is representable)."""
    import lofloat_amd as lof
    m = 448.0
    exact = np.array([m, m / 2, -m / 4, 1.0, 0.5, -2.0, 8.0, 0.0] * 4, dtype=np.float32)
    y = _mx(lof, exact, dev, 32, P3109_E4M3, E8M0)
    assert_bits(y, exact, "exact block")


def test_inplace_and_unaligned(dev):
    import lofloat_amd as lof
    x = c1_values(40003, seed=5)
    base = torch.from_numpy(x).to(dev)
    for off in (0, 1, 2, 3):
        v = base[off:].clone()
        ref = oracle.virtual_mx_round(x[off:], 32, E4M3, E8M0)
        y = lof.virtual_mx_round(v, 32, _fmt(E4M3), _fmt(E8M0), out=v)
        assert_bits(v.cpu().numpy(), ref, f"in-place offset {off}")
        w = base[off:]
        y = lof.virtual_mx_round(w, 32, _fmt(E4M3), _fmt(E8M0))
        assert_bits(y.cpu().numpy(), ref, f"unaligned view offset {off}")


@pytest.mark.parametrize("ef", [E2M1, E4M3])
def test_workload_rows_8192(dev, ef):
    """The 8192 x 8192 workload (per-row 2^U(-8,8) magnitudes, blocks of 32 along rows):
    64 full rows checked bit-exactly, RNE and SR-A with a seeded bits tensor."""
    import lofloat_amd as lof
    x = _rows(64, 8192, seed=8192)
    y = _mx(lof, x, dev, 32, ef, E8M0, 0)
    assert_bits(y, oracle.virtual_mx_round(x, 32, ef, E8M0, 0), "RNE rows")
    bits = random_words(x.size, seed=3)
    y = _mx(lof, x, dev, 32, ef, E8M0, 3, 0, 8, bits=bits)
    assert_bits(y, oracle.virtual_mx_round(x, 32, ef, E8M0, 3, 0, 8, bits=bits), "SR-A rows")


def test_errors(dev):
    import lofloat_amd as lof
    x = torch.ones(64, device=dev)
    with pytest.raises(NotImplementedError):
        lof.virtual_mx_round(x, 48, _fmt(E4M3), _fmt(E8M0))
    with pytest.raises(NotImplementedError):
        lof.virtual_mx_round(x, 32, _fmt(E4M3), _fmt(E8M0), 0, 10)  # TrueSR scale
    with pytest.raises(RuntimeError):
        lof.virtual_mx_round(x.double(), 32, _fmt(E4M3), _fmt(E8M0))
