"""Shared test fixtures.  GPU tests are marked ``gpu``; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")
    config.addinivalue_line("markers", "slow: long-running exhaustive checks")


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible HIP device")
    return torch.device("cuda:0")


def bits_equal(y: np.ndarray, ref: np.ndarray) -> np.ndarray:
    """Elementwise bit equality, NaN compared by class."""
    y = np.ascontiguousarray(y, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    return (y.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(y) & np.isnan(ref))


def assert_bits(y, ref, what=""):
    same = bits_equal(y, ref)
    if not same.all():
        y = np.asarray(y, dtype=np.float32).ravel()
        ref = np.asarray(ref, dtype=np.float32).ravel()
        bad = np.flatnonzero(~same.ravel())
        i = int(bad[0])
        raise AssertionError(
            f"{what}: {bad.size} of {same.size} differ; first at {i}: got {y[i]!r} "
            f"(0x{y.view(np.uint32)[i]:08x}) want {ref[i]!r} (0x{ref.view(np.uint32)[i]:08x})")


def c1_values(n: int, seed: int, specials: float = 0.01) -> np.ndarray:
    """N(0,1) * 2^U(-12,12) with a fraction of +-0 / +-inf / NaN / FP32-subnormal specials."""
    g = np.random.default_rng(seed)
    x = (g.standard_normal(n) * np.exp2(g.uniform(-12, 12, n))).astype(np.float32)
    k = int(n * specials)
    if k:
        idx = g.choice(n, size=k, replace=False)
        kind = g.integers(0, 6, size=k)
        sub = (g.integers(1, 1 << 23, size=k, dtype=np.int64).astype(np.uint32)
               | (g.integers(0, 2, size=k).astype(np.uint32) << 31)).view(np.float32)
        vals = np.select([kind == 0, kind == 1, kind == 2, kind == 3, kind == 4],
                         [np.float32(0.0), np.float32(-0.0), np.float32(np.inf),
                          np.float32(-np.inf), np.float32(np.nan)], sub)
        x[idx] = vals.astype(np.float32)
    return x


def random_words(n: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).integers(0, 2**32, size=n, dtype=np.uint64).astype(
        np.uint32)
