"""GPU parity against the committed golden fixtures (tests/golden/round_fixtures.npz).

The fixtures hold 1024 inputs (random magnitudes over 2^+-12, FP32 specials and format
boundary values), 1024 explicit random words, and the expected outputs of every
deterministic mode into six formats, the stochastic modes with those words, the
mantissa-only overload and MX blocks with an E8M0 scale.  The HIP kernels must reproduce
each array bit for bit (NaN compared by class); no oracle call is made here.
"""
import os
import sys

import numpy as np
import pytest
import torch

from conftest import assert_bits

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import make_fixtures  # noqa: E402


def _fmt(lof, t):
    return lof.FloatFormatDescriptor(t[0], t[1], t[2], t[3], t[4], t[5],
                                     lof.formats.P3109_InfChecker(t[0]),
                                     lof.formats.P3109_NaNChecker(t[0]))


@pytest.fixture(scope="module")
def fixtures():
    return np.load(os.path.join(HERE, "round_fixtures.npz"))


@pytest.mark.parametrize("kind", ["fmt", "mant", "mx"])
def test_kernels_match_golden(dev, fixtures, kind):
    import lofloat_amd as lof
    x = torch.from_numpy(fixtures["x"]).to(dev)
    u = torch.from_numpy(fixtures["u"].view(np.int32)).to(dev)
    e8m0 = _fmt(lof, make_fixtures.FORMATS["e8m0"])
    n = 0
    for key, (k, args) in make_fixtures.cases().items():
        if k != kind:
            continue
        if kind == "fmt":
            ft, mode, L = args
            y = lof.virtual_round(x, _fmt(lof, ft), mode, L, rand_bits=u)
        elif kind == "mant":
            m, mode, L = args
            y = lof.virtual_round(x, m, mode, L, rand_bits=u)
        else:
            bs, ef, mode, L = args
            y = lof.virtual_mx_round(x, bs, _fmt(lof, ef), e8m0, mode, 0, L, rand_bits=u)
        assert_bits(y.cpu().numpy(), fixtures[key], key)
        n += 1
    assert n > 0
