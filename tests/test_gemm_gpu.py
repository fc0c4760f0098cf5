"""GPU parity: emulated-accumulation GEMM against the oracle, bit-exact.

The oracle (oracle/lof_oracle.c) computes every output as the reference leaf
acc = round(RTO_fma(a, b, acc)) in ascending k (cutlass_gemms.cuh:139-165), or the k-chunk
rule for chunk > 1, then D = alpha * acc.  Covered: every tile config, all four operand
layouts, ragged shapes, accumulator mantissas 0..23 and full formats (5e10m saturating and
extended, e4m3, bf16-like), chunks 1/2/4/5/8/16/32, the MFMA chunk-sum configs (exact,
mixed and inexact slices, the forced VALU chain and forced MFMA), the general leaf path, non-RNE
accumulator modes, +-0 / inf / NaN / FP32-subnormal / overflowing inputs, the reference
lof_gemm entry point with scale_a / scale_b, and sampled outputs of a 4096^3 product.
"""
import numpy as np
import pytest
import torch

import oracle
from conftest import assert_bits

pytestmark = pytest.mark.gpu

FP16_SAT = (16, 10, 15, 1, 0, 0)
FP16_EXT = (16, 10, 15, 0, 0, 0)
E4M3 = (8, 3, 7, 1, 0, 0)
BF16_LIKE = (16, 7, 127, 1, 0, 0)


def _fmt(t):
    import lofloat_amd as lof
    return lof.FloatFormatDescriptor(*t, lof.formats.P3109_InfChecker(t[0]),
                                     lof.formats.P3109_NaNChecker(t[0]))


def _mats(M, N, K, seed, kind="normal"):
    g = np.random.default_rng(seed)
    A = (g.standard_normal((M, K)) / np.sqrt(max(K, 1))).astype(np.float32)
    B = (g.standard_normal((K, N))).astype(np.float32)
    if kind == "e4m3":  # pre-rounded inputs: exact products, many exact ties
        A = oracle.virtual_round(A * 16, E4M3).reshape(M, K)
        B = oracle.virtual_round(B, E4M3).reshape(K, N)
    elif kind == "special":
        A.flat[g.choice(A.size, max(1, A.size // 50), replace=False)] = 0.0
        A.flat[g.choice(A.size, max(1, A.size // 50), replace=False)] = -0.0
        B.flat[g.choice(B.size, max(1, B.size // 50), replace=False)] = -0.0
        A.flat[g.choice(A.size, max(1, A.size // 200), replace=False)] = 1e-41
        B.flat[g.choice(B.size, max(1, B.size // 200), replace=False)] = 3e37
        A.flat[g.choice(A.size, 1, replace=False)] = np.inf
        B.flat[g.choice(B.size, 1, replace=False)] = np.nan
    elif kind == "wide":  # magnitudes across the FP32 range
        A = (A * np.exp2(g.uniform(-60, 60, A.shape))).astype(np.float32)
        B = (B * np.exp2(g.uniform(-60, 60, B.shape))).astype(np.float32)
    elif kind == "tiny":  # partial sums around 2^-14 .. 2^-26: the fp16 subnormal range
        A = (A * np.float32(2.0 ** -6)).astype(np.float32)
        B = (B * np.float32(2.0 ** -8)).astype(np.float32)
    elif kind in ("e4m3_small", "mixed"):  # the GEMM workload's inputs: N(0, 1/64) in e4m3
        A = oracle.virtual_round((g.standard_normal((M, K)) / 64).astype(np.float32),
                                 E4M3).reshape(M, K)
        B = oracle.virtual_round((g.standard_normal((K, N)) / 64).astype(np.float32),
                                 E4M3).reshape(K, N)
        if kind == "mixed":  # rows 0-7 scaled by 2^12: tiles holding them fail the exact test
            A[:8] *= np.float32(4096.0)
    return A, B


def _dev_operands(A, B, ta, tb, dev):
    At = torch.from_numpy(np.ascontiguousarray(A.T) if ta else A).to(dev)
    Bt = torch.from_numpy(np.ascontiguousarray(B.T) if tb else B).to(dev)
    return (At.t() if ta else At), (Bt.t() if tb else Bt)


def _run(lof, A, B, accum, dev, ta=False, tb=False, alpha=1.0, config=-1):
    Ad, Bd = _dev_operands(A, B, ta, tb, dev)
    return lof.gemm(Ad, Bd, accum, alpha=alpha, config=config).cpu().numpy()


def _ref(A, B, oacc, alpha=1.0):
    return oracle.gemm(A, B, oacc, alpha=alpha)


def _accums(lof, m=10, chunk=1, fmt=None, mode=0):
    if fmt is None:
        return lof.AccumSpec(mant_bits=m, chunk=chunk, mode=mode), oracle.accum_mant(m, chunk, mode)
    return (lof.AccumSpec(fmt=_fmt(fmt), chunk=chunk, mode=mode),
            oracle.accum_fmt(fmt, chunk, mode))


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 5, 7), (17, 33, 65), (64, 64, 64),
                                   (100, 70, 130), (130, 257, 47), (31, 129, 200)])
def test_shapes_all_configs(dev, shape):
    import lofloat_amd as lof
    M, N, K = shape
    for kind, chunk in (("normal", 1), ("normal", 8), ("e4m3_small", 8)):
        A, B = _mats(M, N, K, seed=M * 7 + N, kind=kind)
        acc, oacc = _accums(lof, 10, chunk)
        ref = _ref(A, B, oacc)
        for cfg in range(lof.gemm_num_configs()):
            if not lof.gemm_config_valid(cfg, chunk):
                with pytest.raises(NotImplementedError):
                    _run(lof, A, B, acc, dev, config=cfg)
                continue
            assert_bits(_run(lof, A, B, acc, dev, config=cfg), ref,
                        f"{shape} {kind} chunk {chunk} config {cfg}")


MFMA_CONFIGS = (5, 6, 7, 8)   # wave tiles 128x128, 128x64, 64x64, 64x128


@pytest.mark.parametrize("kind", ["e4m3_small", "mixed", "normal", "special", "e4m3", "tiny"])
@pytest.mark.parametrize("chunk", [2, 4, 8, 16, 32])
def test_mfma_chunk_paths(dev, kind, chunk):
    """MFMA configs: chunk sums on MFMA where the per-slice exactness test passes, on the
    VALU fmaf chain elsewhere (config | 0x200 forces the chain everywhere); both must equal
    the oracle bit for bit."""
    import lofloat_amd as lof
    A, B = _mats(200, 150, 100, seed=chunk, kind=kind)
    for acc, oacc in (_accums(lof, 10, chunk), _accums(lof, fmt=FP16_SAT, chunk=chunk)):
        ref = _ref(A, B, oacc)
        for cfg in MFMA_CONFIGS:
            assert lof.gemm_config_valid(cfg, chunk)
            assert_bits(_run(lof, A, B, acc, dev, config=cfg), ref, f"{kind} c{chunk} cfg {cfg}")
            assert_bits(_run(lof, A, B, acc, dev, config=cfg | 0x200), ref,
                        f"{kind} c{chunk} cfg {cfg} chain")
            assert_bits(_run(lof, A, B, acc, dev, config=cfg | 0x10000), ref,
                        f"{kind} c{chunk} cfg {cfg} bf16 MFMAs off")


@pytest.mark.parametrize("cfg", MFMA_CONFIGS)
def test_mfma_forced(dev, cfg):
    """config | 0x400 runs every chunk on MFMA without the exactness test: on inputs whose
    chunk sums are exact the MFMA chunk sums must give the oracle's result (this checks the
    MFMA operand and result layouts on every tile position).  The automatic choice must
    match the oracle on general FP32 inputs as well."""
    import lofloat_amd as lof
    for acc, oacc in (_accums(lof, 10, 32), _accums(lof, fmt=FP16_SAT, chunk=4),
                      _accums(lof, 7, 16), _accums(lof, fmt=FP16_EXT, chunk=8),
                      _accums(lof, 10, 2)):
        A, B = _mats(256, 192, 320, seed=cfg, kind="e4m3_small")
        ref = _ref(A, B, oacc)
        # e4m3 values are bfloat16 values: forced chunk sums on the bf16 MFMAs (chunks 8-32)
        # or the f32 MFMA (chunks 2, 4), which pins their operand and result layouts; 0x10000
        # (bf16 MFMAs off) must not change a forced run
        assert_bits(_run(lof, A, B, acc, dev, config=cfg | 0x400), ref,
                    f"forced exact chunk {acc.chunk}")
        assert_bits(_run(lof, A, B, acc, dev, config=cfg | 0x10400), ref,
                    f"forced exact, bf16 flag off, chunk {acc.chunk}")
        A, B = _mats(256, 192, 320, seed=cfg, kind="normal")
        assert_bits(_run(lof, A, B, acc, dev, config=cfg), _ref(A, B, oacc),
                    "auto on inexact inputs")


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_layouts(dev, ta, tb):
    """Transposed operands on every tile kind: a ragged shape (edge loads everywhere) and one
    whose tiles are all interior (the float4 slice loads, whose per-thread coordinates differ
    by layout and by tile shape), at chunks 1, 8 and 32."""
    import lofloat_amd as lof
    for M, N, K, kind in ((96, 80, 112, "normal"), (256, 256, 128, "normal"),
                          (256, 256, 128, "e4m3_small")):
        A, B = _mats(M, N, K, seed=5, kind=kind)
        for chunk in (1, 8, 32):
            acc, oacc = _accums(lof, 10, chunk)
            ref = _ref(A, B, oacc)
            for cfg in (0, 3, 5, 6, 7, 8):
                if not lof.gemm_config_valid(cfg, chunk):
                    continue
                y = _run(lof, A, B, acc, dev, ta, tb, config=cfg)
                assert_bits(y, ref, f"{M}x{N}x{K} {kind} ta={ta} tb={tb} chunk {chunk} cfg {cfg}")


@pytest.mark.parametrize("m", [0, 1, 3, 7, 10, 15, 19, 20, 21, 22, 23])
@pytest.mark.parametrize("kind", ["normal", "e4m3", "wide"])
def test_mantissa_accumulators(dev, m, kind):
    import lofloat_amd as lof
    A, B = _mats(64, 72, 150, seed=m, kind=kind)
    for chunk in (1, 5, 32):
        acc, oacc = _accums(lof, m, chunk)
        assert_bits(_run(lof, A, B, acc, dev), _ref(A, B, oacc), f"m={m} {kind} chunk {chunk}")


@pytest.mark.parametrize("fmt", [FP16_SAT, FP16_EXT, E4M3, BF16_LIKE])
@pytest.mark.parametrize("kind", ["normal", "e4m3", "wide", "tiny"])
def test_format_accumulators(dev, fmt, kind):
    import lofloat_amd as lof
    A, B = _mats(48, 64, 200, seed=fmt[1], kind=kind)
    for chunk in (1, 5, 8, 32):
        acc, oacc = _accums(lof, chunk=chunk, fmt=fmt)
        assert_bits(_run(lof, A, B, acc, dev), _ref(A, B, oacc), f"{fmt} {kind} chunk {chunk}")


@pytest.mark.parametrize("kind", ["normal", "tiny", "e4m3_small"])
@pytest.mark.parametrize("fmt", [FP16_SAT, E4M3])
def test_format_accumulators_every_config(dev, kind, fmt):
    """Every tile config that takes the chunk, full-format accumulators whose subnormal range
    the fast paths cover, on FP32 inputs (products not exact in FP32), subnormal-range sums
    and e4m3 inputs."""
    import lofloat_amd as lof
    A, B = _mats(150, 100, 140, seed=17, kind=kind)
    for chunk in (1, 4, 32):
        acc, oacc = _accums(lof, fmt=fmt, chunk=chunk)
        ref = _ref(A, B, oacc)
        for cfg in range(lof.gemm_num_configs()):
            if lof.gemm_config_valid(cfg, chunk):
                assert_bits(_run(lof, A, B, acc, dev, config=cfg), ref,
                            f"{fmt} {kind} chunk {chunk} cfg {cfg}")


@pytest.mark.parametrize("kind", ["normal", "e4m3", "special", "wide"])
def test_general_path_forced(dev, kind):
    """config | 0x100 sends every leaf through the general (literal RTO) path."""
    import lofloat_amd as lof
    A, B = _mats(40, 56, 90, seed=3, kind=kind)
    for acc, oacc in (_accums(lof, 10), _accums(lof, 7, 8), _accums(lof, fmt=FP16_SAT),
                      _accums(lof, fmt=FP16_EXT, chunk=32)):
        ref = _ref(A, B, oacc)
        assert_bits(_run(lof, A, B, acc, dev, config=0x100), ref, f"general {kind}")
        assert_bits(_run(lof, A, B, acc, dev, config=0), ref, f"fast {kind}")
        if lof.gemm_config_valid(5, acc.chunk):
            assert_bits(_run(lof, A, B, acc, dev, config=5 | 0x100), ref, f"mfma general {kind}")


@pytest.mark.parametrize("mode", [1, 2, 4, 5, 6, 7, 12, 13])
def test_accumulator_modes(dev, mode):
    import lofloat_amd as lof
    A, B = _mats(33, 45, 77, seed=mode, kind="e4m3")
    for acc, oacc in (_accums(lof, 10, mode=mode), _accums(lof, fmt=FP16_SAT, mode=mode, chunk=4)):
        assert_bits(_run(lof, A, B, acc, dev), _ref(A, B, oacc), f"mode {mode}")


def test_special_values(dev):
    import lofloat_amd as lof
    A, B = _mats(70, 66, 129, seed=11, kind="special")
    for acc, oacc in (_accums(lof, 10), _accums(lof, 23), _accums(lof, fmt=FP16_SAT),
                      _accums(lof, fmt=FP16_EXT, chunk=8), _accums(lof, 10, chunk=32)):
        assert_bits(_run(lof, A, B, acc, dev), _ref(A, B, oacc), "special")


def test_zero_sign_rule(dev):
    """An exactly-zero accumulator takes roundTowardNegative's sign (cutlass_gemms.cuh:74-80):
    (+0)+(+0) stays +0, cancellations and -0 products give -0."""
    import lofloat_amd as lof
    A = np.array([[1.0, -1.0], [0.0, 0.0], [-0.0, 0.0], [2.0, 0.0]], dtype=np.float32)
    B = np.array([[1.0], [1.0]], dtype=np.float32)
    acc, oacc = _accums(lof, 10)
    y = _run(lof, A, B, acc, dev)
    ref = _ref(A, B, oacc)
    assert_bits(y, ref, "zero signs")
    assert np.signbit(y[0, 0]) and not np.signbit(y[1, 0]) and np.signbit(y[2, 0])


@pytest.mark.parametrize("chunk", [1, 8, 32])
def test_zero_sums_every_config(dev, chunk):
    """Exact-zero sums on every tile config and both fast leaves: sparse small-integer
    operands with +0 / -0 entries and column pairs that cancel exactly, so outputs hit
    (+0)+(+0), -0 products and cancellations mid-chain; signs compared bit for bit."""
    import lofloat_amd as lof
    g = np.random.default_rng(chunk)
    M, N, K = 70, 66, 96
    A = g.integers(-3, 4, (M, K)).astype(np.float32)
    A[g.random((M, K)) < 0.5] = 0.0
    A[g.random((M, K)) < 0.1] = -0.0
    A[:, 1:48:2] = -A[:, 0:48:2]                   # the first 24 k pairs cancel exactly
    A[5] = 0.0
    A[6] = -0.0
    B = g.integers(-3, 4, (K, N)).astype(np.float32)
    B[g.random((K, N)) < 0.3] = -0.0
    B[1:48:2] = B[0:48:2]
    B[:, 7] = 0.0
    for acc, oacc in (_accums(lof, 10, chunk), _accums(lof, fmt=FP16_EXT, chunk=chunk)):
        ref = _ref(A, B, oacc)
        zeros = ref[ref == 0]
        assert 0 < zeros.size < ref.size and np.signbit(zeros).any() and not np.signbit(zeros).all()
        for cfg in range(lof.gemm_num_configs()):
            if lof.gemm_config_valid(cfg, chunk):
                assert_bits(_run(lof, A, B, acc, dev, config=cfg), ref,
                            f"zero sums chunk {chunk} config {cfg}")
                assert_bits(_run(lof, A, B, acc, dev, config=cfg | 0x100), ref,
                            f"zero sums chunk {chunk} config {cfg} general leaf")


@pytest.mark.parametrize("chunk", [1, 2, 8, 32])
def test_exact_sum_transitions(dev, chunk):
    """Slices and lanes moving between the exact-sum leaves and the general ones: rows 0-7
    start with 2^-12-scaled values (their accumulators stay tiny, with low bits a
    mantissa-only accumulator cannot add exactly to the products that follow) before normal
    e4m3 data, and k 160..191 of rows 16-23 are 2^-14-scaled (those tiles' slice fails the
    test for every lane), so lanes change kind of slice both ways mid-product."""
    import lofloat_amd as lof
    M, N, K = 96, 80, 320
    A, B = _mats(M, N, K, seed=chunk, kind="e4m3_small")
    A[:8, :64] = oracle.virtual_round(A[:8, :64] * np.float32(2.0 ** -12), E4M3).reshape(8, 64)
    A[16:24, 160:192] = oracle.virtual_round(A[16:24, 160:192] * np.float32(2.0 ** -14),
                                             E4M3).reshape(8, 32)
    for acc, oacc in (_accums(lof, 10, chunk), _accums(lof, fmt=FP16_SAT, chunk=chunk),
                      _accums(lof, fmt=FP16_EXT, chunk=chunk)):
        ref = _ref(A, B, oacc)
        for cfg in range(lof.gemm_num_configs()):
            if lof.gemm_config_valid(cfg, chunk):
                assert_bits(_run(lof, A, B, acc, dev, config=cfg), ref,
                            f"transitions chunk {chunk} config {cfg}")
        for cfg in MFMA_CONFIGS:   # 0x800: the same tiles with the exact-sum leaves off
            assert_bits(_run(lof, A, B, acc, dev, config=cfg | 0x800), ref,
                        f"transitions chunk {chunk} config {cfg} without exact-sum leaves")


def _held_corner(seed, M=64, N=128, K=64):
    """Rows 0/1 and 32/33 (one lane's accumulators): slice 0 leaves a large and a tiny
    accumulator with bits 2^-10 below its exponent, slice 1's products lie between them, so
    the exact-sum test passes with S - Lf near 40 for a 10-bit mantissa-only accumulator."""
    g = np.random.default_rng(seed)
    A = np.zeros((M, K), np.float32)
    B = np.zeros((K, N), np.float32)
    B[:32] = oracle.virtual_round((g.standard_normal((32, N)) / 64).astype(np.float32),
                                  E4M3).reshape(32, N)
    B[0] = 0.125
    B[32:] = g.choice([0.125, 0.25, 0.375, -0.125, -0.25, -0.375], (32, N))
    for r in (0, 32):
        A[r, 0] = 2.0 ** int(g.integers(10, 14))
        A[r + 1, 1:32] = oracle.virtual_round_mant(
            (g.standard_normal(31) * 2.0 ** int(g.integers(-20, -14))).astype(np.float32), 3)
        A[r:r + 2, 32:] = g.choice([1.0, 1.5, -1.0, -1.5], (2, 32)) * 2.0 ** int(g.integers(-9, -5))
    return A, B


@pytest.mark.parametrize("case", ["e4m3", "tiny", "rows", "corner"])
def test_binary16_held_leaves(dev, case):
    """Chunk 1's binary16-held leaves (the lane's accumulators held as binary16, scaled by
    2^hs for a 10-bit mantissa-only accumulator) against the oracle, and the same tiles with
    only their packed binary16 form off (0x4000) and with them off (0x2000): e4m3 operands of the GEMM workload; tiny ones (sums in binary16's
    subnormal range); per-row, per-slice scales of 2^-20 .. 2^20 (lanes move between
    passing and failing slices with varying hs); and lanes holding a large and a tiny
    accumulator where S - Lf is near the test's limit of 39."""
    import lofloat_amd as lof
    g = np.random.default_rng(7)
    if case == "corner":
        A, B = _held_corner(11)
    else:
        M, N, K = 96, 144, 320
        A, B = _mats(M, N, K, seed=31, kind="e4m3_small")
        if case == "tiny":
            A = oracle.virtual_round(A * np.float32(2.0 ** -4), E4M3).reshape(M, K)
        elif case == "rows":
            rs = np.repeat(2.0 ** g.integers(-20, 21, (M, K // 32)), 32, axis=1)
            A = oracle.virtual_round_mant((A * rs).astype(np.float32), 3).reshape(M, K)
    for acc, oacc in (_accums(lof, 10, 1), _accums(lof, fmt=FP16_SAT, chunk=1),
                      _accums(lof, fmt=FP16_EXT, chunk=1)):
        ref = _ref(A, B, oacc)
        for cfg in MFMA_CONFIGS:
            for bits in (0, 0x4000, 0x2000):
                assert_bits(_run(lof, A, B, acc, dev, config=cfg | bits), ref,
                            f"held {case} config {cfg} bits {bits:#x}")


def test_packed_rto_leaves(dev):
    """Chunk 1's packed binary16 FMA on RTO lanes (binary16-shaped accumulators, sums
    inexact in FP32): bf16-like 8e7m operands of N(0, 1/64) with every value binary16 holds
    (magnitudes below 2^-16 zeroed), except one tiny value in k 96..127 of A and of B, whose
    slices then take the RTO leaves; against the oracle, and with the packed form off
    (0x4000) and with every short leaf off (0x1800)."""
    import lofloat_amd as lof
    g = np.random.default_rng(808)
    M, N, K = 96, 144, 320
    A = oracle.virtual_round((g.standard_normal((M, K)) / 64).astype(np.float32),
                             BF16_LIKE).reshape(M, K)
    B = oracle.virtual_round((g.standard_normal((K, N)) / 64).astype(np.float32),
                             BF16_LIKE).reshape(K, N)
    A[np.abs(A) < 2.0 ** -16] = 0.0
    B[np.abs(B) < 2.0 ** -16] = 0.0
    assert np.array_equal(A.astype(np.float16).astype(np.float32), A)
    A[5, 100] = np.float32(1.5 * 2.0 ** -30)
    B[110, 7] = np.float32(1.25 * 2.0 ** -28)
    for acc, oacc in (_accums(lof, fmt=FP16_SAT, chunk=1), _accums(lof, fmt=FP16_EXT, chunk=1)):
        ref = _ref(A, B, oacc)
        for cfg in MFMA_CONFIGS:
            for bits in (0, 0x4000, 0x1800):
                assert_bits(_run(lof, A, B, acc, dev, config=cfg | bits), ref,
                            f"packed rto config {cfg} bits {bits:#x}")


@pytest.mark.parametrize("chunk", [1, 2, 8, 32])
def test_rto_leaves(dev, chunk):
    """Inexact sums on the RTO leaves: bf16-like 8e7m operands of N(0, 1/64) (16-bit
    products, exact in FP32; tiny values make acc + p inexact), with mantissa-only and
    binary16-shaped accumulators, against the oracle; then the same tiles with the RTO leaves
    off (0x1000) and with both fast leaf kinds off (0x1800)."""
    import lofloat_amd as lof
    g = np.random.default_rng(100 + chunk)
    M, N, K = 96, 80, 320
    A = oracle.virtual_round((g.standard_normal((M, K)) / 64).astype(np.float32),
                             BF16_LIKE).reshape(M, K)
    B = oracle.virtual_round((g.standard_normal((K, N)) / 64).astype(np.float32),
                             BF16_LIKE).reshape(K, N)
    A[3, :40] = 0.0                                          # zero sums at the start of a row
    B[:, 5] *= np.float32(2.0 ** -20)                        # a column of tiny products
    for acc, oacc in (_accums(lof, 10, chunk), _accums(lof, 7, chunk), _accums(lof, 19, chunk),
                      _accums(lof, fmt=FP16_SAT, chunk=chunk), _accums(lof, fmt=FP16_EXT, chunk=chunk)):
        ref = _ref(A, B, oacc)
        for cfg in MFMA_CONFIGS:
            for bits in (0, 0x1000, 0x1800):
                assert_bits(_run(lof, A, B, acc, dev, config=cfg | bits), ref,
                            f"rto chunk {chunk} config {cfg} bits {bits:#x}")


@pytest.mark.parametrize("chunk", [1, 8, 32])
@pytest.mark.parametrize("kind", ["uniform", "mx", "wide", "overflow", "tiny", "split"])
def test_directed_rounding_leaves(dev, kind, chunk):
    """Products or chunk sums that FP32 cannot hold exactly (24-bit significands): the
    directed-rounding leaves (round-to-odd from the sum rounded down and up) for
    mantissa-only accumulators of every width and for binary16-shaped formats, at chunk 1 and
    on the chunk sums of chunks 8 and 32, against the oracle on every wave tile, and the same
    tiles with those leaves off (0x8000) or with the mixed-precision FMA leaves of the
    binary16-shaped formats off (0x20000).  Kinds: FP32 uniform(-1, 1) operands (the reference
    harness's lof_gemm inputs), MX-rounded e4m3 operands whose overflowing elements pass
    through unrounded, magnitudes across the FP32 range, sums that overflow to FLT_MAX | 1
    and inf, products in the FP32 subnormal range, and columns 2^40 apart (the checked
    binary16-held leaves of the 10-bit accumulator fall back to the directed ones)."""
    import lofloat_amd as lof
    g = np.random.default_rng({"uniform": 1, "mx": 2, "wide": 3, "overflow": 4, "tiny": 5,
                               "split": 6}[kind])
    M, N, K = 80, 96, 192
    if kind == "uniform":
        A = g.uniform(-1, 1, (M, K)).astype(np.float32)
        B = g.uniform(-1, 1, (K, N)).astype(np.float32)
    elif kind == "mx":
        A = oracle.virtual_mx_round((g.standard_normal(M * K) * 0.02).astype(np.float32), 32,
                                    E4M3, (8, 0, 127, 0, 0, 1)).reshape(M, K)
        Bt = oracle.virtual_mx_round((g.standard_normal(N * K) * 0.02).astype(np.float32), 32,
                                     E4M3, (8, 0, 127, 0, 0, 1)).reshape(N, K)
        B = np.ascontiguousarray(Bt.T)
    elif kind == "wide":
        A, B = _mats(M, N, K, seed=31, kind="wide")
    elif kind == "overflow":
        A = (g.standard_normal((M, K)) * 2.0 ** 62).astype(np.float32)
        B = (g.standard_normal((K, N)) * 2.0 ** 64).astype(np.float32)
    elif kind == "tiny":
        A = (g.standard_normal((M, K)) * 2.0 ** -70).astype(np.float32)
        B = (g.standard_normal((K, N)) * 2.0 ** -62).astype(np.float32)
    else:   # columns 2^40 apart within each lane's pair: the small ones scale into binary16's
        # subnormal range, which sends the checked 10-bit leaves back to the directed ones
        A = g.uniform(-1, 1, (M, K)).astype(np.float32)
        B = g.uniform(-1, 1, (K, N)).astype(np.float32)
        B[:, (np.arange(N) // 32) % 2 == 1] *= np.float32(2.0 ** -40)
    accs = [_accums(lof, m, chunk) for m in (0, 5, 10, 19, 20, 21, 22, 23)]
    accs += [_accums(lof, fmt=FP16_SAT, chunk=chunk), _accums(lof, fmt=FP16_EXT, chunk=chunk)]
    for acc, oacc in accs:
        ref = _ref(A, B, oacc)
        what = f"{kind} chunk {chunk} m={acc.mant_bits if acc.fmt is None else 'fp16'}"
        for cfg in MFMA_CONFIGS:
            for bits in (0, 0x8000, 0x20000):
                assert_bits(_run(lof, A, B, acc, dev, config=cfg | bits), ref,
                            f"{what} config {cfg} bits {bits:#x}")
        assert_bits(_run(lof, A, B, acc, dev), ref, f"{what} default config")


def test_alpha_and_reference_entry(dev):
    import lofloat_amd as lof
    A, B = _mats(64, 48, 96, seed=2)
    acc, oacc = _accums(lof, 10)
    assert_bits(_run(lof, A, B, acc, dev, alpha=0.25), _ref(A, B, oacc, alpha=0.25), "alpha")
    Ad, Bd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    # run_scaled: alpha = 1 / (scale_a * scale_b) in float, 1 if the product is 0
    D = lof.lof_gemm(Ad, Bd, 10, lof.RoundingMode.RoundTowardsZero, 3, 4.0, 0.5).cpu().numpy()
    assert_bits(D, _ref(A, B, oacc, alpha=np.float32(1) / (np.float32(4.0) * np.float32(0.5))),
                "scaled")
    D0 = lof.lof_gemm(Ad, Bd, 10, scale_a=0.0).cpu().numpy()
    assert_bits(D0, _ref(A, B, oacc), "scale 0 -> alpha 1")
    D1 = torch.ops.lofloat.lof_gemm(Ad, Bd, 10, 0, 0, 1.0, 1.0).cpu().numpy()
    assert_bits(D1, _ref(A, B, oacc), "torch.ops.lofloat.lof_gemm")


@pytest.mark.parametrize("inputs", ["e4m3", "bf16"])
@pytest.mark.parametrize("chunk", [1, 8, 32])
def test_workload_4096_sampled(dev, inputs, chunk):
    """M = N = K = 4096 with e4m3 or 8e7m pre-rounded N(0, 1/sqrt(K)) inputs and a 5e10m
    (16,10,15) accumulator; 512 sampled outputs checked against the oracle with full K."""
    import lofloat_amd as lof
    n = 4096
    g = np.random.default_rng(4096 + chunk)
    fm = E4M3 if inputs == "e4m3" else BF16_LIKE
    A = oracle.virtual_round((g.standard_normal((n, n)) / 64).astype(np.float32), fm).reshape(n, n)
    B = oracle.virtual_round((g.standard_normal((n, n)) / 64).astype(np.float32), fm).reshape(n, n)
    samples = g.choice(n * n, 512, replace=False)
    for acc, oacc in (_accums(lof, 10, chunk), _accums(lof, fmt=FP16_SAT, chunk=chunk)):
        y = _run(lof, A, B, acc, dev).ravel()[samples]
        ref = oracle.gemm(A, B, oacc, samples=samples)
        assert_bits(y, ref, f"4096^3 {inputs} chunk {chunk}")


def test_errors(dev):
    import lofloat_amd as lof
    a = torch.ones(4, 5, device=dev)
    with pytest.raises(RuntimeError):
        lof.lof_gemm(a, torch.ones(4, 3, device=dev), 10)
    with pytest.raises(NotImplementedError):
        lof.gemm(a, torch.ones(5, 3, device=dev), lof.AccumSpec(mant_bits=24))
    with pytest.raises(NotImplementedError):
        lof.gemm(a, torch.ones(5, 3, device=dev), lof.AccumSpec(mant_bits=10, mode=3))
    with pytest.raises(ValueError):
        lof.gemm(a, torch.ones(5, 3, device=dev), lof.AccumSpec(mant_bits=10, chunk=0))


def _midpoint_traps(g, n, span):
    """n exact sums x = T + d with T a binary16 midpoint in [2^-10, 2^10) and d far below
    T's FP32 resolution, and a binary16 start value h a few ulps from T (x - h is then small
    enough to be an FP32 value when span allows it)."""
    e = g.integers(-10, 10, n)
    q = np.exp2(e - 10.0)
    m = g.integers(1024, 2048, n)
    T = (m + 0.5) * q
    h = (m + g.integers(-span, span + 1, n)) * q
    d = g.choice([-1.0, 1.0], n) * np.exp2(e - 24.0 - g.integers(2, 4, n))
    s = g.choice([-1.0, 1.0], n)
    return s * T, s * h, s * d


@pytest.mark.parametrize("fmt", [FP16_SAT, FP16_EXT])
def test_mixed_fma_leaves_round_once(dev, fmt):
    """Binary16-shaped accumulators: the mixed-precision FMA leaves must round the exact sum
    once.  Each output n is built so that after its first chunk acc = h (a binary16 value)
    and its second chunk brings x = acc + p to T + d, just off a binary16 midpoint T:
    rounding the reference's round-to-odd sum goes to the side of d, while rounding the
    sum to FP32 first would land on T and round to even.  Chunk 1: p = a b with a 24-bit a
    (the product is inexact in FP32); chunk 8: p is an FP32 chunk sum.  Both against the
    oracle, with the mixed FMA leaves on and off (0x20000), on every wave tile; the data
    gives a double-rounding result that differs in a large share of outputs."""
    import lofloat_amd as lof
    from fractions import Fraction
    g = np.random.default_rng(77)
    N = 192
    # chunk 1: K = 2, A = [1, a1], B[0] = h, B[1] = (x - h) / a1 rounded to FP32
    T, h, d = _midpoint_traps(g, N, 3)
    a1 = np.float32(1.0 + 1531 * 2.0 ** -23)
    b1 = ((T + d - h) / np.float64(a1)).astype(np.float32)
    A = np.array([[1.0, a1]], np.float32)
    B = np.stack([h.astype(np.float32), b1])
    # chunk 8: K = 16, acc = h after the first chunk, the second chunk's sum is p = x - h
    T8, h8, d8 = _midpoint_traps(g, N, 3)
    p8 = (T8 + d8 - h8).astype(np.float32)
    assert np.all(p8.astype(np.float64) == T8 + d8 - h8)     # p is an FP32 value
    A8 = np.zeros((1, 16), np.float32)
    A8[0, 0] = 1.0
    A8[0, 8] = 1.0
    B8 = np.zeros((16, N), np.float32)
    B8[0] = h8
    B8[8] = p8
    for chunk, (AA, BB) in ((1, (A, B)), (8, (A8, B8))):
        acc, oacc = _accums(lof, fmt=fmt, chunk=chunk)
        ref = _ref(AA, BB, oacc)
        # the share of outputs where rounding through FP32 first would give other bits
        first = BB[0].astype(np.float64) * float(AA[0, 0])
        if chunk == 1:
            xs = [Fraction(float(f)) + Fraction(float(a1)) * Fraction(float(b)) for f, b in zip(first, BB[1])]
        else:
            xs = [Fraction(float(f)) + Fraction(float(p)) for f, p in zip(first, BB[8])]
        dbl = np.array([np.float16(np.float32(float(x))) for x in xs], np.float16).astype(np.float32)
        assert np.mean(dbl != ref.ravel()) > 0.3
        for cfg in MFMA_CONFIGS:
            for bits in (0, 0x20000):
                assert_bits(_run(lof, AA, BB, acc, dev, config=cfg | bits), ref,
                            f"chunk {chunk} config {cfg} bits {bits:#x}")
