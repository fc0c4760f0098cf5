"""GPU checks of the grouped persistent launch and its HIP-graph capture.

One launch runs many rounding / MX / GEMM jobs of uneven size, each with its own format
and mode; every job's output must equal (bit for bit) both the oracle and the job's own
single-job kernel.  Graph replays are run several times with eager work in between and
each replay is checked (a replay must not depend on state a previous launch left).
"""
import numpy as np
import pytest
import torch

import oracle
from conftest import assert_bits, c1_values, random_words

pytestmark = pytest.mark.gpu


def _fmt(t):
    import lofloat_amd as lof
    return lof.FloatFormatDescriptor(*t, lof.formats.P3109_InfChecker(t[0]),
                                     lof.formats.P3109_NaNChecker(t[0]))


def _make_jobs(lof, dev, seed=0, n_round=24, n_gemm=4, max_log2=18):
    g = np.random.default_rng(seed)
    jobs, refs = [], []
    modes = [0, 12, 4, 1, 3]
    for i in range(n_round):
        n = int(np.exp2(g.uniform(8, max_log2))) + int(g.integers(0, 7))
        x = c1_values(n, seed=seed * 1000 + i)
        e = int(g.integers(2, 9))
        m = int(g.integers(1, 11))
        ft = (1 + e + m, m, (1 << (e - 1)) - 1, int(g.integers(0, 2)), 0, 0)
        mode = modes[i % len(modes)]
        L = min(8, 23 - m) if mode == 3 else 0
        bits = random_words(n, seed=i) if mode == 3 else None
        xt = torch.from_numpy(x).to(dev)
        bt = None if bits is None else torch.from_numpy(bits.view(np.int32)).to(dev)
        kind = i % 4
        if kind == 3:  # mantissa-only job
            jobs.append(lof.RoundJob(xt, m, mode, L, rand_bits=bt))
            refs.append(oracle.virtual_round_mant(x, m, mode, L, bits=bits))
        elif kind == 2 and n >= 64:  # MX job
            bs = 32
            ef = (8, 3, 7, 1, 0, 0)
            jobs.append(lof.MXJob(xt, bs, _fmt(ef), lof.create_e8m0_params(), mode, 0, L,
                                  rand_bits=bt))
            refs.append(oracle.virtual_mx_round(x, bs, ef, (8, 0, 127, 0, 0, 1), mode, 0, L,
                                                bits=bits))
        else:
            sc = 1.0 if i % 5 else 2.5
            jobs.append(lof.RoundJob(xt, _fmt(ft), mode, L, scale=sc, rand_bits=bt))
            refs.append(oracle.virtual_round(x, ft, mode, L, scale=sc, bits=bits))
    for j in range(n_gemm):
        M, N, K = (int(v) for v in g.choice([64, 96, 128, 200], 3))
        A = (g.standard_normal((M, K)) / 8).astype(np.float32)
        B = (g.standard_normal((K, N)) / 8).astype(np.float32)
        if j % 2 == 1:  # e4m3 operands: chunk sums exact, MFMA tiles take them
            A = oracle.virtual_round(A / 8, (8, 3, 7, 1, 0, 0)).reshape(M, K)
            B = oracle.virtual_round(B / 8, (8, 3, 7, 1, 0, 0)).reshape(K, N)
        chunk = [1, 8, 32, 4][j % 4]
        acc = lof.AccumSpec(mant_bits=10, chunk=chunk) if j % 2 == 0 else \
            lof.AccumSpec(fmt=_fmt((16, 10, 15, 1, 0, 0)), chunk=chunk)
        oacc = oracle.accum_mant(10, chunk) if j % 2 == 0 else \
            oracle.accum_fmt((16, 10, 15, 1, 0, 0), chunk)
        Bd = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev).t() if j == 1 else \
            torch.from_numpy(B).to(dev)
        jobs.append(lof.GemmJob(torch.from_numpy(A).to(dev), Bd, acc))
        refs.append(oracle.gemm(A, B, oacc))
    return jobs, refs


def _check(jobs, refs, what):
    for k, (j, r) in enumerate(zip(jobs, refs)):
        assert_bits(j.out.cpu().numpy().ravel(), np.asarray(r).ravel(), f"{what} job {k}")


@pytest.mark.parametrize("config", range(8))
def test_grouped_matches_oracle(dev, config):
    import lofloat_amd as lof
    jobs, refs = _make_jobs(lof, dev, seed=config)
    plan = lof.GroupedPlan(jobs, config=config)
    tiles, grid = plan.info()
    assert tiles > 0 and grid > 0
    for rep in range(3):  # repeated launches reuse the self-resetting work queue
        for j in jobs:
            j.out.fill_(np.nan)
        plan.run()
        torch.cuda.synchronize()
        _check(jobs, refs, f"grouped cfg {config} rep {rep}")


def test_separate_launches_match(dev):
    import lofloat_amd as lof
    jobs, refs = _make_jobs(lof, dev, seed=7)
    plan = lof.GroupedPlan(jobs)
    plan.run_separate()
    torch.cuda.synchronize()
    _check(jobs, refs, "separate")


def test_graph_replays(dev):
    import lofloat_amd as lof
    jobs, refs = _make_jobs(lof, dev, seed=11)
    plan = lof.GroupedPlan(jobs)
    for separate in (False, True):
        graph = plan.capture_graph(separate=separate)
        for rep in range(4):
            for j in jobs:
                j.out.fill_(-1.0)                       # eager work between replays
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            _check(jobs, refs, f"graph separate={separate} replay {rep}")


def test_autotune_and_empty(dev):
    import lofloat_amd as lof
    jobs, refs = _make_jobs(lof, dev, seed=3, n_round=8, n_gemm=1)
    cfg, ms = lof.grouped.autotune(jobs, iters=2)
    assert 0 <= cfg < lof.grouped.num_configs() and ms > 0
    plan = lof.GroupedPlan(jobs)  # uses the tuned default
    plan.run()
    torch.cuda.synchronize()
    _check(jobs, refs, "after autotune")
    empty = lof.GroupedPlan([lof.RoundJob(torch.empty(0, device=dev), 10)])
    assert empty.info()[0] == 0
    empty.run()


def test_device_filling_gemm_runs_alone(dev):
    """A GEMM with at least as many tiles as the persistent grid has workgroups runs as its
    own launch of the single-GEMM kernel, ahead of the persistent launch: the plan's queue
    then holds only the other jobs' tiles, and every output stays bit-identical to the
    per-job launches (the small jobs also to the oracle), through lof_grouped_run, a graph
    replay and lof_grouped_run_noalloc alike."""
    import lofloat_amd as lof
    jobs, refs = _make_jobs(lof, dev, seed=21, n_round=10, n_gemm=2)
    g = np.random.default_rng(21)
    e4m3 = (8, 3, 7, 1, 0, 0)
    A = oracle.virtual_round((g.standard_normal((4096, 64)) / 64).astype(np.float32),
                             e4m3).reshape(4096, 64)
    B = oracle.virtual_round((g.standard_normal((64, 4096)) / 64).astype(np.float32),
                             e4m3).reshape(64, 4096)
    big = lof.GemmJob(torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev),
                      lof.AccumSpec(fmt=_fmt((16, 10, 15, 1, 0, 0)), chunk=32))
    small_tiles = lof.GroupedPlan(jobs).info()[0]
    plan = lof.GroupedPlan(jobs + [big])
    tiles, grid = plan.info()
    assert tiles == small_tiles, (tiles, small_tiles)   # the 4096^2 GEMM left the queue
    assert 4096 * 4096 // (128 * 128) >= grid            # it fills the grid in any config
    plan.run_separate()
    torch.cuda.synchronize()
    want = big.out.cpu().numpy().copy()
    _check(jobs, refs, "separate")
    runs = {"grouped": plan.run, "noalloc": lambda: lof.grouped.run_noalloc(jobs + [big])}
    graph = plan.capture_graph(separate=False)
    runs["graph"] = graph.replay
    for name, run in runs.items():
        for j in jobs + [big]:
            j.out.fill_(np.nan)
        torch.cuda.synchronize()
        run()
        torch.cuda.synchronize()
        _check(jobs, refs, name)
        assert_bits(big.out.cpu().numpy().ravel(), want.ravel(), f"{name} big GEMM")


def test_gemm_autotune(dev):
    import lofloat_amd as lof
    A = torch.randn(256, 192, device=dev)
    B = torch.randn(192, 320, device=dev)
    acc = lof.AccumSpec(mant_bits=10, chunk=8)
    cfg, ms = lof.gemm_autotune(A, B, acc, iters=2)
    assert 0 <= cfg < lof.gemm_num_configs() and ms > 0
    ref = lof.gemm(A, B, acc, config=0)
    assert torch.equal(lof.gemm(A, B, acc), ref)  # every config gives the same bits


def test_gemm_autotune_keeps_aliased_source(dev):
    """lof_gemm_autotune refuses a source C that overlaps D (its timing runs overwrite D), so
    the caller's C survives, and a later lof_gemm_f32 with C == D still reads it."""
    import ctypes
    import lofloat_amd as lof
    from lofloat_amd import _lib
    g = np.random.default_rng(8)
    A = (g.standard_normal((64, 48)) / 8).astype(np.float32)
    B = (g.standard_normal((48, 40)) / 8).astype(np.float32)
    C = g.standard_normal((64, 40)).astype(np.float32)
    Ad, Bd, Dd = (torch.from_numpy(v).to(dev) for v in (A, B, C))
    acc = lof.AccumSpec(mant_bits=10, chunk=8)
    desc = _lib.GemmDesc(64, 40, 48, Ad.data_ptr(), 48, 0, Bd.data_ptr(), 40, 0, Dd.data_ptr(),
                         40, 1.0, acc.c_accum(), Dd.data_ptr(), 40, 0.5)
    best, ms = ctypes.c_int32(-1), ctypes.c_float(0)
    st = _lib.lib.lof_gemm_autotune(ctypes.byref(desc), 2, None, ctypes.byref(best),
                                    ctypes.byref(ms))
    torch.cuda.synchronize()
    assert st == 1 and b"overlaps" in _lib.lib.lof_last_error()
    assert np.array_equal(Dd.cpu().numpy(), C)
    y = lof.gemm(Ad, Bd, acc, out=Dd, C=Dd, beta=0.5).cpu().numpy()
    assert_bits(y, oracle.gemm(A, B, oracle.accum_mant(10, 8), C=C, beta=0.5), "C == D after autotune")
