"""lof_grouped_run_noalloc (the grouped launch with its work list in the library's fixed
device workspace) against lof_grouped_create + lof_grouped_run on the same items, bit for
bit, on two calls in a row, the second with a smaller work list; every output also against
the oracle.  Called through the C ABI directly."""
import ctypes

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


def _jobs(lof, dev, seed, n_round, n_gemm, max_log2):
    g = np.random.default_rng(seed)
    jobs, refs = [], []
    for i in range(n_round):
        n = int(np.exp2(g.uniform(8, max_log2))) + int(g.integers(0, 7))
        x = c1_values(n, seed=seed * 1000 + i)
        e, m = int(g.integers(2, 9)), int(g.integers(1, 11))
        ft = (1 + e + m, m, (1 << (e - 1)) - 1, int(g.integers(0, 2)), 0, 0)
        mode = [0, 12, 4, 1, 3][i % 5]
        L = min(8, 23 - m) if mode == 3 else 0
        bits = random_words(n, seed=seed + i) if mode == 3 else None
        xt = torch.from_numpy(x).to(dev)
        bt = None if bits is None else torch.from_numpy(bits.view(np.int32)).to(dev)
        if i % 3 == 2 and n >= 64:
            jobs.append(lof.MXJob(xt, 32, _fmt((8, 3, 7, 1, 0, 0)), lof.create_e8m0_params(),
                                  mode, 0, L, rand_bits=bt))
            refs.append(oracle.virtual_mx_round(x, 32, (8, 3, 7, 1, 0, 0), (8, 0, 127, 0, 0, 1),
                                                mode, 0, L, bits=bits))
        else:
            jobs.append(lof.RoundJob(xt, _fmt(ft), mode, L, rand_bits=bt))
            refs.append(oracle.virtual_round(x, ft, mode, L, bits=bits))
    for j in range(n_gemm):
        M, N, K = (int(v) for v in g.choice([64, 96, 128, 200], 3))
        A = oracle.virtual_round((g.standard_normal((M, K)) / 64).astype(np.float32),
                                 (8, 3, 7, 1, 0, 0)).reshape(M, K)
        B = oracle.virtual_round((g.standard_normal((K, N)) / 64).astype(np.float32),
                                 (8, 3, 7, 1, 0, 0)).reshape(K, N)
        chunk = [32, 1, 8][j % 3]
        jobs.append(lof.GemmJob(torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev),
                                lof.AccumSpec(mant_bits=10, chunk=chunk)))
        refs.append(oracle.gemm(A, B, oracle.accum_mant(10, chunk)))
    return jobs, refs


def _items(jobs):
    import lofloat_amd as lof
    arr = (lof._lib.GroupItem * len(jobs))()
    for i, j in enumerate(jobs):
        arr[i] = j.item()
    return arr


def test_noalloc_matches_grouped_run(dev):
    import lofloat_amd as lof
    lib = lof._lib.lib
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # second call: fewer and smaller jobs, so stale tasks of the first would show
    for call, (seed, n_round, n_gemm, max_log2) in enumerate(((5, 20, 3, 18), (6, 6, 1, 12))):
        jobs, refs = _jobs(lof, dev, seed, n_round, n_gemm, max_log2)
        plan = lof.GroupedPlan(jobs)
        plan.run()
        torch.cuda.synchronize()
        want = [j.out.clone() for j in jobs]
        for j in jobs:
            j.out.fill_(np.nan)
        items = _items(jobs)
        print(f"noalloc call {call}: {len(jobs)} items", flush=True)
        st = lib.lof_grouped_run_noalloc(items, len(jobs), -1, stream)
        assert st == 0, lib.lof_last_error().decode()
        torch.cuda.synchronize()
        for k, (j, w, r) in enumerate(zip(jobs, want, refs)):
            got = j.out.cpu().numpy().ravel()
            assert_bits(got, w.cpu().numpy().ravel(), f"call {call} job {k} vs grouped run")
            assert_bits(got, np.asarray(r).ravel(), f"call {call} job {k} vs oracle")

