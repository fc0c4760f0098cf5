"""Parity at the full shapes of the LoF_Linear step and the batched workload.

* LoF_Linear: X 8192 x 4096, W 4096 x 4096 (N(0, 0.02)), MX e4m3 / E8M0 operands along K,
  grad_out N(0, 1) rounded to e5m2 by SR-A(8) with counter-based words; the forward, dgrad
  and wgrad GEMMs (K = 4096, 4096, 8192) each checked on 512 sampled outputs, full K, bit for
  bit against the oracle composition, for the 5e10m accumulator at chunk 32 and the 10-bit
  mantissa accumulator at chunk 1.
* The batched workload of bench.py (64 rounding jobs of 2^12 .. 2^24 elements, each with its
  own format and mode, and 12 GEMMs with dimensions up to 8192) in one grouped launch: every
  rounding job compared in full, every GEMM on 512 sampled outputs, against the oracle.
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


@pytest.mark.parametrize("accum", ["fp16_c32", "mant10_c1"])
def test_linear_step_full_shape(dev, accum):
    import lofloat_amd as lof
    M, K, N = 8192, 4096, 4096
    g = torch.Generator(device=dev).manual_seed(2024)
    X = torch.randn(M, K, device=dev, generator=g) * 0.02
    W = torch.randn(N, K, device=dev, generator=g) * 0.02
    G = torch.randn(M, N, device=dev, generator=g)
    seed = 0x5EED
    kw = (dict(accum_format=_fmt(FP16), accum_chunk=32) if accum == "fp16_c32"
          else dict(accum_mant_bits=10, accum_chunk=1))
    lin = lof.LoF_Linear(K, N, act_params=_fmt(E4M3), weight_params=_fmt(E4M3), bias=False,
                         scaling="mx", mx_block_size=32,
                         grad_spec=lof.LinearGradSpec(_fmt(E5M2), 3, 8, seed=seed), **kw).to(dev)
    with torch.no_grad():
        lin.weight.copy_(W)
    x = X.clone().requires_grad_(True)
    out = lin(x)
    out.backward(G)
    torch.cuda.synchronize()
    print(f"linear {accum}: GPU step done", flush=True)

    oacc = (oracle.accum_fmt(FP16, 32) if accum == "fp16_c32" else oracle.accum_mant(10, 1))
    Xh, Wh, Gh = X.cpu().numpy(), W.cpu().numpy(), G.cpu().numpy()
    Xq = oracle.virtual_mx_round(Xh, 32, E4M3, E8M0).reshape(M, K)
    Wq = oracle.virtual_mx_round(Wh, 32, E4M3, E8M0).reshape(N, K)
    Gq = oracle.virtual_round(Gh, E5M2, oracle.SRA, 8, seed=seed).reshape(M, N)
    rng = np.random.default_rng(7)
    s_out = rng.choice(M * N, 512, replace=False)
    s_dx = rng.choice(M * K, 512, replace=False)
    s_dw = rng.choice(N * K, 512, replace=False)
    assert_bits(out.detach().reshape(-1)[torch.from_numpy(s_out).to(dev)].cpu().numpy(),
                oracle.gemm(Xq, Wq, oacc, trans_b=True, samples=s_out), f"{accum} forward")
    assert_bits(x.grad.reshape(-1)[torch.from_numpy(s_dx).to(dev)].cpu().numpy(),
                oracle.gemm(Gq, Wq, oacc, samples=s_dx), f"{accum} dgrad")
    assert_bits(lin.weight.grad.reshape(-1)[torch.from_numpy(s_dw).to(dev)].cpu().numpy(),
                oracle.gemm(Gq, Xq, oacc, trans_a=True, samples=s_dw), f"{accum} wgrad")


def test_batched_workload_full_plan(dev):
    import lofloat_amd as lof
    from bench import c5_jobs
    jobs, _ = c5_jobs(lof, torch, np, dev)
    assert len(jobs) == 76
    plan = lof.GroupedPlan(jobs)
    plan.run()
    torch.cuda.synchronize()
    print(f"batched plan: {plan.info()[0]} tiles run", flush=True)
    rng = np.random.default_rng(9)
    for k, job in enumerate(jobs):
        if isinstance(job, lof.RoundJob):
            x = job.input.cpu().numpy()
            bits = None if job.rand_bits is None else job.rand_bits.cpu().numpy().view(np.uint32)
            ref = oracle.virtual_round(x, job.params.as_tuple(), int(job.mode), job.stoch_len,
                                       bits=bits)
            assert_bits(job.out.cpu().numpy(), ref, f"rounding job {k} (n={x.size})")
        else:
            A, B = job.A.cpu().numpy(), job.B.cpu().numpy()
            M, N = A.shape[0], B.shape[1]
            s = rng.choice(M * N, 512, replace=False)
            ref = oracle.gemm(A, B, oracle.accum_fmt(FP16, 32), samples=s)
            got = job.out.reshape(-1)[torch.from_numpy(s).to(dev)].cpu().numpy()
            assert_bits(got, ref, f"GEMM job {k} ({M}x{N}x{A.shape[1]})")
