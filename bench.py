"""Benchmark of the LoFloat hot path on one MI355X.

Primary line (``value``): MX block rounding of an 8192 x 8192 FP32 matrix, blocks of 32
along rows, e4m3 elements with an E8M0 shared scale, round-to-nearest-even, in GB/s of
algorithmic traffic (read 4 B + write 4 B per element).  One step = one rounding pass over
the matrix, inputs resident in HBM.

Also reported (``extra``): elementwise virtual_round of 2^24 / 2^28 values to (8,3,7); MX
with e2m1 and with SR-A random words; the emulated-accumulation GEMM at M=N=K=4096
(TFLOP/s = 2MNK / t) for 10-bit mantissa and 5e10m accumulators at chunks 1/8/32; one
LoF_Linear forward+backward step (X 8192x4096, W 4096x4096); and the grouped launch of 64
uneven rounding jobs + 12 GEMMs (e4m3 operands) against per-job launches, both captured in
HIP graphs, and the same on a launch-bound batch of 240 small rounding / MX jobs + 16 small
GEMMs; the bit-pack codec encode / decode of 2^28 values to e4m3fn and to an 11-bit
P3109 format; a 3x3 Conv2d layer as an implicit GEMM against unfold + GEMM.

The CPU baseline is the C oracle (a restatement of the reference's scalar semantics) run
with OpenMP on the host cores on the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
VALU_FP32_PEAK_TFLOPS = 157.3
METRIC = ("round GB/s (% HBM peak); emulated-accum GEMM TFLOP/s; Linear fwd+bwd step ms")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed(fn, iters, warmup=1):
    import torch
    for _ in range(warmup):
        fn(0)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for i in range(iters):
        fn(i)
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / iters


def traffic_from_profiles():
    """HBM bytes per launch of the MX kernel from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, calibrated against a copy kernel), if present."""
    p = os.path.join(REPO, "profiles", "mx_round_8192x8192_e4m3_pmc.md")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--quick", action="store_true", help="primary line only")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling)")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="also run N device copies of the 8192x8192 matrix (a plain copy "
                         "kernel of known traffic, to calibrate HBM byte counters)")
    ap.add_argument("--only", choices=["gemm", "linear", "grouped", "pack", "conv", "round"],
                    default=None,
                    help="run one workload alone (for profiling): 'gemm' = the 4096^3 "
                         "emulated-accumulation GEMM with --chunk / --accum / --inputs; "
                         "'linear' = the LoF_Linear step; 'grouped' = the batched launch; "
                         "'pack' = the bit-pack codec encode / decode of 2^28 elements; "
                         "'conv' = the implicit-GEMM Conv2d; 'round' = elementwise virtual_round "
                         "of 2^24 / 2^28 values (with --no-cpu: no oracle check)")
    ap.add_argument("--chunk", type=int, default=32)
    ap.add_argument("--accum", choices=["mant10", "mant23", "5e10m"], default="5e10m")
    ap.add_argument("--inputs", choices=["e4m3", "8e7m", "fp32u"], default="e4m3")
    ap.add_argument("--config", type=int, default=-1, help="GEMM tile config (-1: default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported (rounding "
              "shards by element ranges and the GEMM by output tiles, with no exchange)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import lofloat_amd as lof

    dev = torch.device("cuda:0")
    torch.manual_seed(0)
    e4m3 = lof.make_format(8, 3, 7)
    e2m1 = lof.make_format(4, 1, 1)
    e8m0 = lof.create_e8m0_params()
    if args.only == "gemm":
        gemm_only(lof, torch, dev, args)
        return
    if args.only == "round":
        out = {}
        round_c1(lof, torch, np, dev, out, check=not args.no_cpu)
        print(json.dumps(out), flush=True)
        return
    if args.only in ("linear", "grouped", "pack", "conv"):
        import numpy as np
        out = {}
        {"linear": linear_steps, "grouped": grouped_vs_separate, "pack": pack_codec,
         "conv": conv_layer}[args.only](lof, torch, np, dev, out)
        print(json.dumps(out), flush=True)
        return

    # ---------------------------------------------------------------- primary: MX 8192^2
    R = C = 8192
    n = R * C
    g = torch.Generator(device=dev).manual_seed(1)
    nbuf = 2
    xs = [torch.randn(R, C, device=dev, generator=g) *
          torch.exp2(torch.rand(R, 1, device=dev, generator=g) * 16 - 8) for _ in range(nbuf)]
    ys = [torch.empty_like(xs[0]) for _ in range(nbuf)]

    def mx_step(i):
        lof.virtual_mx_round(xs[i % nbuf], 32, e4m3, e8m0, out=ys[i % nbuf])

    ms = timed(mx_step, args.steps, args.warmup)
    bytes_per = 8.0 * n
    gbs = bytes_per / ms / 1e6
    log(f"mx e4m3 8192^2: {ms * 1e3:.1f} us/step, {gbs:.0f} GB/s")

    for _ in range(args.calib_copy):
        ys[1].copy_(xs[1])
    torch.cuda.synchronize()

    # CPU baseline: the oracle on the host cores, same workload (full matrix)
    cpu_gbs, cores, reps, same = None, 0, 0, None
    if not args.no_cpu:
        import oracle
        xh = xs[0].cpu().numpy().ravel()
        cores = oracle.max_threads()
        best = None
        t_end = time.time() + 10.0
        while reps < 3 or (time.time() < t_end and reps < 10):
            t0 = time.perf_counter()
            ref = oracle.virtual_mx_round(xh, 32, e4m3.as_tuple(), e8m0.as_tuple(),
                                          nthreads=cores)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
            reps += 1
        cpu_gbs = round(bytes_per / best / 1e9, 2)
        same = bool(np.array_equal(ys[0].cpu().numpy().ravel().view(np.uint32),
                                   ref.view(np.uint32)))
        log(f"cpu oracle: {best * 1e3:.1f} ms ({cpu_gbs} GB/s, {cores} threads), parity {same}")

    result = {
        "metric": METRIC,
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: seeded N(0,1) * 2^U(-8,8) per row",
        "config": {"workload": "MX block round 8192x8192 FP32, bs 32 along rows, e4m3 "
                               "(8,3,7) elements, E8M0 scale, RNE", "rows": R, "cols": C,
                   "block_size": 32, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(),
                     "kernel": "lofamd::k_mx<0, 0, 8, true>", "algorithmic_bytes": int(bytes_per)},
        "cpu_baseline": {"value": cpu_gbs, "unit": "GB/s", "cores": cores,
                         "kind": "port",
                         "sample": f"full 8192x8192 matrix, best of {reps} runs of the C oracle "
                                   "(OpenMP)", "parity_bitexact": same},
    }
    del xs, ys
    torch.cuda.empty_cache()
    if not args.quick:
        import oracle
        result["extra"] = extras(lof, torch, np, oracle, dev)
    print(json.dumps(result), flush=True)


def c1_device_values(torch, dev, n, seed, specials=0.01):
    """N(0,1) * 2^U(-12,12) on the device, with about 1% of the entries replaced by +0, -0,
    +inf, -inf, NaN or a random FP32 subnormal (equal shares)."""
    g = torch.Generator(device=dev).manual_seed(seed)
    x = torch.randn(n, device=dev, generator=g) * torch.exp2(
        torch.rand(n, device=dev, generator=g) * 24 - 12)
    k = int(n * specials)
    idx = torch.randint(0, n, (k,), device=dev, generator=g)
    kind = torch.randint(0, 6, (k,), device=dev, generator=g)
    sub = (torch.randint(1, 1 << 23, (k,), device=dev, generator=g, dtype=torch.int32) |
           (torch.randint(0, 2, (k,), device=dev, generator=g, dtype=torch.int32) << 31))
    table = torch.tensor([0.0, -0.0, float("inf"), float("-inf"), float("nan")], device=dev)
    vals = torch.where(kind < 5, table[kind.clamp(max=4)], sub.view(torch.float32))
    x[idx] = vals
    return x


def gemm_inputs(lof, torch, dev, inputs, n=4096):
    """n x n operands: "e4m3" / "8e7m" = N(0, 1/64) pre-rounded to (8,3,7) / (16,7,127);
    "fp32u" = FP32 uniform(-1, 1), the reference harness's lof_gemm inputs."""
    g = torch.Generator(device=dev).manual_seed(3)
    if inputs == "fp32u":
        return tuple(torch.rand(n, n, device=dev, generator=g) * 2 - 1 for _ in range(2))
    fmt = lof.make_format(8, 3, 7) if inputs == "e4m3" else lof.make_format(16, 7, 127)
    A = lof.virtual_round(torch.randn(n, n, device=dev, generator=g) / 64, fmt)
    B = lof.virtual_round(torch.randn(n, n, device=dev, generator=g) / 64, fmt)
    return A, B


def gemm_only(lof, torch, dev, args):
    """One GEMM workload alone: M = N = K = 4096, inputs N(0, 1/64) pre-rounded to e4m3 or
    8e7m, accumulator 10-bit mantissa or 5e10m (16,10,15), chunk --chunk."""
    n = 4096
    A, B = gemm_inputs(lof, torch, dev, args.inputs, n)
    spec = (lof.AccumSpec(fmt=lof.make_format(16, 10, 15), chunk=args.chunk)
            if args.accum == "5e10m" else
            lof.AccumSpec(mant_bits=int(args.accum[4:]), chunk=args.chunk))
    ms = timed(lambda i: lof.gemm(A, B, spec, config=args.config), args.steps, args.warmup)
    tf = 2.0 * n ** 3 / ms / 1e9
    print(json.dumps({"workload": f"gemm {n}^3 {args.inputs} inputs, {args.accum} accumulator, "
                                  f"chunk {args.chunk}", "config": args.config,
                      "ms": round(ms, 3), "TFLOPs": round(tf, 3)}), flush=True)


def extras(lof, torch, np, oracle, dev):
    out = {}
    e4m3 = lof.make_format(8, 3, 7)
    e2m1 = lof.make_format(4, 1, 1)
    e8m0 = lof.create_e8m0_params()
    fp16 = lof.make_format(16, 10, 15)

    round_c1(lof, torch, np, dev, out)
    # C2 variants
    n = 8192 * 8192
    x = torch.randn(8192, 8192, device=dev)
    y = torch.empty_like(x)
    bits = torch.randint(-2**31, 2**31 - 1, (8192, 8192), dtype=torch.int32, device=dev)
    ms = timed(lambda i: lof.virtual_mx_round(x, 32, e2m1, e8m0, out=y), 10)
    out["mx_e2m1_GBs"] = round(8 * n / ms / 1e6, 1)
    ms = timed(lambda i: lof.virtual_mx_round(x, 32, e4m3, e8m0, 3, 0, 8, rand_bits=bits, out=y), 10)
    out["mx_e4m3_srA8_bits_GBs"] = round(12 * n / ms / 1e6, 1)
    log(f"mx variants: {out['mx_e2m1_GBs']} / {out['mx_e4m3_srA8_bits_GBs']} GB/s")
    del x, y, bits

    # C3: emulated GEMM 4096^3 on e4m3 / 8e7m pre-rounded inputs, and the reference
    # harness's FP32 uniform(-1, 1) operands with its 10-bit accumulator at chunk 1 (and the
    # LoF layers' default 23-bit one).  The tile config is autotuned on each workload before
    # it is timed.  Parity: 4096 sampled outputs (full K) against the oracle, bit for bit;
    # the oracle's time on them gives the CPU rate in emulated multiply-adds per second.
    N = 4096
    flop = 2.0 * N ** 3
    samples = np.random.default_rng(11).choice(N * N, 4096, replace=False)
    sample_t = torch.from_numpy(samples).to(dev)
    mant = lambda m: (lambda c: lof.AccumSpec(mant_bits=m, chunk=c),   # noqa: E731
                      lambda c: oracle.accum_mant(m, c))
    accs = {"mant10": mant(10), "mant23": mant(23),
            "5e10m": (lambda c: lof.AccumSpec(fmt=fp16, chunk=c),
                      lambda c: oracle.accum_fmt(fp16.as_tuple(), c))}
    plan = [(inp, name, chunk) for inp in ("e4m3", "8e7m") for name in ("mant10", "5e10m")
            for chunk in (1, 8, 32)] + [("fp32u", "mant10", 1), ("fp32u", "mant23", 1)]
    cur = None
    for inp, name, chunk in plan:
        if cur != inp:
            A, B = gemm_inputs(lof, torch, dev, inp, N)
            Ah, Bh = A.cpu().numpy(), B.cpu().numpy()
            cur = inp
        acc, oacc = accs[name]
        spec = acc(chunk)
        cfg, _ = lof.gemm_autotune(A, B, spec, iters=1)
        it = 2 if chunk == 1 else 4
        ms = timed(lambda i: lof.gemm(A, B, spec), it)
        key = f"gemm4096_{inp}_{name}_c{chunk}"
        out[key + "_TFLOPs"] = round(flop / ms / 1e9, 3)
        out[key + "_config"] = lof.gemm_config_name(cfg)
        got = lof.gemm(A, B, spec).flatten()[sample_t].cpu().numpy()
        t0 = time.perf_counter()
        ref = oracle.gemm(Ah, Bh, oacc(chunk), samples=samples)
        cpu_s = time.perf_counter() - t0
        same = bool(np.array_equal(got.view(np.uint32), np.asarray(ref, np.float32).ravel().view(np.uint32)))
        out[key + "_parity_4096_samples"] = same
        out[key + "_cpu_oracle_GMACs"] = round(samples.size * N / cpu_s / 1e9, 4)
        log(f"{key}: {ms:.2f} ms, {out[key + '_TFLOPs']} TFLOP/s ({out[key + '_config']}), "
            f"sampled parity {same}, oracle {out[key + '_cpu_oracle_GMACs']} G emulated MAC/s")
    del A, B, Ah, Bh
    out["gemm_cpu_oracle_threads"] = oracle.max_threads()
    out["gemm_roofline_peak_TFLOPs"] = VALU_FP32_PEAK_TFLOPS

    linear_steps(lof, torch, np, dev, out)
    grouped_vs_separate(lof, torch, np, dev, out)
    pack_codec(lof, torch, np, dev, out)
    conv_layer(lof, torch, np, dev, out)
    return out


def round_c1(lof, torch, np, dev, out, check=True):
    """Elementwise virtual_round (C1) at 2^24 and 2^28 values; ``check``: also compare every
    output of one buffer with the oracle and time the oracle on it."""
    import oracle
    e4m3 = lof.make_format(8, 3, 7)
    # C1: elementwise virtual_round of N(0,1) * 2^U(-12,12) with 1% +-0 / +-inf / NaN /
    # FP32-subnormal specials to (8,3,7) RNE, rotating buffers past the 256 MB MALL; a plain
    # device copy over the same buffers is the read+write stream ceiling it is compared with;
    # every output of the first buffer is compared with the oracle bit for bit
    for lg in (24, 28):
        n = 1 << lg
        nb = max(2, (1 << 30) // (8 * n) + 1)
        xs = [c1_device_values(torch, dev, n, seed=100 + k) for k in range(nb)]
        ys = [torch.empty(n, device=dev) for _ in range(nb)]
        iters = 50 if lg == 24 else 10
        ms = timed(lambda i: lof.virtual_round(xs[i % nb], e4m3, out=ys[i % nb]), iters)
        ms_copy = timed(lambda i: ys[i % nb].copy_(xs[i % nb]), iters)
        gbs, copy_gbs = 8 * n / ms / 1e6, 8 * n / ms_copy / 1e6
        key = f"virtual_round_2^{lg}_e4m3"
        out[key + "_GBs"] = round(gbs, 1)
        out[key + "_frac_of_8TBs"] = round(gbs / HBM_PEAK_GBS, 4)
        out[key + "_copy_ceiling_GBs"] = round(copy_gbs, 1)
        out[key + "_frac_of_copy"] = round(gbs / copy_gbs, 4)
        if not check:
            log(f"virtual_round 2^{lg}: {ms * 1e3:.1f} us, {gbs:.0f} GB/s (copy {copy_gbs:.0f})")
            del xs, ys
            continue
        lof.virtual_round(xs[0], e4m3, out=ys[0])
        xh = xs[0].cpu().numpy()
        t0 = time.perf_counter()
        ref = oracle.virtual_round(xh, e4m3.as_tuple(), nthreads=oracle.max_threads())
        cpu_s = time.perf_counter() - t0
        got = ys[0].cpu().numpy()
        same = bool(np.array_equal(got.view(np.uint32), ref.view(np.uint32)))
        out[key + "_parity_full"] = same
        out[key + "_cpu_oracle_GBs"] = round(8 * n / cpu_s / 1e9, 2)
        log(f"virtual_round 2^{lg}: {ms * 1e3:.1f} us, {gbs:.0f} GB/s (copy {copy_gbs:.0f}), "
            f"parity {same}, oracle {out[key + '_cpu_oracle_GBs']} GB/s")
        del xs, ys, xh, ref, got
        torch.cuda.empty_cache()


def conv_layer(lof, torch, np, dev, out):
    """Conv2d forward, 3x3 / pad 1, 128 -> 128 channels on 32 x 28 x 28 (a ResNet-50 stage-2
    layer): M = 128, N = 25088, K = 1152, 7.4 GFLOP.  Inputs pre-rounded to e4m3.  Timed:
    the implicit GEMM (input gathered into LDS, no unfold in HBM) for a 5e10m accumulator
    at chunk 32 and for the reference lof_conv2d's 10-bit accumulator at chunk 1; and the
    reference layer's route for the same math, unfold in HBM then the GEMM."""
    e4m3 = lof.make_format(8, 3, 7)
    g = torch.Generator(device=dev).manual_seed(11)
    x = lof.virtual_round(torch.randn(32, 128, 28, 28, device=dev, generator=g) / 8, e4m3)
    w = lof.virtual_round(torch.randn(128, 128, 3, 3, device=dev, generator=g) / 8, e4m3)
    flop = 2.0 * 128 * (32 * 28 * 28) * 1152
    w_col = w.reshape(128, -1)
    for name, spec, it in (("5e10m_c32", lof.AccumSpec(fmt=lof.make_format(16, 10, 15), chunk=32), 10),
                           ("mant10_c1", lof.AccumSpec(mant_bits=10, chunk=1), 3)):
        ms = timed(lambda i: lof.conv2d(x, w, spec, padding=1), it)
        out[f"conv3x3_128_28x28_b32_{name}_TFLOPs"] = round(flop / ms / 1e9, 3)

        def unfold_gemm(i):
            cols = torch.nn.functional.unfold(x, 3, padding=1)            # 32 x 1152 x 784
            cols = cols.permute(1, 0, 2).reshape(1152, -1)                 # K x N, copied
            return lof.gemm(w_col, cols, spec)
        ms_u = timed(unfold_gemm, it)
        out[f"conv3x3_128_28x28_b32_{name}_unfold_gemm_TFLOPs"] = round(flop / ms_u / 1e9, 3)
        log(f"conv {name}: implicit {ms:.2f} ms ({flop / ms / 1e9:.1f} TFLOP/s), "
            f"unfold + gemm {ms_u:.2f} ms ({flop / ms_u / 1e9:.1f} TFLOP/s)")
    y = lof.conv2d(x, w, lof.AccumSpec(fmt=lof.make_format(16, 10, 15), chunk=32), padding=1)
    cols = torch.nn.functional.unfold(x, 3, padding=1).permute(1, 0, 2).reshape(1152, -1)
    y2 = lof.gemm(w_col, cols, lof.AccumSpec(fmt=lof.make_format(16, 10, 15), chunk=32))
    out["conv_implicit_equals_unfold_gemm"] = bool(torch.equal(
        y, y2.view(128, 32, 28, 28).permute(1, 0, 2, 3)))
    del x, w, y, y2, cols
    torch.cuda.empty_cache()


def pack_codec(lof, torch, np, dev, out):
    """Bit-pack codec: encode (round + pack) and decode (unpack + widen) of 2^28 FP32 values
    (1 GiB, four times the MALL) to e4m3fn (8 bits) and to P3109 binary11p5 (11 bits, groups
    of 16 in 22 bytes), RNE.  Algorithmic bytes per element: 4 + n/8 both ways."""
    n = 1 << 28
    g = torch.Generator(device=dev).manual_seed(7)
    x = torch.randn(n, device=dev, generator=g)
    y = torch.empty_like(x)
    fmts = (("e4m3fn", lof.create_ocp_e4m3fn_params()), ("p3109_11p5", lof.create_p3109_params(11, 5)))
    for name, fmt in fmts:
        nbits = fmt.total_bits
        buf = torch.empty(lof.pack.packed_size(nbits, n), dtype=torch.uint8, device=dev)
        algo = n * (4 + nbits / 8)
        ms = timed(lambda i: lof.pack.encode(x, fmt, out=buf), 10)
        out[f"pack_encode_2^28_{name}_GBs"] = round(algo / ms / 1e6, 1)
        ms_d = timed(lambda i: lof.pack.decode(buf, fmt, n, out=y), 10)
        out[f"pack_decode_2^28_{name}_GBs"] = round(algo / ms_d / 1e6, 1)
        log(f"pack {name}: encode {ms * 1e3:.0f} us ({algo / ms / 1e6:.0f} GB/s), "
            f"decode {ms_d * 1e3:.0f} us ({algo / ms_d / 1e6:.0f} GB/s)")
        del buf
    # raw reps (no rounding): the packing alone, 4 + n/8 bytes per element both ways
    reps = torch.randint(0, 1 << 11, (n,), dtype=torch.int32, device=dev)
    ms = timed(lambda i: lof.pack.pack_reps(reps, 11), 10)
    buf = lof.pack.pack_reps(reps, 11)
    ms_d = timed(lambda i: lof.pack.unpack_reps(buf, 11, n), 10)
    algo = n * (4 + 11 / 8)
    out["pack_reps_2^28_n11_GBs"] = round(algo / ms / 1e6, 1)
    out["unpack_reps_2^28_n11_GBs"] = round(algo / ms_d / 1e6, 1)
    log(f"pack_reps n=11: {ms * 1e3:.0f} us, unpack {ms_d * 1e3:.0f} us")
    del reps, buf
    # typed cast (one rep per byte): x -> e4m3fn reps and back, 4 + 1 B per element
    e4 = lof.create_ocp_e4m3fn_params()
    lut = lof.PWLSiLU(R=8.0, lut_bits=4).lut.to(dev)
    r8 = torch.empty(n, dtype=torch.uint8, device=dev)
    ms = timed(lambda i: lof.pack.to_reps(x, e4, out=r8), 10)
    ms_d = timed(lambda i: lof.pack.from_reps(r8, e4, out=y), 10)
    ms_s = timed(lambda i: lof.pwl_silu(x, lut, 8.0), 10)
    out["pwl_silu_2^28_f32_GBs"] = round(8 * n / ms_s / 1e6, 1)
    log(f"pwl_silu 2^28 f32 (lut 16): {ms_s * 1e3:.0f} us ({8 * n / ms_s / 1e6:.0f} GB/s)")
    out["cast_to_reps_2^28_e4m3fn_GBs"] = round(5 * n / ms / 1e6, 1)
    out["cast_from_reps_2^28_e4m3fn_GBs"] = round(5 * n / ms_d / 1e6, 1)
    log(f"typed cast e4m3fn: to {ms * 1e3:.0f} us, from {ms_d * 1e3:.0f} us")
    del x, y, r8, lut
    torch.cuda.empty_cache()


def linear_steps(lof, torch, np, dev, out):
    """LoF_Linear forward + backward: X 8192x4096, W 4096x4096, MX-e4m3 operands, e5m2 SR-A
    gradients, accumulators 5e10m chunk 32, 10-bit mantissa chunk 1 and the layer's default
    23-bit mantissa chunk 1."""
    e4m3 = lof.make_format(8, 3, 7)
    fp16 = lof.make_format(16, 10, 15)
    M, K, Nout = 8192, 4096, 4096
    for label, kw in (("5e10m_c32", dict(accum_format=fp16, accum_chunk=32)),
                      ("mant10_c1", dict(accum_mant_bits=10, accum_chunk=1)),
                      ("mant23_c1", dict(accum_mant_bits=23, accum_chunk=1))):
        lin = lof.LoF_Linear(K, Nout, act_params=e4m3, weight_params=e4m3, bias=False,
                             scaling="mx", mx_block_size=32,
                             grad_spec=lof.LinearGradSpec(lof.make_format(8, 2, 15), 3, 8,
                                                          seed=5), **kw).to(dev)
        with torch.no_grad():
            lin.weight.normal_(0, 0.02)
        X = (torch.randn(M, K, device=dev) * 0.02).requires_grad_(True)
        G = torch.randn(M, Nout, device=dev)
        # the library's GEMM autotuner, on operands of the three products' shapes, layouts
        # and formats (forward X_q W_q^T, dgrad G_q W_q, wgrad G_q^T X_q); LoF_Linear then
        # runs with the tuned tiles
        with torch.no_grad():
            xq, wq, _ = lin.quantized_operands(X.detach())
            gq = lof.virtual_round(G, lof.make_format(8, 2, 15), 3, 8, seed=9)
            for A_, B_ in ((xq, wq.t()), (gq, wq), (gq.t(), xq)):
                lof.gemm_autotune(A_, B_, lin.accum_spec(), iters=1)
            del xq, wq, gq

        def step(i):
            X.grad = None
            lin.weight.grad = None
            lin(X).backward(G)
        ms = timed(step, 2)
        out[f"linear_step_{label}_ms"] = round(ms, 2)
        out[f"linear_step_{label}_TFLOPs"] = round(3 * 2.0 * M * K * Nout / ms / 1e9, 3)
        log(f"linear step {label}: {ms:.1f} ms")
        del lin, X, G
    torch.cuda.empty_cache()


def c5_jobs(lof, torch, np, dev):
    """The batched workload: 64 rounding jobs of log-uniform size 2^12 .. 2^24 elements, each
    with its own format (exponent 2-8, mantissa 1-10 bits) and mode (RNE, round-to-odd,
    round-to-nearest-odd, toward zero, SR-A with explicit random words), and 12 emulated GEMMs
    with dimensions drawn from {1024, 2048, 4096, 8192}, operands N(0, 1/64) pre-rounded to
    e4m3, accumulator 5e10m at chunk 32.  Seeded; returns (jobs, total GEMM flop)."""
    fp16 = lof.make_format(16, 10, 15)
    e4m3 = lof.make_format(8, 3, 7)
    rng = np.random.default_rng(5)
    gen = torch.Generator(device=dev).manual_seed(55)
    jobs = []
    modes = [0, 12, 4, 1, 3]
    for i in range(64):
        nn = int(np.exp2(rng.uniform(12, 24)))
        e, m = int(rng.integers(2, 9)), int(rng.integers(1, 11))
        f = lof.make_format(1 + e + m, m, (1 << (e - 1)) - 1)
        mode = modes[i % 5]
        xi = torch.randn(nn, device=dev, generator=gen)
        bt = (torch.randint(-2**31, 2**31 - 1, (nn,), dtype=torch.int32, device=dev,
                            generator=gen) if mode == 3 else None)
        jobs.append(lof.RoundJob(xi, f, mode, min(8, 23 - m) if mode == 3 else 0, rand_bits=bt))
    dims = [1024, 2048, 4096, 8192]
    gemm_flop = 0.0
    for j in range(12):
        Mg, Ng, Kg = (int(rng.choice(dims)) for _ in range(3))
        A = lof.virtual_round(torch.randn(Mg, Kg, device=dev, generator=gen) / 64, e4m3)
        B = lof.virtual_round(torch.randn(Kg, Ng, device=dev, generator=gen) / 64, e4m3)
        jobs.append(lof.GemmJob(A, B, lof.AccumSpec(fmt=fp16, chunk=32)))
        gemm_flop += 2.0 * Mg * Ng * Kg
    return jobs, gemm_flop


def small_jobs(lof, torch, np, dev):
    """A launch-bound batch: 240 rounding / MX jobs of 2^10 .. 2^16 elements, each with its
    own format (exponent 2-8, mantissa 1-10 bits) and mode, and 16 emulated GEMMs of 64 .. 256
    per side (operands N(0, 1/64) pre-rounded to e4m3, 5e10m accumulator at chunk 32).
    Seeded; returns the jobs."""
    e4m3 = lof.make_format(8, 3, 7)
    fp16 = lof.make_format(16, 10, 15)
    rng = np.random.default_rng(9)
    gen = torch.Generator(device=dev).manual_seed(99)
    jobs = []
    modes = [0, 12, 4, 1, 3]
    for i in range(240):
        nn = int(np.exp2(rng.uniform(10, 16)))
        e, m = int(rng.integers(2, 9)), int(rng.integers(1, 11))
        mode = modes[i % 5]
        xi = torch.randn(nn, device=dev, generator=gen)
        bt = (torch.randint(-2**31, 2**31 - 1, (nn,), dtype=torch.int32, device=dev,
                            generator=gen) if mode == 3 else None)
        L = min(8, 23 - m) if mode == 3 else 0
        if i % 4 == 3 and nn >= 64:
            jobs.append(lof.MXJob(xi, 32, e4m3, lof.create_e8m0_params(), mode, 0, L,
                                  rand_bits=bt))
        else:
            jobs.append(lof.RoundJob(xi, lof.make_format(1 + e + m, m, (1 << (e - 1)) - 1),
                                     mode, L, rand_bits=bt))
    for j in range(16):
        Mg, Ng, Kg = (int(rng.choice([64, 128, 192, 256])) for _ in range(3))
        A = lof.virtual_round(torch.randn(Mg, Kg, device=dev, generator=gen) / 64, e4m3)
        B = lof.virtual_round(torch.randn(Kg, Ng, device=dev, generator=gen) / 64, e4m3)
        jobs.append(lof.GemmJob(A, B, lof.AccumSpec(fmt=fp16, chunk=32)))
    return jobs


def grouped_vs_separate(lof, torch, np, dev, out):
    """The batched workload (c5_jobs): one persistent launch vs per-job launches, both
    captured in HIP graphs; then the same comparison on a launch-bound batch of many small
    jobs (small_jobs)."""
    jobs, gemm_flop = c5_jobs(lof, torch, np, dev)
    cfg, tune_ms = lof.grouped.autotune(jobs, iters=1)
    plan = lof.GroupedPlan(jobs)
    gg = plan.capture_graph(separate=False)
    gs = plan.capture_graph(separate=True)
    ms_g = timed(lambda i: gg.replay(), 3)
    outs_g = [j.out.clone() for j in jobs]
    ms_s = timed(lambda i: gs.replay(), 3)
    same = all(torch.equal(a, j.out) for a, j in zip(outs_g, jobs))
    out.update({"grouped_ms": round(ms_g, 2), "separate_ms": round(ms_s, 2),
                "grouped_speedup": round(ms_s / ms_g, 3), "grouped_config": cfg,
                "grouped_bitexact_vs_separate": bool(same),
                "grouped_round_elems": int(sum(j.input.numel() for j in jobs[:64])),
                "grouped_gemm_TFLOP": round(gemm_flop / 1e12, 3)})
    log(f"grouped {ms_g:.1f} ms vs separate {ms_s:.1f} ms, bit-exact {same}")
    jobs = small_jobs(lof, torch, np, dev)
    plan = lof.GroupedPlan(jobs)
    gg = plan.capture_graph(separate=False)
    gs = plan.capture_graph(separate=True)
    ms_g = timed(lambda i: gg.replay(), 20)
    outs_g = [j.out.clone() for j in jobs]
    ms_s = timed(lambda i: gs.replay(), 20)
    same = all(torch.equal(a, j.out) for a, j in zip(outs_g, jobs))
    out.update({"grouped_small_jobs_ms": round(ms_g, 3), "separate_small_jobs_ms": round(ms_s, 3),
                "grouped_small_jobs_speedup": round(ms_s / ms_g, 2),
                "grouped_small_jobs_bitexact_vs_separate": bool(same),
                "grouped_small_jobs_count": len(jobs)})
    log(f"grouped small jobs ({len(jobs)}) {ms_g:.3f} ms vs separate {ms_s:.3f} ms, "
        f"bit-exact {same}")


if __name__ == "__main__":
    main()
