/*
 * lof_oracle.c -- CPU oracle (restatement of the reference's scalar semantics).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP product path and the
 * CPU baseline that bench.py times beside it.  The product never links it.
 *
 * Reference files followed (SudhanvaKulkarni123/LoFloat):
 *   src/lo_float.h      571-1272  bit rounders + RoundMantissa dispatch
 *   src/lo_float.h      2716-2756 virtual_round(x, ToMantissaBits, ps)
 *   src/lo_float.h      2844-2939 virtual_round(x, FloatingPointParams, ps)
 *   src/Lof_kernel.cu   37-56     elementwise kernels (fused scale)
 *   src/LoPy_bind.cpp   198-214   MX block rounding (CPU binding path)
 *   src/cutlass_gemms.cuh 74-80, 139-165, 1777-1787  RTO FMA, GEMM leaf, alpha
 *   src/blis_lof_gemm.hpp 63-99   host RTO FMA (exact-residual form)
 *
 * Pinning of each function (what checks it independently of this file) is listed
 * in tests/test_oracle_pinning.py; nothing here is "parity unpinned".
 *
 * Build: make -C oracle   (gcc -O2 -fopenmp -frounding-math -ffp-contract=off)
 */
#include "lof_oracle.h"

#include <fenv.h>
#include <math.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t low_mask(int r) { return r >= 32 ? 0xFFFFFFFFu : ((1u << r) - 1u); }

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static int pick_threads(int nthreads) { return nthreads > 0 ? nthreads : orc_max_threads(); }

/* SplitMix64 output function (Steele, Lea, Flood 2014) over a Weyl sequence:
 * the library-wide definition of "random word i of stream seed". */
uint32_t orc_rand_u32(uint64_t seed, uint64_t index) {
    uint64_t z = seed + (index + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z >> 32);
}

/*
 * RoundMantissa (lo_float.h:1232-1272) with the scalar rounders it dispatches to.
 * bits: magnitude bits; r: number of low bits to drop (the caller masks them);
 * pos: sign of the value is + (only RoundUp / RoundDown read it).
 *
 * Random-word mapping (replaces the reference's per-call std::mt19937 draws):
 *   SR-A/B/C : R = u & (2^len - 1)
 *   SR-D     : R = u & (2^len - 1), coin = (u >> len) & 1
 *   TrueSR   : sample = u / 2^32  (uniform in [0,1) on a 2^-32 grid)
 *   Prob     : coin = u & 1
 * r == 0: every rounder returns bits unchanged (the reference's RoundTiesToAway reads
 * bit r-1 there; on magnitude bits that bit is the cleared sign under either the x86
 * or the PTX shift-count rule, so it adds nothing).
 */
uint32_t orc_round_mantissa_bits(uint32_t bits, int r, int mode, int len, int pos, uint32_t u) {
    const uint32_t low = low_mask(r);
    const uint32_t T = bits & ~low;
    const uint32_t tail = bits & low;
    switch (mode) {
    case ORC_RNE:   /* lo_float.h:571-597 */
        return r == 0 ? bits : bits + ((bits >> r) & 1u) + (1u << (r - 1)) - 1u;
    case ORC_RTZ:   /* 947-965 */
        return T;
    case ORC_RAFZ:  /* 1001-1028 */
        return T + (tail != 0 ? (1u << r) : 0u);
    case ORC_SRA: { /* 654-696 */
        if (len <= 0) return bits;
        const uint32_t R = u & low_mask(len);
        return bits + (R << (r - len));
    }
    case ORC_RNO:   /* 1031-1057 */
        return r == 0 ? bits : bits + ((~bits >> r) & 1u) + (1u << (r - 1)) - 1u;
    case ORC_RDOWN: /* 1122-1157 */
        return T + ((!pos && tail != 0) ? (1u << r) : 0u);
    case ORC_RUP:   /* 1062-1100 */
        return T + ((pos && tail != 0) ? (1u << r) : 0u);
    case ORC_RTA:   /* 1177-1201 */
        return r == 0 ? bits : T + (((bits >> (r - 1)) & 1u) << r);
    case ORC_SRB: { /* 700-756, scalar branch 740-755 (no len<=0 early return) */
        const uint32_t R = u & low_mask(len);
        const uint32_t to_add = R + 1u;
        return bits + (tail > 0 ? (to_add << (r - len)) : 0u);
    }
    case ORC_SRC: { /* 759-815, scalar branch 800-814 */
        const uint32_t R = u & low_mask(len);
        const uint32_t coin = R & 1u, rem = R >> 1;
        if (len == 0 || tail == 0) return bits;
        return bits + ((rem + coin) << (r - len + 1));
    }
    case ORC_TRUESR: { /* 866-944, scalar branch 917-943 */
        const double prob = (double)tail / (double)((uint64_t)1 << r);
        const double samp = (double)u * (1.0 / 4294967296.0);
        if (samp < prob) {
            const uint32_t rounded = T + (1u << r);
            return rounded < T ? T : rounded;
        }
        return T;
    }
    case ORC_PROB:  /* 601-651, scalar branch 635-650 */
        return T + (tail != 0 ? ((u & 1u) << r) : 0u);
    case ORC_RTO:   /* 1204-1230 */
        return T | (tail != 0 ? (1u << r) : 0u);
    case ORC_RTTZ:  /* 975-998 */
        return r == 0 ? bits : bits + (1u << (r - 1)) - 1u;
    case ORC_SRD: { /* 821-862 */
        if (len <= 0) return bits;
        const uint32_t R = u & low_mask(len);
        const uint32_t coin = (u >> len) & 1u;
        return bits + ((R + coin) << (r - len));
    }
    default:        /* 1270-1271: no rounding */
        return bits;
    }
}

/* virtual_round(x, FloatingPointParams, ProjSpec), lo_float.h:2844-2939. */
float orc_virtual_round_fmt(float x, const orc_format* F, orc_spec ps, uint32_t u) {
    const int mb = F->mantissa_bits;
    const int to_max_exp = F->signedness == 0                                  /* 2857-2859 */
        ? (1 << (F->bitwidth - mb - 1)) - 1 - F->bias
        : (1 << (F->bitwidth - mb)) - 1 - F->bias;
    const int to_min_exp = 1 - F->bias;                                        /* 2860 */
    const float to_min_val = (float)(pow(2.0, to_min_exp) * pow(2.0, -mb));   /* 2861 */
    const uint32_t sign = f2u(x) & 0x80000000u;                                /* 2855 */
    uint32_t from_bits = f2u(fabsf(x));                                        /* 2864 */
    int from_exp = (int)(from_bits >> 23) - 127;                               /* 2865 */

    if (isnan(x)) return u2f(0x7FC00000u);                                     /* 2869 */

    const float abs_val = fabsf(x);                                            /* 2871 */
    if (abs_val < to_min_val) {                                                /* 2872-2908 */
        const int pos = sign == 0;
        int to_min;
        switch (ps.mode) {
        case ORC_RTZ: to_min = 0; break;
        case ORC_RTO:
        case ORC_RAFZ: to_min = abs_val > 0.0f; break;
        case ORC_RUP: to_min = pos && abs_val > 0.0f; break;
        case ORC_RDOWN: to_min = !pos && abs_val > 0.0f; break;
        case ORC_RNE: case ORC_RNO: case ORC_RTTZ: case ORC_RTA: {
            const float half = to_min_val * 0.5f;
            if (abs_val < half) to_min = 0;
            else if (abs_val > half) to_min = 1;
            else to_min = (ps.mode == ORC_RNO || ps.mode == ORC_RTA);
            break;
        }
        default: to_min = 0; break;
        }
        const uint32_t mag = to_min ? f2u((float)(double)to_min_val) : 0u;
        return u2f(mag | sign);
    }
    if (from_exp < to_min_exp) {                                               /* 2915-2923 */
        const int shift = (23 - mb) - (from_exp - to_min_exp);
        uint32_t sig = (from_bits & 0x7FFFFFu) | 0x800000u;
        sig = orc_round_mantissa_bits(sig, shift, ps.mode, ps.stoch_len, sign == 0, u);
        const uint32_t k = sig >> shift;
        const float v = (float)((double)k * (double)to_min_val);
        return u2f(f2u(v) | sign);
    }
    if (mb < 23) {                                                             /* 2925-2933 */
        from_bits = orc_round_mantissa_bits(from_bits, 23 - mb, ps.mode, ps.stoch_len,
                                            sign == 0, u);
        from_bits &= ~low_mask(23 - mb);
    }
    from_exp = (int)(from_bits >> 23) - 127;                                   /* 2935 */
    if (from_exp > to_max_exp) return F->inf_behavior == 1 ? x : u2f(0x7F800000u);  /* 2936 */
    return u2f(from_bits | sign);                                              /* 2938 */
}

/* virtual_round(x, ToMantissaBits, ProjSpec), lo_float.h:2716-2756. */
float orc_virtual_round_mant(float x, int m, orc_spec ps, uint32_t u) {
    const uint32_t xb = f2u(x);
    const uint32_t sign = xb & 0x80000000u;
    uint32_t from_bits = xb & 0x7FFFFFFFu;
    const int r = 23 - m;
    from_bits = orc_round_mantissa_bits(from_bits, r, ps.mode, ps.stoch_len, sign == 0, u);
    from_bits &= ~low_mask(r);
    return u2f(from_bits | sign);
}

static inline uint32_t word_at(const uint32_t* bits, uint64_t seed, int64_t i) {
    return bits ? bits[i] : orc_rand_u32(seed, (uint64_t)i);
}

/* Lof_kernel.cu:49-56 (round_fp_params_kernel): v = (scale == 1) ? x : scale * x. */
void orc_virtual_round_fmt_array(const float* in, float* out, int64_t n, const orc_format* f,
                                 orc_spec s, float scale, const uint32_t* bits, uint64_t seed,
                                 int nthreads) {
    #pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t i = 0; i < n; ++i) {
        const float v = (scale == 1.0f) ? in[i] : scale * in[i];
        out[i] = orc_virtual_round_fmt(v, f, s, word_at(bits, seed, i));
    }
}

/* Lof_kernel.cu:37-44 (round_mantissa_kernel). */
void orc_virtual_round_mant_array(const float* in, float* out, int64_t n, int mant_bits,
                                  orc_spec s, float scale, const uint32_t* bits, uint64_t seed,
                                  int nthreads) {
    #pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t i = 0; i < n; ++i) {
        const float v = (scale == 1.0f) ? in[i] : scale * in[i];
        out[i] = orc_virtual_round_mant(v, mant_bits, s, word_at(bits, seed, i));
    }
}

/* MX block rounding, CPU binding path LoPy_bind.cpp:198-214:
 *   pmn   = ldexpf(1, me) * (2 - ldexpf(1, -mb))            (float)
 *   amax  = std::max over |x_i|   (a < b ? b : a: NaN never replaces amax)
 *   scale = virtual_round(amax / pmn, scale_fmt, ProjSpec{scale_mode})
 *   x_i   = scale > 0 ? virtual_round(x_i / scale, elem, elem_spec) * scale : 0 */
void orc_virtual_mx_round(const float* in, float* out, int64_t n, int bs,
                          const orc_format* elem, const orc_format* sfmt,
                          orc_spec elem_spec, int scale_mode, const uint32_t* bits,
                          uint64_t seed, int nthreads) {
    const int mb = elem->mantissa_bits;
    const int me = elem->signedness == 0 ? (1 << (elem->bitwidth - mb - 1)) - 1 - elem->bias
                                         : (1 << (elem->bitwidth - mb)) - 1 - elem->bias;
    const float pmn = ldexpf(1.0f, me) * (2.0f - ldexpf(1.0f, -mb));
    const orc_spec sspec = {scale_mode, 0, 0};
    const int64_t nblocks = (n + bs - 1) / bs;
    #pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t b = 0; b < nblocks; ++b) {
        const int64_t b0 = b * bs;
        const int len = (int)((n - b0) < bs ? (n - b0) : bs);
        float amax = 0.0f;
        for (int i = 0; i < len; ++i) {
            const float a = fabsf(in[b0 + i]);
            amax = (amax < a) ? a : amax;
        }
        const float scale = orc_virtual_round_fmt(amax / pmn, sfmt, sspec, 0u);
        if (scale > 0.0f) {
            for (int i = 0; i < len; ++i) {
                const float q = orc_virtual_round_fmt(in[b0 + i] / scale, elem, elem_spec,
                                                      word_at(bits, seed, b0 + i));
                out[b0 + i] = q * scale;
            }
        } else {
            for (int i = 0; i < len; ++i) out[b0 + i] = 0.0f;
        }
    }
}

/* Literal restatement of round_to_odd_fma<float> (cutlass_gemms.cuh:74-80): the two
 * directed FMAs come from the host FPU rounding mode.  Volatile temporaries pin the
 * order of the mode switches and the arithmetic. */
static float fma_in_mode(float a, float b, float c, int mode) {
    volatile float va = a, vb = b, vc = c;
    fesetround(mode);
    volatile float r = fmaf(va, vb, vc);
    fesetround(FE_TONEAREST);
    return r;
}

float orc_rto_fma_dir(float a, float b, float c) {
    const float rd = fma_in_mode(a, b, c, FE_DOWNWARD);
    const float ru = fma_in_mode(a, b, c, FE_UPWARD);
    const uint32_t inexact = (ru != rd) ? 1u : 0u;
    const float rz = (rd >= 0.0f) ? rd : ru;
    return u2f(f2u(rz) | inexact);
}

/* Same definition (rz = rd >= 0 ? rd : ru, OR the inexact flag) with rd / ru found from
 * the exact value a*b + c = s + e (product exact in double, TwoSum as in
 * blis_lof_gemm.hpp:68-76) instead of switching the FPU mode. */
float orc_rto_fma(float a, float b, float c) {
    if (isnan(a) || isnan(b) || isnan(c)) return u2f(0x7FC00001u);
    const double p = (double)a * (double)b;
    const double cd = (double)c;
    const double s = p + cd;
    if (isnan(s)) return u2f(0x7FC00001u);           /* inf - inf: rd, ru are NaN */
    if (isinf(s)) return (float)s;                   /* exact infinity: rd == ru */
    if (s == 0.0) {
        /* exact zero: roundTowardNegative gives +0 only for (+0) + (+0) */
        return (!signbit(p) && !signbit(cd)) ? 0.0f : -0.0f;
    }
    const double bb = s - p;
    const double e = (p - (s - bb)) + (cd - bb);
    const float f = (float)s;                        /* a float next to x */
    float rd, ru;
    if (isinf(f)) {
        /* |x| beyond the float range: the neighbours are +-FLT_MAX and +-inf */
        if (s > 0) { rd = 3.40282347e+38f; ru = INFINITY; }
        else { rd = -INFINITY; ru = -3.40282347e+38f; }
    } else {
        const double d = (s - (double)f) + e;        /* sign(x - f), exact sign */
        if (d == 0.0) return f;
        if (d > 0.0) { rd = f; ru = nextafterf(f, INFINITY); }
        else { ru = f; rd = nextafterf(f, -INFINITY); }
    }
    const float rz = (rd >= 0.0f) ? rd : ru;
    return u2f(f2u(rz) | 1u);
}

static inline float acc_round(float r, const orc_accum* ac) {
    return ac->kind == 0 ? orc_virtual_round_mant(r, ac->mant_bits, ac->spec, 0u)
                         : orc_virtual_round_fmt(r, &ac->fmt, ac->spec, 0u);
}

/* One output of the emulated GEMM.
 * chunk == 1 (reference leaf, cutlass_gemms.cuh:139-165, ascending k):
 *     acc = +0;  acc = round(RTO_fma(a_k, b_k, acc)) for k = 0..K-1
 * chunk c > 1 (k-chunk rule): for each chunk [k0, k0+c) (the last may be short)
 *     p = +0;  p = fmaf(a_k, b_k, p) in ascending k (FP32, round-to-nearest)
 *     acc = round(RTO_fma(p, 1, acc))
 * Epilogue (cutlass LinearCombination, cutlass_gemms.cuh:1686-1690): D = alpha * acc, or
 * with a source term D = alpha * acc + beta * C contracted to fmaf(alpha, acc, beta * C). */
static float gemm_one(int i, int j, int K, const float* A, int64_t lda, int ta,
                      const float* B, int64_t ldb, int tb, float alpha, const orc_accum* ac,
                      const float* C, int64_t ldc, float beta) {
    float acc = 0.0f;
    const int c = ac->chunk > 0 ? ac->chunk : 1;
    for (int k0 = 0; k0 < K; k0 += c) {
        const int k1 = (k0 + c < K) ? k0 + c : K;
        if (c == 1) {
            const float a = ta ? A[(int64_t)k0 * lda + i] : A[(int64_t)i * lda + k0];
            const float b = tb ? B[(int64_t)j * ldb + k0] : B[(int64_t)k0 * ldb + j];
            acc = acc_round(orc_rto_fma(a, b, acc), ac);
        } else {
            float p = 0.0f;
            for (int k = k0; k < k1; ++k) {
                const float a = ta ? A[(int64_t)k * lda + i] : A[(int64_t)i * lda + k];
                const float b = tb ? B[(int64_t)j * ldb + k] : B[(int64_t)k * ldb + j];
                p = fmaf(a, b, p);
            }
            acc = acc_round(orc_rto_fma(p, 1.0f, acc), ac);
        }
    }
    if (beta != 0.0f) return fmaf(alpha, acc, beta * C[(int64_t)i * ldc + j]);
    return alpha * acc;
}

void orc_gemm_f32(int M, int N, int K, const float* A, int64_t lda, int trans_a,
                  const float* B, int64_t ldb, int trans_b, float* D, int64_t ldd,
                  float alpha, const orc_accum* ac, const int64_t* samples,
                  int64_t n_samples, int nthreads, const float* C, int64_t ldc, float beta) {
    if (samples) {
        #pragma omp parallel for schedule(dynamic, 4) num_threads(pick_threads(nthreads))
        for (int64_t s = 0; s < n_samples; ++s) {
            const int i = (int)(samples[s] / N), j = (int)(samples[s] % N);
            D[s] = gemm_one(i, j, K, A, lda, trans_a, B, ldb, trans_b, alpha, ac, C, ldc, beta);
        }
        return;
    }
    #pragma omp parallel for collapse(2) schedule(dynamic, 16) num_threads(pick_threads(nthreads))
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j)
            D[(int64_t)i * ldd + j] = gemm_one(i, j, K, A, lda, trans_a, B, ldb, trans_b,
                                               alpha, ac, C, ldc, beta);
}

/* ============================================================== bit-pack codec */

/* RoundMantissa on a 64-bit word (lo_float.h:1232-1272 with Bits = uint64_t); random
 * word u mapped as in orc_round_mantissa_bits. */
static uint64_t round_bits64(uint64_t bits, int r, int mode, int len, int pos, uint32_t u) {
    const uint64_t low = r >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << r) - 1u);
    const uint64_t T = bits & ~low, tail = bits & low, one_r = (uint64_t)1 << r;
    const uint64_t R = len > 0 ? (uint64_t)(u & low_mask(len)) : 0u;
    switch (mode) {
    case ORC_RNE: return r == 0 ? bits : bits + ((bits >> r) & 1u) + ((uint64_t)1 << (r - 1)) - 1u;
    case ORC_RTZ: return T;
    case ORC_RAFZ: return T + (tail ? one_r : 0u);
    case ORC_SRA: return len <= 0 ? bits : bits + (R << (r - len));
    case ORC_RNO: return r == 0 ? bits : bits + ((~bits >> r) & 1u) + ((uint64_t)1 << (r - 1)) - 1u;
    case ORC_RDOWN: return T + ((!pos && tail) ? one_r : 0u);
    case ORC_RUP: return T + ((pos && tail) ? one_r : 0u);
    case ORC_RTA: return r == 0 ? bits : T + (((bits >> (r - 1)) & 1u) << r);
    case ORC_SRB: return bits + (tail > 0 ? ((R + 1u) << (r - len)) : 0u);
    case ORC_SRC:
        if (len == 0 || tail == 0) return bits;
        return bits + (((R >> 1) + (R & 1u)) << (r - len + 1));
    case ORC_TRUESR: {
        const double prob = (double)tail / ldexp(1.0, r);
        const double samp = (double)u * (1.0 / 4294967296.0);
        return samp < prob ? T + one_r : T;
    }
    case ORC_PROB: return T + (tail ? ((uint64_t)(u & 1u) << r) : 0u);
    case ORC_RTO: return T | (tail ? one_r : 0u);
    case ORC_RTTZ: return r == 0 ? bits : bits + ((uint64_t)1 << (r - 1)) - 1u;
    case ORC_SRD: {
        if (len <= 0) return bits;
        return bits + ((R + ((u >> len) & 1u)) << (r - len));
    }
    default: return bits;
    }
}

/* Stochastic rounding of bits < 2^24 at a shift r > 24 (lo_pack.hpp:199-220): the result is
 * 0 or 1.  The reference shifts a uint64_t by r - len there, undefined once r >= 64; here
 * the result is the exact integer (bits + increment) >> r, i.e. 1 iff
 * k + (bits >> (r - len)) >= 2^len for an increment k * 2^(r - len). */
static uint32_t stoch_wide(uint64_t bits, int r, int mode, int len, uint32_t u) {
    const uint64_t R = len > 0 ? (uint64_t)(u & low_mask(len)) : 0u;
    const int s = r - len;
    const uint64_t hi = s >= 64 ? 0u : (bits >> s);
    const uint64_t full = (uint64_t)1 << len;
    switch (mode) {
    case ORC_SRA: return (len > 0 && R + hi >= full) ? 1u : 0u;
    case ORC_SRB: return (bits > 0 && (R + 1u) + hi >= full) ? 1u : 0u;
    case ORC_SRC: {
        if (len == 0 || bits == 0) return 0u;
        const int s1 = r - len + 1;
        const uint64_t hi1 = s1 >= 64 ? 0u : (bits >> s1);
        return ((R >> 1) + (R & 1u) + hi1 >= ((uint64_t)1 << (len - 1))) ? 1u : 0u;
    }
    case ORC_SRD: return (len > 0 && R + ((u >> len) & 1u) + hi >= full) ? 1u : 0u;
    case ORC_TRUESR: {   /* up iff u * 2^r < bits * 2^32, bits < 2^24 */
        if (bits == 0) return 0u;
        if (r <= 32) return ((uint64_t)u < (bits << (32 - r))) ? 1u : 0u;
        if (r - 32 >= 24) return u == 0 ? 1u : 0u;
        return (((uint64_t)u << (r - 32)) < bits) ? 1u : 0u;
    }
    case ORC_PROB: return bits ? (u & 1u) : 0u;
    default: return 0u;
    }
}

static int chk_match(const orc_checker* c, uint32_t rep, int nan) {
    switch (c->kind) {
    case 1:
        for (int i = 0; i < c->ncodes && i < 3; ++i)
            if (rep == c->codes[i]) return 1;
        return 0;
    case 2: return (rep & c->a) == c->a && (nan ? (rep & c->b) != 0 : (rep & c->b) == 0);
    case 3: return (rep | c->a) == c->codes[0];
    default: return 0;
    }
}

uint32_t orc_float_to_rep(float from, const orc_pack_format* P, orc_spec ps, uint32_t u) {
    const orc_format* F = &P->fmt;
    const int n = F->bitwidth, mb = F->mantissa_bits, bias = F->bias;
    const int is_signed = F->signedness == 0;
    const int exp_bits = is_signed ? n - mb - 1 : n - mb;
    const int exp_off = bias - 127;                       /* kExponentOffset */
    const int digit_shift = mb - 23;                      /* kDigitShift (< 0 here) */
    const int to_min_exp = 1 - bias;
    const int to_max_exp = (1 << exp_bits) - 1 - bias;
    const int has_inf = F->inf_behavior != 1;             /* not Saturating */
    const uint32_t signmask = is_signed ? (1u << (n - 1)) : 0u;
    const uint32_t max_finite = has_inf ? P->is_inf.pat0 - 1u
                              : (is_signed ? (((1u << n) - 1u) >> 1) : ((1u << n) - 1u));
    const uint32_t inf_rep = has_inf ? P->is_inf.pat0 : max_finite;
    const uint32_t xb = f2u(from);
    const int neg = (int)(xb >> 31);
    const uint32_t from_bits = xb & 0x7FFFFFFFu;
#define WITH_SIGN(mag) \
    ((!neg || !is_signed) ? (mag) : (((mag) == 0u && chk_match(&P->is_nan, signmask, 1)) ? 0u : ((mag) | signmask)))

    if (!is_signed && neg) {                                  /* lo_pack.hpp:161-166 */
        if (P->unsigned_behavior == 0) return 0u;
        return F->nan_behavior == 0 ? P->is_nan.pat0 : P->is_nan.pat1;
    }
    if (isinf(from)) {                                        /* 170-174 */
        if (ps.saturation != 1 && has_inf) return WITH_SIGN(inf_rep);
        return WITH_SIGN(max_finite);
    }
    if (isnan(from)) return P->is_nan.pat0;                   /* 175 */
    if (from_bits == 0) return WITH_SIGN(0u);                 /* 176 */
    const int bfe = (int)(from_bits >> 23);

    if (to_min_exp < -125 && bfe == 0) {                      /* 181-198: float subnormal in */
        uint64_t bits = from_bits;
        const int nf = __builtin_clz(from_bits) - (32 - 23) + 1;
        const int be = exp_off - nf + 1;
        if (be <= 0) {
            if (exp_off < 64) bits <<= exp_off;
        } else {
            bits <<= nf;
            bits &= ~((uint64_t)1 << 23);
            bits |= (uint64_t)be << 23;
        }
        bits = round_bits64(bits, -digit_shift, ps.mode, ps.stoch_len, 0, u);
        bits >>= -digit_shift;
        return WITH_SIGN((uint32_t)bits);
    }
    if (to_min_exp > -125) {                                  /* 201-239: to subnormal / zero */
        const int bte = bfe - 127 + bias;
        if (bte <= 0) {
            const uint32_t lead = bfe > 0 ? 1u : 0u;
            const int shift = -digit_shift - bte + (int)lead;
            const uint32_t rfb = (from_bits & 0x7FFFFFu) | (lead << 23);
            uint32_t bits;
            if (shift <= 23 + 1) {
                bits = (uint32_t)(round_bits64(rfb, shift, ps.mode, ps.stoch_len, 0, u) >> shift);
            } else {
                switch (ps.mode) {
                case ORC_RTO: case ORC_RAFZ: bits = rfb > 0 ? 1u : 0u; break;
                case ORC_RTZ: bits = 0u; break;
                case ORC_RUP: bits = (rfb && !neg) ? 1u : 0u; break;
                case ORC_RDOWN: bits = (rfb && neg) ? 1u : 0u; break;
                case ORC_SRA: case ORC_SRB: case ORC_SRC: case ORC_SRD: case ORC_TRUESR: case ORC_PROB:
                    bits = stoch_wide(rfb, shift, ps.mode, ps.stoch_len, u); break;
                default: bits = 0u; break;
                }
            }
            return WITH_SIGN(bits);
        }
    }
    /* normal path, 242-257 */
    uint64_t rb = from_bits;
    rb = round_bits64(rb, -digit_shift, ps.mode, ps.stoch_len, 0, u);
    rb &= ~((((uint64_t)1) << (-digit_shift)) - 1u);
    rb += (uint64_t)(int64_t)exp_off << 23;
    const uint64_t aligned_highest = (uint64_t)max_finite << (-digit_shift);
    const uint32_t bits = (uint32_t)(rb >> (-digit_shift));
    const int narrower = to_max_exp < 128 || (to_max_exp == 128 && mb + 1 < 24);
    if (narrower && rb > aligned_highest) {
        const int emit_inf = ps.saturation == 0 && has_inf;
        return WITH_SIGN(emit_inf ? inf_rep : max_finite);
    }
    return WITH_SIGN(bits);
#undef WITH_SIGN
}

float orc_rep_to_float(uint32_t rep, const orc_pack_format* P) {
    const orc_format* F = &P->fmt;
    const int n = F->bitwidth, mb = F->mantissa_bits, bias = F->bias;
    const int is_signed = F->signedness == 0;
    rep &= n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u);
    const int neg = is_signed ? (int)((rep >> (n - 1)) & 1u) : 0;
    const uint32_t mag = is_signed ? (rep & ((1u << (n - 1)) - 1u)) : rep;
    if (F->inf_behavior != 1 && chk_match(&P->is_inf, rep, 0)) return neg ? -INFINITY : INFINITY;
    if (chk_match(&P->is_nan, rep, 1)) return u2f(0x7FC00000u);
    const uint32_t e = mag >> mb, m = mag & ((1u << mb) - 1u);
    const double v = e == 0 ? ldexp((double)m, 1 - bias - mb)
                            : ldexp(1.0 + (double)m / (double)(1u << mb), (int)e - bias);
    return (float)(neg ? -v : v);
}

static int pk_gcd(int a, int b) { return b == 0 ? a : pk_gcd(b, a % b); }
static int pk_group_bits(int n) { const int al = n <= 8 ? 8 : 16; return n / pk_gcd(n, al) * al; }

int64_t orc_packed_bytes(int n, int64_t count) {
    const int gbits = pk_group_bits(n), gc = gbits / n, gb = gbits / 8;
    return (count + gc - 1) / gc * gb;
}

static inline int64_t slot_bit(int64_t i, int n, int gc, int gbits, int endian) {
    const int64_t g = i / gc;
    const int j = (int)(i % gc);
    const int slot = endian ? gc - 1 - j : j;
    return g * gbits + (int64_t)slot * n;
}

static void put_field(uint8_t* dst, int64_t off, uint32_t v, int n) {
    for (int b = 0; b < n; ++b) {
        const int64_t p = off + b;
        dst[p >> 3] |= (uint8_t)(((v >> b) & 1u) << (p & 7));
    }
}
static uint32_t get_field(const uint8_t* src, int64_t off, int n) {
    uint32_t v = 0;
    for (int b = 0; b < n; ++b) {
        const int64_t p = off + b;
        v |= (uint32_t)((src[p >> 3] >> (p & 7)) & 1u) << b;
    }
    return v;
}

int64_t orc_pack_reps(uint8_t* dst, const uint32_t* reps, int64_t count, int n, int endian) {
    const int gbits = pk_group_bits(n), gc = gbits / n;
    const int64_t bytes = orc_packed_bytes(n, count), padded = (count + gc - 1) / gc * gc;
    const uint32_t mask = (1u << n) - 1u;
    memset(dst, 0, (size_t)bytes);
    for (int64_t i = 0; i < padded; ++i)
        put_field(dst, slot_bit(i, n, gc, gbits, endian), i < count ? (reps[i] & mask) : 0u, n);
    return bytes;
}

int64_t orc_unpack_reps(uint32_t* reps, const uint8_t* src, int64_t count, int n, int endian) {
    const int gbits = pk_group_bits(n), gc = gbits / n;
    for (int64_t i = 0; i < count; ++i) reps[i] = get_field(src, slot_bit(i, n, gc, gbits, endian), n);
    return orc_packed_bytes(n, count);
}

int64_t orc_pack_encode(uint8_t* dst, const float* src, int64_t count,
                        const orc_pack_format* f, orc_spec ps, int endian,
                        const uint32_t* bits, uint64_t seed) {
    const int n = f->fmt.bitwidth;
    const int gbits = pk_group_bits(n), gc = gbits / n;
    const int64_t bytes = orc_packed_bytes(n, count), padded = (count + gc - 1) / gc * gc;
    memset(dst, 0, (size_t)bytes);
    for (int64_t i = 0; i < padded; ++i) {
        /* the final partial group is padded with 0.0f (lo_pack.hpp:427-431) */
        const float x = i < count ? src[i] : 0.0f;
        const uint32_t u = i < count ? word_at(bits, seed, i) : 0u;
        const uint32_t rep = orc_float_to_rep(x, f, ps, u) & ((1u << n) - 1u);
        put_field(dst, slot_bit(i, n, gc, gbits, endian), rep, n);
    }
    return bytes;
}

int64_t orc_pack_decode(float* dst, const uint8_t* src, int64_t count,
                        const orc_pack_format* f, int endian) {
    const int n = f->fmt.bitwidth;
    const int gbits = pk_group_bits(n), gc = gbits / n;
    for (int64_t i = 0; i < count; ++i)
        dst[i] = orc_rep_to_float(get_field(src, slot_bit(i, n, gc, gbits, endian), n), f);
    return orc_packed_bytes(n, count);
}

void orc_float_to_rep_array(const float* x, uint32_t* out, int64_t n, const orc_pack_format* f,
                            orc_spec ps, const uint32_t* bits, uint64_t seed) {
    for (int64_t i = 0; i < n; ++i) out[i] = orc_float_to_rep(x[i], f, ps, word_at(bits, seed, i));
}
void orc_rep_to_float_array(const uint32_t* rep, float* out, int64_t n, const orc_pack_format* f) {
    for (int64_t i = 0; i < n; ++i) out[i] = orc_rep_to_float(rep[i], f);
}

/* ---------------------------------------------------------------- piecewise-linear SiLU */
/* pwl_silu_kernel (Lof_kernel.cu:240-258) for one element, in float as the kernel computes:
 * clamp to [-R, R] (fmaxf first, so NaN becomes -R), scale by inv_step = N / (2R) computed
 * as in pwl_silu (Lof_kernel.cu:261), truncate to the segment index capped at N-1, fma
 * interpolation, relu outside [-R, R]. */
float orc_pwl_silu(float x, const float* lut, int N, float R) {
    const float inv_step = (float)N / (2.0f * R);
    const float xc = fminf(fmaxf(x, -R), R);
    const float u = (xc + R) * inv_step;
    int i = (int)u;
    if (i > N - 1) i = N - 1;
    const float frac = u - (float)i;
    const float y0 = lut[i], y1 = lut[i + 1];
    const float y = fmaf(frac, y1 - y0, y0);
    if (fabsf(x) > R) return x > 0.0f ? x : 0.0f;
    return y;
}

void orc_pwl_silu_array(const float* x, float* y, int64_t n, const float* lut, int N, float R) {
    for (int64_t k = 0; k < n; ++k) y[k] = orc_pwl_silu(x[k], lut, N, R);
}
