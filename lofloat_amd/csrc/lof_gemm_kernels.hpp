// lof_gemm_kernels.hpp -- emulated-accumulation GEMM tile for gfx950 (VALU).
//
// Per output, in ascending k (the reference leaf, cutlass_gemms.cuh:139-165):
//     acc <- round(RTO_fma(a, b, acc))                                  chunk == 1
//     p <- fmaf(a, b, p) over a chunk, then acc <- round(RTO_fma(p,1,acc))  chunk > 1
// round() is virtual_round(., mant_bits) or virtual_round(., fmt).
//
// How one leaf is computed exactly:
//   p = a*b is exact in double (24x24-bit significands); s = RN53(p + acc) and the TwoSum
//   residual e = (p + acc) - s are exact double operations, so x = s + e exactly.
//   Fast path (round-to-nearest-even, target mantissa m <= 19, |s| inside the format's
//   normal range): RN_m(RTO24(x)) == RN_m(x) because round-to-odd to 24 bits followed by
//   round-to-nearest to m+1 <= 22 bits is a correct rounding.  RN_m(x) is the round of
//   s's 52-bit significand at bit 52-m (which sits in the high word for m <= 19), with an
//   exact tie on s broken by the sign of e.  Below 2^elo the format's grid has the fixed
//   quantum 2^(elo - m) and the rounding position moves up accordingly.  ~20 VALU ops; the
//   leaves of a thread's outputs run branch-free and the rare general-path ones afterwards.
//   General path (any mode, any m, zeros, infinities, NaN, FP32 subnormal or out-of-range
//   accumulators): build RTO24(x) literally (neighbours rd/ru of x, rz = rd >= 0 ? rd : ru,
//   OR the inexact bit; an exact zero gets roundTowardNegative's sign) and apply the
//   reference's virtual_round to it.
// The accumulator lives in registers as a double holding a float value.
#pragma once

#include <hip/hip_runtime.h>

#include "lof_core.hpp"

namespace lofamd {

// Division of a 32-bit unsigned n by an invariant d >= 1 (Granlund & Montgomery, "Division
// by invariant integers using multiplication", 1994, Fig. 4.1): l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, t = mulhi(m, n), q = (t + ((n - t) >> s1)) >> s2 with
// s1 = min(l, 1), s2 = max(l - 1, 0) (so d = 1 gives m = 1, q = n).
struct FastDiv {
    uint32_t m;
    int32_t s1, s2;
};
inline FastDiv make_fast_div(uint32_t d) {
    int l = 0;
    while (l < 32 && ((uint64_t)1 << l) < d) ++l;
    FastDiv f;
    f.m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1u);
    f.s1 = l < 1 ? l : 1;
    f.s2 = l > 1 ? l - 1 : 0;
    return f;
}
__device__ __forceinline__ uint32_t fast_div(uint32_t n, const FastDiv& f) {
    const uint32_t t = __umulhi(n, f.m);
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

// Implicit-GEMM convolution (cross-correlation): B(k, n) is gathered from an activation
// tensor and D(m, n) / C(m, n) live at strided output positions.  k = (d1, d2, d3) with
// d1 = k / D1, d2 = (k % D1) / D2, d3 = k % D2: (c, r, s) for an OIHW filter, (r, s, c) for
// KRSC.  n = (b, p, q) = output image b, row p, column q.  conv = 0 for a plain GEMM.
struct ConvGeom {
    int32_t conv;
    int32_t krsc;                 // 0: k = (c, r, s) (OIHW), 1: k = (r, s, c) (KRSC)
    int32_t H, W, P, Q;
    int32_t ph, pw, sh, sw, dh, dw;
    int32_t D2;                   // S (OIHW) or C (KRSC)
    int32_t sxc, sxh, sxw;        // activation strides (elements); batch stride sxb
    int64_t sxb;
    int64_t sdb, sdm, sdpq;       // output strides: batch, output channel m, position p*Q+q
    FastDiv div_pq, div_q, div_d1, div_d2;
    uint32_t PQ, D1;
};

// Global-memory access for the GEMM tiles.  Their operands arrive as generic pointers
// (members of GemmArgs), which the compiler would read and write with FLAT instructions;
// those count against the LDS wait counter as well, so every wait for an LDS read would also
// wait for the next slice's prefetch.  Through the global address space they are
// global_load / global_store.
template <class T>
__device__ __forceinline__ T gld(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
    *(__attribute__((address_space(1))) T*)p = v;
}
// float4 (a class type, whose copy would go through a generic reference) as a native vector
__device__ __forceinline__ float4 gld4(const float* p) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 v = *(const __attribute__((address_space(1))) v4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

struct GemmArgs {
    int32_t M, N, K;
    const float* A;
    int64_t lda;
    int32_t ta;
    const float* B;
    int64_t ldb;
    int32_t tb;
    float* D;
    int64_t ldd;
    float alpha;
    const float* C;      // read when beta != 0
    int64_t ldc;
    float beta;
    int32_t acc_kind;    // 0 mantissa-only, 1 full format
    int32_t acc_r;       // mantissa-only: dropped bits 23 - mant_bits
    int32_t acc_mode;
    FmtDev accF;         // full format (mode included)
    int32_t chunk;
    int32_t fast;        // fast path enabled
    uint32_t half_m1;    // fast path: (1 << (rh - 1)) - 1, rh = 20 - m
    uint32_t keep_mask;  // fast path: ~((1 << rh) - 1)
    int32_t rh;
    uint32_t lo_bound;   // fast path domain on |s|'s high word: [lo_bound, lo_bound + range)
    uint32_t range;      // 0 when the fast path is off
    // the same fast path on a float partial sum s (chunk leaf of the MFMA tile)
    int32_t rf;          // 23 - m
    uint32_t half_m1f;   // (1 << (rf - 1)) - 1
    uint32_t keep_maskf; // ~((1 << rf) - 1)
    uint32_t lo_f;       // domain on |s|'s float bits: [lo_f, lo_f + range_f)
    uint32_t range_f;    // 0 when the fast path is off
    int32_t elo_b;       // smallest normal exponent of the format, FP32-biased
    int32_t elo_d;       // the same, FP64-biased
    int32_t subn;        // the float fast path covers part of the format's subnormal range
    // exact-sum leaves of the MFMA tile (used when range_f != 0)
    float sum_g;         // (1 + 2^-(m+1))^n rounded up, n = leaves per K slice (32 / chunk)
    float sum_n;         // n
    int32_t sum_ehi;     // |s| must stay below 2^sum_ehi (the format's top binade excluded)
    int32_t sum_q;       // every format value is a multiple of 2^sum_q
    int32_t sum_lreq;    // products must be multiples of 2^sum_lreq = max(sum_q, -126)
    int32_t sum_m;       // the format's mantissa bits
    int32_t sum_f16;     // the format's grid is binary16's below 2^15: round by conversion
    int32_t sum_rto;     // RTO leaves round this accumulator exactly (mantissa-only, binary16)
    int32_t sum_h10;     // mantissa-only with m = 10: binary16's significand, any exponent
    int32_t sum_off;     // testing: bit 0 exact-sum leaves off, bit 1 RTO leaves off,
                         // bit 2 binary16-held chunk-1 leaves off, bit 3 their packed form,
                         // bit 4 the directed-rounding leaves, bit 5 the mixed-precision
                         // FMA leaves of binary16-shaped accumulators
    // directed-rounding chunk-1 leaves (any product, exact or not): 0 off, 1 mantissa-only
    // accumulator (round-to-nearest-even at bit dr_rf of the pattern), 2 binary16-shaped
    int32_t dr_kind;
    int32_t dr_rf;       // 23 - mantissa bits
    uint32_t dr_half;    // (1 << (dr_rf - 1)) - 1, 0 when dr_rf == 0
    uint32_t dr_lsb;     // 1, 0 when dr_rf == 0 (rounding to 23 bits is the identity)
    uint32_t dr_mask;    // ~((1 << dr_rf) - 1)
    int32_t vec;         // A, B 16-byte aligned with lda, ldb multiples of 4 (float4 loads)
    int32_t mfma_mode;   // MFMA tile: 0 exactness test, 1 every chunk on the VALU chain,
                         // 2 every chunk on MFMA (testing only: skips the test)
    int32_t mfma_bf16;   // exact chunk sums of bfloat16-valued slices on the bf16 MFMAs
    ConvGeom cv;         // wave tiles only: B gathered from an activation (cv.conv != 0)
};

__device__ __forceinline__ uint32_t next_up_bits(uint32_t b) {
    if (b == 0x80000000u) return 1u;
    return (b >> 31) ? b - 1u : b + 1u;
}
__device__ __forceinline__ uint32_t next_down_bits(uint32_t b) {
    if (b == 0u) return 0x80000001u;
    return (b >> 31) ? b + 1u : b - 1u;
}

// RTO24(p + acc) as float bits, given s = RN53(p + acc) and e = (p + acc) - s.
// Mirrors round_to_odd_fma (cutlass_gemms.cuh:74-80): rd/ru are the float neighbours of
// the exact value (FLT_MAX / inf beyond the range), rz = rd >= 0 ? rd : ru, | inexact.
__device__ __noinline__ uint32_t rto24_bits(double p, double acc, double s, double e) {
    if (__builtin_isnan(s)) return 0x7FC00001u;
    if (__builtin_isinf(s)) return __float_as_uint((float)s);
    if (s == 0.0) {
        const bool pz = __double_as_longlong(p) == 0;
        const bool az = __double_as_longlong(acc) == 0;
        return (pz && az) ? 0u : 0x80000000u;   // (+0)+(+0) = +0, any other exact 0 is -0
    }
    const float f = (float)s;
    if (__builtin_isinf(f)) return s > 0.0 ? 0x7F7FFFFFu : 0xFF7FFFFFu;
    const double d = (s - (double)f) + e;      // exact sign of x - f
    const uint32_t fb = __float_as_uint(f);
    if (d == 0.0) return fb;
    uint32_t rd, ru;
    if (d > 0.0) { rd = fb; ru = next_up_bits(fb); }
    else { ru = fb; rd = next_down_bits(fb); }
    const uint32_t rz = (__uint_as_float(rd) >= 0.0f) ? rd : ru;
    return rz | 1u;
}

// What the out-of-line general leaf reads, for passing it by value.  The general leaf and
// leaf_slow take either this or a reference to the kernel's GemmArgs (P).  A reference gives
// GemmArgs a private-memory copy whose fields the kernel reads back after each call instead
// of holding them in registers across it: fewer registers (the VALU tiles, the 128-row MFMA
// tiles and the short-chunk tiles keep their occupancy that way), but scratch reads in the
// tile loop, whose waits (the vector-memory counter) also wait for the next slice's
// in-flight prefetch.
struct LeafGen {
    int32_t acc_kind, acc_mode, acc_r;
    uint32_t range;   // nonzero while the fast path is on (leaf_zero's condition)
    FmtDev accF;
};

template <class P>   // const GemmArgs& or LeafGen
__device__ __noinline__ double leaf_general(double p, double acc, double s, double e, P g) {
    const uint32_t F = rto24_bits(p, acc, s, e);
    const uint32_t R = g.acc_kind == 0 ? vr_mant_rt(F, g.acc_mode, g.acc_r, 0, 0u)
                                       : vr_fmt_rt(F, g.accF, 0u);
    return (double)__uint_as_float(R);
}

__device__ __forceinline__ LeafGen leaf_gen(const GemmArgs& g) {
    return LeafGen{g.acc_kind, g.acc_mode, g.acc_r, g.range, g.accF};
}

// One emulated accumulation step with exact product/partial sum p, fast path only: returns
// false (result undefined) where the general path is needed.  With SUBN the rounding
// position follows the format's fixed subnormal quantum 2^(elo - m) below 2^elo:
// P = 20 - m + max(0, elo - E) in s's high word, valid while P <= 19, i.e. for
// |s| >= 2^(elo - m + 1) (the domain's lower bound then).
template <bool SUBN>
__device__ __forceinline__ bool leaf_fast(double p, double acc, const GemmArgs& g, double& out) {
    const double s = p + acc;
    const double bb = s - p;
    const double e = (p - (s - bb)) + (acc - bb);
    const uint32_t hi = (uint32_t)__double2hiint(s);
    const uint32_t lo = (uint32_t)__double2loint(s);
    const uint32_t same_sign = (((uint32_t)__double2hiint(e) ^ hi) >> 31) ^ 1u;
    uint32_t hi2;
    if constexpr (SUBN) {
        const int eb = (int)((hi >> 20) & 0x7FFu);
        const uint32_t P = (uint32_t)(g.rh + max(g.elo_d - eb, 0));
        const uint32_t lsb = (hi >> P) & 1u;
        const uint32_t up = (e == 0.0) ? lsb : same_sign;
        const uint32_t c = (lo != 0u) ? 1u : up;
        const uint32_t unit = 1u << P;
        hi2 = (hi + (unit >> 1) - 1u + c) & (0u - unit);
    } else {
        const uint32_t lsb = (hi >> g.rh) & 1u;
        const uint32_t up = (e == 0.0) ? lsb : same_sign;
        const uint32_t c = (lo != 0u) ? 1u : up;
        hi2 = (hi + g.half_m1 + c) & g.keep_mask;
    }
    out = __hiloint2double((int)hi2, 0);
    return ((hi & 0x7FFFFFFFu) - g.lo_bound) < g.range;
}

// The leaf of an exactly-zero sum while the fast path is on (round to nearest even, which
// keeps a zero): RTO24 gives +0 for (+0) + (+0) and -0 for any other exact zero
// (roundTowardNegative).  Outside the fast domain, so checked in the slow branch, where it
// spares the general path the zero sums of padded tile rows / columns and zero data.
__device__ __forceinline__ bool leaf_zero(double p, double acc, const GemmArgs& g, double& out) {
    if (g.range == 0u || p + acc != 0.0) return false;
    out = (__double_as_longlong(p) | __double_as_longlong(acc)) == 0 ? 0.0 : -0.0;
    return true;
}

template <class P>   // const GemmArgs& or LeafGen
__device__ __noinline__ double leaf_slow(double p, double acc, P g) {
    const double s = p + acc;
    const double bb = s - p;
    const double e = (p - (s - bb)) + (acc - bb);
    return leaf_general<P>(p, acc, s, e, g);
}

// One emulated accumulation step (fast path, else the general path).  The fast path's
// domain reaches into the format's subnormal range when g.subn is set, so the rounding
// position must follow it there.  BYVAL: the general path's arguments by value (LeafGen).
template <bool BYVAL = false>
__device__ __forceinline__ double leaf(double p, double acc, const GemmArgs& g) {
    double r;
    const bool ok = g.subn ? leaf_fast<true>(p, acc, g, r) : leaf_fast<false>(p, acc, g, r);
    if (__builtin_expect(!ok, 0) && !leaf_zero(p, acc, g, r)) {
        if constexpr (BYVAL) r = leaf_slow<LeafGen>(p, acc, leaf_gen(g));
        else r = leaf_slow<const GemmArgs&>(p, acc, g);
    }
    return r;
}

// Leaves of N outputs, in groups that are branch-free in the common case: fast results are
// committed where valid, outputs needing the general path are recorded in a bit mask, and
// one wave-uniform branch runs the general path for them (acc still holds their old value).
template <bool SUBN, int N>
__device__ __forceinline__ void leaves_d(const double* p, double* acc, const GemmArgs& g) {
    constexpr int G = 8;   // outputs per branch-free group (bounds the live registers)
    static_assert(N % G == 0 || N < G, "group size");
#pragma unroll
    for (int g0 = 0; g0 < N; g0 += G) {
        uint32_t slow = 0;
#pragma unroll
        for (int o = g0; o < g0 + G && o < N; ++o) {
            double r;
            const bool ok = leaf_fast<SUBN>(p[o], acc[o], g, r);
            acc[o] = ok ? r : acc[o];
            slow |= (uint32_t)(!ok) << (o - g0);
        }
        if (__builtin_expect(__any(slow != 0), 0)) {
#pragma unroll
            for (int o = g0; o < g0 + G && o < N; ++o) {
                double z;
                if ((slow >> (o - g0)) & 1u)
                    acc[o] = leaf_zero(p[o], acc[o], g, z) ? z : leaf_slow<const GemmArgs&>(p[o], acc[o], g);
            }
        }
    }
}

// One BM x BN output tile with TM x TN outputs per thread (256 threads), K staged through
// LDS in BK slices.  tile_m / tile_n index the tile grid.
template <int BM, int BN, int TM, int TN, int BK, bool SUBN>
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, int tile_m, int tile_n, int tid,
                                          float* smem) {
    static_assert((BM / TM) * (BN / TN) == 256, "256 threads per tile");
    constexpr int LDA_S = BM + 4;   // padded LDS rows
    constexpr int LDB_S = BN + 4;
    float* As = smem;               // [BK][LDA_S]: As[k][m]
    float* Bs = smem + BK * LDA_S;  // [BK][LDB_S]: Bs[k][n]
    const int tx = tid % (BN / TN);
    const int ty = tid / (BN / TN);
    const int m0 = tile_m * BM, n0 = tile_n * BN;

    double acc[TM][TN];
    float part[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) { acc[i][j] = 0.0; part[i][j] = 0.0f; }
    int in_chunk = 0;

    for (int k0 = 0; k0 < g.K; k0 += BK) {
        // stage A[m0:m0+BM, k0:k0+BK] and B[k0:k0+BK, n0:n0+BN] (zero outside the matrix)
#pragma unroll
        for (int it = 0; it < (BM * BK) / 256; ++it) {
            const int idx = tid + it * 256;
            int m, k;
            if (g.ta) { m = idx % BM; k = idx / BM; }   // A stored K x M: contiguous in m
            else { k = idx % BK; m = idx / BK; }        // A row-major: contiguous in k
            const int gm = m0 + m, gk = k0 + k;
            float v = 0.0f;
            if (gm < g.M && gk < g.K)
                v = gld(g.ta ? g.A + (int64_t)gk * g.lda + gm : g.A + (int64_t)gm * g.lda + gk);
            As[k * LDA_S + m] = v;
        }
#pragma unroll
        for (int it = 0; it < (BN * BK) / 256; ++it) {
            const int idx = tid + it * 256;
            int n, k;
            if (g.tb) { k = idx % BK; n = idx / BK; }   // B stored N x K: contiguous in k
            else { n = idx % BN; k = idx / BN; }        // B row-major: contiguous in n
            const int gn = n0 + n, gk = k0 + k;
            float v = 0.0f;
            if (gn < g.N && gk < g.K)
                v = gld(g.tb ? g.B + (int64_t)gn * g.ldb + gk : g.B + (int64_t)gk * g.ldb + gn);
            Bs[k * LDB_S + n] = v;
        }
        __syncthreads();
        const int kmax = min(BK, g.K - k0);
        for (int kk = 0; kk < kmax; ++kk) {
            float a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = As[kk * LDA_S + ty * TM + i];
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = Bs[kk * LDB_S + tx * TN + j];
            if (g.chunk == 1) {
                double ad[TM], bd[TN], pd[TM * TN];
#pragma unroll
                for (int i = 0; i < TM; ++i) ad[i] = (double)a[i];
#pragma unroll
                for (int j = 0; j < TN; ++j) bd[j] = (double)b[j];
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j) pd[i * TN + j] = ad[i] * bd[j];
                leaves_d<SUBN, TM * TN>(pd, &acc[0][0], g);
            } else {
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j) part[i][j] = __builtin_fmaf(a[i], b[j], part[i][j]);
                if (++in_chunk == g.chunk) {
                    in_chunk = 0;
                    double pd[TM * TN];
#pragma unroll
                    for (int i = 0; i < TM; ++i)
#pragma unroll
                        for (int j = 0; j < TN; ++j) {
                            pd[i * TN + j] = (double)part[i][j];
                            part[i][j] = 0.0f;
                        }
                    leaves_d<SUBN, TM * TN>(pd, &acc[0][0], g);
                }
            }
        }
        __syncthreads();
    }
    if (in_chunk != 0) {   // short last chunk
        double pd[TM * TN];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) pd[i * TN + j] = (double)part[i][j];
        leaves_d<SUBN, TM * TN>(pd, &acc[0][0], g);
    }
    // epilogue (cutlass LinearCombination): D = alpha * acc, or with a source term
    // D = fma(alpha, acc, beta * C)
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        const int gm = m0 + ty * TM + i;
        if (gm >= g.M) continue;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int gn = n0 + tx * TN + j;
            if (gn >= g.N) continue;
            const float a = (float)acc[i][j];
            float d;
            if (g.beta != 0.0f) d = __builtin_fmaf(g.alpha, a, g.beta * gld(g.C + (int64_t)gm * g.ldc + gn));
            else d = g.alpha * a;
            gst(g.D + (int64_t)gm * g.ldd + gn, d);
        }
    }
}

template <int BM, int BN, int BK>
constexpr int gemm_smem_floats() { return BK * (BM + 4) + BK * (BN + 4); }

}  // namespace lofamd
