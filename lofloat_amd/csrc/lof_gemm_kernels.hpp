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
//   exact tie on s broken by the sign of e.  ~20 VALU ops, no branches taken.
//   General path (any mode, any m, zeros, infinities, NaN, FP32 subnormal or out-of-range
//   accumulators): build RTO24(x) literally (neighbours rd/ru of x, rz = rd >= 0 ? rd : ru,
//   OR the inexact bit; an exact zero gets roundTowardNegative's sign) and apply the
//   reference's virtual_round to it.
// The accumulator lives in registers as a double holding a float value.
#pragma once

#include <hip/hip_runtime.h>

#include "lof_core.hpp"

namespace lofamd {

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

__device__ __noinline__ double leaf_general(double p, double acc, double s, double e,
                                            const GemmArgs& g) {
    const uint32_t F = rto24_bits(p, acc, s, e);
    const uint32_t R = g.acc_kind == 0 ? vr_mant_rt(F, g.acc_mode, g.acc_r, 0, 0u)
                                       : vr_fmt_rt(F, g.accF, 0u);
    return (double)__uint_as_float(R);
}

// One emulated accumulation step with exact product/partial sum p.  The fast-path result
// is computed unconditionally; lanes outside its domain (or every lane when the fast path
// is off: range == 0) redo the step on the general path.
__device__ __forceinline__ double leaf(double p, double acc, const GemmArgs& g) {
    const double s = p + acc;
    const double bb = s - p;
    const double e = (p - (s - bb)) + (acc - bb);
    const uint32_t hi = (uint32_t)__double2hiint(s);
    const uint32_t lo = (uint32_t)__double2loint(s);
    const uint32_t lsb = (hi >> g.rh) & 1u;
    const uint32_t same_sign = (((uint32_t)__double2hiint(e) ^ hi) >> 31) ^ 1u;
    const uint32_t up = (e == 0.0) ? lsb : same_sign;
    const uint32_t c = (lo != 0u) ? 1u : up;
    const uint32_t hi2 = (hi + g.half_m1 + c) & g.keep_mask;
    double r = __hiloint2double((int)hi2, 0);
    const bool ok = ((hi & 0x7FFFFFFFu) - g.lo_bound) < g.range;
    if (__builtin_expect(!ok, 0)) r = leaf_general(p, acc, s, e, g);
    return r;
}

// One BM x BN output tile with TM x TN outputs per thread (256 threads), K staged through
// LDS in BK slices.  tile_m / tile_n index the tile grid.
template <int BM, int BN, int TM, int TN, int BK>
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
                v = g.ta ? g.A[(int64_t)gk * g.lda + gm] : g.A[(int64_t)gm * g.lda + gk];
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
                v = g.tb ? g.B[(int64_t)gn * g.ldb + gk] : g.B[(int64_t)gk * g.ldb + gn];
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
                double ad[TM], bd[TN];
#pragma unroll
                for (int i = 0; i < TM; ++i) ad[i] = (double)a[i];
#pragma unroll
                for (int j = 0; j < TN; ++j) bd[j] = (double)b[j];
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j) acc[i][j] = leaf(ad[i] * bd[j], acc[i][j], g);
            } else {
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j) part[i][j] = __builtin_fmaf(a[i], b[j], part[i][j]);
                if (++in_chunk == g.chunk) {
                    in_chunk = 0;
#pragma unroll
                    for (int i = 0; i < TM; ++i)
#pragma unroll
                        for (int j = 0; j < TN; ++j) {
                            acc[i][j] = leaf((double)part[i][j], acc[i][j], g);
                            part[i][j] = 0.0f;
                        }
                }
            }
        }
        __syncthreads();
    }
    if (in_chunk != 0) {   // short last chunk
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = leaf((double)part[i][j], acc[i][j], g);
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
            if (g.beta != 0.0f) d = __builtin_fmaf(g.alpha, a, g.beta * g.C[(int64_t)gm * g.ldc + gn]);
            else d = g.alpha * a;
            g.D[(int64_t)gm * g.ldd + gn] = d;
        }
    }
}

template <int BM, int BN, int BK>
constexpr int gemm_smem_floats() { return BK * (BM + 4) + BK * (BN + 4); }

}  // namespace lofamd
