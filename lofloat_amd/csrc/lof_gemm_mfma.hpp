// lof_gemm_mfma.hpp -- emulated-accumulation GEMM with MFMA chunk sums (gfx950).
//
// For chunk c > 1 the accumulation rule is: p = FP32 fmaf chain over the chunk's c products
// (ascending k, starting at +0), then acc <- round(RTO(p + acc)).  When every partial sum of
// the chunk is exactly representable in FP32, p is the exact sum of the products and any
// summation order gives it; v_mfma_f32_32x32x2f32 then computes the chunk sums.
//
// Exactness test, per workgroup tile and 32-wide K slice (every chunk of the slice is a
// subset of it): with amax_A, amax_B the largest magnitudes in the A and B slice tiles and
// lsb_A, lsb_B the smallest exponents of a set significand bit among their nonzero values,
// every product is a multiple of 2^L, L = lsb_A + lsb_B, and every partial sum is bounded by
// c * amax_A * amax_B < 2^(log2 c + e_A + e_B + 2).  The slice is exact when
//     L >= -126                      (no FP32-subnormal product or sum; no subnormal input)
//     log2 c + e_A + e_B + 2 <= L + 24   (the sums fit a 24-bit significand)
//     log2 c + e_A + e_B + 2 <= 127      (no overflow)
// and no input is inf or NaN.  A slice that fails runs the fmaf chain on VALU from the same
// LDS tiles, in the MFMA register layout, so the result is bit-identical either way.
//
// Tile: 256 lanes = 4 waves in 2 x 2; each wave owns (32*WMB) x (32*WNB) outputs as WMB x WNB
// blocks of 32 x 32.  MFMA 32x32x2 f32 register layout (wave64): lane l supplies
// A[l % 32][l / 32] and B[l / 32][l % 32]; result register r of lane l is
// D[8 * (r / 4) + 4 * (l / 32) + r % 4][l % 32].
//
// Pipelining: the next K slice is loaded from HBM into registers while the current slice
// computes from LDS; the chunk length is a template parameter so a chunk's LDS reads are
// issued together and its MFMAs run back to back.
#pragma once

#include <hip/hip_runtime.h>

#include "lof_gemm_kernels.hpp"

namespace lofamd {

typedef float f32x16 __attribute__((ext_vector_type(16)));

// Smallest exponent of a set significand bit of a nonzero float; FP32-subnormal values
// report -1000 so that a slice holding one never takes the MFMA path.
__device__ __forceinline__ int lsb_exp(uint32_t absbits) {
    const uint32_t be = absbits >> 23;
    if (be == 0u) return -1000;
    const uint32_t sig = (absbits & 0x7FFFFFu) | 0x800000u;
    return (int)be - 150 + __builtin_ctz(sig);
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __noinline__ float leaf_f_slow(float p, float acc, const GemmArgs& g) {
    return (float)leaf((double)p, (double)acc, g);
}

// acc <- round(RTO(p + acc)) for float p, acc (the chunk leaf), fast path only: returns
// false (result undefined) where the general path is needed.  s = RN24(p + acc) and its
// TwoSum residual e give x = s + e exactly.  For round-to-nearest-even into m <= 19 bits,
// RN(RTO24(x)) onto the format's grid equals RN(x) onto it (round-to-odd keeps >= 2 guard
// bits), and that grid has quantum 2^(E - m) in the normal range and the fixed 2^(elo - m)
// in the format's subnormal range.  So the result is s's bit pattern rounded at bit
// P = 23 - m + max(0, elo - E), an exact tie broken by e's sign.  SUBN = false when the
// format has no subnormal range above FP32's normal range (P is then always 23 - m).
// The fast domain is |s| in [2^lo_e, 2^ehi), lo_e = max(elo - m + 1, -126), which keeps P
// inside the mantissa field and leaves the top binade (overflow rule) to the general path.
template <bool SUBN>
__device__ __forceinline__ bool leaf_f_fast(float p, float acc, const GemmArgs& g, float& out) {
    const float s = p + acc;
    const float bb = s - p;
    const float e = (p - (s - bb)) + (acc - bb);
    const uint32_t sb = __float_as_uint(s);
    uint32_t r;
    if constexpr (SUBN) {
        const int eb = (int)((sb >> 23) & 0xFFu);
        const uint32_t P = (uint32_t)(g.rf + max(g.elo_b - eb, 0));
        const uint32_t lsb = (sb >> P) & 1u;
        const uint32_t same_sign = ((__float_as_uint(e) ^ sb) >> 31) ^ 1u;
        const uint32_t c = (e == 0.0f) ? lsb : same_sign;
        const uint32_t unit = 1u << P;
        r = (sb + (unit >> 1) - 1u + c) & (0u - unit);
    } else {
        const uint32_t lsb = (sb >> g.rf) & 1u;
        const uint32_t same_sign = ((__float_as_uint(e) ^ sb) >> 31) ^ 1u;
        const uint32_t c = (e == 0.0f) ? lsb : same_sign;
        r = (sb + g.half_m1f + c) & g.keep_maskf;
    }
    out = __uint_as_float(r);
    return ((sb & 0x7FFFFFFFu) - g.lo_f) < g.range_f;
}

// Leaves of one 32x32 block's 16 outputs per lane, branch-free in the common case: fast results are committed
// where valid, lanes needing the general path are recorded in a bit mask, and one
// wave-uniform branch runs the general path for them (acc still holds their old value).
template <bool SUBN>
__device__ __forceinline__ void leaves_f16(const f32x16& p, float* acc, const GemmArgs& g) {
    uint32_t slow = 0;
#pragma unroll
    for (int o = 0; o < 16; ++o) {
        float r;
        const bool ok = leaf_f_fast<SUBN>(p[o], acc[o], g, r);
        acc[o] = ok ? r : acc[o];
        slow |= (uint32_t)(!ok) << o;
    }
    if (__builtin_expect(__any(slow != 0), 0)) {
#pragma unroll
        for (int o = 0; o < 16; ++o)
            if ((slow >> o) & 1u) acc[o] = leaf_f_slow(p[o], acc[o], g);
    }
}

template <int WMB, int WNB>
struct MfmaTile {
    static constexpr int BM = 64 * WMB;
    static constexpr int BN = 64 * WNB;
    static constexpr int BK = 32;
    static constexpr int LDA_S = BM + 4;   // As[k][m], rows padded to keep float4 alignment
    static constexpr int LDB_S = BN + 4;   // Bs[k][n]
    static constexpr int SMEM_FLOATS = BK * LDA_S + BK * LDB_S + 16;
    static constexpr int A_PER_THREAD = BM * BK / 256;
    static constexpr int B_PER_THREAD = BN * BK / 256;
};

template <int WMB, int WNB, int CHUNK, bool SUBN>
__device__ __forceinline__ void gemm_tile_mfma(const GemmArgs& g, int tile_m, int tile_n,
                                               int tid, float* smem) {
    using T = MfmaTile<WMB, WNB>;
    constexpr int BM = T::BM, BN = T::BN, BK = T::BK, LDA_S = T::LDA_S, LDB_S = T::LDB_S;
    constexpr int NA = T::A_PER_THREAD, NB = T::B_PER_THREAD;
    constexpr int LOG2C = CHUNK == 2 ? 1 : CHUNK == 4 ? 2 : CHUNK == 8 ? 3 : CHUNK == 16 ? 4 : 5;
    static_assert(CHUNK >= 2 && CHUNK <= BK && (BK % CHUNK) == 0, "chunk must divide 32");
    float* As = smem;
    float* Bs = smem + BK * LDA_S;
    uint32_t* stat = (uint32_t*)(Bs + BK * LDB_S);   // [4 waves][amaxA, lsbA, amaxB, lsbB]

    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int l32 = lane & 31, lhi = lane >> 5;
    const int m0 = tile_m * BM, n0 = tile_n * BN;
    const int am0 = wm * 32 * WMB, bn0 = wn * 32 * WNB;   // this wave's sub-tile

    float acc[WMB][WNB][16];
    f32x16 part[WMB][WNB];
#pragma unroll
    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
        for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[bi][bj][r] = 0.0f;

    // global -> registers for the slice at k0, zero outside the matrix.  The loads are
    // unconditional from clamped (always valid) addresses, so they issue without branches;
    // M, N, K >= 1 whenever a tile runs.
    float va[NA], vb[NB];
    auto load_slice = [&](int k0) {
#pragma unroll
        for (int it = 0; it < NA; ++it) {
            const int idx = tid + it * 256;
            int m, k;
            if (g.ta) { m = idx % BM; k = idx / BM; }   // A stored K x M: contiguous in m
            else { k = idx % BK; m = idx / BK; }        // A row-major: contiguous in k
            const int gm = m0 + m, gk = k0 + k;
            const int cm = min(gm, g.M - 1), ck = min(gk, g.K - 1);
            const float v = g.ta ? g.A[(int64_t)ck * g.lda + cm] : g.A[(int64_t)cm * g.lda + ck];
            va[it] = (gm < g.M && gk < g.K) ? v : 0.0f;
        }
#pragma unroll
        for (int it = 0; it < NB; ++it) {
            const int idx = tid + it * 256;
            int n, k;
            if (g.tb) { k = idx % BK; n = idx / BK; }   // B stored N x K: contiguous in k
            else { n = idx % BN; k = idx / BN; }        // B row-major: contiguous in n
            const int gn = n0 + n, gk = k0 + k;
            const int cn = min(gn, g.N - 1), ck = min(gk, g.K - 1);
            const float v = g.tb ? g.B[(int64_t)cn * g.ldb + ck] : g.B[(int64_t)ck * g.ldb + cn];
            vb[it] = (gn < g.N && gk < g.K) ? v : 0.0f;
        }
    };
    // registers -> LDS, with this slice's exactness statistics written to `stat`
    auto store_slice = [&]() {
        uint32_t amaxA = 0u, amaxB = 0u;
        int lsbA = 1000, lsbB = 1000;
#pragma unroll
        for (int it = 0; it < NA; ++it) {
            const int idx = tid + it * 256;
            int m, k;
            if (g.ta) { m = idx % BM; k = idx / BM; }
            else { k = idx % BK; m = idx / BK; }
            As[k * LDA_S + m] = va[it];
            const uint32_t ab = __float_as_uint(va[it]) & 0x7FFFFFFFu;
            amaxA = max(amaxA, ab);
            if (ab) lsbA = min(lsbA, lsb_exp(ab));
        }
#pragma unroll
        for (int it = 0; it < NB; ++it) {
            const int idx = tid + it * 256;
            int n, k;
            if (g.tb) { k = idx % BK; n = idx / BK; }
            else { n = idx % BN; k = idx / BN; }
            Bs[k * LDB_S + n] = vb[it];
            const uint32_t ab = __float_as_uint(vb[it]) & 0x7FFFFFFFu;
            amaxB = max(amaxB, ab);
            if (ab) lsbB = min(lsbB, lsb_exp(ab));
        }
        amaxA = wave_max_u32(amaxA);
        amaxB = wave_max_u32(amaxB);
        lsbA = wave_min_i32(lsbA);
        lsbB = wave_min_i32(lsbB);
        if (lane == 0) {
            stat[wave * 4 + 0] = amaxA;
            stat[wave * 4 + 1] = (uint32_t)lsbA;
            stat[wave * 4 + 2] = amaxB;
            stat[wave * 4 + 3] = (uint32_t)lsbB;
        }
    };

    if (g.K > 0) {
        load_slice(0);
        store_slice();
    }
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        __syncthreads();   // slice k0 and its statistics are in LDS
        bool exact;
        {
            const uint32_t amaxA = max(max(stat[0], stat[4]), max(stat[8], stat[12]));
            const uint32_t amaxB = max(max(stat[2], stat[6]), max(stat[10], stat[14]));
            const int lsbA = min(min((int)stat[1], (int)stat[5]), min((int)stat[9], (int)stat[13]));
            const int lsbB = min(min((int)stat[3], (int)stat[7]), min((int)stat[11], (int)stat[15]));
            if (amaxA == 0u || amaxB == 0u) {
                exact = true;                              // every product is zero
            } else if (amaxA >= 0x7F800000u || amaxB >= 0x7F800000u) {
                exact = false;                             // inf or NaN in the slice
            } else {
                const int L = lsbA + lsbB;
                const int top = LOG2C + (int)(amaxA >> 23) - 127 + (int)(amaxB >> 23) - 127 + 2;
                exact = L >= -126 && top <= L + 24 && top <= 127;
            }
            if (g.mfma_mode != 0) exact = g.mfma_mode == 2;
        }
        const bool has_next = k0 + BK < g.K;
        if (has_next) load_slice(k0 + BK);   // in flight while this slice computes

        const int kmax = min(BK, g.K - k0);
#pragma unroll
        for (int q = 0; q < BK / CHUNK; ++q) {
            const int kb = q * CHUNK;
            if (kb >= kmax) break;
            if (exact) {
                float a[CHUNK / 2][WMB], b[CHUNK / 2][WNB];
#pragma unroll
                for (int kp = 0; kp < CHUNK / 2; ++kp) {
                    const int kk = kb + 2 * kp + lhi;
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi) a[kp][bi] = As[kk * LDA_S + am0 + bi * 32 + l32];
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj) b[kp][bj] = Bs[kk * LDB_S + bn0 + bj * 32 + l32];
                }
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                        for (int r = 0; r < 16; ++r) part[bi][bj][r] = 0.0f;
                // all of the chunk's LDS reads are issued before its MFMAs
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int kp = 0; kp < CHUNK / 2; ++kp)
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj)
                            part[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                                a[kp][bi], b[kp][bj], part[bi][bj], 0, 0, 0);
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                        for (int r = 0; r < 16; ++r)   // an exact zero sum is +0, as the chain gives
                            part[bi][bj][r] = part[bi][bj][r] + 0.0f;
            } else {
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                        for (int r = 0; r < 16; ++r) part[bi][bj][r] = 0.0f;
#pragma unroll
                for (int kq = 0; kq < CHUNK; ++kq) {
                    const int kk = kb + kq;
                    if (kk >= kmax) break;
                    float4 a4[WMB][4];
                    float b[WNB];
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                        for (int rq = 0; rq < 4; ++rq)
                            a4[bi][rq] = *(const float4*)&As[kk * LDA_S + am0 + bi * 32 + rq * 8 + lhi * 4];
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj) b[bj] = Bs[kk * LDB_S + bn0 + bj * 32 + l32];
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                            for (int rq = 0; rq < 4; ++rq) {
                                part[bi][bj][rq * 4 + 0] = __builtin_fmaf(a4[bi][rq].x, b[bj], part[bi][bj][rq * 4 + 0]);
                                part[bi][bj][rq * 4 + 1] = __builtin_fmaf(a4[bi][rq].y, b[bj], part[bi][bj][rq * 4 + 1]);
                                part[bi][bj][rq * 4 + 2] = __builtin_fmaf(a4[bi][rq].z, b[bj], part[bi][bj][rq * 4 + 2]);
                                part[bi][bj][rq * 4 + 3] = __builtin_fmaf(a4[bi][rq].w, b[bj], part[bi][bj][rq * 4 + 3]);
                            }
                }
            }
#pragma unroll
            for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                for (int bj = 0; bj < WNB; ++bj) leaves_f16<SUBN>(part[bi][bj], acc[bi][bj], g);
        }
        __syncthreads();   // every wave is done reading slice k0
        if (has_next) store_slice();
    }

    // epilogue (cutlass LinearCombination): D = alpha * acc, or fma(alpha, acc, beta * C)
#pragma unroll
    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
        for (int bj = 0; bj < WNB; ++bj) {
            const int gn = n0 + bn0 + bj * 32 + l32;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gm = m0 + am0 + bi * 32 + (r >> 2) * 8 + lhi * 4 + (r & 3);
                if (gm >= g.M || gn >= g.N) continue;
                const float a = acc[bi][bj][r];
                float d;
                if (g.beta != 0.0f) d = __builtin_fmaf(g.alpha, a, g.beta * g.C[(int64_t)gm * g.ldc + gn]);
                else d = g.alpha * a;
                g.D[(int64_t)gm * g.ldd + gn] = d;
            }
        }
}

}  // namespace lofamd
