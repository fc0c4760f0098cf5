// lof_core.hpp -- device rounding core for gfx950 (and its host-side parameter setup).
//
// Bit-exact restatement of the reference's scalar rounding (SudhanvaKulkarni123/LoFloat):
//   round_bits<MODE>   RoundMantissa + the 15 bit rounders   lo_float.h:571-1272
//   vr_fmt<MODE>       virtual_round(x, FloatingPointParams)  lo_float.h:2844-2939
//   vr_mant<MODE>      virtual_round(x, ToMantissaBits)       lo_float.h:2716-2756
// Everything is 32-bit integer work on the float's bit pattern, so it runs on the
// VALU at full rate; the kernels built on it are HBM-bound.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lofloat.h"

namespace lofamd {

enum Mode : int {
    RNE = 0, RTZ = 1, RAFZ = 2, SRA = 3, RNO = 4, RDOWN = 5, RUP = 6, RTA = 7,
    SRB = 8, SRC = 9, TRUESR = 10, PROB = 11, RTO = 12, RTTZ = 13, SRD = 14
};

__host__ __device__ constexpr bool is_stochastic(int m) {
    return m == SRA || m == SRB || m == SRC || m == TRUESR || m == PROB || m == SRD;
}
// Modes whose rounder reads stoch_len and must have stoch_len <= dropped bits.
__host__ __device__ constexpr bool uses_len(int m) {
    return m == SRA || m == SRB || m == SRC || m == SRD;
}

// Everything a kernel needs for virtual_round(x, FloatingPointParams, ProjSpec),
// precomputed on the host (make_fmt_dev).
struct FmtDev {
    int32_t r;             // dropped mantissa bits, 23 - mantissa_bits
    int32_t emin_b;        // ToMin_exp + 127: (mag >> 23) < emin_b  <=>  from_exp < ToMin_exp
    int32_t emax_b;        // ToMax_exp + 127: (mag >> 23) > emax_b  <=>  from_exp > ToMax_exp
    uint32_t min_bits;     // bits of To_min_val = (float)(2^ToMin_exp * 2^-mb)
    uint32_t half_bits;    // bits of To_min_val * 0.5f
    float min_val;
    int32_t saturating;    // Inf_Behaviors::Saturating
    int32_t stoch_len;
    int32_t mode;
};

// SplitMix64 output function over a Weyl sequence (Steele, Lea, Flood 2014).
__host__ __device__ __forceinline__ uint32_t rand_u32(uint64_t seed, uint64_t index) {
    uint64_t z = seed + (index + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z >> 32);
}

__device__ __forceinline__ uint32_t lowmask(int r) { return (1u << r) - 1u; }  // r in [0, 31]

// RoundMantissa(bits, r, ps, pos) for r >= 1 (every caller guarantees it: the normal path
// only rounds when mantissa_bits < 23, the subnormal path drops >= 1 bit, and the
// mantissa-only overload short-circuits r == 0).  Stochastic modes read one random word u:
//   SR-A/B/C: R = u & (2^len-1);  SR-D: R = u & (2^len-1), coin = bit len of u;
//   TrueSR: sample u / 2^32;  Prob: coin = u & 1.
template <int MODE>
__device__ __forceinline__ uint32_t round_bits(uint32_t bits, int r, int len, bool pos,
                                               uint32_t u) {
    const uint32_t low = lowmask(r);
    const uint32_t T = bits & ~low;
    const uint32_t tail = bits & low;
    const uint32_t one = 1u << r;
    if constexpr (MODE == RNE) {            // lo_float.h:571-597
        return bits + ((bits >> r) & 1u) + (one >> 1) - 1u;
    } else if constexpr (MODE == RTZ) {     // 947-965
        return T;
    } else if constexpr (MODE == RAFZ) {    // 1001-1028
        return T + (tail ? one : 0u);
    } else if constexpr (MODE == SRA) {     // 654-696
        if (len <= 0) return bits;
        return bits + ((u & lowmask(len)) << (r - len));
    } else if constexpr (MODE == RNO) {     // 1031-1057
        return bits + ((~bits >> r) & 1u) + (one >> 1) - 1u;
    } else if constexpr (MODE == RDOWN) {   // 1122-1157
        return T + ((!pos && tail) ? one : 0u);
    } else if constexpr (MODE == RUP) {     // 1062-1100
        return T + ((pos && tail) ? one : 0u);
    } else if constexpr (MODE == RTA) {     // 1177-1201
        return T + (((bits >> (r - 1)) & 1u) << r);
    } else if constexpr (MODE == SRB) {     // 700-756 (scalar branch)
        return bits + (tail ? (((u & lowmask(len)) + 1u) << (r - len)) : 0u);
    } else if constexpr (MODE == SRC) {     // 759-815 (scalar branch)
        const uint32_t R = u & lowmask(len);
        if (len == 0 || tail == 0) return bits;
        return bits + (((R >> 1) + (R & 1u)) << (r - len + 1));
    } else if constexpr (MODE == TRUESR) {  // 866-944: up iff u/2^32 < tail/2^r
        const bool up = ((uint64_t)u << r) < ((uint64_t)tail << 32);
        return up ? T + one : T;
    } else if constexpr (MODE == PROB) {    // 601-651
        return T + (tail ? ((u & 1u) << r) : 0u);
    } else if constexpr (MODE == RTO) {     // 1204-1230
        return T | (tail ? one : 0u);
    } else if constexpr (MODE == RTTZ) {    // 975-998
        return bits + (one >> 1) - 1u;
    } else if constexpr (MODE == SRD) {     // 821-862
        if (len <= 0) return bits;
        return bits + (((u & lowmask(len)) + ((u >> len) & 1u)) << (r - len));
    } else {
        return bits;
    }
}

// virtual_round(x, FloatingPointParams, ProjSpec) on the bit pattern xb.
template <int MODE>
__device__ __forceinline__ uint32_t vr_fmt(uint32_t xb, const FmtDev& P, uint32_t u) {
    const uint32_t mag = xb & 0x7FFFFFFFu;
    const uint32_t sign = xb ^ mag;
    if (mag > 0x7F800000u) return 0x7FC00000u;          // isnan -> +qNaN (2869)
    if (mag < P.min_bits) {                             // |x| < To_min_val (2871-2908)
        bool to_min;                                    // non-negative floats order like uints
        if constexpr (MODE == RTZ) {
            to_min = false;
        } else if constexpr (MODE == RTO || MODE == RAFZ) {
            to_min = mag != 0u;
        } else if constexpr (MODE == RUP) {
            to_min = sign == 0u && mag != 0u;
        } else if constexpr (MODE == RDOWN) {
            to_min = sign != 0u && mag != 0u;
        } else if constexpr (MODE == RNE || MODE == RNO || MODE == RTTZ || MODE == RTA) {
            to_min = mag > P.half_bits ||
                     ((MODE == RNO || MODE == RTA) && mag == P.half_bits);
        } else {
            to_min = false;
        }
        return sign | (to_min ? P.min_bits : 0u);
    }
    const int eb = (int)(mag >> 23);
    if (eb < P.emin_b) {                                // target-subnormal grid (2915-2923)
        const int shift = P.r + (P.emin_b - eb);
        uint32_t sig = (mag & 0x7FFFFFu) | 0x800000u;
        sig = round_bits<MODE>(sig, shift, P.stoch_len, sign == 0u, u);
        const uint32_t k = sig >> shift;
        // k * To_min_val is exact in float (k <= 2^24, To_min_val a power of two, 0 or
        // inf), so this equals the reference's (float)((double)k * (double)To_min_val).
        const float v = (float)k * P.min_val;
        return sign | __float_as_uint(v);
    }
    uint32_t m = mag;
    if (P.r > 0) {                                      // 2925-2933
        m = round_bits<MODE>(m, P.r, P.stoch_len, sign == 0u, u);
        m &= ~lowmask(P.r);
    }
    if ((int)(m >> 23) > P.emax_b) return P.saturating ? xb : 0x7F800000u;  // 2935-2936
    return sign | m;
}

// virtual_round(x, ToMantissaBits, ProjSpec) on the bit pattern xb; r = 23 - bits >= 1.
template <int MODE>
__device__ __forceinline__ uint32_t vr_mant(uint32_t xb, int r, int len, uint32_t u) {
    const uint32_t sign = xb & 0x80000000u;
    uint32_t m = round_bits<MODE>(xb & 0x7FFFFFFFu, r, len, sign == 0u, u);
    m &= ~lowmask(r);
    return sign | m;
}

// Runtime-mode dispatchers (for paths where the mode is not a template argument).
__device__ inline uint32_t vr_fmt_rt(uint32_t xb, const FmtDev& P, uint32_t u) {
    switch (P.mode) {
    case RNE: return vr_fmt<RNE>(xb, P, u);
    case RTZ: return vr_fmt<RTZ>(xb, P, u);
    case RAFZ: return vr_fmt<RAFZ>(xb, P, u);
    case SRA: return vr_fmt<SRA>(xb, P, u);
    case RNO: return vr_fmt<RNO>(xb, P, u);
    case RDOWN: return vr_fmt<RDOWN>(xb, P, u);
    case RUP: return vr_fmt<RUP>(xb, P, u);
    case RTA: return vr_fmt<RTA>(xb, P, u);
    case SRB: return vr_fmt<SRB>(xb, P, u);
    case SRC: return vr_fmt<SRC>(xb, P, u);
    case TRUESR: return vr_fmt<TRUESR>(xb, P, u);
    case PROB: return vr_fmt<PROB>(xb, P, u);
    case RTO: return vr_fmt<RTO>(xb, P, u);
    case RTTZ: return vr_fmt<RTTZ>(xb, P, u);
    case SRD: return vr_fmt<SRD>(xb, P, u);
    default: return xb;
    }
}

__device__ inline uint32_t vr_mant_rt(uint32_t xb, int mode, int r, int len, uint32_t u) {
    if (r == 0) return xb;
    switch (mode) {
    case RNE: return vr_mant<RNE>(xb, r, len, u);
    case RTZ: return vr_mant<RTZ>(xb, r, len, u);
    case RAFZ: return vr_mant<RAFZ>(xb, r, len, u);
    case SRA: return vr_mant<SRA>(xb, r, len, u);
    case RNO: return vr_mant<RNO>(xb, r, len, u);
    case RDOWN: return vr_mant<RDOWN>(xb, r, len, u);
    case RUP: return vr_mant<RUP>(xb, r, len, u);
    case RTA: return vr_mant<RTA>(xb, r, len, u);
    case SRB: return vr_mant<SRB>(xb, r, len, u);
    case SRC: return vr_mant<SRC>(xb, r, len, u);
    case TRUESR: return vr_mant<TRUESR>(xb, r, len, u);
    case PROB: return vr_mant<PROB>(xb, r, len, u);
    case RTO: return vr_mant<RTO>(xb, r, len, u);
    case RTTZ: return vr_mant<RTTZ>(xb, r, len, u);
    case SRD: return vr_mant<SRD>(xb, r, len, u);
    default: return xb;
    }
}

}  // namespace lofamd
