// lof_pack.hpp -- device side of the n-bit pack codec (quantise to an n-bit rep and bit-pack,
// and the inverse), 1 <= n <= 16.
//
// Restates lo_pack.hpp (SudhanvaKulkarni123/LoFloat):
//   float_to_rep<MODE>   convert_float_to_rep   lo_pack.hpp:108-258
//   rep_to_float         convert_rep_to_float   lo_pack.hpp:268-304
//   group geometry       pack_group_* / bit_offset / put_bits / get_bits   lo_pack.hpp:58-100
// Elements are n-bit fields, LSB-first within bytes, in groups of lcm(n, 8) (n <= 8) or
// lcm(n, 16) bits; Forward puts element i of a group in slot i, Reverse in slot gc-1-i.  A
// group has no padding bits (its bit count is a multiple of n), so the packed buffer is a
// stream of slots: slot S holds element S (Forward) or g*gc + gc-1-s for S = g*gc + s.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lof_core.hpp"

namespace lofamd {

// An Inf or NaN checker functor as data (see lof_checker in lofloat.h).
struct CheckDev {
    int32_t kind, ncodes;
    uint32_t codes[3];
    uint32_t a, b;
};

__host__ __device__ __forceinline__ bool check_match(const CheckDev& c, uint32_t rep, bool nan) {
    switch (c.kind) {
    case 1: {
        bool m = false;
        for (int i = 0; i < 3; ++i) m |= (i < c.ncodes) && rep == c.codes[i];
        return m;
    }
    case 2: return (rep & c.a) == c.a && (nan ? (rep & c.b) != 0u : (rep & c.b) == 0u);
    case 3: return (rep | c.a) == c.codes[0];
    default: return false;
    }
}

// Everything the codec needs, precomputed on the host (make_pack_dev).
struct PackDev {
    int32_t n, mb, bias;
    int32_t is_signed;
    int32_t r;               // 23 - mb: dropped float mantissa bits
    int32_t exp_off;         // bias - 127
    int32_t branch;          // 1: bias > 126 (float subnormals normalised), 2: bias < 126, 0
    int32_t has_inf;
    int32_t narrower;        // (ToMax_exp, digits) < (128, 24)
    int32_t unsigned_nan;    // negative into unsigned: 0 -> 0, 1 -> qNaN, 2 -> sNaN
    int32_t saturation;      // Saturation_Mode: 0 OvfInf, 1 SatFinite, 2 SatPropagate
    int32_t stoch_len;
    uint32_t signmask, max_finite, inf_rep, qnan, snan, mask;
    int32_t nan_signmask;    // IsNaN(signmask): no -0 code
    uint64_t aligned_highest;
    uint32_t off2;           // exp_off * 2^mb (mod 2^32): the rebias after the shift
    uint32_t ovf_rep;        // magnitude of an overflowing value: saturation 0 with Inf -> Inf
    int32_t fast32;          // exp_off <= 255: the normal path fits 32-bit arithmetic
    CheckDev is_inf, is_nan; // decode
    int32_t gc, gbits;       // group geometry
};

// RoundMantissa on a 64-bit word (r in [1, 63]); random words mapped as in round_bits.
template <int MODE>
__device__ __forceinline__ uint64_t round_bits64(uint64_t bits, int r, int len, uint32_t u) {
    const uint64_t one = (uint64_t)1 << r;
    const uint64_t low = one - 1u;
    const uint64_t T = bits & ~low, tail = bits & low;
    const uint64_t R = len > 0 ? (uint64_t)(u & (uint32_t)(((uint64_t)1 << len) - 1u)) : 0u;
    if constexpr (MODE == RNE) return bits + ((bits >> r) & 1u) + (one >> 1) - 1u;
    else if constexpr (MODE == RTZ) return T;
    else if constexpr (MODE == RAFZ) return T + (tail ? one : 0u);
    else if constexpr (MODE == SRA) return len <= 0 ? bits : bits + (R << (r - len));
    else if constexpr (MODE == RNO) return bits + ((~bits >> r) & 1u) + (one >> 1) - 1u;
    else if constexpr (MODE == RDOWN) return T + (tail ? one : 0u);   // pos = false
    else if constexpr (MODE == RUP) return T;                          // pos = false
    else if constexpr (MODE == RTA) return T + (((bits >> (r - 1)) & 1u) << r);
    else if constexpr (MODE == SRB) return bits + (tail > 0 ? ((R + 1u) << (r - len)) : 0u);
    else if constexpr (MODE == SRC)
        return (len == 0 || tail == 0) ? bits : bits + (((R >> 1) + (R & 1u)) << (r - len + 1));
    else if constexpr (MODE == TRUESR)   // up iff u * 2^r < tail * 2^32 (r <= 24 here)
        return ((uint64_t)u << r) < (tail << 32) ? T + one : T;
    else if constexpr (MODE == PROB) return T + (tail ? ((uint64_t)(u & 1u) << r) : 0u);
    else if constexpr (MODE == RTO) return T | (tail ? one : 0u);
    else if constexpr (MODE == RTTZ) return bits + (one >> 1) - 1u;
    else if constexpr (MODE == SRD)
        return len <= 0 ? bits : bits + ((R + ((u >> len) & 1u)) << (r - len));
    else return bits;
}

// Stochastic rounding of bits < 2^24 at a shift r > 24: 0 or 1, the exact integer
// (bits + increment) >> r (the reference's uint64_t shift is undefined once r >= 64).
template <int MODE>
__device__ __forceinline__ uint32_t stoch_wide(uint32_t bits, int r, int len, uint32_t u) {
    const uint64_t R = len > 0 ? (uint64_t)(u & (uint32_t)(((uint64_t)1 << len) - 1u)) : 0u;
    const int s = r - len;
    const uint64_t hi = s >= 32 ? 0u : (uint64_t)(bits >> s);
    const uint64_t full = (uint64_t)1 << len;
    if constexpr (MODE == SRA) return (len > 0 && R + hi >= full) ? 1u : 0u;
    else if constexpr (MODE == SRB) return (bits > 0 && R + 1u + hi >= full) ? 1u : 0u;
    else if constexpr (MODE == SRC) {
        if (len == 0 || bits == 0) return 0u;
        const int s1 = s + 1;
        const uint64_t hi1 = s1 >= 32 ? 0u : (uint64_t)(bits >> s1);
        return ((R >> 1) + (R & 1u) + hi1 >= ((uint64_t)1 << (len - 1))) ? 1u : 0u;
    } else if constexpr (MODE == SRD) {
        return (len > 0 && R + ((u >> len) & 1u) + hi >= full) ? 1u : 0u;
    } else if constexpr (MODE == TRUESR) {
        if (bits == 0) return 0u;
        if (r <= 32) return ((uint64_t)u < ((uint64_t)bits << (32 - r))) ? 1u : 0u;
        if (r - 32 >= 24) return u == 0u ? 1u : 0u;
        return (((uint64_t)u << (r - 32)) < (uint64_t)bits) ? 1u : 0u;
    } else if constexpr (MODE == PROB) {
        return bits ? (u & 1u) : 0u;
    } else {
        return 0u;
    }
}

// convert_float_to_rep (lo_pack.hpp:108-258) with SOURCE float and the parameters in P.
// RoundMantissa is called there without a sign, i.e. pos = false, so RoundUp truncates and
// RoundDown rounds magnitudes away from zero; only the deep-underflow branch reads the sign.
template <int MODE>
__device__ __forceinline__ uint32_t float_to_rep(float from, const PackDev& P, uint32_t u) {
    const uint32_t xb = __float_as_uint(from);
    const bool neg = (xb >> 31) != 0u;
    const uint32_t from_bits = xb & 0x7FFFFFFFu;
    const bool sgn = neg && P.is_signed;
    auto with_sign = [&](uint32_t mag) -> uint32_t {   // :151-155
        if (!sgn) return mag;
        if (mag == 0u && P.nan_signmask) return 0u;
        return mag | P.signmask;
    };
    if (!P.is_signed && neg)                                                 // :161-166
        return P.unsigned_nan == 0 ? 0u : (P.unsigned_nan == 1 ? P.qnan : P.snan);
    if (from_bits == 0x7F800000u)                                            // :170-174
        return with_sign((P.saturation != 1 && P.has_inf) ? P.inf_rep : P.max_finite);
    if (from_bits > 0x7F800000u) return P.qnan;                             // :175
    if (from_bits == 0u) return with_sign(0u);                              // :176
    const int bfe = (int)(from_bits >> 23);

    if (P.branch == 1 && bfe == 0) {                                         // :181-198
        uint64_t bits = from_bits;
        const int nf = __builtin_clz(from_bits) - 9 + 1;
        const int be = P.exp_off - nf + 1;
        if (be <= 0) {
            if (P.exp_off < 64) bits <<= P.exp_off;
        } else {
            bits <<= nf;
            bits &= ~((uint64_t)1 << 23);
            bits |= (uint64_t)be << 23;
        }
        bits = round_bits64<MODE>(bits, P.r, P.stoch_len, u) >> P.r;
        return with_sign((uint32_t)bits);
    }
    if (P.branch == 2) {                                                     // :201-239
        const int bte = bfe - 127 + P.bias;
        if (bte <= 0) {
            const uint32_t lead = bfe > 0 ? 1u : 0u;
            const int shift = P.r - bte + (int)lead;
            const uint32_t rfb = (from_bits & 0x7FFFFFu) | (lead << 23);
            uint32_t bits;
            if (shift <= 24) {
                bits = (uint32_t)(round_bits64<MODE>(rfb, shift, P.stoch_len, u) >> shift);
            } else if constexpr (MODE == RTO || MODE == RAFZ) {
                bits = rfb ? 1u : 0u;
            } else if constexpr (MODE == RUP) {
                bits = (rfb && !neg) ? 1u : 0u;
            } else if constexpr (MODE == RDOWN) {
                bits = (rfb && neg) ? 1u : 0u;
            } else if constexpr (is_stochastic(MODE)) {
                bits = stoch_wide<MODE>(rfb, shift, P.stoch_len, u);
            } else {
                bits = 0u;
            }
            return with_sign(bits);
        }
    }
    uint64_t rb = round_bits64<MODE>((uint64_t)from_bits, P.r, P.stoch_len, u);   // :242-257
    rb &= ~((((uint64_t)1) << P.r) - 1u);
    rb += (uint64_t)(int64_t)P.exp_off << 23;
    if (P.narrower && rb > P.aligned_highest)
        return with_sign((P.saturation == 0 && P.has_inf) ? P.inf_rep : P.max_finite);
    return with_sign((uint32_t)(rb >> P.r));
}

// round_bits64 for operands where bits + 2^r < 2^32 (r in [1, 24]): the same integers in
// 32-bit arithmetic.
template <int MODE>
__device__ __forceinline__ uint32_t round_bits32(uint32_t bits, int r, int len, uint32_t u) {
    const uint32_t one = 1u << r, low = one - 1u;
    const uint32_t T = bits & ~low, tail = bits & low;
    const uint32_t R = len > 0 ? (u & (uint32_t)(((uint64_t)1 << len) - 1u)) : 0u;
    if constexpr (MODE == RNE) return bits + ((bits >> r) & 1u) + (one >> 1) - 1u;
    else if constexpr (MODE == RTZ) return T;
    else if constexpr (MODE == RAFZ) return T + (tail ? one : 0u);
    else if constexpr (MODE == SRA) return len <= 0 ? bits : bits + (R << (r - len));
    else if constexpr (MODE == RNO) return bits + ((~bits >> r) & 1u) + (one >> 1) - 1u;
    else if constexpr (MODE == RDOWN) return T + (tail ? one : 0u);   // pos = false
    else if constexpr (MODE == RUP) return T;                          // pos = false
    else if constexpr (MODE == RTA) return T + (((bits >> (r - 1)) & 1u) << r);
    else if constexpr (MODE == SRB) return bits + (tail > 0 ? ((R + 1u) << (r - len)) : 0u);
    else if constexpr (MODE == SRC)
        return (len == 0 || tail == 0) ? bits : bits + (((R >> 1) + (R & 1u)) << (r - len + 1));
    else if constexpr (MODE == TRUESR)   // u * 2^r < tail * 2^32  <=>  u < tail * 2^(32-r)
        return (uint64_t)u < ((uint64_t)tail << (32 - r)) ? T + one : T;
    else if constexpr (MODE == PROB) return T + (tail ? ((u & 1u) << r) : 0u);
    else if constexpr (MODE == RTO) return T | (tail ? one : 0u);
    else if constexpr (MODE == RTTZ) return bits + (one >> 1) - 1u;
    else if constexpr (MODE == SRD)
        return len <= 0 ? bits : bits + ((R + ((u >> len) & 1u)) << (r - len));
    else return bits;
}

// float_to_rep without branches for the inputs that need no special handling: finite,
// not negative into an unsigned format, and either zero, a normal FP32 value on the normal
// path, or (bias < 126) a value below the format's normal range (for the stochastic modes,
// only where its alignment shift is <= 24).  Sets `ok` false for every other input, whose rep the caller takes from
// float_to_rep; where ok is true the result equals float_to_rep's.
//   normal path (lo_pack.hpp:242-257): ((round(fb, r) & ~low) + exp_off * 2^23) >> r
//       = (round(fb, r) >> r) + exp_off * 2^mb, exact in 32 bits here (the rebias cannot
//       wrap for these inputs, P.fast32 bounds it above), and its overflow test
//       rb > max_finite * 2^r is the same as (rb >> r) > max_finite;
//   below the normal range (lo_pack.hpp:201-239): round(rfb, shift) >> shift, taken only by
//       the lanes that need it (its shift is per lane; the normal path's is uniform); for
//       shift > 24 (rfb < 2^24 <= half the unit) the deterministic modes give 0, except the
//       directed ones, which the reference evaluates with the sign.
template <int MODE>
__device__ __forceinline__ uint32_t float_to_rep_fast(float from, const PackDev& P, uint32_t u,
                                                      bool& ok) {
    const uint32_t xb = __float_as_uint(from);
    const bool neg = (xb >> 31) != 0u;
    const uint32_t fb = xb & 0x7FFFFFFFu;
    const int bfe = (int)(fb >> 23);
    const int bte = bfe + P.exp_off;              // bfe - 127 + bias
    const bool sub = P.branch == 2 && bte <= 0;
    const uint32_t lead = bfe > 0 ? 1u : 0u;
    const int shs = min(P.r - bte + (int)lead, 31);
    uint32_t mag = (round_bits32<MODE>(fb, P.r, P.stoch_len, u) >> P.r) + P.off2;
    mag = (P.narrower && mag > P.max_finite) ? P.ovf_rep : mag;
    if (sub) {
        const int s = min(shs, 24);
        mag = round_bits32<MODE>((fb & 0x7FFFFFu) | (lead << 23), s, P.stoch_len, u) >> s;
        if constexpr (!is_stochastic(MODE)) {
            // shift > 24 with a nonzero significand below 2^24: only the directed modes
            // give a 1 (lo_pack.hpp:214-222); every other deterministic mode gives 0
            const uint32_t deep = (MODE == RTO || MODE == RAFZ) ? 1u
                                  : MODE == RUP ? (neg ? 0u : 1u)
                                  : MODE == RDOWN ? (neg ? 1u : 0u) : 0u;
            mag = shs > 24 ? deep : mag;
        }
    }
    mag = fb == 0u ? 0u : mag;
    ok = P.fast32 && fb < 0x7F800000u && !(neg && !P.is_signed) &&
         (fb == 0u || (sub ? (!is_stochastic(MODE) || shs <= 24) : bfe > 0));
    const bool sgn = neg && P.is_signed;
    return !sgn ? mag : ((mag == 0u && P.nan_signmask) ? 0u : (mag | P.signmask));
}

// convert_rep_to_float (lo_pack.hpp:268-304): Inf checked before NaN, the value formed in
// double from the fields, then rounded to float.  (1 + m / 2^mb) * 2^(e - bias) is formed
// as (2^mb + m) * 2^(e - bias - mb): the same double, without a division.
__device__ __forceinline__ float rep_to_float(uint32_t rep, const PackDev& P) {
    rep &= P.mask;
    const bool neg = P.is_signed && ((rep >> (P.n - 1)) & 1u);
    const uint32_t mag = P.is_signed ? (rep & (P.signmask - 1u)) : rep;
    const bool inf = P.has_inf && check_match(P.is_inf, rep, false);
    const bool nan = check_match(P.is_nan, rep, true);
    const uint32_t e = mag >> P.mb, m = mag & ((1u << P.mb) - 1u);
    const uint32_t sig = e == 0u ? m : (m | (1u << P.mb));
    const int k = (e == 0u ? 1 : (int)e) - P.bias - P.mb;
    uint32_t vb = __float_as_uint((float)ldexp((double)sig, k));
    vb = inf ? 0x7F800000u : (nan ? 0x7FC00000u : vb);
    return __uint_as_float(vb | ((neg && !nan) || (neg && inf) ? 0x80000000u : 0u));
}

// The element held by slot S of the packed stream (S-th n-bit field).
__host__ __device__ __forceinline__ int64_t slot_element(int64_t S, int gc, int endian) {
    if (!endian) return S;
    const int64_t g = S / gc;
    const int s = (int)(S - g * gc);
    return g * gc + (gc - 1 - s);
}

}  // namespace lofamd
