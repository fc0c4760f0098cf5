// lof_round_kernels.hpp -- per-tile device routines of the rounding kernels.  A "tile" is
// what one 256-thread workgroup does in one step: VPT * 1024 consecutive floats for the
// elementwise rounding, 2048 for MX.  The single-tensor kernels run one tile per
// workgroup; the grouped persistent kernel pulls tiles of many tensors from a work queue
// and calls the same routines.
#pragma once

#include <hip/hip_runtime.h>

#include "lof_core.hpp"

namespace lofamd {

struct RoundArgs {
    const float* in;
    float* out;
    int64_t n;
    FmtDev P;            // full-format path
    int32_t r;           // mantissa-only path: dropped bits
    int32_t len;         // mantissa-only path: stoch_len
    float scale;
    int32_t has_scale;
    const uint32_t* bits;
    uint64_t seed;
    int32_t vec_ok;      // in/out/bits 16-byte aligned
};

struct MxArgs {
    const float* in;
    float* out;
    int64_t n;
    int32_t lanes_log2;  // log2(block_size / 4)
    FmtDev E;            // element format + element spec
    FmtDev S;            // scale format + scale mode (stoch_len 0)
    float pmn;           // largest finite normal of the element format
    const uint32_t* bits;
    uint64_t seed;
    int32_t vec_ok;      // aligned and n % 4 == 0
};

template <int RSRC>
__device__ __forceinline__ uint32_t word(const uint32_t* bits, uint64_t seed, int64_t i) {
    if constexpr (RSRC == 1) return bits[i];
    else if constexpr (RSRC == 2) return rand_u32(seed, (uint64_t)i);
    else return 0u;
}

template <int MODE, bool MANT>
__device__ __forceinline__ float round_one(float x, const RoundArgs& a, uint32_t u) {
    const float v = a.has_scale ? a.scale * x : x;      // Lof_kernel.cu:41, 53
    const uint32_t vb = __float_as_uint(v);
    uint32_t y;
    if constexpr (MANT) y = a.r == 0 ? vb : vr_mant<MODE>(vb, a.r, a.len, u);
    else y = vr_fmt<MODE>(vb, a.P, u);
    return __uint_as_float(y);
}

// One tile = TPB lanes x VPT float4 slots, TPB * VPT * 4 consecutive floats.
template <int MODE, bool MANT, int RSRC, int VPT, int TPB = 256>
__device__ __forceinline__ void round_tile(const RoundArgs& a, int64_t tile, int tid) {
    if (a.vec_ok) {
        const int64_t nvec = a.n >> 2;
        const int64_t v0 = tile * (TPB * VPT) + tid;
        float4 x[VPT];
        uint4 w[VPT];
#pragma unroll
        for (int j = 0; j < VPT; ++j) {
            const int64_t vi = v0 + (int64_t)j * TPB;
            if (vi < nvec) {
                x[j] = reinterpret_cast<const float4*>(a.in)[vi];
                if constexpr (RSRC == 1) w[j] = reinterpret_cast<const uint4*>(a.bits)[vi];
            }
        }
#pragma unroll
        for (int j = 0; j < VPT; ++j) {
            const int64_t vi = v0 + (int64_t)j * TPB;
            if (vi < nvec) {
                uint32_t u0 = 0, u1 = 0, u2 = 0, u3 = 0;
                if constexpr (RSRC == 1) { u0 = w[j].x; u1 = w[j].y; u2 = w[j].z; u3 = w[j].w; }
                if constexpr (RSRC == 2) {
                    u0 = rand_u32(a.seed, 4 * vi + 0); u1 = rand_u32(a.seed, 4 * vi + 1);
                    u2 = rand_u32(a.seed, 4 * vi + 2); u3 = rand_u32(a.seed, 4 * vi + 3);
                }
                float4 y;
                y.x = round_one<MODE, MANT>(x[j].x, a, u0);
                y.y = round_one<MODE, MANT>(x[j].y, a, u1);
                y.z = round_one<MODE, MANT>(x[j].z, a, u2);
                y.w = round_one<MODE, MANT>(x[j].w, a, u3);
                reinterpret_cast<float4*>(a.out)[vi] = y;
            }
        }
        // the n % 4 trailing floats: first lanes of tile 0
        if (tile == 0) {
            const int64_t i = (nvec << 2) + tid;
            if (i < a.n) a.out[i] = round_one<MODE, MANT>(a.in[i], a, word<RSRC>(a.bits, a.seed, i));
        }
    } else {
        const int64_t e0 = tile * (TPB * VPT * 4) + tid;
#pragma unroll
        for (int j = 0; j < 4 * VPT; ++j) {
            const int64_t i = e0 + (int64_t)j * TPB;
            if (i < a.n) a.out[i] = round_one<MODE, MANT>(a.in[i], a, word<RSRC>(a.bits, a.seed, i));
        }
    }
}

// NaN-ignoring max as std::max(amax, v) with amax never NaN (LoPy_bind.cpp:207).
__device__ __forceinline__ float max_keep(float m, float v) { return (m < v) ? v : m; }

__device__ __forceinline__ float dpp_shuffle(float v, int ctrl) {
    switch (ctrl) {  // the control word must be a compile-time constant
    case 0xB1: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
    case 0x4E: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
    case 0x141: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
    default: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, false));
    }
}

// Max over groups of `lanes` consecutive lanes (a power of two <= 64, groups aligned),
// result in every lane of the group.  Inputs are never NaN (partial maxima of max_keep),
// so fmaxf is the plain maximum.  Offsets 1/2 use DPP quad permutes, 4 the half-row
// mirror and 8 the row mirror (each lane meets a lane of the other half of its group);
// 16 and 32 cross DPP rows and go through ds_bpermute.
__device__ __forceinline__ float group_max(float m, int lanes) {
    if (lanes > 1) m = fmaxf(m, dpp_shuffle(m, 0xB1));   // quad_perm [1,0,3,2]
    if (lanes > 2) m = fmaxf(m, dpp_shuffle(m, 0x4E));   // quad_perm [2,3,0,1]
    if (lanes > 4) m = fmaxf(m, dpp_shuffle(m, 0x141));  // row_half_mirror
    if (lanes > 8) m = fmaxf(m, dpp_shuffle(m, 0x140));  // row_mirror
    if (lanes > 16) m = fmaxf(m, __shfl_xor(m, 16, 64));
    if (lanes > 32) m = fmaxf(m, __shfl_xor(m, 32, 64));
    return m;
}

// Shared MX scale of a block from its amax: virtual_round(amax / pmn, scale format).
template <bool SRNE>
__device__ __forceinline__ float mx_scale(float amax, const MxArgs& a) {
    const uint32_t t = __float_as_uint(amax / a.pmn);   // division, as the CPU binding
    return __uint_as_float(SRNE ? vr_fmt<RNE>(t, a.S, 0u) : vr_fmt_rt(t, a.S, 0u));
}

template <int MODE, int EPL>
__device__ __forceinline__ void mx_round_lane(float (&x)[EPL], const uint32_t (&u)[EPL],
                                              float scale, const FmtDev& E) {
    const uint32_t sb = __float_as_uint(scale);
    const uint32_t sexp = sb >> 23;  // sign is 0 for every scale that reaches here
    if (!(scale > 0.0f)) {           // zero / NaN scale zeroes the block
#pragma unroll
        for (int e = 0; e < EPL; ++e) x[e] = 0.0f;
    } else if ((sb & 0x7FFFFFu) == 0u && sexp <= 254u) {
        // power-of-two scale: its reciprocal is exact, so x * (1/scale) rounds the same
        // real number as x / scale
        const float inv = __uint_as_float(sexp <= 253u ? (254u - sexp) << 23 : 0x00400000u);
#pragma unroll
        for (int e = 0; e < EPL; ++e)
            x[e] = __uint_as_float(vr_fmt<MODE>(__float_as_uint(x[e] * inv), E, u[e])) * scale;
    } else {
#pragma unroll
        for (int e = 0; e < EPL; ++e)
            x[e] = __uint_as_float(vr_fmt<MODE>(__float_as_uint(x[e] / scale), E, u[e])) * scale;
    }
}

// One MX tile = TPB * 8 consecutive floats (2048 for 256 lanes).  EPL = floats per lane: 8
// (two 16-byte accesses, block_size >= 8) or 4 (block_size 4, two passes).  A block is
// block_size / EPL consecutive lanes of one wave (TPB >= 64, block_size <= 256).
constexpr int64_t kMxTileElems = 2048;

template <int MODE, int RSRC, int EPL, bool SRNE, int TPB = 256>
__device__ __forceinline__ void mx_tile(const MxArgs& a, int64_t tile, int tid) {
    static_assert(TPB >= 64 && TPB % 64 == 0, "whole waves");
    constexpr int64_t TILE = (int64_t)TPB * 8;
    constexpr int PASSES = 8 / EPL;
    const int lanes = (1 << a.lanes_log2) * 4 / EPL;
#pragma unroll
    for (int pass = 0; pass < PASSES; ++pass) {
        const int64_t e0 = tile * TILE + (int64_t)pass * (TPB * EPL) + (int64_t)tid * EPL;
        float x[EPL];
        uint32_t u[EPL];
        if (a.vec_ok) {  // 16-byte aligned, n % 4 == 0
            const int64_t nv = a.n >> 2;
#pragma unroll
            for (int h = 0; h < EPL / 4; ++h) {
                const int64_t vi = (e0 >> 2) + h;
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                uint4 w = make_uint4(0u, 0u, 0u, 0u);
                if (vi < nv) {
                    v = reinterpret_cast<const float4*>(a.in)[vi];
                    if constexpr (RSRC == 1) w = reinterpret_cast<const uint4*>(a.bits)[vi];
                }
                x[4 * h + 0] = v.x; x[4 * h + 1] = v.y; x[4 * h + 2] = v.z; x[4 * h + 3] = v.w;
                u[4 * h + 0] = w.x; u[4 * h + 1] = w.y; u[4 * h + 2] = w.z; u[4 * h + 3] = w.w;
            }
        } else {
#pragma unroll
            for (int e = 0; e < EPL; ++e) {
                const int64_t i = e0 + e;
                x[e] = i < a.n ? a.in[i] : 0.0f;
                u[e] = 0u;
                if constexpr (RSRC == 1) u[e] = i < a.n ? a.bits[i] : 0u;
            }
        }
        if constexpr (RSRC == 2) {
#pragma unroll
            for (int e = 0; e < EPL; ++e) u[e] = rand_u32(a.seed, (uint64_t)(e0 + e));
        }
        float m = 0.0f;
#pragma unroll
        for (int e = 0; e < EPL; ++e) m = max_keep(m, fabsf(x[e]));
        m = group_max(m, lanes);
        const float scale = mx_scale<SRNE>(m, a);
        mx_round_lane<MODE, EPL>(x, u, scale, a.E);
        if (a.vec_ok) {
            const int64_t nv = a.n >> 2;
#pragma unroll
            for (int h = 0; h < EPL / 4; ++h) {
                const int64_t vi = (e0 >> 2) + h;
                if (vi < nv)
                    reinterpret_cast<float4*>(a.out)[vi] =
                        make_float4(x[4 * h + 0], x[4 * h + 1], x[4 * h + 2], x[4 * h + 3]);
            }
        } else {
#pragma unroll
            for (int e = 0; e < EPL; ++e)
                if (e0 + e < a.n) a.out[e0 + e] = x[e];
        }
    }
}

}  // namespace lofamd
