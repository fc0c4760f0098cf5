// lof_round_kernels.hpp -- per-tile device routines of the rounding kernels.  A "tile" is
// what one 256-thread workgroup does in one step: VPT * 1024 consecutive floats.  The
// single-tensor kernels run one tile per workgroup; the grouped persistent kernel pulls
// tiles of many tensors from a work queue and calls the same routines.
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

template <int MODE, bool MANT, int RSRC, int VPT>
__device__ __forceinline__ void round_tile(const RoundArgs& a, int64_t tile, int tid) {
    constexpr int TPB = 256;
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

template <int MODE>
__device__ __forceinline__ float mx_elem(float x, float scale, float inv, bool use_inv,
                                         const FmtDev& E, uint32_t u) {
    const float q = use_inv ? x * inv : x / scale;      // identical values when inv is exact
    const float r = __uint_as_float(vr_fmt<MODE>(__float_as_uint(q), E, u));
    return r * scale;
}

template <int MODE, int RSRC, int VPT>
__device__ __forceinline__ void mx_tile(const MxArgs& a, int64_t tile, int tid) {
    constexpr int TPB = 256;
    const int64_t nslots = (a.n + 3) >> 2;               // 4-float slots
    const int64_t v0 = tile * (TPB * VPT) + tid;
    const int lanes = 1 << a.lanes_log2;
    float4 x[VPT];
    uint4 w[VPT];
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
        const int64_t vi = v0 + (int64_t)j * TPB;
        if (a.vec_ok) {
            if (vi < nslots) {
                x[j] = reinterpret_cast<const float4*>(a.in)[vi];
                if constexpr (RSRC == 1) w[j] = reinterpret_cast<const uint4*>(a.bits)[vi];
            } else {
                x[j] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        } else {
            const int64_t i = 4 * vi;
            x[j].x = i + 0 < a.n ? a.in[i + 0] : 0.f;
            x[j].y = i + 1 < a.n ? a.in[i + 1] : 0.f;
            x[j].z = i + 2 < a.n ? a.in[i + 2] : 0.f;
            x[j].w = i + 3 < a.n ? a.in[i + 3] : 0.f;
            if constexpr (RSRC == 1) {
                w[j].x = i + 0 < a.n ? a.bits[i + 0] : 0u;
                w[j].y = i + 1 < a.n ? a.bits[i + 1] : 0u;
                w[j].z = i + 2 < a.n ? a.bits[i + 2] : 0u;
                w[j].w = i + 3 < a.n ? a.bits[i + 3] : 0u;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
        const int64_t vi = v0 + (int64_t)j * TPB;
        // block amax: lane max, then butterfly over the block's lanes (all inside one wave)
        float m = 0.0f;
        m = max_keep(m, fabsf(x[j].x));
        m = max_keep(m, fabsf(x[j].y));
        m = max_keep(m, fabsf(x[j].z));
        m = max_keep(m, fabsf(x[j].w));
        for (int off = 1; off < lanes; off <<= 1) m = max_keep(m, __shfl_xor(m, off, 64));
        // scale = virtual_round(amax / pmn, scale format, scale mode)
        const float t = m / a.pmn;
        const float scale = __uint_as_float(vr_fmt_rt(__float_as_uint(t), a.S, 0u));
        const uint32_t sb = __float_as_uint(scale);
        const uint32_t sexp = (sb >> 23) & 0xFFu;
        // power-of-two scale with an exactly representable reciprocal: x * (1/scale)
        // rounds the same real number as x / scale
        const bool use_inv = (sb & 0x7FFFFFu) == 0u && sexp >= 1u && sexp <= 254u;
        const float inv = use_inv ? 1.0f / scale : 0.0f;
        if (vi >= nslots) continue;
        const bool live = scale > 0.0f;
        uint32_t u0 = 0, u1 = 0, u2 = 0, u3 = 0;
        if constexpr (RSRC == 1) { u0 = w[j].x; u1 = w[j].y; u2 = w[j].z; u3 = w[j].w; }
        if constexpr (RSRC == 2) {
            u0 = rand_u32(a.seed, 4 * vi + 0); u1 = rand_u32(a.seed, 4 * vi + 1);
            u2 = rand_u32(a.seed, 4 * vi + 2); u3 = rand_u32(a.seed, 4 * vi + 3);
        }
        float4 y;
        if (live) {
            y.x = mx_elem<MODE>(x[j].x, scale, inv, use_inv, a.E, u0);
            y.y = mx_elem<MODE>(x[j].y, scale, inv, use_inv, a.E, u1);
            y.z = mx_elem<MODE>(x[j].z, scale, inv, use_inv, a.E, u2);
            y.w = mx_elem<MODE>(x[j].w, scale, inv, use_inv, a.E, u3);
        } else {
            y = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (a.vec_ok) {
            reinterpret_cast<float4*>(a.out)[vi] = y;
        } else {
            const int64_t i = 4 * vi;
            if (i + 0 < a.n) a.out[i + 0] = y.x;
            if (i + 1 < a.n) a.out[i + 1] = y.y;
            if (i + 2 < a.n) a.out[i + 2] = y.z;
            if (i + 3 < a.n) a.out[i + 3] = y.w;
        }
    }
}

}  // namespace lofamd
