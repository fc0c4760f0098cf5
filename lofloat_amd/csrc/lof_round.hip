// lof_round.hip -- HBM-streaming rounding kernels for gfx950 and their C-ABI entry points:
//   lof_virtual_round_f32       virtual_round(x, FloatingPointParams)  per element
//   lof_virtual_round_mant_f32  virtual_round(x, ToMantissaBits)       per element
//   lof_virtual_mx_round_f32    MX block rounding with a shared scale per block
//   lof_fill_random_u32         counter-based random words
//
// Layout: elementwise, one lane owns 4 consecutive floats (one 16-byte global_load_dwordx4
// / store) and a 256-thread workgroup covers 1024 floats, so every wave moves 64 * 16 B per
// access, fully coalesced; one float4 per lane with one workgroup per 1024 floats measured
// fastest on gfx950 HBM.  The arithmetic is 32-bit integer work on the bit patterns
// (lof_core.hpp), which leaves the kernels bound by HBM bandwidth.
//
// MX: one lane owns 8 consecutive floats (4 when block_size is 4) and a block of bs floats
// is bs/8 consecutive lanes; blocks never straddle a wave, so the block amax is a DPP
// butterfly over those lanes and the shared scale lives in registers.
#include <hip/hip_runtime.h>

#include "lof_core.hpp"
#include "lof_host.hpp"
#include "lof_round_kernels.hpp"

namespace lofamd {

namespace {

constexpr int kTPB = 256;

template <int MODE, bool MANT, int RSRC, int VPT>
__global__ __launch_bounds__(kTPB) void k_round(RoundArgs a) {
    round_tile<MODE, MANT, RSRC, VPT>(a, (int64_t)blockIdx.x, threadIdx.x);
}

// MX workgroups of 512 lanes (tiles of 4096 floats): measured 2-3 % above 256 lanes on the
// 8192 x 8192 matrix, in rounds interleaved with 128, 256 and 1024 lanes on one box.
constexpr int kMxTPB = 512;

template <int MODE, int RSRC, int EPL, bool SRNE>
__global__ __launch_bounds__(kMxTPB) void k_mx(MxArgs a) {
    mx_tile<MODE, RSRC, EPL, SRNE, kMxTPB>(a, (int64_t)blockIdx.x, threadIdx.x);
}

__global__ __launch_bounds__(kTPB) void k_fill_random(uint32_t* bits, int64_t n, uint64_t seed,
                                                      uint64_t offset) {
    const int64_t stride = (int64_t)gridDim.x * kTPB;
    for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += stride)
        bits[i] = rand_u32(seed, offset + (uint64_t)i);
}

constexpr int kRoundVPT = 1;  // one float4 per lane: measured best on gfx950 HBM

template <bool MANT, int RSRC>
lof_status launch_round_mode(const RoundArgs& a, int mode, hipStream_t st) {
    const int64_t per_block = (int64_t)kTPB * kRoundVPT * 4;
    const int64_t blocks = (a.n + per_block - 1) / per_block;
    if (blocks > 0x7FFFFFFF) return set_error(LOF_ERR_UNSUPPORTED, "tensor too large");
    dim3 grid((unsigned)blocks), block(kTPB);
#define LOF_CASE(M)                                                                     \
    case M:                                                                             \
        hipLaunchKernelGGL((k_round<M, MANT, RSRC, kRoundVPT>), grid, block, 0, st, a); \
        break;
    if constexpr (RSRC == 0) {
        switch (mode) {
            LOF_CASE(RNE) LOF_CASE(RTZ) LOF_CASE(RAFZ) LOF_CASE(RNO) LOF_CASE(RDOWN)
            LOF_CASE(RUP) LOF_CASE(RTA) LOF_CASE(RTO) LOF_CASE(RTTZ)
            // stochastic modes that never read a random word (stoch_len 0 for A/C/D;
            // B reads R = u & 0 = 0) are routed here by the caller
            LOF_CASE(SRA) LOF_CASE(SRB) LOF_CASE(SRC) LOF_CASE(SRD)
            default: return set_error(LOF_ERR_INVALID_ARGUMENT, "bad mode %d", mode);
        }
    } else {
        switch (mode) {
            LOF_CASE(SRA) LOF_CASE(SRB) LOF_CASE(SRC) LOF_CASE(TRUESR) LOF_CASE(PROB)
            LOF_CASE(SRD)
            default: return set_error(LOF_ERR_INVALID_ARGUMENT, "bad stochastic mode %d", mode);
        }
    }
#undef LOF_CASE
    return check_hip(hipGetLastError(), "virtual_round launch");
}

template <int RSRC, int EPL, bool SRNE>
lof_status launch_mx_mode(const MxArgs& a, int mode, hipStream_t st) {
    constexpr int64_t tile = (int64_t)kMxTPB * 8;
    const int64_t blocks = (a.n + tile - 1) / tile;
    if (blocks > 0x7FFFFFFF) return set_error(LOF_ERR_UNSUPPORTED, "tensor too large");
    dim3 grid((unsigned)blocks), block(kMxTPB);
#define LOF_CASE(M)                                                                 \
    case M:                                                                         \
        hipLaunchKernelGGL((k_mx<M, RSRC, EPL, SRNE>), grid, block, 0, st, a);      \
        break;
    if constexpr (RSRC == 0) {
        switch (mode) {
            LOF_CASE(RNE) LOF_CASE(RTZ) LOF_CASE(RAFZ) LOF_CASE(RNO) LOF_CASE(RDOWN)
            LOF_CASE(RUP) LOF_CASE(RTA) LOF_CASE(RTO) LOF_CASE(RTTZ)
            LOF_CASE(SRA) LOF_CASE(SRB) LOF_CASE(SRC) LOF_CASE(SRD)
            default: return set_error(LOF_ERR_INVALID_ARGUMENT, "bad mode %d", mode);
        }
    } else {
        switch (mode) {
            LOF_CASE(SRA) LOF_CASE(SRB) LOF_CASE(SRC) LOF_CASE(TRUESR) LOF_CASE(PROB)
            LOF_CASE(SRD)
            default: return set_error(LOF_ERR_INVALID_ARGUMENT, "bad stochastic mode %d", mode);
        }
    }
#undef LOF_CASE
    return check_hip(hipGetLastError(), "virtual_mx_round launch");
}

}  // namespace

// Stochastic modes that actually consume a random word for this spec.
bool needs_random(lof_round_spec s) {
    switch (s.mode) {
    case SRA: case SRC: case SRD: return s.stoch_len > 0;
    case SRB: return s.stoch_len > 0;
    case TRUESR: case PROB: return true;
    default: return false;
    }
}

lof_status launch_round(const RoundArgs& a, bool mant, int mode, int rsrc, hipStream_t st) {
    if (mant) {
        if (rsrc == 0) return launch_round_mode<true, 0>(a, mode, st);
        if (rsrc == 1) return launch_round_mode<true, 1>(a, mode, st);
        return launch_round_mode<true, 2>(a, mode, st);
    }
    if (rsrc == 0) return launch_round_mode<false, 0>(a, mode, st);
    if (rsrc == 1) return launch_round_mode<false, 1>(a, mode, st);
    return launch_round_mode<false, 2>(a, mode, st);
}

template <int EPL, bool SRNE>
lof_status launch_mx_epl(const MxArgs& a, int mode, int rsrc, hipStream_t st) {
    if (rsrc == 0) return launch_mx_mode<0, EPL, SRNE>(a, mode, st);
    if (rsrc == 1) return launch_mx_mode<1, EPL, SRNE>(a, mode, st);
    return launch_mx_mode<2, EPL, SRNE>(a, mode, st);
}

lof_status launch_mx(const MxArgs& a, int mode, int rsrc, hipStream_t st) {
    const bool srne = a.S.mode == RNE;
    if (a.lanes_log2 == 0) {  // block_size 4: one float4 per lane
        return srne ? launch_mx_epl<4, true>(a, mode, rsrc, st)
                    : launch_mx_epl<4, false>(a, mode, rsrc, st);
    }
    return srne ? launch_mx_epl<8, true>(a, mode, rsrc, st)
                : launch_mx_epl<8, false>(a, mode, rsrc, st);
}

lof_status prepare_round(const float* in, float* out, int64_t n, const lof_format* fmt,
                         int mant_bits, bool mant, lof_round_spec spec, float scale,
                         const uint32_t* rand_bits, uint64_t seed, RoundArgs* a, int* rsrc) {
    const char* what = mant ? "lof_virtual_round_mant_f32" : "lof_virtual_round_f32";
    if (n < 0) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: n < 0", what);
    if (n > 0 && (!in || !out)) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: NULL data", what);
    int dropped;
    if (mant) {
        if (mant_bits < 0 || mant_bits > 23)
            return set_error(LOF_ERR_UNSUPPORTED, "%s: mant_bits %d outside [0, 23]", what,
                             mant_bits);
        dropped = 23 - mant_bits;
    } else {
        lof_status s = validate_format(fmt, what);
        if (s != LOF_OK) return s;
        dropped = 23 - fmt->mantissa_bits;
    }
    lof_status s = validate_spec(spec, dropped, what);
    if (s != LOF_OK) return s;
    *a = RoundArgs{};
    a->in = in;
    a->out = out;
    a->n = n;
    if (!mant) a->P = make_fmt_dev(*fmt, spec);
    a->r = dropped;
    a->len = spec.stoch_len;
    a->scale = scale;
    a->has_scale = scale != 1.0f;
    a->bits = rand_bits;
    a->seed = seed;
    a->vec_ok = aligned16(in) && aligned16(out) && (!rand_bits || aligned16(rand_bits));
    *rsrc = needs_random(spec) ? (rand_bits ? 1 : 2) : 0;
    return LOF_OK;
}

lof_status prepare_mx(const float* in, float* out, int64_t n, int32_t bs,
                      const lof_format* elem_fmt, const lof_format* scale_fmt,
                      lof_round_spec elem_spec, int32_t scale_mode, const uint32_t* rand_bits,
                      uint64_t seed, MxArgs* a, int* rsrc) {
    const char* what = "lof_virtual_mx_round_f32";
    if (n < 0) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: n < 0", what);
    if (n > 0 && (!in || !out)) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: NULL data", what);
    if (bs < 4 || bs > 256 || (bs & (bs - 1)) != 0)
        return set_error(LOF_ERR_UNSUPPORTED,
                         "%s: unsupported block_size %d (use a power of two in [4, 256])", what,
                         bs);
    lof_status s = validate_format(elem_fmt, what);
    if (s != LOF_OK) return s;
    s = validate_format(scale_fmt, what);
    if (s != LOF_OK) return s;
    s = validate_spec(elem_spec, 23 - elem_fmt->mantissa_bits, what);
    if (s != LOF_OK) return s;
    if (!mode_valid(scale_mode))
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad scale mode %d", what, scale_mode);
    if (scale_mode == TRUESR || scale_mode == PROB)
        return set_error(LOF_ERR_UNSUPPORTED,
                         "%s: scale_round_mode %d needs random bits for the scale; use a "
                         "deterministic mode or SR-A/B/C/D (run with stoch_len 0)",
                         what, scale_mode);
    *a = MxArgs{};
    a->in = in;
    a->out = out;
    a->n = n;
    int l2 = 0;
    while ((4 << l2) < bs) ++l2;
    a->lanes_log2 = l2;
    a->E = make_fmt_dev(*elem_fmt, elem_spec);
    lof_round_spec ss{scale_mode, 0, 0};
    a->S = make_fmt_dev(*scale_fmt, ss);
    a->pmn = mx_private_max_normal(*elem_fmt);
    a->bits = rand_bits;
    a->seed = seed;
    a->vec_ok = aligned16(in) && aligned16(out) && (!rand_bits || aligned16(rand_bits)) &&
                (n % 4 == 0);
    *rsrc = needs_random(elem_spec) ? (rand_bits ? 1 : 2) : 0;
    return LOF_OK;
}

}  // namespace lofamd

using namespace lofamd;

extern "C" {

LOF_API uint32_t lof_rand_u32(uint64_t seed, uint64_t index) { return rand_u32(seed, index); }

LOF_API lof_status lof_fill_random_u32(uint32_t* bits, int64_t n, uint64_t seed,
                                       uint64_t offset, lof_stream_t stream) {
    if (n < 0 || (n > 0 && !bits))
        return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_fill_random_u32: bad arguments");
    if (n == 0) return LOF_OK;
    int64_t blocks = (n + kTPB - 1) / kTPB;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_random, dim3((unsigned)blocks), dim3(kTPB), 0,
                       (hipStream_t)stream, bits, n, seed, offset);
    return check_hip(hipGetLastError(), "lof_fill_random_u32 launch");
}

LOF_API lof_status lof_virtual_round_f32(const float* in, float* out, int64_t n,
                                         const lof_format* fmt, lof_round_spec spec, float scale,
                                         const uint32_t* rand_bits, uint64_t seed,
                                         lof_stream_t stream) {
    RoundArgs a;
    int rsrc;
    lof_status s = prepare_round(in, out, n, fmt, 0, false, spec, scale, rand_bits, seed, &a,
                                 &rsrc);
    if (s != LOF_OK || n == 0) return s;
    return launch_round(a, false, spec.mode, rsrc, (hipStream_t)stream);
}

LOF_API lof_status lof_virtual_round_mant_f32(const float* in, float* out, int64_t n,
                                              int32_t mant_bits, lof_round_spec spec, float scale,
                                              const uint32_t* rand_bits, uint64_t seed,
                                              lof_stream_t stream) {
    RoundArgs a;
    int rsrc;
    lof_status s = prepare_round(in, out, n, nullptr, mant_bits, true, spec, scale, rand_bits,
                                 seed, &a, &rsrc);
    if (s != LOF_OK || n == 0) return s;
    return launch_round(a, true, spec.mode, rsrc, (hipStream_t)stream);
}

LOF_API lof_status lof_virtual_mx_round_f32(const float* in, float* out, int64_t n,
                                            int32_t block_size, const lof_format* elem_fmt,
                                            const lof_format* scale_fmt, lof_round_spec elem_spec,
                                            int32_t scale_mode, const uint32_t* rand_bits,
                                            uint64_t seed, lof_stream_t stream) {
    MxArgs a;
    int rsrc;
    lof_status s = prepare_mx(in, out, n, block_size, elem_fmt, scale_fmt, elem_spec, scale_mode,
                              rand_bits, seed, &a, &rsrc);
    if (s != LOF_OK || n == 0) return s;
    return launch_mx(a, elem_spec.mode, rsrc, (hipStream_t)stream);
}

}  // extern "C"
