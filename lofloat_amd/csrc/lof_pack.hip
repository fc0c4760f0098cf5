// lof_pack.hip -- n-bit pack codec kernels and C-ABI (lof_packed_size, lof_pack_encode_f32,
// lof_pack_decode_f32, lof_pack_reps, lof_unpack_reps).
//
// The packed buffer is a stream of n-bit slots (see lof_pack.hpp).  Each thread owns one
// 32-bit word of that stream:
//   encode: it converts the elements whose slots overlap its word (at most 32/n + 1, an
//           element straddling two words is converted by both) and stores the word once;
//   decode: it reads its word and the next one and produces the elements whose slots start
//           in its word.
// Words are accessed as 32-bit loads/stores when the buffer is 4-byte aligned and the word
// lies inside it, else byte by byte (the buffer's byte length is groups * group_bytes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "lof_host.hpp"
#include "lof_pack.hpp"

namespace lofamd {

namespace {

constexpr int kPackTPB = 256;

struct PackArgs {
    const float* src;        // encode input / decode output (as dstf)
    float* dstf;
    const uint32_t* reps;    // raw reps in (pack_reps) / out (unpack_reps: repsout)
    uint32_t* repsout;
    uint8_t* dst;            // packed out (encode)
    const uint8_t* psrc;     // packed in (decode)
    int64_t count;           // elements
    int64_t slots;           // groups * gc
    int64_t bytes;           // groups * group_bytes
    int64_t words;           // ceil(bytes / 4)
    const uint32_t* bits;    // random words (stochastic encode), or NULL
    uint64_t seed;
    int32_t endian;
    int32_t aligned;         // the packed buffer is 4-byte aligned
    PackDev P;
};

template <int RSRC>
__device__ __forceinline__ uint32_t pack_word_for(const PackArgs& a, int64_t e) {
    if constexpr (RSRC == 1) return a.bits[e];
    else if constexpr (RSRC == 2) return rand_u32(a.seed, (uint64_t)e);
    else return 0u;
}

// RAW: pack reps[] as given (masked to n bits); else quantise src[] with float_to_rep.
template <int MODE, int RSRC, bool RAW>
__global__ __launch_bounds__(kPackTPB) void k_pack_encode(PackArgs a) {
    const int64_t w = (int64_t)blockIdx.x * kPackTPB + threadIdx.x;
    if (w >= a.words) return;
    const int n = a.P.n;
    const int64_t b0 = w * 32;
    const int64_t s_end = min((b0 + 31) / n, a.slots - 1);
    uint32_t word = 0u;
    for (int64_t S = b0 / n; S <= s_end; ++S) {
        const int64_t e = slot_element(S, a.P.gc, a.endian);
        uint32_t rep = 0u;   // a partial final group is padded with encodings of +0.0f: 0
        if (e < a.count) {
            if constexpr (RAW) rep = a.reps[e];
            else rep = float_to_rep<MODE>(a.src[e], a.P, pack_word_for<RSRC>(a, e));
        }
        rep &= a.P.mask;
        const int64_t p = S * n - b0;
        word |= p >= 0 ? (rep << p) : (rep >> (-p));
    }
    const int64_t byte0 = w * 4;
    if (a.aligned && byte0 + 4 <= a.bytes) {
        reinterpret_cast<uint32_t*>(a.dst)[w] = word;
    } else {
        for (int k = 0; k < 4; ++k)
            if (byte0 + k < a.bytes) a.dst[byte0 + k] = (uint8_t)(word >> (8 * k));
    }
}

__device__ __forceinline__ uint32_t load_word(const PackArgs& a, int64_t w) {
    const int64_t byte0 = w * 4;
    if (a.aligned && byte0 + 4 <= a.bytes) return reinterpret_cast<const uint32_t*>(a.psrc)[w];
    uint32_t v = 0u;
    for (int k = 0; k < 4; ++k)
        if (byte0 + k < a.bytes) v |= (uint32_t)a.psrc[byte0 + k] << (8 * k);
    return v;
}

template <bool RAW>
__global__ __launch_bounds__(kPackTPB) void k_pack_decode(PackArgs a) {
    const int64_t w = (int64_t)blockIdx.x * kPackTPB + threadIdx.x;
    if (w >= a.words) return;
    const int n = a.P.n;
    const int64_t b0 = w * 32;
    const int64_t s0 = (b0 + n - 1) / n;
    const int64_t s1 = min((b0 + 31) / n, a.slots - 1);
    if (s0 > s1) return;
    uint64_t win = load_word(a, w);
    if (s1 * n + n > b0 + 32) win |= (uint64_t)load_word(a, w + 1) << 32;   // straddles
    for (int64_t S = s0; S <= s1; ++S) {
        const int64_t e = slot_element(S, a.P.gc, a.endian);
        if (e >= a.count) continue;
        const uint32_t rep = (uint32_t)(win >> (S * n - b0)) & a.P.mask;
        if constexpr (RAW) a.repsout[e] = rep;
        else a.dstf[e] = rep_to_float(rep, a.P);
    }
}

int gcd_i(int a, int b) { return b == 0 ? a : gcd_i(b, a % b); }

void geometry(int n, int* gc, int* gbits) {
    const int al = n <= 8 ? 8 : 16;
    *gbits = n / gcd_i(n, al) * al;
    *gc = *gbits / n;
}

lof_status validate_checker(const lof_checker& c, const char* what) {
    if (c.kind < 0 || c.kind > 3)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: checker kind %d not in [0, 3]", what, c.kind);
    if (c.ncodes < 0 || c.ncodes > 3 || (c.kind == 1 && c.ncodes == 0) || (c.kind == 3 && c.ncodes < 1))
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad checker code count %d", what, c.ncodes);
    return LOF_OK;
}

CheckDev to_dev(const lof_checker& c) {
    CheckDev d{};
    d.kind = c.kind;
    d.ncodes = c.ncodes;
    for (int i = 0; i < 3; ++i) d.codes[i] = c.codes[i];
    d.a = c.a;
    d.b = c.b;
    return d;
}

lof_status make_pack_dev(const lof_pack_format* pf, lof_round_spec spec, PackDev* P,
                         const char* what) {
    if (!pf) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: format is NULL", what);
    const lof_format& f = pf->fmt;
    const int n = f.bitwidth, mb = f.mantissa_bits;
    if (n < 1 || n > 16)
        return set_error(LOF_ERR_UNSUPPORTED, "%s: bitwidth %d outside [1, 16]", what, n);
    const bool is_signed = f.signedness == LOF_SIGNED;
    const int exp_bits = is_signed ? n - mb - 1 : n - mb;
    if (mb < 0 || exp_bits < 0 || exp_bits > 30)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad mantissa width %d for %d bits", what,
                         mb, n);
    if (f.bias < -100 || f.bias > 100000)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bias %d out of range", what, f.bias);
    if (pf->unsigned_behavior != 0 && pf->unsigned_behavior != 1)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad unsigned_behavior", what);
    lof_status s = validate_checker(pf->is_inf, what);
    if (s != LOF_OK) return s;
    if ((s = validate_checker(pf->is_nan, what)) != LOF_OK) return s;
    if (!mode_valid(spec.mode))
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad rounding mode %d", what, spec.mode);
    if (spec.saturation < 0 || spec.saturation > 2)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad saturation mode %d", what,
                         spec.saturation);
    if (uses_len(spec.mode) && (spec.stoch_len < 0 || spec.stoch_len > 23 - mb))
        return set_error(LOF_ERR_INVALID_ARGUMENT,
                         "%s: stoch_len %d outside [0, %d] (dropped mantissa bits)", what,
                         spec.stoch_len, 23 - mb);
    *P = PackDev{};
    P->n = n;
    P->mb = mb;
    P->bias = f.bias;
    P->is_signed = is_signed ? 1 : 0;
    P->r = 23 - mb;
    P->exp_off = f.bias - 127;
    const int to_min_exp = 1 - f.bias;
    P->branch = to_min_exp < -125 ? 1 : (to_min_exp > -125 ? 2 : 0);
    P->has_inf = f.inf_behavior != LOF_INF_SATURATING ? 1 : 0;
    const long long to_max_exp = (1LL << exp_bits) - 1 - f.bias;
    P->narrower = (to_max_exp < 128 || (to_max_exp == 128 && mb + 1 < 24)) ? 1 : 0;
    P->unsigned_nan = pf->unsigned_behavior == 0 ? 0 : (f.nan_behavior == LOF_NAN_3109 ? 1 : 2);
    P->saturation = spec.saturation;
    P->stoch_len = spec.stoch_len;
    P->mask = n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1u);
    P->signmask = is_signed ? (1u << (n - 1)) : 0u;
    P->max_finite = P->has_inf ? pf->is_inf.pat0 - 1u
                               : (is_signed ? (((1u << n) - 1u) >> 1) : ((1u << n) - 1u));
    P->inf_rep = P->has_inf ? pf->is_inf.pat0 : P->max_finite;
    P->qnan = pf->is_nan.pat0;
    P->snan = pf->is_nan.pat1;
    P->is_inf = to_dev(pf->is_inf);
    P->is_nan = to_dev(pf->is_nan);
    P->nan_signmask = check_match(P->is_nan, P->signmask, true) ? 1 : 0;
    P->aligned_highest = (uint64_t)P->max_finite << P->r;
    geometry(n, &P->gc, &P->gbits);
    return LOF_OK;
}

void size_args(PackArgs* a, int64_t count) {
    const int64_t groups = (count + a->P.gc - 1) / a->P.gc;
    a->count = count;
    a->slots = groups * a->P.gc;
    a->bytes = groups * (a->P.gbits / 8);
    a->words = (a->bytes + 3) / 4;
}

dim3 grid_for(int64_t words) { return dim3((unsigned)((words + kPackTPB - 1) / kPackTPB)); }

template <int MODE, int RSRC>
void launch_enc(const PackArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((k_pack_encode<MODE, RSRC, false>), grid_for(a.words), dim3(kPackTPB), 0, st, a);
}

lof_status launch_encode(const PackArgs& a, int mode, int rsrc, hipStream_t st) {
    if (a.words == 0) return LOF_OK;
    if (a.words / kPackTPB >= (int64_t)1 << 31)
        return set_error(LOF_ERR_UNSUPPORTED, "lof_pack_encode_f32: input too large");
#define LOF_PE(M, R) if (mode == M && rsrc == R) { launch_enc<M, R>(a, st); goto done; }
    LOF_PE(RNE, 0) LOF_PE(RTZ, 0) LOF_PE(RAFZ, 0) LOF_PE(RNO, 0) LOF_PE(RDOWN, 0)
    LOF_PE(RUP, 0) LOF_PE(RTA, 0) LOF_PE(RTO, 0) LOF_PE(RTTZ, 0)
    LOF_PE(SRA, 0) LOF_PE(SRB, 0) LOF_PE(SRC, 0) LOF_PE(SRD, 0)
    LOF_PE(SRA, 1) LOF_PE(SRB, 1) LOF_PE(SRC, 1) LOF_PE(TRUESR, 1) LOF_PE(PROB, 1) LOF_PE(SRD, 1)
    LOF_PE(SRA, 2) LOF_PE(SRB, 2) LOF_PE(SRC, 2) LOF_PE(TRUESR, 2) LOF_PE(PROB, 2) LOF_PE(SRD, 2)
#undef LOF_PE
    return set_error(LOF_ERR_UNSUPPORTED, "lof_pack_encode_f32: mode %d not supported", mode);
done:
    return check_hip(hipGetLastError(), "lof_pack_encode_f32 launch");
}

}  // namespace

}  // namespace lofamd

using namespace lofamd;

extern "C" {

LOF_API int64_t lof_packed_size(int32_t bitwidth, int64_t count) {
    if (bitwidth < 1 || bitwidth > 16 || count < 0) return -1;
    int gc, gbits;
    geometry(bitwidth, &gc, &gbits);
    return (count + gc - 1) / gc * (gbits / 8);
}

LOF_API lof_status lof_pack_encode_f32(uint8_t* dst, const float* src, int64_t count,
                                       const lof_pack_format* fmt, lof_round_spec spec,
                                       int32_t endian, const uint32_t* rand_bits, uint64_t seed,
                                       lof_stream_t stream) {
    const char* what = "lof_pack_encode_f32";
    PackArgs a{};
    lof_status s = make_pack_dev(fmt, spec, &a.P, what);
    if (s != LOF_OK) return s;
    if (count < 0) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: count < 0", what);
    if (endian != 0 && endian != 1)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: endian must be 0 or 1", what);
    if (count > 0 && (!dst || !src)) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: NULL buffer", what);
    // Deterministic modes ignore random words; TrueSR and Prob have no length
    int rsrc = 0;
    if (is_stochastic(spec.mode)) rsrc = rand_bits ? 1 : 2;
    if (uses_len(spec.mode) && spec.stoch_len == 0) rsrc = 0;   // degenerate: no draw used
    size_args(&a, count);
    a.src = src;
    a.dst = dst;
    a.bits = rand_bits;
    a.seed = seed;
    a.endian = endian;
    a.aligned = ((uintptr_t)dst & 3u) == 0 ? 1 : 0;
    return launch_encode(a, spec.mode, rsrc, (hipStream_t)stream);
}

LOF_API lof_status lof_pack_decode_f32(float* dst, const uint8_t* src, int64_t count,
                                       const lof_pack_format* fmt, int32_t endian,
                                       lof_stream_t stream) {
    const char* what = "lof_pack_decode_f32";
    PackArgs a{};
    lof_status s = make_pack_dev(fmt, lof_round_spec{0, 0, 0}, &a.P, what);
    if (s != LOF_OK) return s;
    if (count < 0) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: count < 0", what);
    if (endian != 0 && endian != 1)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: endian must be 0 or 1", what);
    if (count > 0 && (!dst || !src)) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: NULL buffer", what);
    size_args(&a, count);
    if (a.words == 0) return LOF_OK;
    a.dstf = dst;
    a.psrc = src;
    a.endian = endian;
    a.aligned = ((uintptr_t)src & 3u) == 0 ? 1 : 0;
    hipLaunchKernelGGL((k_pack_decode<false>), grid_for(a.words), dim3(kPackTPB), 0,
                       (hipStream_t)stream, a);
    return check_hip(hipGetLastError(), "lof_pack_decode_f32 launch");
}

LOF_API lof_status lof_pack_reps(uint8_t* dst, const uint32_t* reps, int64_t count,
                                 int32_t bitwidth, int32_t endian, lof_stream_t stream) {
    const char* what = "lof_pack_reps";
    if (bitwidth < 1 || bitwidth > 16)
        return set_error(LOF_ERR_UNSUPPORTED, "%s: bitwidth %d outside [1, 16]", what, bitwidth);
    if (count < 0 || (endian != 0 && endian != 1) || (count > 0 && (!dst || !reps)))
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad arguments", what);
    PackArgs a{};
    a.P.n = bitwidth;
    a.P.mask = (1u << bitwidth) - 1u;
    geometry(bitwidth, &a.P.gc, &a.P.gbits);
    size_args(&a, count);
    if (a.words == 0) return LOF_OK;
    a.reps = reps;
    a.dst = dst;
    a.endian = endian;
    a.aligned = ((uintptr_t)dst & 3u) == 0 ? 1 : 0;
    hipLaunchKernelGGL((k_pack_encode<RNE, 0, true>), grid_for(a.words), dim3(kPackTPB), 0,
                       (hipStream_t)stream, a);
    return check_hip(hipGetLastError(), "lof_pack_reps launch");
}

LOF_API lof_status lof_unpack_reps(uint32_t* reps, const uint8_t* src, int64_t count,
                                   int32_t bitwidth, int32_t endian, lof_stream_t stream) {
    const char* what = "lof_unpack_reps";
    if (bitwidth < 1 || bitwidth > 16)
        return set_error(LOF_ERR_UNSUPPORTED, "%s: bitwidth %d outside [1, 16]", what, bitwidth);
    if (count < 0 || (endian != 0 && endian != 1) || (count > 0 && (!reps || !src)))
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad arguments", what);
    PackArgs a{};
    a.P.n = bitwidth;
    a.P.mask = (1u << bitwidth) - 1u;
    geometry(bitwidth, &a.P.gc, &a.P.gbits);
    size_args(&a, count);
    if (a.words == 0) return LOF_OK;
    a.repsout = reps;
    a.psrc = src;
    a.endian = endian;
    a.aligned = ((uintptr_t)src & 3u) == 0 ? 1 : 0;
    hipLaunchKernelGGL((k_pack_decode<true>), grid_for(a.words), dim3(kPackTPB), 0,
                       (hipStream_t)stream, a);
    return check_hip(hipGetLastError(), "lof_unpack_reps launch");
}

}  // extern "C"
