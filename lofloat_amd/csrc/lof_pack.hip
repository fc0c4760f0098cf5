// lof_pack.hip -- n-bit pack codec kernels and C-ABI (lof_packed_size, lof_pack_encode_f32,
// lof_pack_decode_f32, lof_pack_reps, lof_unpack_reps).
//
// The packed buffer is a stream of n-bit slots (see lof_pack.hpp).  The group count gc is a
// power of two (1..16) for every n in [1, 16], so slot S holds element S ^ (gc - 1) in
// Reverse order, and a tile of kTile = 8192 elements is a whole number of groups whose
// slots fill exactly 256 * n 32-bit words.  One 256-thread workgroup owns one tile:
//   encode: phase 1 reads the tile's inputs as float4 (coalesced), converts each element to
//           its rep and writes the reps into LDS (16-bit, in slot order); phase 2 packs 32
//           slots per thread into n words, puts the tile's 256 * n words back into LDS and
//           stores them coalesced;
//   decode: phase 1 loads the tile's words into LDS (coalesced); phase 2 extracts four
//           fields per thread per step from a 64-bit LDS window, widens them and stores a
//           float4.
// Words are accessed as 32-bit loads/stores when the buffer is 4-byte aligned and the word
// lies inside it, else byte by byte (the buffer's byte length is groups * group_bytes); a
// partial tile falls back to guarded scalar element accesses.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "lof_host.hpp"
#include "lof_pack.hpp"

namespace lofamd {

namespace {

constexpr int kPackTPB = 256;
constexpr int kTile = kPackTPB * 32;   // elements (slots) per workgroup

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
    int64_t tiles;           // ceil(count / kTile)
    const uint32_t* bits;    // random words (stochastic encode), or NULL
    uint64_t seed;
    int32_t xm;              // slot of element e within the tile: e ^ xm (Reverse: gc - 1)
    int32_t aligned;         // the packed buffer is 4-byte aligned
    int32_t vec;             // the element-side buffers (src / dstf / reps / bits) are 16-byte aligned
    int32_t rep_bytes;       // typed cast: bytes per rep container (1 for n <= 8, else 2)
    PackDev P;
};

template <int RSRC>
__device__ __forceinline__ uint32_t pack_word_for(const PackArgs& a, int64_t e) {
    if constexpr (RSRC == 1) return a.bits[e];
    else if constexpr (RSRC == 2) return rand_u32(a.seed, (uint64_t)e);
    else return 0u;
}

// The rep of one element: the branch-free fast path, and float_to_rep for the lanes it
// does not cover (a wave skips that call when none of its lanes needs it).
template <int MODE, int RSRC, bool RAW>
__device__ __forceinline__ uint32_t elem_rep(const PackArgs& a, uint32_t xbits, uint32_t u) {
    if constexpr (RAW) {
        return xbits & a.P.mask;
    } else {
        bool ok;
        uint32_t rep = float_to_rep_fast<MODE>(__uint_as_float(xbits), a.P, u, ok);
        if (!ok) rep = float_to_rep<MODE>(__uint_as_float(xbits), a.P, u);
        return rep & a.P.mask;
    }
}

// RAW: pack reps[] as given (masked to n bits); else quantise src[] with float_to_rep.
template <int MODE, int RSRC, bool RAW>
__global__ __launch_bounds__(kPackTPB) void k_pack_encode(PackArgs a) {
    __shared__ __align__(16) uint16_t slot_rep[kTile];
    const int tid = threadIdx.x;
    const int64_t tbase = (int64_t)blockIdx.x * kTile;
    const uint32_t* xin = RAW ? a.reps : reinterpret_cast<const uint32_t*>(a.src);
    const bool full = a.vec && tbase + kTile <= a.count;
    const int xq = a.xm & ~3, xi = a.xm & 3;
    // slots of elements 4q+i are (4q ^ xq) + (i ^ xi): one aligned quad, permuted
    auto put = [&](int q, const uint32_t (&r)[4]) {
        uint2 packed;
        packed.x = r[0 ^ xi] | (r[1 ^ xi] << 16);
        packed.y = r[2 ^ xi] | (r[3 ^ xi] << 16);
        *reinterpret_cast<uint2*>(&slot_rep[(4 * q) ^ xq]) = packed;
    };
    if (full) {
        // all eight loads are issued before the first conversion waits on one
        uint4 x[8], u[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            x[k] = reinterpret_cast<const uint4*>(xin + tbase)[k * kPackTPB + tid];
            if constexpr (RSRC == 1) u[k] = reinterpret_cast<const uint4*>(a.bits + tbase)[k * kPackTPB + tid];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int q = k * kPackTPB + tid;       // quad of elements 4q .. 4q+3 in the tile
            if constexpr (RSRC == 2) {
                const uint64_t e0 = (uint64_t)(tbase + 4 * q);
                u[k] = make_uint4(rand_u32(a.seed, e0), rand_u32(a.seed, e0 + 1),
                                  rand_u32(a.seed, e0 + 2), rand_u32(a.seed, e0 + 3));
            } else if constexpr (RSRC == 0) {
                u[k] = make_uint4(0u, 0u, 0u, 0u);
            }
            const uint32_t r[4] = {elem_rep<MODE, RSRC, RAW>(a, x[k].x, u[k].x),
                                   elem_rep<MODE, RSRC, RAW>(a, x[k].y, u[k].y),
                                   elem_rep<MODE, RSRC, RAW>(a, x[k].z, u[k].z),
                                   elem_rep<MODE, RSRC, RAW>(a, x[k].w, u[k].w)};
            put(q, r);
        }
    } else {
        // a partial final group is padded with encodings of +0.0f, i.e. rep 0
        for (int k = 0; k < 8; ++k) {
            const int q = k * kPackTPB + tid;
            const int64_t e0 = tbase + 4 * q;
            uint32_t r[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                r[i] = e0 + i < a.count
                           ? elem_rep<MODE, RSRC, RAW>(a, xin[e0 + i], pack_word_for<RSRC>(a, e0 + i))
                           : 0u;
            put(q, r);
        }
    }
    __syncthreads();
    // Thread t packs slots 32t .. 32t+31 (whole groups: gc divides 32), i.e. words
    // t*n .. t*n+n-1 of the tile.  The bit cursor is the same in every lane, so its branches
    // are scalar.  The words go back into the same LDS (after a barrier), then out coalesced.
    const int n = a.P.n;
    uint4 v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = reinterpret_cast<const uint4*>(slot_rep)[tid * 4 + i];
    __syncthreads();
    uint32_t* wbuf = reinterpret_cast<uint32_t*>(slot_rep);   // 4096 words >= 256 * n
    uint32_t cur = 0u;
    int filled = 0, wo = tid * n;
#pragma unroll
    for (int sl = 0; sl < 32; ++sl) {
        const uint4& q4 = v[sl >> 3];
        const uint32_t d = ((sl >> 1) & 3) == 0 ? q4.x : ((sl >> 1) & 3) == 1 ? q4.y
                         : ((sl >> 1) & 3) == 2 ? q4.z : q4.w;
        const uint32_t rep = (sl & 1) ? (d >> 16) : (d & 0xFFFFu);
        cur |= rep << filled;
        filled += n;
        if (filled >= 32) {
            wbuf[wo++] = cur;
            filled -= 32;
            cur = rep >> (n - filled);   // the bits of rep past the word's end (0 if none)
        }
    }
    __syncthreads();
    const int64_t wbase = (int64_t)blockIdx.x * (kPackTPB * n);
    for (int j = 0; j < n; ++j) {
        const int wl = j * kPackTPB + tid;          // word of the tile
        const int64_t byte0 = (wbase + wl) * 4;
        if (byte0 >= a.bytes) break;
        const uint32_t word = wbuf[wl];
        if (a.aligned && byte0 + 4 <= a.bytes) {
            reinterpret_cast<uint32_t*>(a.dst)[wbase + wl] = word;
        } else {
            for (int k = 0; k < 4; ++k)
                if (byte0 + k < a.bytes) a.dst[byte0 + k] = (uint8_t)(word >> (8 * k));
        }
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
__device__ __forceinline__ uint32_t widen(const PackArgs& a, uint32_t rep) {
    if constexpr (RAW) return rep;
    else return __float_as_uint(rep_to_float(rep, a.P));
}

template <bool RAW>
__global__ __launch_bounds__(kPackTPB) void k_pack_decode(PackArgs a) {
    __shared__ uint32_t words[kPackTPB * 16 + 1];
    const int tid = threadIdx.x;
    const int n = a.P.n;
    const int64_t wbase = (int64_t)blockIdx.x * (kPackTPB * n);
    if (a.aligned && (wbase + kPackTPB * n) * 4 <= a.bytes) {
        // whole tile inside the buffer: every load issued before the first LDS write
        uint32_t w[16];
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (j < n) w[j] = reinterpret_cast<const uint32_t*>(a.psrc)[wbase + j * kPackTPB + tid];
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (j < n) words[j * kPackTPB + tid] = w[j];
    } else {
        for (int j = 0; j < n; ++j) {
            const int wl = j * kPackTPB + tid;
            words[wl] = (wbase + wl) * 4 < a.bytes ? load_word(a, wbase + wl) : 0u;
        }
    }
    if (tid == 0) words[kPackTPB * n] = 0u;
    __syncthreads();
    const int64_t tbase = (int64_t)blockIdx.x * kTile;
    uint32_t* out = RAW ? a.repsout : reinterpret_cast<uint32_t*>(a.dstf);
    const bool full = a.vec && tbase + kTile <= a.count;
    const uint32_t mask = a.P.mask;
#pragma unroll 2
    for (int k = 0; k < 8; ++k) {
        const int q = k * kPackTPB + tid;
        uint32_t v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int bit = ((4 * q + i) ^ a.xm) * n;
            const uint64_t win = (uint64_t)words[bit >> 5] | ((uint64_t)words[(bit >> 5) + 1] << 32);
            v[i] = widen<RAW>(a, (uint32_t)(win >> (bit & 31)) & mask);
        }
        const int64_t e0 = tbase + 4 * q;
        if (full) {
            reinterpret_cast<uint4*>(out + tbase)[q] = make_uint4(v[0], v[1], v[2], v[3]);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (e0 + i < a.count) out[e0 + i] = v[i];
        }
    }
}

// ---------------------------------------------------------------------------------------
// Typed cast: one rep per 1- or 2-byte container (the storage of a Templated_Float<Fp>
// array), no bit packing.  A block covers kCastTile elements; thread t handles the quads
// k * 256 + t (k < 4), so every load and store instruction is coalesced across lanes.
constexpr int kCastTile = kPackTPB * 16;

__device__ __forceinline__ void store_rep_quad(const PackArgs& a, int64_t e0, const uint32_t (&r)[4],
                                               bool whole) {
    if (whole && a.aligned) {
        if (a.rep_bytes == 1)
            *reinterpret_cast<uint32_t*>(a.dst + e0) = r[0] | (r[1] << 8) | (r[2] << 16) | (r[3] << 24);
        else
            *reinterpret_cast<uint2*>(a.dst + 2 * e0) = make_uint2(r[0] | (r[1] << 16), r[2] | (r[3] << 16));
        return;
    }
    for (int i = 0; i < 4; ++i) {
        if (e0 + i >= a.count) break;
        if (a.rep_bytes == 1) {
            a.dst[e0 + i] = (uint8_t)r[i];
        } else {
            a.dst[2 * (e0 + i)] = (uint8_t)r[i];
            a.dst[2 * (e0 + i) + 1] = (uint8_t)(r[i] >> 8);
        }
    }
}

template <int MODE, int RSRC>
__global__ __launch_bounds__(kPackTPB) void k_cast_to_rep(PackArgs a) {
    const int tid = threadIdx.x;
    const int64_t bbase = (int64_t)blockIdx.x * kCastTile;
    const uint32_t* xin = reinterpret_cast<const uint32_t*>(a.src);
    if (a.vec && bbase + kCastTile <= a.count) {
        uint4 x[4], u[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            x[k] = reinterpret_cast<const uint4*>(xin + bbase)[k * kPackTPB + tid];
            if constexpr (RSRC == 1) u[k] = reinterpret_cast<const uint4*>(a.bits + bbase)[k * kPackTPB + tid];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t e0 = bbase + 4 * (k * kPackTPB + tid);
            if constexpr (RSRC == 2)
                u[k] = make_uint4(rand_u32(a.seed, (uint64_t)e0), rand_u32(a.seed, (uint64_t)e0 + 1),
                                  rand_u32(a.seed, (uint64_t)e0 + 2), rand_u32(a.seed, (uint64_t)e0 + 3));
            else if constexpr (RSRC == 0)
                u[k] = make_uint4(0u, 0u, 0u, 0u);
            const uint32_t r[4] = {elem_rep<MODE, RSRC, false>(a, x[k].x, u[k].x),
                                   elem_rep<MODE, RSRC, false>(a, x[k].y, u[k].y),
                                   elem_rep<MODE, RSRC, false>(a, x[k].z, u[k].z),
                                   elem_rep<MODE, RSRC, false>(a, x[k].w, u[k].w)};
            store_rep_quad(a, e0, r, true);
        }
    } else {
        for (int k = 0; k < 4; ++k) {
            const int64_t e0 = bbase + 4 * (k * kPackTPB + tid);
            if (e0 >= a.count) break;
            uint32_t r[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (e0 + i < a.count)
                    r[i] = elem_rep<MODE, RSRC, false>(a, xin[e0 + i], pack_word_for<RSRC>(a, e0 + i));
            store_rep_quad(a, e0, r, e0 + 4 <= a.count);
        }
    }
}

__device__ __forceinline__ uint32_t load_rep(const PackArgs& a, int64_t e) {
    return a.rep_bytes == 1 ? (uint32_t)a.psrc[e]
                            : (uint32_t)a.psrc[2 * e] | ((uint32_t)a.psrc[2 * e + 1] << 8);
}

__global__ __launch_bounds__(kPackTPB) void k_cast_from_rep(PackArgs a) {
    const int tid = threadIdx.x;
    const int64_t bbase = (int64_t)blockIdx.x * kCastTile;
    uint32_t* out = reinterpret_cast<uint32_t*>(a.dstf);
    if (a.vec && a.aligned && bbase + kCastTile <= a.count) {
        uint32_t w[4][2];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int q = k * kPackTPB + tid;
            if (a.rep_bytes == 1) {
                w[k][0] = reinterpret_cast<const uint32_t*>(a.psrc + bbase)[q];
                w[k][1] = 0u;
            } else {
                const uint2 v = reinterpret_cast<const uint2*>(a.psrc + 2 * bbase)[q];
                w[k][0] = v.x;
                w[k][1] = v.y;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t r[4];
            if (a.rep_bytes == 1) {
#pragma unroll
                for (int i = 0; i < 4; ++i) r[i] = (w[k][0] >> (8 * i)) & 0xFFu;
            } else {
                r[0] = w[k][0] & 0xFFFFu; r[1] = w[k][0] >> 16;
                r[2] = w[k][1] & 0xFFFFu; r[3] = w[k][1] >> 16;
            }
            reinterpret_cast<uint4*>(out + bbase)[k * kPackTPB + tid] =
                make_uint4(__float_as_uint(rep_to_float(r[0], a.P)), __float_as_uint(rep_to_float(r[1], a.P)),
                           __float_as_uint(rep_to_float(r[2], a.P)), __float_as_uint(rep_to_float(r[3], a.P)));
        }
    } else {
        for (int k = 0; k < 4; ++k) {
            const int64_t e0 = bbase + 4 * (k * kPackTPB + tid);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (e0 + i < a.count) out[e0 + i] = __float_as_uint(rep_to_float(load_rep(a, e0 + i), a.P));
        }
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
    P->off2 = (uint32_t)((int64_t)P->exp_off * ((int64_t)1 << mb));
    P->ovf_rep = (P->saturation == 0 && P->has_inf) ? P->inf_rep : P->max_finite;
    P->fast32 = P->exp_off <= 255 ? 1 : 0;
    geometry(n, &P->gc, &P->gbits);
    return LOF_OK;
}

// Sizes of a call on `count` elements; false if the grid would exceed 2^31 - 1 workgroups.
bool size_args(PackArgs* a, int64_t count, int32_t endian) {
    const int64_t groups = (count + a->P.gc - 1) / a->P.gc;
    a->count = count;
    a->slots = groups * a->P.gc;
    a->bytes = groups * (a->P.gbits / 8);
    a->tiles = (count + kTile - 1) / kTile;
    a->xm = endian ? a->P.gc - 1 : 0;
    return a->tiles < ((int64_t)1 << 31);
}

dim3 grid_for(const PackArgs& a) { return dim3((unsigned)a.tiles); }

template <int MODE, int RSRC>
void launch_enc(const PackArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((k_pack_encode<MODE, RSRC, false>), grid_for(a), dim3(kPackTPB), 0, st, a);
}

template <int MODE, int RSRC>
void launch_cast(const PackArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((k_cast_to_rep<MODE, RSRC>),
                       dim3((unsigned)((a.count + kCastTile - 1) / kCastTile)), dim3(kPackTPB), 0,
                       st, a);
}

lof_status launch_cast_to_rep(const PackArgs& a, int mode, int rsrc, hipStream_t st) {
#define LOF_PC(M, R) if (mode == M && rsrc == R) { launch_cast<M, R>(a, st); goto done; }
    LOF_PC(RNE, 0) LOF_PC(RTZ, 0) LOF_PC(RAFZ, 0) LOF_PC(RNO, 0) LOF_PC(RDOWN, 0)
    LOF_PC(RUP, 0) LOF_PC(RTA, 0) LOF_PC(RTO, 0) LOF_PC(RTTZ, 0)
    LOF_PC(SRA, 0) LOF_PC(SRB, 0) LOF_PC(SRC, 0) LOF_PC(SRD, 0)
    LOF_PC(SRA, 1) LOF_PC(SRB, 1) LOF_PC(SRC, 1) LOF_PC(TRUESR, 1) LOF_PC(PROB, 1) LOF_PC(SRD, 1)
    LOF_PC(SRA, 2) LOF_PC(SRB, 2) LOF_PC(SRC, 2) LOF_PC(TRUESR, 2) LOF_PC(PROB, 2) LOF_PC(SRD, 2)
#undef LOF_PC
    return set_error(LOF_ERR_UNSUPPORTED, "lof_cast_to_rep_f32: mode %d not supported", mode);
done:
    return check_hip(hipGetLastError(), "lof_cast_to_rep_f32 launch");
}

lof_status launch_encode(const PackArgs& a, int mode, int rsrc, hipStream_t st) {
    if (a.tiles == 0) return LOF_OK;
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
    if (!size_args(&a, count, endian)) return set_error(LOF_ERR_UNSUPPORTED, "%s: input too large", what);
    a.src = src;
    a.dst = dst;
    a.bits = rand_bits;
    a.seed = seed;
    a.vec = aligned16(src) && (rsrc != 1 || aligned16(rand_bits)) ? 1 : 0;
    if (a.P.n == 8 || a.P.n == 16) {
        // one element per group of whole bytes (either order): the packed stream is the
        // little-endian rep containers, written directly without the LDS slot stage
        if (count == 0) return LOF_OK;
        a.rep_bytes = a.P.n / 8;
        a.aligned = ((uintptr_t)dst & (uintptr_t)(4 * a.rep_bytes - 1)) == 0 ? 1 : 0;
        return launch_cast_to_rep(a, spec.mode, rsrc, (hipStream_t)stream);
    }
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
    if (!size_args(&a, count, endian)) return set_error(LOF_ERR_UNSUPPORTED, "%s: input too large", what);
    if (a.tiles == 0) return LOF_OK;
    a.dstf = dst;
    a.psrc = src;
    a.vec = aligned16(dst) ? 1 : 0;
    if (a.P.n == 8 || a.P.n == 16) {   // byte-aligned single-element groups: rep containers
        a.rep_bytes = a.P.n / 8;
        a.aligned = ((uintptr_t)src & (uintptr_t)(4 * a.rep_bytes - 1)) == 0 ? 1 : 0;
        hipLaunchKernelGGL(k_cast_from_rep, dim3((unsigned)((count + kCastTile - 1) / kCastTile)),
                           dim3(kPackTPB), 0, (hipStream_t)stream, a);
        return check_hip(hipGetLastError(), "lof_pack_decode_f32 launch");
    }
    a.aligned = ((uintptr_t)src & 3u) == 0 ? 1 : 0;
    hipLaunchKernelGGL((k_pack_decode<false>), grid_for(a), dim3(kPackTPB), 0,
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
    if (!size_args(&a, count, endian)) return set_error(LOF_ERR_UNSUPPORTED, "%s: input too large", what);
    if (a.tiles == 0) return LOF_OK;
    a.reps = reps;
    a.dst = dst;
    a.aligned = ((uintptr_t)dst & 3u) == 0 ? 1 : 0;
    a.vec = aligned16(reps) ? 1 : 0;
    hipLaunchKernelGGL((k_pack_encode<RNE, 0, true>), grid_for(a), dim3(kPackTPB), 0,
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
    if (!size_args(&a, count, endian)) return set_error(LOF_ERR_UNSUPPORTED, "%s: input too large", what);
    if (a.tiles == 0) return LOF_OK;
    a.repsout = reps;
    a.psrc = src;
    a.aligned = ((uintptr_t)src & 3u) == 0 ? 1 : 0;
    a.vec = aligned16(reps) ? 1 : 0;
    hipLaunchKernelGGL((k_pack_decode<true>), grid_for(a), dim3(kPackTPB), 0,
                       (hipStream_t)stream, a);
    return check_hip(hipGetLastError(), "lof_unpack_reps launch");
}

LOF_API int32_t lof_rep_bytes(int32_t bitwidth) {
    return bitwidth < 1 || bitwidth > 16 ? -1 : (bitwidth <= 8 ? 1 : 2);
}

LOF_API lof_status lof_cast_to_rep_f32(void* dst, const float* src, int64_t count,
                                       const lof_pack_format* fmt, lof_round_spec spec,
                                       const uint32_t* rand_bits, uint64_t seed,
                                       lof_stream_t stream) {
    const char* what = "lof_cast_to_rep_f32";
    PackArgs a{};
    lof_status s = make_pack_dev(fmt, spec, &a.P, what);
    if (s != LOF_OK) return s;
    if (count < 0) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: count < 0", what);
    if (count == 0) return LOF_OK;
    if (!dst || !src) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: NULL buffer", what);
    if ((count + kCastTile - 1) / kCastTile >= ((int64_t)1 << 31))
        return set_error(LOF_ERR_UNSUPPORTED, "%s: input too large", what);
    int rsrc = 0;
    if (is_stochastic(spec.mode)) rsrc = rand_bits ? 1 : 2;
    if (uses_len(spec.mode) && spec.stoch_len == 0) rsrc = 0;   // degenerate: no draw used
    a.count = count;
    a.src = src;
    a.dst = static_cast<uint8_t*>(dst);
    a.bits = rand_bits;
    a.seed = seed;
    a.rep_bytes = a.P.n <= 8 ? 1 : 2;
    a.aligned = ((uintptr_t)dst & (uintptr_t)(4 * a.rep_bytes - 1)) == 0 ? 1 : 0;
    a.vec = aligned16(src) && (rsrc != 1 || aligned16(rand_bits)) ? 1 : 0;
    return launch_cast_to_rep(a, spec.mode, rsrc, (hipStream_t)stream);
}

LOF_API lof_status lof_cast_from_rep_f32(float* dst, const void* src, int64_t count,
                                         const lof_pack_format* fmt, lof_stream_t stream) {
    const char* what = "lof_cast_from_rep_f32";
    PackArgs a{};
    lof_status s = make_pack_dev(fmt, lof_round_spec{0, 0, 0}, &a.P, what);
    if (s != LOF_OK) return s;
    if (count < 0) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: count < 0", what);
    if (count == 0) return LOF_OK;
    if (!dst || !src) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: NULL buffer", what);
    if ((count + kCastTile - 1) / kCastTile >= ((int64_t)1 << 31))
        return set_error(LOF_ERR_UNSUPPORTED, "%s: input too large", what);
    a.count = count;
    a.dstf = dst;
    a.psrc = static_cast<const uint8_t*>(src);
    a.rep_bytes = a.P.n <= 8 ? 1 : 2;
    a.aligned = ((uintptr_t)src & (uintptr_t)(4 * a.rep_bytes - 1)) == 0 ? 1 : 0;
    a.vec = aligned16(dst) ? 1 : 0;
    hipLaunchKernelGGL(k_cast_from_rep, dim3((unsigned)((count + kCastTile - 1) / kCastTile)),
                       dim3(kPackTPB), 0, (hipStream_t)stream, a);
    return check_hip(hipGetLastError(), "lof_cast_from_rep_f32 launch");
}

}  // extern "C"
