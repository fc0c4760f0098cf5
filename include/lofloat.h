/*
 * lofloat.h -- C-ABI of the MI355X (gfx950) LoFloat hot-path library
 *              (liblofloat_hip.so).
 *
 * Plain pointers, sizes and PODs only: no torch, no C++ types, no exceptions.
 * Every entry point is stream-ordered, performs no host synchronisation and no
 * device allocation, and may be captured into a HIP graph.  The one exception
 * is lof_grouped_create(), which allocates the plan's device descriptors and
 * copies them synchronously; create plans before capture.
 *
 * Device pointers are HIP device memory on the current device.  `stream` is a
 * hipStream_t (NULL = the legacy default stream).  Errors come back as a
 * lof_status; lof_last_error() gives the message for the calling thread.
 *
 * Numerics: results are bit-identical to the reference library's scalar
 * semantics (SudhanvaKulkarni123/LoFloat, src/lo_float.h virtual_round),
 * quirks included (saturating overflow returns the unrounded value, extended
 * overflow returns +inf, NaN becomes +qNaN, ...).  Stochastic modes consume
 * one uint32 random word per element: from `rand_bits[i]` when given, else
 * the counter-based word lof_rand_u32(seed, i).
 */
#ifndef LOFLOAT_H
#define LOFLOAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LOF_API __attribute__((visibility("default")))

typedef void* lof_stream_t; /* hipStream_t */

typedef enum lof_status {
    LOF_OK = 0,
    LOF_ERR_INVALID_ARGUMENT = 1,
    LOF_ERR_UNSUPPORTED = 2,
    LOF_ERR_HIP = 3,
    LOF_ERR_ALLOC = 4
} lof_status;

/* Rounding_Mode, in the reference's X-macro order (fp_tools.hpp:13-28). */
enum {
    LOF_RM_ROUND_TO_NEAREST_EVEN = 0,
    LOF_RM_ROUND_TOWARDS_ZERO = 1,
    LOF_RM_ROUND_AWAY_FROM_ZERO = 2,
    LOF_RM_STOCHASTIC_A = 3,
    LOF_RM_ROUND_TO_NEAREST_ODD = 4,
    LOF_RM_ROUND_DOWN = 5,
    LOF_RM_ROUND_UP = 6,
    LOF_RM_ROUND_TIES_TO_AWAY = 7,
    LOF_RM_STOCHASTIC_B = 8,
    LOF_RM_STOCHASTIC_C = 9,
    LOF_RM_TRUE_STOCHASTIC = 10,
    LOF_RM_PROBABILISTIC = 11,
    LOF_RM_ROUND_TO_ODD = 12,
    LOF_RM_ROUND_TIES_TOWARDS_ZERO = 13,
    LOF_RM_STOCHASTIC_D = 14
};

/* Inf_Behaviors / NaN_Behaviors / Signedness (fp_tools.hpp:30-50). */
enum { LOF_INF_EXTENDED = 0, LOF_INF_SATURATING = 1 };
enum { LOF_NAN_3109 = 0, LOF_NAN_754 = 1 };
enum { LOF_SIGNED = 0, LOF_UNSIGNED = 1 };

/* FloatingPointParams without the checker functors, which virtual_round never
 * reads (fp_tools.hpp:187-252; LoPy_bind.cpp:178-185). */
typedef struct lof_format {
    int32_t bitwidth;      /* total bits incl. sign */
    int32_t mantissa_bits; /* 0..23 */
    int32_t bias;
    int32_t inf_behavior;  /* LOF_INF_* */
    int32_t nan_behavior;  /* LOF_NAN_* (not read by rounding) */
    int32_t signedness;    /* LOF_SIGNED / LOF_UNSIGNED */
} lof_format;

/* ProjSpec (fp_tools.hpp:78-88). */
typedef struct lof_round_spec {
    int32_t mode;       /* LOF_RM_* */
    int32_t saturation; /* Saturation_Mode; not read by virtual_round */
    int32_t stoch_len;  /* random bits for SR-A/B/C/D; 0 <= stoch_len <= 23 - mantissa_bits */
} lof_round_spec;

LOF_API const char* lof_last_error(void);
LOF_API int32_t lof_version(void);

/* Random word `index` of stream `seed` (SplitMix64 output over a Weyl sequence):
 * what the kernels use when no rand_bits array is passed.  Host-callable. */
LOF_API uint32_t lof_rand_u32(uint64_t seed, uint64_t index);

/* Fill bits[0..n) with lof_rand_u32(seed, offset + i) on the device. */
LOF_API lof_status lof_fill_random_u32(uint32_t* bits, int64_t n, uint64_t seed, uint64_t offset,
                                       lof_stream_t stream);

/* out[i] = virtual_round(v_i, *fmt, spec), v_i = (scale == 1) ? in[i] : scale * in[i].
 * Replaces lo_float::round_fp_params (Lof_kernel.cu:49-60; Lof_kernels.h via
 * LoPy_bind.cpp:17-18) and the CUDA branch of virtual_round_params
 * (LoPy_bind.cpp:122-139).  in == out is allowed.  rand_bits: NULL or n words. */
LOF_API lof_status lof_virtual_round_f32(const float* in, float* out, int64_t n,
                                         const lof_format* fmt, lof_round_spec spec, float scale,
                                         const uint32_t* rand_bits, uint64_t seed,
                                         lof_stream_t stream);

/* out[i] = virtual_round(v_i, mant_bits, spec) (mantissa-only overload,
 * lo_float.h:2716-2756).  Replaces lo_float::round_mantissa (Lof_kernel.cu:37-48;
 * LoPy_bind.cpp:14-15, 80-120).  0 <= mant_bits <= 23. */
LOF_API lof_status lof_virtual_round_mant_f32(const float* in, float* out, int64_t n,
                                              int32_t mant_bits, lof_round_spec spec, float scale,
                                              const uint32_t* rand_bits, uint64_t seed,
                                              lof_stream_t stream);

/* MX block rounding over contiguous blocks of block_size elements (the last block may be
 * short): scale = virtual_round(amax / pmn, *scale_fmt, scale_mode) with
 * pmn = ldexpf(1, emax) * (2 - 2^-mb) of the element format; out = scale > 0 ?
 * virtual_round(x / scale, *elem_fmt, elem_spec) * scale : 0.
 * Replaces lo_float::run_virtual_mx_round (Lof_kernel.cu:70-94, Lof_kernels.h:44-48) and
 * virtual_mx_round (LoPy_bind.cpp:165-217).  block_size: power of two in [4, 256].
 * scale_mode: deterministic, or SR-A/B/C/D (which the reference runs with stoch_len 0).
 * in == out is allowed. */
LOF_API lof_status lof_virtual_mx_round_f32(const float* in, float* out, int64_t n,
                                            int32_t block_size, const lof_format* elem_fmt,
                                            const lof_format* scale_fmt, lof_round_spec elem_spec,
                                            int32_t scale_mode, const uint32_t* rand_bits,
                                            uint64_t seed, lof_stream_t stream);

/* ---------------------------------------------------------------------------------
 * Emulated-accumulation GEMM.
 *
 * D[i][j] = alpha * acc_ij (or fmaf(alpha, acc_ij, beta * C[i][j]) when beta != 0, the
 * CUTLASS LinearCombination epilogue) where, in ascending k,
 *   chunk == 1 : acc = round(RTO_fma(a_ik, b_kj, acc))   (the reference leaf,
 *                cutlass_gemms.cuh:139-165: round_to_odd_fma then virtual_round)
 *   chunk  > 1 : per chunk of `chunk` k values (the last may be short)
 *                p = fmaf chain of the chunk's products from +0 (FP32, round to nearest),
 *                acc = round(RTO_fma(p, 1, acc))
 * starting from acc = +0.  round() is virtual_round(., mant_bits) (kind
 * LOF_ACCUM_MANTISSA, what the reference's lof_gemm does) or virtual_round(., fmt)
 * (kind LOF_ACCUM_FORMAT) with the deterministic mode in accum.spec.
 * RTO_fma(a, b, c) rounds a*b + c to odd in FP32; an exact zero gets the sign
 * roundTowardNegative gives it.
 * A: M x K row-major (trans_a = 0, lda >= K) or stored K x M (trans_a = 1, lda >= M).
 * B: K x N row-major (trans_b = 0, ldb >= N) or stored N x K (trans_b = 1, ldb >= K).
 * D: M x N row-major, ldd >= N.  C: M x N row-major, ldc >= N (read only when beta != 0;
 * may alias D).
 * ------------------------------------------------------------------------------- */
enum { LOF_ACCUM_MANTISSA = 0, LOF_ACCUM_FORMAT = 1 };

typedef struct lof_accum {
    int32_t kind;        /* LOF_ACCUM_* */
    int32_t mant_bits;   /* LOF_ACCUM_MANTISSA: 0..23 */
    lof_format fmt;      /* LOF_ACCUM_FORMAT */
    lof_round_spec spec; /* deterministic mode; the reference kernel always uses RNE */
    int32_t chunk;       /* >= 1 */
} lof_accum;

typedef struct lof_gemm_desc {
    int32_t M, N, K;
    const float* A;
    int64_t lda;
    int32_t trans_a;
    const float* B;
    int64_t ldb;
    int32_t trans_b;
    float* D;
    int64_t ldd;
    float alpha;
    lof_accum accum;
    const float* C;
    int64_t ldc;
    float beta;
} lof_gemm_desc;

/* Run one GEMM.  config: -1 = autotuned choice (or the default tile), else an index
 * < lof_gemm_num_configs() that is valid for the chunk (lof_gemm_config_valid).  Testing
 * bits OR-ed into config:
 *   0x100   forces the literal (general) leaf path;
 *   0x200   makes an MFMA config compute every chunk sum with the VALU fmaf chain;
 *   0x400   makes it compute every chunk sum on MFMA without the exactness test (this one
 *           can change results);
 *   0x800   turns the MFMA configs' exact-sum leaves off;
 *   0x1000  turns their RTO leaves off;
 *   0x2000  turns their binary16-held chunk-1 leaves off;
 *   0x4000  turns off only the packed binary16 form of those leaves;
 *   0x8000  turns off the directed-rounding leaves (round-to-odd from the sum rounded down
 *           and up, for mantissa-only and binary16-shaped accumulators at chunk 1, and on
 *           the 64x64 and 64x128 wave tiles for the sums of chunks 8 to 32 into
 *           mantissa-only accumulators of 20 or more bits);
 *   0x10000 turns off the bf16 MFMA chunk sums (exact chunks of 8 to 32 then sum on the
 *           VALU chain; chunks of 2 and 4 always use the f32 MFMA);
 *   0x20000 turns off the mixed-precision FMA leaves of binary16-shaped accumulators
 *           (v_fma_mixlo_f16, one rounding of the exact sum: at chunk 1 on lanes whose
 *           products or sums are inexact in FP32, and for every chunk sum of a slice that
 *           fails the exact-sum test).
 * Chunk sums run on MFMA only where every partial sum of the chunk is exact in FP32 (chunks of
 * 8 to 32 on the bf16 MFMAs, where the slice's values are all bfloat16 values; chunks of 2
 * and 4 on the f32 MFMA), and the
 * short leaves only where they provably equal the reference leaf, both checked per tile and
 * k-slice at run time, so all configs give bit-identical results.  Replaces lo_float::LoF_gemm
 * (Lof_kernel.cu:96-154) and lo_float::Gemm::operator() (cutlass_gemms.cuh:1789-1822;
 * alpha = 1/(scale_a*scale_b) per run_scaled, 1777-1787, is the caller's job). */
LOF_API lof_status lof_gemm_f32(const lof_gemm_desc* desc, int32_t config, lof_stream_t stream);

LOF_API int32_t lof_gemm_num_configs(void);
/* 1 if tile config `config` can run a GEMM with k-chunk `chunk` (MFMA configs need chunk
 * in {1, 2, 4, 8, 16, 32}; chunk 1 runs on VALU), else 0.  No reference counterpart: the
 * reference has one kernel. */
LOF_API int32_t lof_gemm_config_valid(int32_t config, int32_t chunk);
/* Writes a short "BMxBN/TMxTN" style name; returns LOF_ERR_INVALID_ARGUMENT if out of range. */
LOF_API lof_status lof_gemm_config_name(int32_t config, char* buf, int32_t buflen);

/* Autotuner: times the candidate tile configs on this problem -- the wave tiles for chunks
 * they take, else the VALU tiles -- (desc->D is overwritten; C and beta are
 * not used, and a C that overlaps D is rejected with LOF_ERR_INVALID_ARGUMENT), records the
 * winner for problems of the same shape class and accumulator (kind, mantissa bits, format
 * bias, chunk) and returns it.  The operands' values are not part of the key. */
LOF_API lof_status lof_gemm_autotune(const lof_gemm_desc* desc, int32_t iters,
                                     lof_stream_t stream, int32_t* best_config, float* best_ms);

/* ---------------------------------------------------------------------------------
 * Conv2d forward (cross-correlation, groups = 1) as an implicit GEMM on the same
 * emulated-accumulation leaf: M = K (output channels), N = batch * P * Q (output
 * positions), reduction over the filter taps, B gathered from x with zero padding, never
 * materialised.  P = (H + 2 pad_h - dil_h (R - 1) - 1) / stride_h + 1, likewise Q.
 *   LOF_CONV_NCHW: x N x C x H x W, w K x C x R x S (OIHW), y and c N x K x P x Q; the
 *     reduction index is (c, r, s) with s fastest -- w viewed as a K x CRS matrix times
 *     unfold(x), the order of LoF_Conv2d's im2col + lof_gemm path (layers.py:687-737).
 *   LOF_CONV_NHWC: x N x H x W x C, w K x R x S x C (KRSC), y and c N x P x Q x K; the
 *     reduction index is (r, s, c) with c fastest.
 * Each output is the ascending-k leaf chain of lof_gemm_f32 over its reduction.  Convolutions
 * run on two wave tiles, configs 7 (64 x 64) and 6 (128 x 64) (others return
 * LOF_ERR_UNSUPPORTED), with chunk in {1, 2, 4, 8, 16, 32}; x must have fewer than 2^30
 * elements.
 * ------------------------------------------------------------------------------- */
enum { LOF_CONV_NCHW = 0, LOF_CONV_NHWC = 1 };

typedef struct lof_conv2d_desc {
    int32_t N, H, W, C;          /* input batch, height, width, channels */
    int32_t K, R, S;             /* output channels, filter height, width */
    int32_t pad_h, pad_w, stride_h, stride_w, dil_h, dil_w;
    int32_t layout;              /* LOF_CONV_NCHW or LOF_CONV_NHWC */
    const float* x;
    const float* w;
    float* y;
    float alpha;
    lof_accum accum;
    const float* c;              /* read when beta != 0; y's layout */
    float beta;
} lof_conv2d_desc;

/* y = alpha * conv(x, w) [+ beta * c].  config: -1 for the default wave tile, else 6 or 7,
 * with lof_gemm_f32's testing bits.  Replaces lo_float::LoF_conv2d
 * (Lof_kernel.cu:156-228; NCHW/OIHW, alpha = 1 / (weight_scale * input_scale), beta = 0)
 * and lo_float::Conv2d::operator() (cutlass_conv.cuh:118-181; NHWC/KRSC).  The reference's
 * CUTLASS mainloop fixes its own k order; the order here is the one stated above. */
LOF_API lof_status lof_conv2d_f32(const lof_conv2d_desc* desc, int32_t config,
                                  lof_stream_t stream);

/* ---------------------------------------------------------------------------------
 * Grouped persistent launch: many rounding jobs and GEMMs, each with its own format
 * and mode, in ONE kernel launch that walks a device work queue.
 * ------------------------------------------------------------------------------- */
enum {
    LOF_ITEM_ROUND = 0,      /* lof_virtual_round_f32 semantics */
    LOF_ITEM_ROUND_MANT = 1, /* lof_virtual_round_mant_f32 semantics */
    LOF_ITEM_MX = 2,         /* lof_virtual_mx_round_f32 semantics */
    LOF_ITEM_GEMM = 3        /* lof_gemm_f32 semantics */
};

typedef struct lof_group_item {
    int32_t kind;
    /* rounding items */
    const float* in;
    float* out;
    int64_t n;
    lof_format fmt;       /* ROUND: target format; MX: element format */
    int32_t mant_bits;    /* ROUND_MANT */
    lof_round_spec spec;  /* element rounding */
    float scale;          /* ROUND / ROUND_MANT pre-scale */
    const uint32_t* rand_bits;
    uint64_t seed;
    int32_t block_size;   /* MX */
    lof_format scale_fmt; /* MX */
    int32_t scale_mode;   /* MX */
    /* GEMM items */
    lof_gemm_desc gemm;
} lof_group_item;

typedef struct lof_grouped lof_grouped;

/* Validate the items and upload the work list.  config: -1 = default (or the tuned
 * choice from lof_grouped_autotune), else < lof_grouped_num_configs().  A GEMM item with
 * at least as many tiles as the persistent grid has workgroups fills the device by itself:
 * lof_grouped_run launches it through the single-GEMM kernel (whose tile is faster than
 * the combined kernel's) ahead of the persistent launch that runs every other item.  The
 * outputs are bit-identical either way. */
LOF_API lof_status lof_grouped_create(const lof_group_item* items, int32_t count, int32_t config,
                                      lof_grouped** plan);
LOF_API lof_status lof_grouped_run(lof_grouped* plan, lof_stream_t stream);
/* Launch each item with its own per-item entry point (the baseline the grouped launch
 * is compared with). */
LOF_API lof_status lof_grouped_run_separate(lof_grouped* plan, lof_stream_t stream);
LOF_API lof_status lof_grouped_destroy(lof_grouped* plan);
LOF_API int32_t lof_grouped_num_configs(void);
LOF_API lof_status lof_grouped_config_name(int32_t config, char* buf, int32_t buflen);
/* Work tiles of the persistent launch (device-filling GEMMs, launched on their own, are
 * not counted) and its grid size. */
LOF_API lof_status lof_grouped_info(const lof_grouped* plan, int64_t* tiles, int32_t* grid);
/* Autotuner for the grouped launch: times every config on these items (outputs are
 * overwritten) and returns the best. */
LOF_API lof_status lof_grouped_autotune(const lof_group_item* items, int32_t count, int32_t iters,
                                        lof_stream_t stream, int32_t* best_config,
                                        float* best_ms);

/* The grouped launch of lof_grouped_create + lof_grouped_run in one call that allocates
 * nothing.  No counterpart in the reference: it is for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  The work queue, the job
 * start table and the task records live in one fixed device workspace of the library
 * (about 170 KB per GPU, held until the process ends), so count <= 256 items
 * (LOF_ERR_UNSUPPORTED beyond).  Device-filling GEMMs are launched on their own first, as
 * in lof_grouped_run; the task table of the rest is copied from host memory into the
 * workspace on `stream` (a pageable copy, which HIP performs synchronously), then one
 * kernel runs it; outputs are bit-identical to lof_grouped_run on the same items.  Every
 * call on a GPU shares that GPU's workspace: calls that overlap on one GPU (other streams,
 * other threads' streams) race and must be ordered by the caller.  config as in
 * lof_grouped_create. */
LOF_API lof_status lof_grouped_run_noalloc(const lof_group_item* items, int32_t count,
                                           int32_t config, lof_stream_t stream);

/* HIP graph helpers: capture plan runs into a graph and replay it. */
typedef struct lof_graph lof_graph;
LOF_API lof_status lof_graph_capture_grouped(lof_grouped* plan, int32_t separate,
                                             lof_stream_t stream, lof_graph** graph);
LOF_API lof_status lof_graph_launch(lof_graph* graph, lof_stream_t stream);
LOF_API lof_status lof_graph_destroy(lof_graph* graph);

/* ---------------------------------------------------------------------------------
 * Bit-pack codec: quantise FP32 to an n-bit rep (1 <= n <= 16) and pack the reps tightly,
 * or unpack and widen back to FP32 (lo_pack.hpp).  A group of lcm(n, 8) bits (n <= 8) or
 * lcm(n, 16) bits holds group_count = bits / n elements; fields are LSB-first within
 * bytes; endian 0 (Forward) puts element i of a group in slot i, 1 (Reverse) in slot
 * group_count-1-i.  A partial final group is padded with encodings of +0.0f.
 * ------------------------------------------------------------------------------- */
enum { LOF_PACK_FORWARD = 0, LOF_PACK_REVERSE = 1 };
enum { LOF_NEG_TO_ZERO = 0, LOF_NEG_TO_NAN = 1 };   /* Unsigned_behavior (fp_tools.hpp:30-50) */

/* An IsInf / IsNaN checker functor of FloatingPointParams, as data.  The codec reads them
 * (unlike virtual_round).  kind: 0 matches nothing; 1 rep equals one of codes[0..ncodes);
 * 2 (rep & a) == a and (rep & b) != 0 (NaN) or == 0 (Inf), the IEEE-style form of
 * DeviceInfChecker / DeviceNaNChecker (lo_float.h:4386-4398) and the FP16 / BF16 checkers
 * (fp_tools.hpp:331-396); 3 (rep | a) == codes[0], the P3109 Inf checker
 * (lo_float.h:1712-1759).  pat0 / pat1: qNanBitPattern / sNanBitPattern of a NaN checker,
 * minPosInf / minNegInf of an Inf checker. */
typedef struct lof_checker {
    int32_t kind;
    int32_t ncodes;
    uint32_t codes[3];
    uint32_t a, b;
    uint32_t pat0, pat1;
} lof_checker;

/* FloatingPointParams<IsInf, IsNaN> with the fields the codec reads (fp_tools.hpp:187-252). */
typedef struct lof_pack_format {
    lof_format fmt;             /* bitwidth 1..16 */
    int32_t unsigned_behavior;  /* LOF_NEG_TO_ZERO / LOF_NEG_TO_NAN */
    lof_checker is_inf, is_nan;
} lof_pack_format;

/* Bytes a packed buffer of `count` n-bit elements occupies; -1 if n is outside [1, 16].
 * Replaces LoF_packed_size (lo_pack.hpp:392-394). */
LOF_API int64_t lof_packed_size(int32_t bitwidth, int64_t count);

/* dst[0 .. lof_packed_size) <- pack(convert_float_to_rep(src[i], fmt, spec)) on the device.
 * spec.saturation is read (OvfInf / SatFinite / SatPropagate); stochastic modes take one
 * random word per element (rand_bits, or lof_rand_u32(seed, i) when NULL).  Replaces
 * LoF_encode_cuda (lo_pack.hpp:504-517) and the CPU LoF_encode(float) (lo_pack.hpp:414-429). */
LOF_API lof_status lof_pack_encode_f32(uint8_t* dst, const float* src, int64_t count,
                                       const lof_pack_format* fmt, lof_round_spec spec,
                                       int32_t endian, const uint32_t* rand_bits, uint64_t seed,
                                       lof_stream_t stream);

/* dst[i] <- convert_rep_to_float(field i of src).  Replaces LoF_decode_cuda
 * (lo_pack.hpp:519-531) and the CPU LoF_decode(float) (lo_pack.hpp:448-463). */
LOF_API lof_status lof_pack_decode_f32(float* dst, const uint8_t* src, int64_t count,
                                       const lof_pack_format* fmt, int32_t endian,
                                       lof_stream_t stream);

/* Pack / unpack already-quantised reps (the low n bits of each uint32).  Replace the typed
 * LoF_encode / LoF_decode of Templated_Float arrays (lo_pack.hpp:397-412, 431-446). */
LOF_API lof_status lof_pack_reps(uint8_t* dst, const uint32_t* reps, int64_t count,
                                 int32_t bitwidth, int32_t endian, lof_stream_t stream);
LOF_API lof_status lof_unpack_reps(uint32_t* reps, const uint8_t* src, int64_t count,
                                   int32_t bitwidth, int32_t endian, lof_stream_t stream);

/* ---------------------------------------------------------------------------------
 * Piecewise-linear SiLU: y = linear interpolation of an (N+1)-knot table over [-R, R]
 * for |x| <= R, relu(x) outside, computed in float for every element type.
 * ------------------------------------------------------------------------------- */
enum { LOF_DTYPE_F32 = 0, LOF_DTYPE_F16 = 1, LOF_DTYPE_F64 = 2 };

/* out[i] = pwl(in[i]) with lut[0..segments] on the device; segments a power of two in
 * [2, 4096]; in / out of element type dtype (they may alias).  Replaces pwl_silu_kernel /
 * pwl_silu (Lof_kernel.cu:231-265) and pwl_silu_forward (LoPy_bind.cpp:228-277). */
LOF_API lof_status lof_pwl_silu(void* out, const void* in, int64_t n, int32_t dtype,
                                const float* lut, int32_t segments, float R,
                                lof_stream_t stream);

/* ---------------------------------------------------------------------------------
 * Typed cast: FP32 <-> arrays of n-bit reps, one per container (uint8 for n <= 8, uint16
 * little-endian for 9 <= n <= 16), i.e. the storage of a Templated_Float<Fp> array.  The
 * conversion is the scalar ConvertImpl<float, Templated_Float<Fp>>::run of lo_float.h
 * (lo_float.h:3057-3364; the runtime twin the codec restates), with the Saturation_Mode of
 * spec, not the SIMD array run<arch> (lo_float.h:3497+), which drops the special values.
 * ------------------------------------------------------------------------------- */
/* Bytes per rep container: 1 (n <= 8), 2 (n <= 16), -1 outside [1, 16]. */
LOF_API int32_t lof_rep_bytes(int32_t bitwidth);

/* dst[i] <- the rep of src[i] in fmt under spec (random words as lof_pack_encode_f32).
 * Replaces lo_float::Project<Templated_Float<Fp>>(float*, Templated_Float<Fp>*, n, ps)
 * (lo_float.h:4014-4022) and the scalar Round<T> / Project<T> (lo_float.h:3979-3988). */
LOF_API lof_status lof_cast_to_rep_f32(void* dst, const float* src, int64_t count,
                                       const lof_pack_format* fmt, lof_round_spec spec,
                                       const uint32_t* rand_bits, uint64_t seed,
                                       lof_stream_t stream);
/* dst[i] <- the FP32 value of rep src[i] (ConvertImpl<Templated_Float<Fp>, float>::run,
 * the codec's convert_rep_to_float). */
LOF_API lof_status lof_cast_from_rep_f32(float* dst, const void* src, int64_t count,
                                         const lof_pack_format* fmt, lof_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* LOFLOAT_H */
