/*
 * lof_oracle.h -- CPU oracle for the LoFloat hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only to compare against / time beside the HIP product path.  Nothing in
 * lofloat_amd/ links, loads or calls it.
 *
 * It restates, in plain C, the scalar semantics of the reference library
 * (SudhanvaKulkarni123/LoFloat).  Each function cites the reference file:line
 * it follows.  Stochastic modes take one explicit uint32 random word per
 * element instead of the reference's per-call mt19937 / per-thread SplitMix64
 * draws; the word -> random-integer mapping is documented at
 * orc_round_mantissa_bits().
 */
#ifndef LOF_ORACLE_H
#define LOF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Enum values follow the reference X-macro order (fp_tools.hpp:13-50). */
enum {
    ORC_RNE = 0, ORC_RTZ = 1, ORC_RAFZ = 2, ORC_SRA = 3, ORC_RNO = 4,
    ORC_RDOWN = 5, ORC_RUP = 6, ORC_RTA = 7, ORC_SRB = 8, ORC_SRC = 9,
    ORC_TRUESR = 10, ORC_PROB = 11, ORC_RTO = 12, ORC_RTTZ = 13, ORC_SRD = 14
};

/* FloatingPointParams minus the (unused) checker functors (fp_tools.hpp:187-252). */
typedef struct orc_format {
    int32_t bitwidth;
    int32_t mantissa_bits;
    int32_t bias;
    int32_t inf_behavior;   /* 0 Extended, 1 Saturating */
    int32_t nan_behavior;   /* 0 _3109, 1 _754 (never read by virtual_round) */
    int32_t signedness;     /* 0 Signed, 1 Unsigned */
} orc_format;

/* ProjSpec (fp_tools.hpp:78-88). */
typedef struct orc_spec {
    int32_t mode;
    int32_t saturation;     /* never read by virtual_round */
    int32_t stoch_len;
} orc_spec;

/* Accumulator description for the emulated GEMM. */
typedef struct orc_accum {
    int32_t kind;           /* 0: mantissa-only (virtual_round(x, bits)), 1: full format */
    int32_t mant_bits;      /* kind 0 */
    orc_format fmt;         /* kind 1 */
    orc_spec spec;          /* rounding of the accumulator (deterministic modes) */
    int32_t chunk;          /* k-chunk length c >= 1 */
} orc_accum;

/* Counter-based random word used when no explicit bits array is given. */
uint32_t orc_rand_u32(uint64_t seed, uint64_t index);

/* RoundMantissa (lo_float.h:1232-1272) on magnitude bits, explicit random word u. */
uint32_t orc_round_mantissa_bits(uint32_t bits, int roundoff, int mode, int stoch_len,
                                 int pos, uint32_t u);

/* Scalar virtual_round overloads. */
float orc_virtual_round_fmt(float x, const orc_format* f, orc_spec s, uint32_t u);
float orc_virtual_round_mant(float x, int mant_bits, orc_spec s, uint32_t u);

/* Elementwise launch semantics (Lof_kernel.cu:37-56): v = (scale == 1) ? x : scale*x,
 * then round.  bits == NULL -> random words from orc_rand_u32(seed, i). */
void orc_virtual_round_fmt_array(const float* in, float* out, int64_t n, const orc_format* f,
                                 orc_spec s, float scale, const uint32_t* bits, uint64_t seed,
                                 int nthreads);
void orc_virtual_round_mant_array(const float* in, float* out, int64_t n, int mant_bits,
                                  orc_spec s, float scale, const uint32_t* bits, uint64_t seed,
                                  int nthreads);

/* MX block rounding (LoPy_bind.cpp:198-214). */
void orc_virtual_mx_round(const float* in, float* out, int64_t n, int block_size,
                          const orc_format* elem, const orc_format* scale_fmt,
                          orc_spec elem_spec, int scale_mode, const uint32_t* bits,
                          uint64_t seed, int nthreads);

/* Round-to-odd FMA.  _dir: literal fma_rd / fma_ru restatement via the host FPU rounding
 * mode (cutlass_gemms.cuh:74-80).  Plain: exact-residual form of the host twin
 * (blis_lof_gemm.hpp:63-99) with the device zero-sign and overflow rules. */
float orc_rto_fma_dir(float a, float b, float c);
float orc_rto_fma(float a, float b, float c);

/* Emulated-accumulation GEMM.  D = alpha * acc (beta == 0) or fmaf(alpha, acc,
 * beta * C), acc per output as documented in lof_oracle.c.  trans_a: A stored K x M
 * (lda >= M); trans_b: B stored N x K (ldb >= K).  samples != NULL: compute only outputs
 * samples[s] = i*N + j, written to D[s]. */
void orc_gemm_f32(int M, int N, int K, const float* A, int64_t lda, int trans_a,
                  const float* B, int64_t ldb, int trans_b, float* D, int64_t ldd,
                  float alpha, const orc_accum* acc, const int64_t* samples,
                  int64_t n_samples, int nthreads, const float* C, int64_t ldc, float beta);

/* ---------------------------------------------------------------- bit-pack codec
 * lo_pack.hpp: n-bit floats (1 <= n <= 16) packed in groups of lcm(n, 8) (n <= 8) or
 * lcm(n, 16) bits, LSB-first within bytes; endian 0 = Forward (element i in slot i),
 * 1 = Reverse (slot G-1-i). */

/* An Inf or NaN checker functor of FloatingPointParams, as data (lo_float.h:1411-1800,
 * fp_tools.hpp:255-400, lo_float.h:4386-4398).  kind: 0 never matches; 1 rep equals one
 * of codes[0 .. ncodes); 2 (rep & a) == a and (rep & b) != 0 (NaN) / == 0 (Inf);
 * 3 (rep | a) == codes[0].  pat0 / pat1: qNanBitPattern / sNanBitPattern of a NaN
 * checker, minPosInf / minNegInf of an Inf checker. */
typedef struct orc_checker {
    int32_t kind;
    int32_t ncodes;
    uint32_t codes[3];
    uint32_t a, b;
    uint32_t pat0, pat1;
} orc_checker;

typedef struct orc_pack_format {
    orc_format fmt;
    int32_t unsigned_behavior;   /* 0 NegtoZero, 1 NegtoNaN (fp_tools.hpp:30-50) */
    orc_checker is_inf, is_nan;
} orc_pack_format;

/* convert_float_to_rep (lo_pack.hpp:108-258); ps.saturation is read here. */
uint32_t orc_float_to_rep(float x, const orc_pack_format* f, orc_spec ps, uint32_t u);
/* convert_rep_to_float (lo_pack.hpp:268-304). */
float orc_rep_to_float(uint32_t rep, const orc_pack_format* f);
/* LoF_packed_size (lo_pack.hpp:392-394). */
int64_t orc_packed_bytes(int n, int64_t count);
/* LoF_encode(float) / LoF_decode(float) (lo_pack.hpp:414-463); bits == NULL -> words
 * orc_rand_u32(seed, i) for element i.  Return bytes written / consumed. */
int64_t orc_pack_encode(uint8_t* dst, const float* src, int64_t count,
                        const orc_pack_format* f, orc_spec ps, int endian,
                        const uint32_t* bits, uint64_t seed);
int64_t orc_pack_decode(float* dst, const uint8_t* src, int64_t count,
                        const orc_pack_format* f, int endian);
/* LoF_encode / LoF_decode of typed values (lo_pack.hpp:397-412, 431-446): raw reps. */
int64_t orc_pack_reps(uint8_t* dst, const uint32_t* reps, int64_t count, int n, int endian);
int64_t orc_unpack_reps(uint32_t* reps, const uint8_t* src, int64_t count, int n, int endian);

/* Elementwise forms of the two codec functions (random words as for orc_pack_encode). */
void orc_float_to_rep_array(const float* x, uint32_t* out, int64_t n, const orc_pack_format* f,
                            orc_spec ps, const uint32_t* bits, uint64_t seed);
void orc_rep_to_float_array(const uint32_t* rep, float* out, int64_t n, const orc_pack_format* f);

int orc_max_threads(void);

/* pwl_silu_kernel (Lof_kernel.cu:240-258): one element, and an array. */
float orc_pwl_silu(float x, const float* lut, int N, float R);
void orc_pwl_silu_array(const float* x, float* y, int64_t n, const float* lut, int N, float R);

#ifdef __cplusplus
}
#endif
#endif
