// lof_gemm_mfma.hpp -- emulated-accumulation GEMM with MFMA chunk sums (gfx950).
//
// For chunk c > 1 the accumulation rule is: p = FP32 fmaf chain over the chunk's c products
// (ascending k, starting at +0), then acc <- round(RTO(p + acc)).  The same tile also runs
// chunk 1 (acc <- round(RTO_fma(a, b, acc)) per k) on VALU, with float products where a
// slice's products are exact in FP32 (operand significand widths summing to <= 24).  When every partial sum of
// the chunk is exactly representable in FP32, p is the exact sum of the products and any
// summation order gives it; v_mfma_f32_32x32x2f32 then computes the chunk sums.
//
// Exactness test, per workgroup tile and 32-wide K slice (every chunk of the slice is a
// subset of it): with amax_A, amax_B the largest magnitudes in the A and B slice tiles and
// lsb_A, lsb_B lower bounds on the exponents of the lowest set significand bits of their
// nonzero values (smallest exponent plus smallest trailing-zero count of the significands),
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
// computes from one LDS buffer, then stored to the other (one barrier per slice); interior
// slices use float4 loads.  The chunk length is a template parameter so a chunk's LDS reads
// are issued together and its MFMAs run back to back.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "lof_gemm_kernels.hpp"

namespace lofamd {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Two floats that are bfloat16 values (their low 16 bits are zero) as one packed pair: the
// high halves of lo and hi.
__device__ __forceinline__ uint32_t bf16_pair(float lo, float hi) {
    return __builtin_amdgcn_perm(__float_as_uint(hi), __float_as_uint(lo), 0x07060302u);
}

// p[0], p[stride], ..., p[7 * stride] (LDS floats holding bfloat16 values) as an MFMA operand
__device__ __forceinline__ bf16x8 lds_bf16x8(const float* p, int stride) {
    const u32x4 v = {bf16_pair(p[0], p[stride]), bf16_pair(p[2 * stride], p[3 * stride]),
                     bf16_pair(p[4 * stride], p[5 * stride]), bf16_pair(p[6 * stride], p[7 * stride])};
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v |= (uint32_t)__shfl_xor((int)v, o, 64);
    return v;
}

// The float leaf's slow path: the exact-zero leaf, else the general leaf (the double fast
// leaf, which gives the same value where it applies, is left out: it reads more of GemmArgs).
// The second form is the same through a reference to the kernel's arguments (see leaf_slow).
__device__ __noinline__ float leaf_f_slow(float p, float acc, LeafGen L) {
    const double dp = p, da = acc;
    if (L.range != 0u && dp + da == 0.0)   // leaf_zero: RTO24's sign of an exact zero
        return (__double_as_longlong(dp) | __double_as_longlong(da)) == 0 ? 0.0f : -0.0f;
    return (float)leaf_slow<LeafGen>(dp, da, L);
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

// leaf_zero for a float sum (float addition cannot underflow to zero, so s == 0 is exact)
__device__ __forceinline__ bool leaf_f_zero(float p, float acc, const GemmArgs& g, float& out) {
    if (g.range_f == 0u || p + acc != 0.0f) return false;
    out = (__float_as_uint(p) | __float_as_uint(acc)) == 0u ? 0.0f : -0.0f;
    return true;
}

// Exact-sum leaf on the negated accumulator (the tile keeps nacc = -acc): t = nacc - p =
// -(acc + p), computed exactly in FP32 (the caller has proved it), so RTO24 is the identity
// and the leaf is round(acc + p) = -round(t): round-to-nearest-even at bit rf of t's
// pattern, which is right in the format's normal range and leaves a value in its subnormal
// range unchanged (the caller has also proved such sums lie on the format's grid).  The
// rounding keeps the sign, so the result is the new nacc.  An exact zero t gets IEEE's
// round-to-nearest sign, which for t = -(acc + p) is roundTowardNegative's sign of acc + p:
// the reference's zero rule ((+0) + (+0) = +0, any other exact zero is -0).
// F16: the accumulator's grid is IEEE binary16's below 2^15 (m = 10, bias 15) and the
// caller has bounded |t| below 2^15, so the hardware conversion pair (round-to-nearest-even,
// binary16 subnormals kept) computes the same rounding in two instructions.
template <bool F16>
__device__ __forceinline__ float leaf_exact(float t, const GemmArgs& g) {
    if constexpr (F16) {
        return (float)(_Float16)t;
    } else {
        const uint32_t b = __float_as_uint(t);
        return __uint_as_float((b + g.half_m1f + ((b >> g.rf) & 1u)) & g.keep_maskf);
    }
}

// RTO leaf on the negated accumulator, for any (finite, in-range) sum: t = RN24(nacc - p) and
// its TwoSum residual e give x = -(acc + p) = t + e exactly, so RTO24(x) -- the reference's
// round-to-odd FMA result -- is t when e == 0, else the magnitude truncation of x with its
// last bit set: t | 1 when e has t's sign (|x| > |t|), (t - 1 ulp) | 1 otherwise.  Rounding
// that as leaf_exact does is round(RTO24(acc + p)) negated: exact wherever leaf_exact's
// rounding is the accumulator's for every float, i.e. for a mantissa-only accumulator
// (bitwise round-to-nearest-even) and a binary16-shaped one below 2^15 (the conversion).
template <bool F16>
__device__ __forceinline__ float leaf_rto(float p, float na, const GemmArgs& g) {
#pragma clang fp contract(off)
    const float t = na - p;
    const float bb = t - na;
    const float e = (na - (t - bb)) + (-p - bb);
    const uint32_t tb = __float_as_uint(t);
    const bool same = (int32_t)(__float_as_uint(e) ^ tb) >= 0;
    const uint32_t odd = (same ? tb : tb - 1u) | 1u;
    return leaf_exact<F16>(__uint_as_float(e == 0.0f ? tb : odd), g);
}

// binary16(h - a b) for FP32 a, b and a binary16 h: v_fma_mixlo_f16 rounds the exact result
// once (the product is not rounded first; FP32-subnormal products and binary16 subnormals
// included; an exact zero takes round-to-nearest's sign).  For a binary16-shaped accumulator
// whose sums stay below 2^15 this is the reference leaf round(RTO24(a b + acc)) on the negated
// accumulator h = -acc: round-to-odd keeps 13 more bits than binary16, so rounding it again
// gives the single rounding of the exact sum.
__device__ __forceinline__ _Float16 mix_fnma16(float a, float b, _Float16 h) {
    _Float16 r;
    asm("v_fma_mixlo_f16 %0, -%1, %2, %3 op_sel_hi:[0,0,1]" : "=v"(r) : "v"(a), "v"(b), "v"(h));
    return r;
}

// na[o] <- binary16(na[o] - p[o]) for FP32 na and p, one rounding of the exact difference,
// for 16 outputs in place: the chunk leaf of a binary16-shaped accumulator (na = -acc holds a
// binary16 value) on the same argument.  One asm block with tied operands, so the leaves
// add no registers.
__device__ __forceinline__ void mix_sub16x16(const f32x16& p, float (&na)[16]) {
    asm(
        "v_fma_mixlo_f16 %0, %16, -1.0, %0 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %1, %17, -1.0, %1 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %2, %18, -1.0, %2 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %3, %19, -1.0, %3 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %4, %20, -1.0, %4 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %5, %21, -1.0, %5 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %6, %22, -1.0, %6 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %7, %23, -1.0, %7 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %8, %24, -1.0, %8 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %9, %25, -1.0, %9 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %10, %26, -1.0, %10 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %11, %27, -1.0, %11 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %12, %28, -1.0, %12 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %13, %29, -1.0, %13 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %14, %30, -1.0, %14 op_sel_hi:[0,0,0]\n\t"
        "v_fma_mixlo_f16 %15, %31, -1.0, %15 op_sel_hi:[0,0,0]\n\t"
        "v_cvt_f32_f16 %0, %0\n\t"
        "v_cvt_f32_f16 %1, %1\n\t"
        "v_cvt_f32_f16 %2, %2\n\t"
        "v_cvt_f32_f16 %3, %3\n\t"
        "v_cvt_f32_f16 %4, %4\n\t"
        "v_cvt_f32_f16 %5, %5\n\t"
        "v_cvt_f32_f16 %6, %6\n\t"
        "v_cvt_f32_f16 %7, %7\n\t"
        "v_cvt_f32_f16 %8, %8\n\t"
        "v_cvt_f32_f16 %9, %9\n\t"
        "v_cvt_f32_f16 %10, %10\n\t"
        "v_cvt_f32_f16 %11, %11\n\t"
        "v_cvt_f32_f16 %12, %12\n\t"
        "v_cvt_f32_f16 %13, %13\n\t"
        "v_cvt_f32_f16 %14, %14\n\t"
        "v_cvt_f32_f16 %15, %15"
        : "+v"(na[0]), "+v"(na[1]), "+v"(na[2]), "+v"(na[3]), "+v"(na[4]), "+v"(na[5]), "+v"(na[6]), "+v"(na[7]), "+v"(na[8]), "+v"(na[9]), "+v"(na[10]), "+v"(na[11]), "+v"(na[12]), "+v"(na[13]), "+v"(na[14]), "+v"(na[15])
        : "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3]), "v"(p[4]), "v"(p[5]), "v"(p[6]), "v"(p[7]), "v"(p[8]), "v"(p[9]), "v"(p[10]), "v"(p[11]), "v"(p[12]), "v"(p[13]), "v"(p[14]), "v"(p[15]));
}

// na[o] <- binary16(na[o] - p[o]) when every difference is exact in FP32 (the exact-sum leaf
// of a binary16-shaped accumulator: one subtraction and the conversion pair, all full rate),
// for 16 outputs in place, tied like mix_sub16x16.
__device__ __forceinline__ void exact_sub16x16(const f32x16& p, float (&na)[16]) {
    asm(
        "v_sub_f32 %0, %0, %16\n\t"
        "v_sub_f32 %1, %1, %17\n\t"
        "v_sub_f32 %2, %2, %18\n\t"
        "v_sub_f32 %3, %3, %19\n\t"
        "v_sub_f32 %4, %4, %20\n\t"
        "v_sub_f32 %5, %5, %21\n\t"
        "v_sub_f32 %6, %6, %22\n\t"
        "v_sub_f32 %7, %7, %23\n\t"
        "v_sub_f32 %8, %8, %24\n\t"
        "v_sub_f32 %9, %9, %25\n\t"
        "v_sub_f32 %10, %10, %26\n\t"
        "v_sub_f32 %11, %11, %27\n\t"
        "v_sub_f32 %12, %12, %28\n\t"
        "v_sub_f32 %13, %13, %29\n\t"
        "v_sub_f32 %14, %14, %30\n\t"
        "v_sub_f32 %15, %15, %31\n\t"
        "v_cvt_f16_f32 %0, %0\n\t"
        "v_cvt_f16_f32 %1, %1\n\t"
        "v_cvt_f16_f32 %2, %2\n\t"
        "v_cvt_f16_f32 %3, %3\n\t"
        "v_cvt_f16_f32 %4, %4\n\t"
        "v_cvt_f16_f32 %5, %5\n\t"
        "v_cvt_f16_f32 %6, %6\n\t"
        "v_cvt_f16_f32 %7, %7\n\t"
        "v_cvt_f16_f32 %8, %8\n\t"
        "v_cvt_f16_f32 %9, %9\n\t"
        "v_cvt_f16_f32 %10, %10\n\t"
        "v_cvt_f16_f32 %11, %11\n\t"
        "v_cvt_f16_f32 %12, %12\n\t"
        "v_cvt_f16_f32 %13, %13\n\t"
        "v_cvt_f16_f32 %14, %14\n\t"
        "v_cvt_f16_f32 %15, %15\n\t"
        "v_cvt_f32_f16 %0, %0\n\t"
        "v_cvt_f32_f16 %1, %1\n\t"
        "v_cvt_f32_f16 %2, %2\n\t"
        "v_cvt_f32_f16 %3, %3\n\t"
        "v_cvt_f32_f16 %4, %4\n\t"
        "v_cvt_f32_f16 %5, %5\n\t"
        "v_cvt_f32_f16 %6, %6\n\t"
        "v_cvt_f32_f16 %7, %7\n\t"
        "v_cvt_f32_f16 %8, %8\n\t"
        "v_cvt_f32_f16 %9, %9\n\t"
        "v_cvt_f32_f16 %10, %10\n\t"
        "v_cvt_f32_f16 %11, %11\n\t"
        "v_cvt_f32_f16 %12, %12\n\t"
        "v_cvt_f32_f16 %13, %13\n\t"
        "v_cvt_f32_f16 %14, %14\n\t"
        "v_cvt_f32_f16 %15, %15"
        : "+v"(na[0]), "+v"(na[1]), "+v"(na[2]), "+v"(na[3]), "+v"(na[4]), "+v"(na[5]), "+v"(na[6]), "+v"(na[7]), "+v"(na[8]), "+v"(na[9]), "+v"(na[10]), "+v"(na[11]), "+v"(na[12]), "+v"(na[13]), "+v"(na[14]), "+v"(na[15])
        : "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3]), "v"(p[4]), "v"(p[5]), "v"(p[6]), "v"(p[7]), "v"(p[8]), "v"(p[9]), "v"(p[10]), "v"(p[11]), "v"(p[12]), "v"(p[13]), "v"(p[14]), "v"(p[15]));
}

// RD / RU: a[o] * b + c[o] rounded down and up, for 16 outputs.  The wave's FP32 rounding
// mode (MODE.FP_ROUND[1:0]: 2 toward -inf, 1 toward +inf, 0 nearest even) is switched only
// inside this one asm block, which ends in round-to-nearest-even again, so no FP operation
// the compiler schedules ever runs under a directed mode.  The FMA is correctly rounded in
// every mode, FP32 subnormals included (denormals are not flushed), so rd / ru are the two
// directed roundings of the exact a * b + c.
__device__ __forceinline__ void fma_down_up16(const float (&a)[16], float b, const float (&c)[16],
                                              float (&rd)[16], float (&ru)[16]) {
    asm volatile(
        "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 2\n\t"
        "s_nop 1\n\t"
        "v_fma_f32 %0, %32, %48, %49\n\t"
        "v_fma_f32 %1, %33, %48, %50\n\t"
        "v_fma_f32 %2, %34, %48, %51\n\t"
        "v_fma_f32 %3, %35, %48, %52\n\t"
        "v_fma_f32 %4, %36, %48, %53\n\t"
        "v_fma_f32 %5, %37, %48, %54\n\t"
        "v_fma_f32 %6, %38, %48, %55\n\t"
        "v_fma_f32 %7, %39, %48, %56\n\t"
        "v_fma_f32 %8, %40, %48, %57\n\t"
        "v_fma_f32 %9, %41, %48, %58\n\t"
        "v_fma_f32 %10, %42, %48, %59\n\t"
        "v_fma_f32 %11, %43, %48, %60\n\t"
        "v_fma_f32 %12, %44, %48, %61\n\t"
        "v_fma_f32 %13, %45, %48, %62\n\t"
        "v_fma_f32 %14, %46, %48, %63\n\t"
        "v_fma_f32 %15, %47, %48, %64\n\t"
        "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 1\n\t"
        "s_nop 1\n\t"
        "v_fma_f32 %16, %32, %48, %49\n\t"
        "v_fma_f32 %17, %33, %48, %50\n\t"
        "v_fma_f32 %18, %34, %48, %51\n\t"
        "v_fma_f32 %19, %35, %48, %52\n\t"
        "v_fma_f32 %20, %36, %48, %53\n\t"
        "v_fma_f32 %21, %37, %48, %54\n\t"
        "v_fma_f32 %22, %38, %48, %55\n\t"
        "v_fma_f32 %23, %39, %48, %56\n\t"
        "v_fma_f32 %24, %40, %48, %57\n\t"
        "v_fma_f32 %25, %41, %48, %58\n\t"
        "v_fma_f32 %26, %42, %48, %59\n\t"
        "v_fma_f32 %27, %43, %48, %60\n\t"
        "v_fma_f32 %28, %44, %48, %61\n\t"
        "v_fma_f32 %29, %45, %48, %62\n\t"
        "v_fma_f32 %30, %46, %48, %63\n\t"
        "v_fma_f32 %31, %47, %48, %64\n\t"
        "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 0\n\t"
        "s_nop 1"
        : "=&v"(rd[0]), "=&v"(rd[1]), "=&v"(rd[2]), "=&v"(rd[3]), "=&v"(rd[4]), "=&v"(rd[5]),
          "=&v"(rd[6]), "=&v"(rd[7]), "=&v"(rd[8]), "=&v"(rd[9]), "=&v"(rd[10]), "=&v"(rd[11]),
          "=&v"(rd[12]), "=&v"(rd[13]), "=&v"(rd[14]), "=&v"(rd[15]),
          "=&v"(ru[0]), "=&v"(ru[1]), "=&v"(ru[2]), "=&v"(ru[3]), "=&v"(ru[4]), "=&v"(ru[5]),
          "=&v"(ru[6]), "=&v"(ru[7]), "=&v"(ru[8]), "=&v"(ru[9]), "=&v"(ru[10]), "=&v"(ru[11]),
          "=&v"(ru[12]), "=&v"(ru[13]), "=&v"(ru[14]), "=&v"(ru[15])
        : "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(a[6]),
          "v"(a[7]), "v"(a[8]), "v"(a[9]), "v"(a[10]), "v"(a[11]), "v"(a[12]), "v"(a[13]),
          "v"(a[14]), "v"(a[15]), "v"(b),
          "v"(c[0]), "v"(c[1]), "v"(c[2]), "v"(c[3]), "v"(c[4]), "v"(c[5]), "v"(c[6]),
          "v"(c[7]), "v"(c[8]), "v"(c[9]), "v"(c[10]), "v"(c[11]), "v"(c[12]), "v"(c[13]),
          "v"(c[14]), "v"(c[15]));
}

// Directed-rounding leaves of 16 outputs: acc[o] <- round(RTO24(a[o] * b + acc[o])) for any
// operands, the products exact or not.  This is round_to_odd_fma's construction
// (cutlass_gemms.cuh:74-80) from the directed roundings rd / ru of x = a * b + acc: the
// truncation rz = (rd >= 0 ? rd : ru) with its last bit set when rd != ru.  So an exact x is
// rd, an exact zero with roundTowardNegative's sign ((+0) + (+0) = +0, any other exact zero
// -0), and an overflowing x gives FLT_MAX | 1 with its sign.  When rd != ru, rd and ru have
// the same sign bit (FP32 subnormals are kept, so only an exact zero straddles it) and the
// truncation is the one of smaller magnitude, i.e. the smaller bit pattern: rz | 1 is
// min(rd, ru) | 1 on the patterns (the larger pattern is +inf / -inf on overflow).
// round() is round-to-nearest-even at bit dr_rf of the pattern (a mantissa-only accumulator:
// bitwise virtual_round, lo_float.h:2716-2756) or, F16, the binary16 conversion (a
// binary16-shaped format, for sums the caller has bounded below 2^15).
template <bool F16>
__device__ __forceinline__ void leaves_dr16(const float (&a)[16], float b, float (&acc)[16],
                                            const GemmArgs& g) {
    float rd[16], ru[16];
    fma_down_up16(a, b, acc, rd, ru);
#pragma unroll
    for (int o = 0; o < 16; ++o) {
        const uint32_t d = __float_as_uint(rd[o]), u = __float_as_uint(ru[o]);
        const uint32_t t = rd[o] == ru[o] ? d : (min(d, u) | 1u);
        if constexpr (F16)
            acc[o] = (float)(_Float16)__uint_as_float(t);
        else   // round to nearest even at bit dr_rf: the kept lsb (none when dr_rf is 0)
            acc[o] = __uint_as_float((t + g.dr_half + __builtin_amdgcn_ubfe(t, g.dr_rf, g.dr_lsb)) &
                                     g.dr_mask);
    }
}

// RD / RU of na[o] - p[o] for 16 outputs, under the same mode switch as fma_down_up16.
__device__ __forceinline__ void sub_down_up16(const float (&na)[16], const f32x16& p,
                                              float (&rd)[16], float (&ru)[16]) {
    asm volatile(
        "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 2\n\t"
        "s_nop 1\n\t"
        "v_sub_f32 %0, %32, %48\n\t"
        "v_sub_f32 %1, %33, %49\n\t"
        "v_sub_f32 %2, %34, %50\n\t"
        "v_sub_f32 %3, %35, %51\n\t"
        "v_sub_f32 %4, %36, %52\n\t"
        "v_sub_f32 %5, %37, %53\n\t"
        "v_sub_f32 %6, %38, %54\n\t"
        "v_sub_f32 %7, %39, %55\n\t"
        "v_sub_f32 %8, %40, %56\n\t"
        "v_sub_f32 %9, %41, %57\n\t"
        "v_sub_f32 %10, %42, %58\n\t"
        "v_sub_f32 %11, %43, %59\n\t"
        "v_sub_f32 %12, %44, %60\n\t"
        "v_sub_f32 %13, %45, %61\n\t"
        "v_sub_f32 %14, %46, %62\n\t"
        "v_sub_f32 %15, %47, %63\n\t"
        "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 1\n\t"
        "s_nop 1\n\t"
        "v_sub_f32 %16, %32, %48\n\t"
        "v_sub_f32 %17, %33, %49\n\t"
        "v_sub_f32 %18, %34, %50\n\t"
        "v_sub_f32 %19, %35, %51\n\t"
        "v_sub_f32 %20, %36, %52\n\t"
        "v_sub_f32 %21, %37, %53\n\t"
        "v_sub_f32 %22, %38, %54\n\t"
        "v_sub_f32 %23, %39, %55\n\t"
        "v_sub_f32 %24, %40, %56\n\t"
        "v_sub_f32 %25, %41, %57\n\t"
        "v_sub_f32 %26, %42, %58\n\t"
        "v_sub_f32 %27, %43, %59\n\t"
        "v_sub_f32 %28, %44, %60\n\t"
        "v_sub_f32 %29, %45, %61\n\t"
        "v_sub_f32 %30, %46, %62\n\t"
        "v_sub_f32 %31, %47, %63\n\t"
        "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 0, 2), 0\n\t"
        "s_nop 1"
        : "=&v"(rd[0]), "=&v"(rd[1]), "=&v"(rd[2]), "=&v"(rd[3]), "=&v"(rd[4]), "=&v"(rd[5]),
          "=&v"(rd[6]), "=&v"(rd[7]), "=&v"(rd[8]), "=&v"(rd[9]), "=&v"(rd[10]), "=&v"(rd[11]),
          "=&v"(rd[12]), "=&v"(rd[13]), "=&v"(rd[14]), "=&v"(rd[15]),
          "=&v"(ru[0]), "=&v"(ru[1]), "=&v"(ru[2]), "=&v"(ru[3]), "=&v"(ru[4]), "=&v"(ru[5]),
          "=&v"(ru[6]), "=&v"(ru[7]), "=&v"(ru[8]), "=&v"(ru[9]), "=&v"(ru[10]), "=&v"(ru[11]),
          "=&v"(ru[12]), "=&v"(ru[13]), "=&v"(ru[14]), "=&v"(ru[15])
        : "v"(na[0]), "v"(na[1]), "v"(na[2]), "v"(na[3]), "v"(na[4]), "v"(na[5]), "v"(na[6]),
          "v"(na[7]), "v"(na[8]), "v"(na[9]), "v"(na[10]), "v"(na[11]), "v"(na[12]),
          "v"(na[13]), "v"(na[14]), "v"(na[15]),
          "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3]), "v"(p[4]), "v"(p[5]), "v"(p[6]),
          "v"(p[7]), "v"(p[8]), "v"(p[9]), "v"(p[10]), "v"(p[11]), "v"(p[12]), "v"(p[13]),
          "v"(p[14]), "v"(p[15]));
}

// The chunk leaf on the negated accumulator: nacc[o] <- -round(RTO24(p[o] + acc[o])) from the
// directed roundings of y = nacc - p = -(p + acc), for a mantissa-only accumulator.
// Round-to-odd and both roundings are odd functions, so the construction of leaves_dr16
// applied to y gives the negated result; only an exact zero changes: -(RTO zero of x) is
// roundTowardPositive's zero of y, i.e. ru.
__device__ __forceinline__ void leaves_dr16_neg(const f32x16& p, float (&nacc)[16],
                                                const GemmArgs& g) {
    float rd[16], ru[16];
    sub_down_up16(nacc, p, rd, ru);
#pragma unroll
    for (int o = 0; o < 16; ++o) {
        const uint32_t d = __float_as_uint(rd[o]), u = __float_as_uint(ru[o]);
        const uint32_t t = rd[o] == ru[o] ? u : (min(d, u) | 1u);
        nacc[o] = __uint_as_float((t + g.dr_half + __builtin_amdgcn_ubfe(t, g.dr_rf, g.dr_lsb)) &
                                  g.dr_mask);
    }
}

// Leaves of one 32x32 block's 16 outputs per lane, branch-free in the common case: fast
// results are committed where valid, outputs needing the general path are recorded in a bit
// mask, and one wave-uniform branch runs the general path for them (acc still holds their
// old value).  BYVAL passes the general path what it needs by value (LeafGen), so the
// kernel's GemmArgs is never referenced from out-of-line code (see leaf_slow).
template <bool SUBN, bool BYVAL = false>
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
        for (int o = 0; o < 16; ++o) {
            float z;
            if ((slow >> o) & 1u)
                if (leaf_f_zero(p[o], acc[o], g, z)) acc[o] = z;
                else if constexpr (BYVAL) acc[o] = leaf_f_slow(p[o], acc[o], leaf_gen(g));
                else acc[o] = leaf_f_slow(p[o], acc[o], g);
        }
    }
}

template <int WMB, int WNB>
struct MfmaTile {
    static constexpr int BM = 64 * WMB;
    static constexpr int BN = 64 * WNB;
    static constexpr int BK = 32;
    static constexpr int LDA_S = BM + 4;   // As[k][m], rows padded to keep float4 alignment
    static constexpr int LDB_S = BN + 4;   // Bs[k][n]
    static constexpr int BUF_FLOATS = BK * LDA_S + BK * LDB_S;
    static constexpr int STAT_WORDS = 4 * 8;   // [wave][amaxA, minA, orA, amaxB, minB, orB, keyA, keyB]
    static constexpr int SMEM_FLOATS = 2 * (BUF_FLOATS + STAT_WORDS);
    static constexpr int NA4 = BM * BK / 1024;   // float4 of the A slice per thread
    static constexpr int NB4 = BN * BK / 1024;
};

// Exactness statistics of a slice operand, accumulated per element v (bit pattern):
//   amax = max(|v| << 1)           the largest magnitude (sign shifted out)
//   mnk  = min((|v| << 1) - 1)     the smallest nonzero magnitude (zeros wrap to ~0)
//   orb  = OR of the patterns      its low 23 bits give the smallest trailing-zero count
// so every nonzero value's lowest set bit is at exponent >= e_min - 150 + ctz(orb | 2^23).
// With KEY, also
//   key  = min(E + min(ctz(v), 24)), E = ((|v| << 1) - 1) >> 24
// For an FP32-normal v with a nonzero fraction, E is its biased exponent and ctz(v) the
// trailing-zero count of the fraction, so key - 150 is the exponent of its lowest set bit;
// for a power of two E is one less and min(ctz, 24) is 24 (even exponent) or 23 (odd), so
// key - 150 is that exponent or one below it.  Either way key - 150 bounds every normal
// value's lowest set bit from below; zeros give 279 and drop out of the minimum.
template <bool KEY>
struct SliceStat {
    uint32_t amax = 0u, mnk = 0xFFFFFFFFu, orb = 0u, key = 0xFFFFFFFFu;
    __device__ __forceinline__ void add(float v) {
        const uint32_t b = __float_as_uint(v);
        const uint32_t sh = b << 1;
        const uint32_t nz = sh - 1u;
        amax = max(amax, sh);
        mnk = min(mnk, nz);
        orb |= b;
        if constexpr (KEY) key = min(key, (nz >> 24) + min((uint32_t)__builtin_ctz(b | 0x80000000u), 24u));
    }
    __device__ __forceinline__ void add(const float4& v) { add(v.x); add(v.y); add(v.z); add(v.w); }
};

// lowest-set-bit exponent bound and largest exponent of an operand; false if the operand
// holds an FP32-subnormal value (excluded from the MFMA path)
__device__ __forceinline__ bool slice_bounds(uint32_t amax, uint32_t mnk, uint32_t orb, int& lsb,
                                             int& emax) {
    const uint32_t mn = (mnk + 1u) >> 1;   // smallest nonzero |v| bits
    const int emin_b = (int)(mn >> 23);
    lsb = emin_b - 150 + __builtin_ctz((orb & 0x7FFFFFu) | 0x800000u);
    emax = (int)(amax >> 24) - 127;
    return emin_b != 0;
}

// CONV: B is gathered from g.cv's activation and D / C are stored at its strided output
// positions (implicit-GEMM convolution); a separate instantiation, so the plain GEMM
// kernels carry none of its registers.
// DRC: chunk sums (chunks 8-32) take the directed-rounding leaves in place of the TwoSum
// round-to-odd leaves: the fast leaf for mantissa-only accumulators of 20 or more bits, which
// no other leaf covers.  An instantiation of its own, since both kinds in one kernel spill
// and slow the exact-sum workloads by a fifth.
// GMEM: g lives in global memory (a task of the grouped kernel), not in the kernel's
// arguments, so a reference to it costs no private copy.
template <int WMB, int WNB, int CHUNK, bool SUBN, bool CONV = false, bool DRC = false,
          bool GMEM = false>
__device__ __forceinline__ void gemm_tile_mfma(const GemmArgs& g, int tile_m, int tile_n,
                                               int tid, float* smem) {
    using T = MfmaTile<WMB, WNB>;
    constexpr int BM = T::BM, BN = T::BN, BK = T::BK, LDA_S = T::LDA_S, LDB_S = T::LDB_S;
    constexpr int NA4 = T::NA4, NB4 = T::NB4;
    constexpr int LOG2C = CHUNK == 1 ? 0 : CHUNK == 2 ? 1 : CHUNK == 4 ? 2 : CHUNK == 8 ? 3
                        : CHUNK == 16 ? 4 : 5;
    static_assert(CHUNK >= 1 && CHUNK <= BK && (BK % CHUNK) == 0, "chunk must divide 32");
    // per-value lowest-bit statistics for the short chunks, where the exact-sum leaves
    // carry most of the work (measured: they cost more than they win from chunk 8 up)
    constexpr bool kKey = CHUNK <= 4;
    // exact chunk sums on the bf16 MFMAs (chunks 8-32) or the f32 MFMA (chunks 2, 4)
    constexpr bool kBf16 = CHUNK >= 8;
    // directed-rounding leaves at chunk 1, and for the sums of chunks 8-32 in the DRC
    // instantiation; otherwise chunk sums keep the TwoSum round-to-odd leaves
    constexpr bool kDr = CHUNK == 1 || (DRC && CHUNK >= 8);
    // checked binary16-held chunk-1 leaves of the 10-bit accumulator (below): the 64-row tiles
    // only (the 128-row tiles hold twice the accumulators and are left as they are), and not
    // the SUBN kernels, which no mantissa-only accumulator runs
    constexpr bool kChk = CHUNK == 1 && WMB == 1 && !SUBN;
    // binary16-shaped accumulators run the SUBN kernels (their subnormal range lies above
    // FP32's), so the other kernels leave out their chunk-sum leaves -- except at chunk 32,
    // whose kernel measured 129 registers of spills without them (a different schedule)
    constexpr bool kF16Leaves = (SUBN || CHUNK >= 32) && !DRC;
    // the general leaf's arguments by value: the tile loop then holds GemmArgs in registers
    // instead of reading it back from scratch (whose waits also wait for the in-flight
    // prefetch; measured 30% faster on the 64x128 tile at chunk 32, 10-25% at chunk 1, 24%
    // on the 128-row tiles at chunk 32).  Tiles where the values cost more registers than
    // the copy keep the reference (measured: by value the 128-row tiles at chunk 1 lose a
    // wave per SIMD, the 64x128 tile at chunks 2 and 4 runs 3-20% slower).
    constexpr bool kLeafByVal = !GMEM && !(WMB == 2 && CHUNK == 1) &&
                                !(WMB == 1 && WNB == 2 && (CHUNK == 2 || CHUNK == 4));

    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int l32 = lane & 31, lhi = lane >> 5;
    const int m0 = tile_m * BM, n0 = tile_n * BN;
    const int am0 = wm * 32 * WMB, bn0 = wn * 32 * WNB;   // this wave's sub-tile
    const bool interior_mn = m0 + BM <= g.M && n0 + BN <= g.N;

    // The lane's accumulators, held as acc (negd false) for the general leaves or as -acc
    // (negd true) for the exact-sum leaves, which work on the negation directly; a lane
    // flips them only when it changes from one kind of slice to the other.
    float nacc[WMB][WNB][16];
    bool negd = false;
    f32x16 part[WMB][WNB];
#pragma unroll
    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
        for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
            for (int r = 0; r < 16; ++r) nacc[bi][bj][r] = 0.0f;
    auto set_form = [&](bool neg) {
        if (neg != negd) {
#pragma unroll
            for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                    for (int r = 0; r < 16; ++r) nacc[bi][bj][r] = -nacc[bi][bj][r];
            negd = neg;
        }
    };

    // Slice element mapping: float4 f = tid + it * 256 of the A slice holds
    //   ta = 0 (A row-major, k contiguous): row m = f / 8, k = 4 * (f % 8) .. + 3
    //   ta = 1 (A stored K x M, m contiguous): k = f / (BM / 4), m = 4 * (f % (BM / 4)) .. + 3
    // and likewise for B (tb = 0: n contiguous; tb = 1: k contiguous).
    float4 va[NA4], vb[NB4];
    // Float4 `it` lies 256 float4s past float4 0, which moves it by a whole number of rows
    // (256 is a multiple of 8, BM / 4 and BN / 4): per thread, the coordinates of float4 0,
    // and per float4 a wave-uniform step.  (Assigned through references in two branches, the
    // coordinate pair was kept in private memory, and its scratch reads -- whose waits also
    // wait for the in-flight prefetch -- sat in every slice's store.)
    // t: tid, recomputed per slice from an opaque copy (otid) so that the compiler does not
    // hold the per-float4 coordinates and addresses across the loop in registers.
    auto otid = [&]() { int t = tid; asm volatile("" : "+v"(t)); return t; };
    auto a_coord = [&](int it, int j, int t) -> int2 {   // (m, k)
        const int a_mb = g.ta ? 4 * (t % (BM / 4)) : t / 8;
        const int a_kb = g.ta ? t / (BM / 4) : 4 * (t % 8);
        return make_int2(a_mb + (g.ta ? j : 32 * it), a_kb + (g.ta ? it * (1024 / BM) : j));
    };
    auto b_coord = [&](int it, int j, int t) -> int2 {   // (n, k)
        const int b_nb = g.tb ? t / 8 : 4 * (t % (BN / 4));
        const int b_kb = g.tb ? 4 * (t % 8) : t / (BN / 4);
        return make_int2(b_nb + (g.tb ? 32 * it : j), b_kb + (g.tb ? j : it * (1024 / BN)));
    };
    // The edge loads' element coordinates, from a copy of tid the compiler cannot see
    // through: otherwise it hoists their loop-invariant parts out of the tile loop for
    // every element, and the registers they hold spill into the interior slices' path.
    auto edge_coord = [&](bool isa, int it, int j) -> int2 {
        int t = tid;
        asm volatile("" : "+v"(t));
        const int f = t + it * 256;
        if (isa) return g.ta ? make_int2(4 * (f % (BM / 4)) + j, f / (BM / 4))
                             : make_int2(f / 8, 4 * (f % 8) + j);
        return g.tb ? make_int2(f / 8, 4 * (f % 8) + j)
                    : make_int2(4 * (f % (BN / 4)) + j, f / (BN / 4));
    };
    // Convolution: this thread's four B columns n = n0 + 4 * (tid % (BN / 4)) + j are the
    // same in every slice (256 is a multiple of BN / 4); their output image, first input row
    // and column and base offset are computed once.  A column past N gets a row that puts
    // every tap outside the image.
    int cv_ih0[4], cv_iw0[4], cv_base[4];
    if constexpr (CONV) {
        const ConvGeom& cv = g.cv;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gn = n0 + 4 * (tid % (BN / 4)) + j;
            const uint32_t b = fast_div((uint32_t)gn, cv.div_pq);
            const uint32_t pq = (uint32_t)gn - b * cv.PQ;
            const uint32_t p = fast_div(pq, cv.div_q);
            const uint32_t q = pq - p * (uint32_t)cv.Q;
            const bool valid = gn < g.N;
            cv_ih0[j] = valid ? (int)p * cv.sh - cv.ph : -(1 << 30);
            cv_iw0[j] = (int)q * cv.sw - cv.pw;
            cv_base[j] = valid ? (int)((int64_t)b * cv.sxb) + cv_ih0[j] * cv.sxh + cv_iw0[j] * cv.sxw
                               : 0;
        }
    }
    // B(k, n) for the slice rows k = k0 + f / (BN / 4): the tap (c, r, s) of k, then one
    // element per column j, zero where the tap falls in the padding or k >= K
    auto gather_b = [&](int k0) {
        const int t = otid();
        const ConvGeom& cv = g.cv;
#pragma unroll
        for (int it = 0; it < NB4; ++it) {
            const int2 nk = b_coord(it, 0, t);
            const int n = nk.x, k = nk.y;
            const int gk = k0 + k;
            const uint32_t d1 = fast_div((uint32_t)gk, cv.div_d1);
            const uint32_t rem = (uint32_t)gk - d1 * cv.D1;
            const uint32_t d2 = fast_div(rem, cv.div_d2);
            const uint32_t d3 = rem - d2 * (uint32_t)cv.D2;
            const int c = cv.krsc ? (int)d3 : (int)d1;
            const int r = cv.krsc ? (int)d1 : (int)d2;
            const int sx = cv.krsc ? (int)d2 : (int)d3;
            const int rr = r * cv.dh, ss = sx * cv.dw;
            const int koff = c * cv.sxc + rr * cv.sxh + ss * cv.sxw;
            float e[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int ih = cv_ih0[j] + rr, iw = cv_iw0[j] + ss;
                const bool in = gk < g.K && (unsigned)ih < (unsigned)cv.H &&
                                (unsigned)iw < (unsigned)cv.W;
                const float v = gld(g.B + (in ? cv_base[j] + koff : 0));   // offset 0 when outside
                e[j] = in ? v : 0.0f;
            }
            vb[it] = make_float4(e[0], e[1], e[2], e[3]);
        }
    };
    auto load_slice = [&](int k0) {
        const int t = otid();
        const bool whole = g.vec && interior_mn && k0 + BK <= g.K;
        if (whole) {   // whole slice inside: float4 loads
#pragma unroll
            for (int it = 0; it < NA4; ++it) {
                const int2 mk = a_coord(it, 0, t);
                const int m = mk.x, k = mk.y;
                const float* p = g.ta ? g.A + (int64_t)(k0 + k) * g.lda + (m0 + m)
                                      : g.A + (int64_t)(m0 + m) * g.lda + (k0 + k);
                va[it] = gld4(p);
            }
        } else {   // edge: element loads from clamped addresses, zero outside the matrix
#pragma unroll
            for (int it = 0; it < NA4; ++it) {
                float e[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int2 mk = edge_coord(true, it, j);
                    const int m = mk.x, k = mk.y;
                    const int gm = m0 + m, gk = k0 + k;
                    const int cm = min(gm, g.M - 1), ck = min(gk, g.K - 1);
                    const float v = gld(g.ta ? g.A + (int64_t)ck * g.lda + cm : g.A + (int64_t)cm * g.lda + ck);
                    e[j] = (gm < g.M && gk < g.K) ? v : 0.0f;
                }
                va[it] = make_float4(e[0], e[1], e[2], e[3]);
            }
        }
        if constexpr (CONV) {
            gather_b(k0);
        } else if (whole) {
#pragma unroll
            for (int it = 0; it < NB4; ++it) {
                const int2 nk = b_coord(it, 0, t);
                const int n = nk.x, k = nk.y;
                const float* p = g.tb ? g.B + (int64_t)(n0 + n) * g.ldb + (k0 + k)
                                      : g.B + (int64_t)(k0 + k) * g.ldb + (n0 + n);
                vb[it] = gld4(p);
            }
        } else {
#pragma unroll
            for (int it = 0; it < NB4; ++it) {
                float e[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int2 nk = edge_coord(false, it, j);
                    const int n = nk.x, k = nk.y;
                    const int gn = n0 + n, gk = k0 + k;
                    const int cn = min(gn, g.N - 1), ck = min(gk, g.K - 1);
                    const float v = gld(g.tb ? g.B + (int64_t)cn * g.ldb + ck : g.B + (int64_t)ck * g.ldb + cn);
                    e[j] = (gn < g.N && gk < g.K) ? v : 0.0f;
                }
                vb[it] = make_float4(e[0], e[1], e[2], e[3]);
            }
        }
    };
    // registers -> LDS buffer `buf`, with the slice's exactness statistics
    auto store_slice = [&](int buf) {
        const int t = otid();
        float* As = smem + buf * (T::BUF_FLOATS + T::STAT_WORDS);
        float* Bs = As + BK * LDA_S;
        uint32_t* stat = (uint32_t*)(Bs + BK * LDB_S);
        SliceStat<kKey> sa, sb;
#pragma unroll
        for (int it = 0; it < NA4; ++it) {
            const int2 mk = a_coord(it, 0, t);
            const int m = mk.x, k = mk.y;
            if (g.ta) {
                *(float4*)&As[k * LDA_S + m] = va[it];
            } else {
                As[(k + 0) * LDA_S + m] = va[it].x;
                As[(k + 1) * LDA_S + m] = va[it].y;
                As[(k + 2) * LDA_S + m] = va[it].z;
                As[(k + 3) * LDA_S + m] = va[it].w;
            }
            sa.add(va[it]);
        }
#pragma unroll
        for (int it = 0; it < NB4; ++it) {
            const int2 nk = b_coord(it, 0, t);
            const int n = nk.x, k = nk.y;
            if (g.tb) {
                Bs[(k + 0) * LDB_S + n] = vb[it].x;
                Bs[(k + 1) * LDB_S + n] = vb[it].y;
                Bs[(k + 2) * LDB_S + n] = vb[it].z;
                Bs[(k + 3) * LDB_S + n] = vb[it].w;
            } else {
                *(float4*)&Bs[k * LDB_S + n] = vb[it];
            }
            sb.add(vb[it]);
        }
        sa.amax = wave_max_u32(sa.amax);
        sb.amax = wave_max_u32(sb.amax);
        sa.mnk = wave_min_u32(sa.mnk);
        sb.mnk = wave_min_u32(sb.mnk);
        sa.orb = wave_or_u32(sa.orb);
        sb.orb = wave_or_u32(sb.orb);
        if constexpr (kKey) {
            sa.key = wave_min_u32(sa.key);
            sb.key = wave_min_u32(sb.key);
        }
        if (lane == 0) {
            uint32_t* st = stat + wave * 8;
            st[0] = sa.amax; st[1] = sa.mnk; st[2] = sa.orb;
            st[3] = sb.amax; st[4] = sb.mnk; st[5] = sb.orb;
            st[6] = sa.key; st[7] = sb.key;
        }
    };

    if (g.K > 0) {
        load_slice(0);
        store_slice(0);
    }
    int buf = 0;
    for (int k0 = 0; k0 < g.K; k0 += BK, buf ^= 1) {
        // Slice k0 and its statistics are in LDS buffer `buf`; every wave is done with the
        // other buffer (slice k0 - 32), which the store below refills.
        __syncthreads();
        const float* As = smem + buf * (T::BUF_FLOATS + T::STAT_WORDS);
        const float* Bs = As + BK * LDA_S;
        const uint32_t* stat = (const uint32_t*)(Bs + BK * LDB_S);
        bool exact;
        // for the exact-sum test: every product (chunk 1) or chunk sum of the slice is a
        // multiple of 2^sL and below 2^sTop in magnitude; sfin: no inf or NaN in the slice
        int sL = 0, sTop = 0;
        bool sfin = false;
        // chunk 1: every nonzero B value of the slice is a multiple of 2^bLsb (so at least
        // 2^bLsb) and below 2^(bTop + 1), with at most bW significant bits; a16: every A
        // value is a binary16 value; left at values no test passes when a slice is zero
        int bLsb = -1000, bTop = 1000, bW = 99;
        bool a16 = false;
        // chunk > 1: every A and B value of the slice is a bfloat16 value (<= 8 significant
        // bits; FP32-subnormal values excluded), so exact chunk sums can use the bf16 MFMAs
        bool bf = false;
        {
            uint32_t amA = 0u, mnA = 0xFFFFFFFFu, orA = 0u, amB = 0u, mnB = 0xFFFFFFFFu, orB = 0u;
            uint32_t kyA = 0xFFFFFFFFu, kyB = 0xFFFFFFFFu;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                amA = max(amA, stat[w * 8 + 0]); mnA = min(mnA, stat[w * 8 + 1]); orA |= stat[w * 8 + 2];
                amB = max(amB, stat[w * 8 + 3]); mnB = min(mnB, stat[w * 8 + 4]); orB |= stat[w * 8 + 5];
                if constexpr (kKey) { kyA = min(kyA, stat[w * 8 + 6]); kyB = min(kyB, stat[w * 8 + 7]); }
            }
            if (amA == 0u || amB == 0u) {
                exact = true;                              // every product is zero
                sfin = amA < 0xFF000000u && amB < 0xFF000000u;
                sL = 1000;
                sTop = -1000;
            } else if (amA >= 0xFF000000u || amB >= 0xFF000000u) {
                exact = false;                             // inf or NaN in the slice
            } else {
                int lsbA, lsbB, eA, eB;
                const bool normA = slice_bounds(amA, mnA, orA, lsbA, eA);
                const bool normB = slice_bounds(amB, mnB, orB, lsbB, eB);
                if constexpr (kKey) {   // the per-value bound, tighter than slice_bounds'
                    lsbA = max(lsbA, (int)kyA - 150);
                    lsbB = max(lsbB, (int)kyB - 150);
                }
                const int L = lsbA + lsbB;
                if constexpr (CHUNK == 1) {
                    // every product a*b exact in FP32: significand widths sum to <= 24, and
                    // no product overflows or falls below 2^-149
                    const int wA = 24 - __builtin_ctz((orA & 0x7FFFFFu) | 0x800000u);
                    const int wB = 24 - __builtin_ctz((orB & 0x7FFFFFu) | 0x800000u);
                    exact = normA && normB && wA + wB <= 24 && eA + eB + 2 <= 127 && L >= -149;
                    sTop = eA + eB + 2;
                    bLsb = lsbB;
                    bTop = eB;
                    bW = wB;
                    // <= 11 significant bits, lowest bit >= 2^-24, below 2^16: binary16
                    // holds it (normal from 2^-14, a multiple of 2^-24 below)
                    a16 = wA <= 11 && lsbA >= -24 && eA <= 15;
                } else {
                    const int top = LOG2C + eA + eB + 2;
                    exact = normA && normB && L >= -126 && top <= L + 24 && top <= 127;
                    sTop = top;
                    bf = g.mfma_bf16 && normA && normB &&
                         __builtin_ctz((orA & 0x7FFFFFu) | 0x800000u) >= 16 &&
                         __builtin_ctz((orB & 0x7FFFFFu) | 0x800000u) >= 16;
                }
                sL = L;
                sfin = true;
            }
            if (g.mfma_mode != 0) {
                exact = g.mfma_mode == 2;
                sfin = false;
            }
        }
        const bool has_next = k0 + BK < g.K;
        if (has_next) load_slice(k0 + BK);   // in flight while this slice computes

        // Exact-sum test, per lane: is every sum s = acc + p of this slice exact in FP32?
        // With |acc| < 2^ea and |p| < 2^sTop at the slice start, a leaf grows |acc| to at
        // most (|acc| + 2^sTop)(1 + 2^-(m+1)), so over the slice's n leaves every s is below
        // g (2^ea + n 2^sTop) + 2^sTop < 2^S, g = (1 + 2^-(m+1))^n (rounded up on the host;
        // the float evaluation gets a 2^-16 margin, which also covers the RTO leaves below).
        // Every acc is a format value, a multiple of 2^max(E(acc) - m, q), and at the start a
        // multiple of 2^Lf, Lf = min(sL, lowest
        // such bound over the lane's nonzero acc); an exact sum keeps both properties, so by
        // induction every s is exact when
        //   S <= sL + 24       (acc a multiple of 2^sL: s is one too, with < 24 bits)
        //   sTop <= Lf + 23    (a finer acc: s < 2^(max(E(acc) + 2, sTop + 1)) has < 24 bits
        //                       above acc's lowest bit, as m <= 19)
        // and the leaf formula applies when S <= sum_ehi (no top binade), Lf >= -126 (s is
        // zero or FP32-normal) and sL >= q (sums in the format's subnormal range are on its
        // grid).
        //
        // Where that fails, RTO leaves (leaf_rto) still apply to a lane whose accumulator format
        // they round exactly (mantissa-only, or binary16-shaped), when the slice's products are
        // exact (chunk 1) or its chunk sums are FP32 values, nothing is inf / NaN and every sum
        // stays below 2^sum_ehi.  The margin allows the RTO24 steps ((1 + 2^-23)^n) and, for a
        // chain's inexact chunk sums, the bound takes |p| < 2^(sTop + 1).
        //
        // Chunk 1 holds a passing lane's accumulators as binary16 values h = -acc 2^hs, so
        // one mixed-precision FMA, h <- binary16(fma(-a, b 2^hs, h)), is the whole leaf (the
        // FP32 fma is exact, so rounding it to binary16 is rounding the sum).  That is the
        // accumulator's rounding, scaled by 2^hs, when every scaled sum is below 2^15 (no
        // overflow) and a multiple of 2^-24 (binary16 keeps any such sum below 2^-14, which
        // has at most 10 significant bits, so the accumulator keeps it too; from 2^-14 up both
        // round to nearest-even at 11 significant bits):
        //   binary16-shaped format: hs = 0 (its own grid; sum_ehi is 15);
        //   10-bit mantissa-only:   hs = 15 - S, which needs Lf + hs >= -24, i.e.
        //                           S - Lf <= 39, and b 2^hs FP32-normal for every nonzero b.
        //
        // When every A value and every scaled B value is a binary16 value as well, the
        // binary16 FMA h <- binary16(-a (b 2^hs) + h) computes the same leaf, two outputs'
        // worth per packed instruction.
        //
        // Chunk 1, products exact or not: the directed-rounding leaves (leaves_dr16) round the
        // reference's RTO24 sum exactly for a mantissa-only accumulator (any sum, any width)
        // and for a binary16-shaped one while every sum stays below 2^15; a lane whose
        // accumulators hold a NaN keeps the general leaf.
        bool sum_ok = false, rto_ok = false, h_ok = false, pk_ok = false, dr_ok = false;
        bool h_chk = false;   // chunk 1: the checked held leaves (kChk)
        int hs = 0;
        const bool f16 = g.sum_f16;
        const bool try_fast = sfin && g.range_f != 0u && (exact || (CHUNK > 1 && g.sum_rto)) &&
                              (g.sum_off & 3) != 3;
        // chunk sums: mantissa-only accumulators only (the DRC kernels' accumulators)
        const bool try_dr = kDr && sfin && (CHUNK == 1 ? g.dr_kind != 0 : g.dr_kind == 1) &&
                            !(g.sum_off & 16);
        if (try_fast || try_dr) {
            uint32_t amax = 0u, mnk = 0xFFFFFFFFu;
#pragma unroll
            for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const uint32_t b = __float_as_uint(nacc[bi][bj][r]) & 0x7FFFFFFFu;
                        amax = max(amax, b);
                        mnk = min(mnk, b - 1u);            // zeros wrap to ~0
                    }
            const int ea = amax != 0u ? (int)(amax >> 23) - 126 : -1000;   // inf / NaN: 129
            const int lsb_acc = mnk != 0xFFFFFFFFu
                                    ? max(g.sum_q, (int)((mnk + 1u) >> 23) - 127 - g.sum_m)
                                    : 1000;
            const int Lf = min(sL, lsb_acc);
            const int pt = exact ? sTop : sTop + 1;
            const float pa = ea > -1000 ? __builtin_ldexpf(1.0f, max(ea, -126)) : 0.0f;
            const float pp = sTop > -1000 ? __builtin_ldexpf(1.0f, max(pt, -126)) : 0.0f;
            const float bnd = __builtin_fmaf(g.sum_g, __builtin_fmaf(g.sum_n, pp, pa), pp) *
                              1.0000152587890625f;   // 1 + 2^-16
            const int S = (int)(__float_as_uint(bnd) >> 23) - 126;
            if (try_fast) {
            sum_ok = exact && sL >= g.sum_lreq && S <= min(sL + 24, g.sum_ehi) &&
                     sTop <= Lf + 23 && Lf >= -126 && !(g.sum_off & 1);
            // (chunk sums into a binary16-shaped accumulator take mix_sub16x16 for this leaf;
            // with those off, the general leaves)
            rto_ok = !sum_ok && g.sum_rto && S <= g.sum_ehi && !(g.sum_off & 2) &&
                     !(CHUNK >= 8 && !CONV && f16 && (g.sum_off & 32));
            }
            if constexpr (CHUNK == 1) {
              if (try_fast) {
                if (sum_ok && !(g.sum_off & 4)) {
                    if (f16) {
                        h_ok = true;
                    } else if (g.sum_h10) {
                        hs = 15 - S;
                        h_ok = S - Lf <= 39 && bLsb + hs >= -126 && bTop + 1 + hs <= 127;
                    }
                    pk_ok = h_ok && a16 && bW <= 11 && bLsb + hs >= -24 && bTop + hs <= 15 &&
                            !(g.sum_off & 8);
                }
                // binary16-shaped format on RTO lanes (sums below 2^15, maybe inexact in
                // FP32): the binary16 FMA rounds the exact sum once, which is what rounding
                // the reference's RTO24 sum to 11 significant bits gives
                if (rto_ok && f16 && a16 && bW <= 11 && bLsb >= -24 && bTop <= 15 &&
                    !(g.sum_off & 12)) {
                    pk_ok = true;
                    hs = 0;
                }
              }
            }
            dr_ok = try_dr && !sum_ok && !pk_ok && amax <= 0x7F800000u &&
                    (g.dr_kind == 1 || S <= g.sum_ehi);
            // chunk 1, binary16-shaped accumulator, every sum below 2^15: the held leaves with
            // the mixed FMA (mix_fnma16) round the exact a b + h once, so they take the lanes
            // whose products or sums are inexact in FP32 as well
            if constexpr (CHUNK == 1) {
                if (dr_ok && g.dr_kind == 2 && !(g.sum_off & 32)) {
                    h_ok = true;
                    hs = 0;
                    dr_ok = false;
                }
            }
            if constexpr (kChk) {
                // 10-bit mantissa-only accumulator, products or sums inexact in FP32: held
                // leaves on h = nacc 2^hs, checked.  Every scaled sum is below 2^15; where one
                // is at least 2^-14 binary16 rounds it at 11 significant bits as the
                // accumulator does, and a leaf result at most 2^-14 in magnitude (the only
                // ones whose exact sum can lie below, underflows to zero included) sends the
                // lanes of the wave on this path back over the slice with the directed-
                // rounding leaves.  The start values must convert exactly (each nonzero one
                // at least 2^-14 scaled) and be nonzero (zero start values would flag the
                // slice's leading zero products).  The accumulator rounds FP32-subnormal
                // values on a fixed grid and overflows past FP32's range, so every sum must
                // stay below 2^127 (S <= 127) and the check must also catch every sum below
                // 2^-126 (2^-14 scaled is at least 2^-126 when hs <= 112).
                if (dr_ok && g.sum_h10 && amax < 0x7F800000u && !(g.sum_off & 32)) {
                    const int h10 = 15 - S;
                    const int emin = mnk != 0xFFFFFFFFu ? (int)((mnk + 1u) >> 23) - 127 : 1000;
                    uint32_t amin = 0xFFFFFFFFu;
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                            for (int r = 0; r < 16; ++r)
                                amin = min(amin, __float_as_uint(nacc[bi][bj][r]) & 0x7FFFFFFFu);
                    if (amin != 0u && S <= g.sum_ehi && h10 <= 112 && bLsb + h10 >= -126 &&
                        bTop + 1 + h10 <= 127 && emin + h10 >= -14) {
                        h_chk = true;
                        hs = h10;
                        dr_ok = false;
                    }
                }
            }
        }

        const int kmax = min(BK, g.K - k0);
        if constexpr (CHUNK == 1) {
          set_form(sum_ok || pk_ok || h_ok || h_chk || (rto_ok && !dr_ok));
          // nacc <- -round(a * b + acc): exact sums (fma(-a, b, nacc) is -(a * b + acc)
          // exactly) or RTO leaves on the exact product a * b
          auto leaves_c1 = [&](auto f16, auto rto) {
            constexpr bool F16 = decltype(f16)::value, RTO = decltype(rto)::value;
            for (int kk = 0; kk < kmax; ++kk) {
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
                            float* na = &nacc[bi][bj][rq * 4];
                            const float av[4] = {a4[bi][rq].x, a4[bi][rq].y, a4[bi][rq].z,
                                                 a4[bi][rq].w};
#pragma unroll
                            for (int x = 0; x < 4; ++x) {
                                if constexpr (RTO)
                                    na[x] = leaf_rto<F16>(av[x] * b[bj], na[x], g);
                                else
                                    na[x] = leaf_exact<F16>(__builtin_fmaf(-av[x], b[bj], na[x]), g);
                            }
                        }
            }
          };
          // pairs of rows: hp <- binary16 fma(-a, b 2^hs, hp), hp = nacc 2^hs; SCALED false
          // for the binary16-shaped format, whose hs is always 0
          auto pk_leaves = [&](auto scaled) {
            constexpr bool SCALED = decltype(scaled)::value;
            auto sc = [&](float v, int e) {
                if constexpr (SCALED) return __builtin_ldexpf(v, e);
                else return v;
            };
            f16x2 hp[WMB][WNB][8];
#pragma unroll
            for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                    for (int r = 0; r < 8; ++r)
                        hp[bi][bj][r] = f16x2{(_Float16)sc(nacc[bi][bj][2 * r], hs),
                                              (_Float16)sc(nacc[bi][bj][2 * r + 1], hs)};
            for (int kk = 0; kk < kmax; ++kk) {
                float4 a4[WMB][4];
                f16x2 ah[WMB][8], bh[WNB];
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int rq = 0; rq < 4; ++rq)
                        a4[bi][rq] = *(const float4*)&As[kk * LDA_S + am0 + bi * 32 + rq * 8 + lhi * 4];
#pragma unroll
                for (int bj = 0; bj < WNB; ++bj) {
                    const _Float16 v = (_Float16)sc(Bs[kk * LDB_S + bn0 + bj * 32 + l32], hs);
                    bh[bj] = f16x2{v, v};
                }
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int rq = 0; rq < 4; ++rq) {
                        ah[bi][2 * rq] = f16x2{(_Float16)-a4[bi][rq].x, (_Float16)-a4[bi][rq].y};
                        ah[bi][2 * rq + 1] = f16x2{(_Float16)-a4[bi][rq].z, (_Float16)-a4[bi][rq].w};
                    }
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                        for (int r = 0; r < 8; ++r)
                            hp[bi][bj][r] = __builtin_elementwise_fma(ah[bi][r], bh[bj], hp[bi][bj][r]);
            }
#pragma unroll
            for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        nacc[bi][bj][2 * r] = sc((float)hp[bi][bj][r].x, -hs);
                        nacc[bi][bj][2 * r + 1] = sc((float)hp[bi][bj][r].y, -hs);
                    }
          };
          if (pk_ok) {
            if (f16) pk_leaves(std::false_type{});
            else pk_leaves(std::true_type{});
          } else if (kChk && h_chk) {   // the held leaves below, checked (see the lane test)
            _Float16 h[WMB][WNB][16];
            bool bad = false;
#pragma unroll
            for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        h[bi][bj][r] = (_Float16)__builtin_ldexpf(nacc[bi][bj][r], hs);
            for (int kk = 0; kk < kmax; ++kk) {
                float4 a4[WMB][4];
                float b[WNB];
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int rq = 0; rq < 4; ++rq)
                        a4[bi][rq] = *(const float4*)&As[kk * LDA_S + am0 + bi * 32 + rq * 8 + lhi * 4];
#pragma unroll
                for (int bj = 0; bj < WNB; ++bj)
                    b[bj] = __builtin_ldexpf(Bs[kk * LDB_S + bn0 + bj * 32 + l32], hs);
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                        for (int rq = 0; rq < 4; ++rq) {
                            _Float16* hv = &h[bi][bj][rq * 4];
                            const float av[4] = {a4[bi][rq].x, a4[bi][rq].y, a4[bi][rq].z,
                                                 a4[bi][rq].w};
#pragma unroll
                            for (int x = 0; x < 4; ++x) {
                                hv[x] = mix_fnma16(av[x], b[bj], hv[x]);
                                bad |= (hv[x] <= (_Float16)6.103515625e-05f) &
                                       (hv[x] >= (_Float16)-6.103515625e-05f);
                            }
                        }
            }
            if (__builtin_amdgcn_ballot_w64(bad) == 0) {
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                        for (int r = 0; r < 16; ++r)
                            nacc[bi][bj][r] = __builtin_ldexpf((float)h[bi][bj][r], -hs);
            } else {   // a result within binary16's subnormal range: the directed-rounding
                       // leaves over the slice from the unchanged start values, plain form
                set_form(false);
                for (int kk = 0; kk < kmax; ++kk) {
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
                    for (int bi = 0; bi < WMB; ++bi) {
                        const float av[16] = {a4[bi][0].x, a4[bi][0].y, a4[bi][0].z, a4[bi][0].w,
                                              a4[bi][1].x, a4[bi][1].y, a4[bi][1].z, a4[bi][1].w,
                                              a4[bi][2].x, a4[bi][2].y, a4[bi][2].z, a4[bi][2].w,
                                              a4[bi][3].x, a4[bi][3].y, a4[bi][3].z, a4[bi][3].w};
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj) leaves_dr16<false>(av, b[bj], nacc[bi][bj], g);
                    }
                }
            }
          } else if (h_ok) {   // h <- binary16(h - a b 2^hs), h = nacc 2^hs (exact both ways)
            _Float16 h[WMB][WNB][16];
#pragma unroll
            for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        h[bi][bj][r] = (_Float16)__builtin_ldexpf(nacc[bi][bj][r], hs);
            for (int kk = 0; kk < kmax; ++kk) {
                float4 a4[WMB][4];
                float b[WNB];
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int rq = 0; rq < 4; ++rq)
                        a4[bi][rq] = *(const float4*)&As[kk * LDA_S + am0 + bi * 32 + rq * 8 + lhi * 4];
#pragma unroll
                for (int bj = 0; bj < WNB; ++bj)
                    b[bj] = __builtin_ldexpf(Bs[kk * LDB_S + bn0 + bj * 32 + l32], hs);
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                        for (int rq = 0; rq < 4; ++rq) {
                            _Float16* hv = &h[bi][bj][rq * 4];
                            const float av[4] = {a4[bi][rq].x, a4[bi][rq].y, a4[bi][rq].z,
                                                 a4[bi][rq].w};
#pragma unroll
                            for (int x = 0; x < 4; ++x)
                                hv[x] = mix_fnma16(av[x], b[bj], hv[x]);
                        }
            }
#pragma unroll
            for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        nacc[bi][bj][r] = __builtin_ldexpf((float)h[bi][bj][r], -hs);
          } else if (sum_ok) {
            if (f16) leaves_c1(std::true_type{}, std::false_type{});
            else leaves_c1(std::false_type{}, std::false_type{});
          } else if (dr_ok) {   // acc <- round(RTO24(a * b + acc)), plain form
            auto dr_leaves = [&](auto f16) {
              constexpr bool F16 = decltype(f16)::value;
              for (int kk = 0; kk < kmax; ++kk) {
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
                for (int bi = 0; bi < WMB; ++bi) {
                    const float av[16] = {a4[bi][0].x, a4[bi][0].y, a4[bi][0].z, a4[bi][0].w,
                                          a4[bi][1].x, a4[bi][1].y, a4[bi][1].z, a4[bi][1].w,
                                          a4[bi][2].x, a4[bi][2].y, a4[bi][2].z, a4[bi][2].w,
                                          a4[bi][3].x, a4[bi][3].y, a4[bi][3].z, a4[bi][3].w};
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj) leaves_dr16<F16>(av, b[bj], nacc[bi][bj], g);
                }
              }
            };
            if (g.dr_kind == 2) dr_leaves(std::true_type{});
            else dr_leaves(std::false_type{});
          } else if (rto_ok) {
            if (f16) leaves_c1(std::true_type{}, std::true_type{});
            else leaves_c1(std::false_type{}, std::true_type{});
          } else {
            float (&acc)[WMB][WNB][16] = nacc;
            // acc <- round(RTO_fma(a, b, acc)) for every k: float products where the slice's
            // products are exact in FP32, double products otherwise
            for (int kk = 0; kk < kmax; ++kk) {
                float4 a4[WMB][4];
                float b[WNB];
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int rq = 0; rq < 4; ++rq)
                        a4[bi][rq] = *(const float4*)&As[kk * LDA_S + am0 + bi * 32 + rq * 8 + lhi * 4];
#pragma unroll
                for (int bj = 0; bj < WNB; ++bj) b[bj] = Bs[kk * LDB_S + bn0 + bj * 32 + l32];
                if (exact) {
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj) {
                            f32x16 pv;
#pragma unroll
                            for (int rq = 0; rq < 4; ++rq) {
                                pv[rq * 4 + 0] = __fmul_rn(a4[bi][rq].x, b[bj]);
                                pv[rq * 4 + 1] = __fmul_rn(a4[bi][rq].y, b[bj]);
                                pv[rq * 4 + 2] = __fmul_rn(a4[bi][rq].z, b[bj]);
                                pv[rq * 4 + 3] = __fmul_rn(a4[bi][rq].w, b[bj]);
                            }
                            leaves_f16<SUBN, kLeafByVal>(pv, acc[bi][bj], g);
                        }
                } else {   // products wider than FP32: exact in double, one leaf at a time
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj) {
                            const double bd = (double)b[bj];
                            const float av[16] = {a4[bi][0].x, a4[bi][0].y, a4[bi][0].z, a4[bi][0].w,
                                                  a4[bi][1].x, a4[bi][1].y, a4[bi][1].z, a4[bi][1].w,
                                                  a4[bi][2].x, a4[bi][2].y, a4[bi][2].z, a4[bi][2].w,
                                                  a4[bi][3].x, a4[bi][3].y, a4[bi][3].z, a4[bi][3].w};
#pragma unroll
                            for (int o = 0; o < 16; ++o)
                                acc[bi][bj][o] = (float)leaf<kLeafByVal>((double)av[o] * bd,
                                                             (double)acc[bi][bj][o], g);
                        }
                }
            }
          }
        } else {
        // chunk sums: exact-sum leaves, else the directed-rounding leaves (DRC) or the TwoSum
        // round-to-odd leaves, else the general leaves
        set_form(sum_ok || dr_ok || (!kDr && rto_ok));
        float (&acc)[WMB][WNB][16] = nacc;
#pragma unroll
        for (int q = 0; q < BK / CHUNK; ++q) {
            const int kb = q * CHUNK;
            if (kb >= kmax) break;
            // Chunks of 8 and more sum on the bf16 MFMAs only (an exact slice whose values are
            // not all bfloat16 values takes the chain); chunks of 2 and 4 on the f32 MFMA.  One
            // kernel holding both MFMA kinds spills heavily at three waves per SIMD.
            if (exact && (!kBf16 || bf || g.mfma_mode == 2)) {
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                        for (int r = 0; r < 16; ++r) part[bi][bj][r] = 0.0f;
                if constexpr (kBf16) {
                    // Every A and B value of the slice is a bfloat16 value and every partial sum of
                    // the chunk is exact in FP32, so the bf16 MFMAs' products and sums are exact
                    // too: 16 k per v_mfma_f32_32x32x16_bf16 (lane l: row / column l % 32, k
                    // 8 (l / 32) .. + 7), a chunk of 8 on one v_mfma_f32_32x32x8_bf16 (k
                    // 4 (l / 32) .. + 3).
                    if constexpr (CHUNK >= 16) {
#pragma unroll
                        for (int kq = 0; kq < CHUNK / 16; ++kq) {
                            __builtin_amdgcn_sched_barrier(0);   // one k group's operands live at a time
                            const int kk = kb + 16 * kq + 8 * lhi;
                            bf16x8 av[WMB], bv[WNB];
#pragma unroll
                            for (int bi = 0; bi < WMB; ++bi)
                                av[bi] = lds_bf16x8(&As[kk * LDA_S + am0 + bi * 32 + l32], LDA_S);
#pragma unroll
                            for (int bj = 0; bj < WNB; ++bj)
                                bv[bj] = lds_bf16x8(&Bs[kk * LDB_S + bn0 + bj * 32 + l32], LDB_S);
#pragma unroll
                            for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                                for (int bj = 0; bj < WNB; ++bj)
                                    part[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                        av[bi], bv[bj], part[bi][bj], 0, 0, 0);
                        }
                    } else {
                        static_assert(CHUNK == 8, "one 32x32x8 MFMA per chunk");
                        float ea[4][WMB], eb[4][WNB];
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int kk = kb + 4 * lhi + j;
#pragma unroll
                            for (int bi = 0; bi < WMB; ++bi) ea[j][bi] = As[kk * LDA_S + am0 + bi * 32 + l32];
#pragma unroll
                            for (int bj = 0; bj < WNB; ++bj) eb[j][bj] = Bs[kk * LDB_S + bn0 + bj * 32 + l32];
                        }
                        bf16x4 av[WMB], bv[WNB];
#pragma unroll
                        for (int bi = 0; bi < WMB; ++bi)
                            av[bi] = __builtin_bit_cast(bf16x4, u32x2{bf16_pair(ea[0][bi], ea[1][bi]),
                                                                      bf16_pair(ea[2][bi], ea[3][bi])});
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj)
                            bv[bj] = __builtin_bit_cast(bf16x4, u32x2{bf16_pair(eb[0][bj], eb[1][bj]),
                                                                      bf16_pair(eb[2][bj], eb[3][bj])});
#pragma unroll
                        for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                            for (int bj = 0; bj < WNB; ++bj)
                                part[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(
                                    av[bi], bv[bj], part[bi][bj], 0, 0, 0);
                    }
                } else {
                    float a[CHUNK / 2][WMB], b[CHUNK / 2][WNB];
#pragma unroll
                    for (int kp = 0; kp < CHUNK / 2; ++kp) {
                        const int kk = kb + 2 * kp + lhi;
#pragma unroll
                        for (int bi = 0; bi < WMB; ++bi) a[kp][bi] = As[kk * LDA_S + am0 + bi * 32 + l32];
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj) b[kp][bj] = Bs[kk * LDB_S + bn0 + bj * 32 + l32];
                    }
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
                }
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                        for (int r = 0; r < 16; ++r)   // an exact zero sum is +0, as the chain gives
                            part[bi][bj][r] = part[bi][bj][r] + 0.0f;
            } else {
                // the fmaf chain from +0: its first k is fma(a, b, +0), which needs no zeroed
                // registers (kb < kmax here, so that k exists)
                auto chain_k = [&](int kk, auto first) {
                    constexpr bool FIRST = decltype(first)::value;
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
                                const float av[4] = {a4[bi][rq].x, a4[bi][rq].y, a4[bi][rq].z,
                                                     a4[bi][rq].w};
#pragma unroll
                                for (int x = 0; x < 4; ++x)
                                    part[bi][bj][rq * 4 + x] = __builtin_fmaf(
                                        av[x], b[bj], FIRST ? 0.0f : part[bi][bj][rq * 4 + x]);
                            }
                };
                chain_k(kb, std::true_type{});
#pragma unroll
                for (int kq = 1; kq < CHUNK; ++kq) {
                    const int kk = kb + kq;
                    if (kk >= kmax) break;
                    chain_k(kk, std::false_type{});
                }
            }
            if constexpr (CHUNK >= 8 && !CONV) {
            if (sum_ok) {   // nacc <- -round(p + acc) = round(nacc - p), the sum exact
                if (kF16Leaves && f16) {   // one tied asm block per 16 outputs: no new registers
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj) exact_sub16x16(part[bi][bj], nacc[bi][bj]);
                } else {
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                            for (int r = 0; r < 16; ++r)
                                nacc[bi][bj][r] = leaf_exact<false>(nacc[bi][bj][r] - part[bi][bj][r], g);
                }
            } else if (kF16Leaves && f16 && rto_ok) {
                // binary16-shaped, every sum below 2^15, sums inexact in FP32: nacc <-
                // binary16(nacc - p), one rounding of the exact difference
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj) mix_sub16x16(part[bi][bj], nacc[bi][bj]);
            } else if (kDr && dr_ok) {   // nacc <- -round(RTO24(p + acc)), directed roundings
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj) leaves_dr16_neg(part[bi][bj], nacc[bi][bj], g);
            } else if (!kDr && rto_ok) {   // nacc <- -round(RTO24(p + acc)), mantissa-only
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                        for (int r = 0; r < 16; ++r)
                            nacc[bi][bj][r] = leaf_rto<false>(part[bi][bj][r], nacc[bi][bj][r], g);
            } else {
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj) leaves_f16<SUBN, kLeafByVal>(part[bi][bj], acc[bi][bj], g);
            }
            } else {
            // chunks of 2 and 4 (16 and 8 chunks per slice, unrolled), which measured faster
            // with the compiler's own code for both binary16 leaves, and the convolution tiles
            // (with the asm leaves the 64x64 chunk-8 one spills 146 registers)
            if (sum_ok) {   // nacc <- -round(p + acc) = round(nacc - p), the sum exact
                if (f16) {
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                            for (int r = 0; r < 16; ++r)
                                nacc[bi][bj][r] = leaf_exact<true>(nacc[bi][bj][r] - part[bi][bj][r], g);
                } else {
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                            for (int r = 0; r < 16; ++r)
                                nacc[bi][bj][r] = leaf_exact<false>(nacc[bi][bj][r] - part[bi][bj][r], g);
                }
            } else if (rto_ok) {   // nacc <- -round(RTO24(p + acc)) for any in-range sum
                if (f16) {
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                            for (int r = 0; r < 16; ++r)
                                nacc[bi][bj][r] = leaf_rto<true>(part[bi][bj][r], nacc[bi][bj][r], g);
                } else {
#pragma unroll
                    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                        for (int bj = 0; bj < WNB; ++bj)
#pragma unroll
                            for (int r = 0; r < 16; ++r)
                                nacc[bi][bj][r] = leaf_rto<false>(part[bi][bj][r], nacc[bi][bj][r], g);
                }
            } else {
#pragma unroll
                for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < WNB; ++bj) leaves_f16<SUBN, kLeafByVal>(part[bi][bj], acc[bi][bj], g);
            }
            }
        }
        }
        if (has_next) store_slice(buf ^ 1);
    }

    // epilogue (cutlass LinearCombination): D = alpha * acc, or fma(alpha, acc, beta * C);
    // a convolution's D(m, n) / C(m, n) sit at b * sdb + m * sdm + (p * Q + q) * sdpq
#pragma unroll
    for (int bi = 0; bi < WMB; ++bi)
#pragma unroll
        for (int bj = 0; bj < WNB; ++bj) {
            const int gn = n0 + bn0 + bj * 32 + l32;
            int64_t col = gn, row_stride_d = g.ldd, row_stride_c = g.ldc;
            if constexpr (CONV) {
                const uint32_t b = fast_div((uint32_t)gn, g.cv.div_pq);
                col = (int64_t)b * g.cv.sdb + (int64_t)((uint32_t)gn - b * g.cv.PQ) * g.cv.sdpq;
                row_stride_d = row_stride_c = g.cv.sdm;
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gm = m0 + am0 + bi * 32 + (r >> 2) * 8 + lhi * 4 + (r & 3);
                if (gm >= g.M || gn >= g.N) continue;
                const float a = negd ? -nacc[bi][bj][r] : nacc[bi][bj][r];
                float d;
                if (g.beta != 0.0f) d = __builtin_fmaf(g.alpha, a, g.beta * gld(g.C + (int64_t)gm * row_stride_c + col));
                else d = g.alpha * a;
                gst(g.D + (int64_t)gm * row_stride_d + col, d);
            }
        }
}

}  // namespace lofamd
