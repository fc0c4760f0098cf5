// lof_host.hpp -- host-side helpers shared by the C-ABI translation units:
// thread-local error text, argument validation and FmtDev precomputation.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "lof_core.hpp"
#include "lofloat.h"

namespace lofamd {

lof_status set_error(lof_status st, const char* fmt, ...);
lof_status check_hip(hipError_t e, const char* what);

inline float bits_to_float(uint32_t b) { float f; std::memcpy(&f, &b, 4); return f; }
inline uint32_t float_to_bits(float f) { uint32_t b; std::memcpy(&b, &f, 4); return b; }

inline bool mode_valid(int m) { return m >= 0 && m <= 14; }

// Exponent of the largest finite binade of a format (lo_float.h:2857-2859).
inline int fmt_max_exp(const lof_format& f) {
    return f.signedness == LOF_SIGNED ? (1 << (f.bitwidth - f.mantissa_bits - 1)) - 1 - f.bias
                                      : (1 << (f.bitwidth - f.mantissa_bits)) - 1 - f.bias;
}

inline lof_status validate_format(const lof_format* f, const char* what) {
    if (!f) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: format is NULL", what);
    if (f->mantissa_bits < 0 || f->mantissa_bits > 23)
        return set_error(LOF_ERR_UNSUPPORTED, "%s: mantissa_bits %d outside [0, 23]", what,
                         f->mantissa_bits);
    const int ew = f->bitwidth - f->mantissa_bits - (f->signedness == LOF_SIGNED ? 1 : 0);
    if (ew < 0 || ew > 30)
        return set_error(LOF_ERR_UNSUPPORTED, "%s: exponent field width %d outside [0, 30]",
                         what, ew);
    if (f->bias < -100 || f->bias > 100000)
        return set_error(LOF_ERR_UNSUPPORTED, "%s: bias %d outside [-100, 100000]", what,
                         f->bias);
    if (f->inf_behavior != LOF_INF_EXTENDED && f->inf_behavior != LOF_INF_SATURATING)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad inf_behavior %d", what,
                         f->inf_behavior);
    if (f->signedness != LOF_SIGNED && f->signedness != LOF_UNSIGNED)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad signedness %d", what, f->signedness);
    return LOF_OK;
}

inline lof_status validate_spec(lof_round_spec s, int dropped_bits, const char* what) {
    if (!mode_valid(s.mode))
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: rounding mode %d outside [0, 14]", what,
                         s.mode);
    if (uses_len(s.mode) && (s.stoch_len < 0 || s.stoch_len > dropped_bits))
        return set_error(LOF_ERR_INVALID_ARGUMENT,
                         "%s: stoch_len %d must be in [0, %d] (the dropped mantissa bits)", what,
                         s.stoch_len, dropped_bits);
    return LOF_OK;
}

// FloatingPointParams + ProjSpec -> FmtDev (lo_float.h:2857-2861).
inline FmtDev make_fmt_dev(const lof_format& f, lof_round_spec s) {
    FmtDev P{};
    const int mb = f.mantissa_bits;
    const int to_max = fmt_max_exp(f);
    const int to_min = 1 - f.bias;
    // (float)(2^ToMin * 2^-mb): correctly rounded power of two (0 or inf at the extremes)
    const float min_val = (float)std::ldexp(1.0, to_min - mb);
    const float half = (float)((double)min_val * 0.5);  // == min_val * 0.5f (one rounding)
    P.r = 23 - mb;
    P.emin_b = to_min + 127;
    P.emax_b = to_max + 127;
    P.min_bits = float_to_bits(min_val);
    P.half_bits = float_to_bits(half);
    P.min_val = min_val;
    P.saturating = f.inf_behavior == LOF_INF_SATURATING;
    P.stoch_len = s.stoch_len;
    P.mode = s.mode;
    return P;
}

// Largest finite normal of an MX element format as the reference's CPU binding computes
// it in float: ldexpf(1, emax) * (2 - ldexpf(1, -mb)) (LoPy_bind.cpp:199-202).
inline float mx_private_max_normal(const lof_format& f) {
    const int me = fmt_max_exp(f);
    const float a = (float)std::ldexp(1.0, me);
    const float b = (float)(2.0 - std::ldexp(1.0, -f.mantissa_bits));
    return (float)((double)a * (double)b);
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace lofamd
