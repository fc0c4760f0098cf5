// lof_silu.hip -- piecewise-linear SiLU (C-ABI lof_pwl_silu).
//
// Replaces pwl_silu_kernel / pwl_silu (Lof_kernel.cu:231-265) behind pwl_silu_forward
// (LoPy_bind.cpp:228-277): y = lerp of an (N+1)-knot table on [-R, R] for |x| <= R, relu(x)
// outside, in float arithmetic whatever the element type:
//   xc = fminf(fmaxf(x, -R), R); u = (xc + R) * inv_step; i = min((int)u, N - 1);
//   y = fmaf(u - i, lut[i + 1] - lut[i], lut[i]);  y = |x| > R ? relu(x) : y
// with inv_step = (float)N / (2 * R), every operation rounded as written (no contraction:
// the reference's CPU loop, the literal reading of its kernel).  (fmaxf returns -R for a
// NaN x, so NaN maps to lut[0].)
// HBM-bound: one workgroup per kTile elements stages the table (<= 4097 floats) in LDS once,
// then every lane reads and writes 16-byte vectors of the element type.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lof_host.hpp"
#include "lofloat.h"

namespace lofamd {

namespace {

constexpr int kSiluTPB = 256;
constexpr int kSiluPer = 16;                       // elements per thread
constexpr int kSiluTile = kSiluTPB * kSiluPer;
constexpr int kMaxSegments = 4096;

struct SiluArgs {
    const void* in;
    void* out;
    const float* lut;
    int64_t n;
    int32_t segments;
    float R, inv_step;
    int32_t vec;                                   // in / out 16-byte aligned
};

// Each operation rounds on its own, as written: fusing u's product into frac = u - i
// (which contraction would do) changes the results.
__device__ __forceinline__ float pwl(float x, const float* t, int N, float R, float inv_step) {
#pragma clang fp contract(off)
    const float xc = fminf(fmaxf(x, -R), R);
    const float u = (xc + R) * inv_step;
    int i = (int)u;
    i = i > N - 1 ? N - 1 : i;
    const float frac = u - (float)i;
    const float y0 = t[i], y1 = t[i + 1];
    const float y = __builtin_fmaf(frac, y1 - y0, y0);
    return fabsf(x) > R ? (x > 0.0f ? x : 0.0f) : y;
}

template <typename T> __device__ __forceinline__ float to_f(T v) { return (float)v; }
template <> __device__ __forceinline__ float to_f<__half>(__half v) { return __half2float(v); }
template <typename T> __device__ __forceinline__ T from_f(float v) { return (T)v; }
template <> __device__ __forceinline__ __half from_f<__half>(float v) { return __float2half_rn(v); }

template <typename T>
__global__ __launch_bounds__(kSiluTPB) void k_pwl_silu(SiluArgs a) {
    __shared__ float tab[kMaxSegments + 1];
    const int tid = threadIdx.x;
    const int N = a.segments;
    for (int j = tid; j <= N; j += kSiluTPB) tab[j] = a.lut[j];
    __syncthreads();
    const T* in = static_cast<const T*>(a.in);
    T* out = static_cast<T*>(a.out);
    constexpr int V = 16 / (int)sizeof(T);         // elements per 16-byte vector
    constexpr int NV = kSiluPer / V;               // vectors per thread
    const int64_t base = (int64_t)blockIdx.x * kSiluTile;
    if (a.vec && base + kSiluTile <= a.n) {
        uint4 v[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k)               // vector k * 256 + tid of the tile
            v[k] = reinterpret_cast<const uint4*>(in + base)[k * kSiluTPB + tid];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            T e[V];
            __builtin_memcpy(e, &v[k], 16);
#pragma unroll
            for (int i = 0; i < V; ++i) e[i] = from_f<T>(pwl(to_f<T>(e[i]), tab, N, a.R, a.inv_step));
            __builtin_memcpy(&v[k], e, 16);
            reinterpret_cast<uint4*>(out + base)[k * kSiluTPB + tid] = v[k];
        }
    } else {
        for (int64_t e = base + tid; e < a.n && e < base + kSiluTile; e += kSiluTPB)
            out[e] = from_f<T>(pwl(to_f<T>(in[e]), tab, N, a.R, a.inv_step));
    }
}

}  // namespace

}  // namespace lofamd

using namespace lofamd;

extern "C" {

LOF_API lof_status lof_pwl_silu(void* out, const void* in, int64_t n, int32_t dtype,
                                const float* lut, int32_t segments, float R,
                                lof_stream_t stream) {
    const char* what = "lof_pwl_silu";
    if (segments < 2 || segments > kMaxSegments || (segments & (segments - 1)) != 0)
        return set_error(LOF_ERR_INVALID_ARGUMENT,
                         "%s: lut length %d is not a supported power of 2 in [2, 4096]", what,
                         segments);
    if (!(R > 0.0f)) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: R must be positive", what);
    if (n < 0) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: n < 0", what);
    if (n == 0) return LOF_OK;
    if (!out || !in || !lut) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: NULL buffer", what);
    if ((n + kSiluTile - 1) / kSiluTile >= ((int64_t)1 << 31))
        return set_error(LOF_ERR_UNSUPPORTED, "%s: input too large", what);
    SiluArgs a{};
    a.in = in;
    a.out = out;
    a.lut = lut;
    a.n = n;
    a.segments = segments;
    a.R = R;
    a.inv_step = (float)segments / (2.0f * R);     // as pwl_silu (Lof_kernel.cu:261)
    a.vec = (((uintptr_t)in | (uintptr_t)out) & 15u) == 0 ? 1 : 0;
    const dim3 grid((unsigned)((n + kSiluTile - 1) / kSiluTile));
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
    case LOF_DTYPE_F32: hipLaunchKernelGGL(k_pwl_silu<float>, grid, dim3(kSiluTPB), 0, st, a); break;
    case LOF_DTYPE_F16: hipLaunchKernelGGL(k_pwl_silu<__half>, grid, dim3(kSiluTPB), 0, st, a); break;
    case LOF_DTYPE_F64: hipLaunchKernelGGL(k_pwl_silu<double>, grid, dim3(kSiluTPB), 0, st, a); break;
    default: return set_error(LOF_ERR_UNSUPPORTED, "%s: dtype %d not supported", what, dtype);
    }
    return check_hip(hipGetLastError(), "lof_pwl_silu launch");
}

}  // extern "C"
