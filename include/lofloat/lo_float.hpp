// lo_float.hpp -- C++ host API of the MI355X LoFloat hot path, over the C-ABI in
// lofloat.h (link liblofloat_hip.so).
//
// Mirrors the reference's C++ interface for this path (SudhanvaKulkarni123/LoFloat):
//   enums + ProjSpec + FloatingPointParams      src/fp_tools.hpp:13-124, 187-252
//   round_mantissa / round_fp_params             src/LoPy_bind.cpp:13-20, Lof_kernel.cu:37-60
//   run_virtual_mx_round                          src/Lof_kernels.h:44-48, Lof_kernel.cu:70-94
//   virtual_round over arrays (both overloads)    src/lo_float.h:4264-4268, 4276-4281
//   virtual_mx_round                              src/lo_float.h:4311-4383
//   Layout, Matrix<T, idx, Layout>                src/layouts.h:16-19, src/Matrix.h:21-78
//   Gemm<MatA, MatB, MatC, MatD>                  src/cutlass_gemms.cuh:1629-1896
//   Conv2d<...>, Conv2dProblemSize                src/cutlass_conv.cuh:20-184
//   Project over rep arrays (typed cast)          src/lo_float.h:3979-3988, 4014-4022
//   PackEndian, LoF_packed_size, LoF_encode_cuda, LoF_decode_cuda and the checker
//   functors they read                           src/lo_pack.hpp:50, 392-531; lo_float.h:1411-1789
// Same names, argument meanings and enum values.  Differences: every launcher takes an
// optional hipStream_t (default: the legacy default stream, as the reference launches), an
// optional per-element random-word array and seed for the stochastic modes, and reports
// failures by throwing lo_float::Error (the reference's launchers report nothing); the
// Gemm never allocates a workspace and never synchronises the host.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "../lofloat.h"

namespace lo_float {

enum class Rounding_Mode : uint8_t {
    RoundToNearestEven, RoundTowardsZero, RoundAwayFromZero, StochasticRoundingA,
    RoundToNearestOdd, RoundDown, RoundUp, RoundTiesToAway, StochasticRoundingB,
    StochasticRoundingC, True_StochasticRounding, ProbabilisticRounding, RoundToOdd,
    RoundTiesTowardsZero, StochasticRoundingD
};
enum Saturation_Mode : uint8_t { OvfInf, SatFinite, SatPropagate };
enum Signedness : uint8_t { Signed, Unsigned };
enum Unsigned_behavior : uint8_t { NegtoZero, NegtoNaN };
enum Inf_Behaviors : uint8_t { Extended, Saturating };
enum NaN_Behaviors : uint8_t { _3109, _754 };
enum Layout : uint8_t { ColMajor = 0, RowMajor = 1 };

struct ProjSpec {
    Rounding_Mode rounding_mode;
    Saturation_Mode saturation_mode;
    int stoch_length;
    constexpr explicit ProjSpec(Rounding_Mode rm = Rounding_Mode::RoundToNearestEven,
                                Saturation_Mode sm = Saturation_Mode::OvfInf, int stoch_len = 0)
        : rounding_mode(rm), saturation_mode(sm), stoch_length(stoch_len) {}
};

// Placeholder checker type: rounding never consults the Inf/NaN checkers.
struct NoChecker {};

template <class IsInfFunctor = NoChecker, class IsNaNFunctor = NoChecker>
struct FloatingPointParams {
    int bitwidth;
    int mantissa_bits;
    int bias;
    Inf_Behaviors OV_behavior;
    NaN_Behaviors NA_behavior;
    Signedness is_signed;
    IsInfFunctor IsInf;
    IsNaNFunctor IsNaN;
    Unsigned_behavior unsigned_behavior;

    constexpr FloatingPointParams(int bw, int mb, int b, Inf_Behaviors ovb, NaN_Behaviors nab,
                                  Signedness sgn, IsInfFunctor inf = IsInfFunctor{},
                                  IsNaNFunctor nan = IsNaNFunctor{},
                                  Unsigned_behavior ub = Unsigned_behavior::NegtoZero)
        : bitwidth(bw), mantissa_bits(mb), bias(b), OV_behavior(ovb), NA_behavior(nab),
          is_signed(sgn), IsInf(inf), IsNaN(nan), unsigned_behavior(ub) {}
};

class Error : public std::runtime_error {
public:
    Error(lof_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    lof_status status;
};

namespace detail {
inline void check(lof_status s, const char* what) {
    if (s != LOF_OK) throw Error(s, std::string(what) + ": " + lof_last_error());
}
template <class I, class N>
inline lof_format to_c(const FloatingPointParams<I, N>& p) {
    return lof_format{p.bitwidth, p.mantissa_bits, p.bias, (int32_t)p.OV_behavior,
                      (int32_t)p.NA_behavior, (int32_t)p.is_signed};
}
inline lof_round_spec to_c(const ProjSpec& ps) {
    return lof_round_spec{(int32_t)ps.rounding_mode, (int32_t)ps.saturation_mode,
                          ps.stoch_length};
}
}  // namespace detail

// out[i] = virtual_round(scale * in[i], mantissa_bits, ps)  (Lof_kernel.cu:37-48).
template <typename T>
void round_mantissa(const T* in, T* out, int64_t n, int mantissa_bits, ProjSpec ps, T scale,
                    hipStream_t stream = nullptr, const uint32_t* rand_bits = nullptr,
                    uint64_t seed = 0) {
    static_assert(std::is_same_v<T, float>, "the gfx950 path rounds FP32 tensors");
    detail::check(lof_virtual_round_mant_f32(in, out, n, mantissa_bits, detail::to_c(ps), scale,
                                             rand_bits, seed, (lof_stream_t)stream),
                  "round_mantissa");
}

// inout[i] = virtual_round(scale * inout[i], params, ps)  (Lof_kernel.cu:49-60).
template <typename T, class I, class N>
void round_fp_params(T* inout, int64_t n, FloatingPointParams<I, N> params, ProjSpec ps, T scale,
                     hipStream_t stream = nullptr, const uint32_t* rand_bits = nullptr,
                     uint64_t seed = 0) {
    static_assert(std::is_same_v<T, float>, "the gfx950 path rounds FP32 tensors");
    const lof_format f = detail::to_c(params);
    detail::check(lof_virtual_round_f32(inout, inout, n, &f, detail::to_c(ps), scale, rand_bits,
                                        seed, (lof_stream_t)stream),
                  "round_fp_params");
}

// In-place MX block rounding (Lof_kernel.cu:70-94); block_size a power of two in [4, 256].
template <class I1, class N1, class I2, class N2>
void run_virtual_mx_round(float* inout, int64_t n, int block_size,
                          FloatingPointParams<I1, N1> params_public,
                          FloatingPointParams<I2, N2> params_private, ProjSpec ps_public,
                          ProjSpec ps_private, hipStream_t stream = nullptr,
                          const uint32_t* rand_bits = nullptr, uint64_t seed = 0) {
    const lof_format pub = detail::to_c(params_public), priv = detail::to_c(params_private);
    detail::check(lof_virtual_mx_round_f32(inout, inout, n, block_size, &priv, &pub,
                                           detail::to_c(ps_private),
                                           (int32_t)ps_public.rounding_mode, rand_bits, seed,
                                           (lof_stream_t)stream),
                  "run_virtual_mx_round");
}

// lo_float::virtual_round over arrays (lo_float.h:4264-4268, mantissa-only overload;
// 4276-4281, format overload), with the reference's argument order: values / results are
// device arrays of n floats (they may be the same array).  The arithmetic is the scalar
// virtual_round of every element (lo_float.h:2716-2756, 2844-2939), which the reference's
// host SIMD array loop departs from (it drops signs and leaves a tail unwritten); stochastic
// modes take one random word per element (rand_bits, or lof_rand_u32(seed, i)).
template <typename From>
void virtual_round(From* values, From* results, int ToMantissaBits, int n,
                   ProjSpec ps = ProjSpec{}, hipStream_t stream = nullptr,
                   const uint32_t* rand_bits = nullptr, uint64_t seed = 0) {
    static_assert(std::is_same_v<From, float>, "the gfx950 path rounds FP32 arrays");
    if (n <= 0) return;
    detail::check(lof_virtual_round_mant_f32(values, results, n, ToMantissaBits,
                                             detail::to_c(ps), 1.0f, rand_bits, seed,
                                             (lof_stream_t)stream),
                  "virtual_round");
}

template <typename From, class I, class N>
void virtual_round(From* values, From* results, int n, FloatingPointParams<I, N> ToFp,
                   ProjSpec ps = ProjSpec{}, hipStream_t stream = nullptr,
                   const uint32_t* rand_bits = nullptr, uint64_t seed = 0) {
    static_assert(std::is_same_v<From, float>, "the gfx950 path rounds FP32 arrays");
    if (n <= 0) return;
    const lof_format f = detail::to_c(ToFp);
    detail::check(lof_virtual_round_f32(values, results, n, &f, detail::to_c(ps), 1.0f,
                                        rand_bits, seed, (lof_stream_t)stream),
                  "virtual_round");
}

// lo_float::virtual_mx_round (lo_float.h:4311-4383): in-place MX block rounding of a device
// array of n floats in contiguous blocks of block_size (the last block may be short).  Each
// block's scale is virtual_round(amax / max_normal(private), params_public) under
// ps_public's mode with no random bits; elements round into params_private under
// ps_private.  block_size must be a power of two in [4, 256] (else Error); n <= 0 or
// block_size <= 0 does nothing, as in the reference.
template <typename From, class PubInf, class PubNaN, class PrivInf, class PrivNaN>
void virtual_mx_round(From* inout, int n, int block_size,
                      FloatingPointParams<PubInf, PubNaN> params_public,
                      FloatingPointParams<PrivInf, PrivNaN> params_private,
                      ProjSpec ps_public = ProjSpec{}, ProjSpec ps_private = ProjSpec{},
                      hipStream_t stream = nullptr, const uint32_t* rand_bits = nullptr,
                      uint64_t seed = 0) {
    static_assert(std::is_same_v<From, float>, "the gfx950 path rounds FP32 arrays");
    if (block_size <= 0 || n <= 0) return;
    const lof_format pub = detail::to_c(params_public), priv = detail::to_c(params_private);
    detail::check(lof_virtual_mx_round_f32(inout, inout, n, block_size, &priv, &pub,
                                           detail::to_c(ps_private),
                                           (int32_t)ps_public.rounding_mode, rand_bits, seed,
                                           (lof_stream_t)stream),
                  "virtual_mx_round");
}

// Dense matrix view (Matrix.h:21-78): device pointer, m x n, leading dimension ld.
template <typename T, typename idx = int, Layout L = ColMajor>
class Matrix {
public:
    idx m;
    idx n;
    idx ld;
    T* data;
    static constexpr Layout layout = L;
    using scalar_type = T;
    using idx_type = idx;

    Matrix(T* data_, idx m_, idx n_, idx ld_) : m(m_), n(n_), ld(ld_), data(data_) {}
    constexpr idx get_idx(idx row, idx col) const {
        return L == ColMajor ? col * ld + row : row * ld + col;
    }
    constexpr idx rows() const { return m; }
    constexpr idx cols() const { return n; }
};

// Emulated-accumulation GEMM (cutlass_gemms.cuh:1629-1896): D = alpha * A * B + beta * C
// with the FP32 accumulator re-rounded to accum_mantissa_bits after every FMA (the
// reference leaf), in ascending k.  The CUTLASS tile parameters are accepted for source
// compatibility; tiles are chosen by the library's autotuned table instead.  Extensions:
// set_accum_format() rounds the accumulator to a full format, set_chunk() accumulates
// `chunk` products in FP32 between roundings.
template <typename MatrixA_, typename MatrixB_, typename MatrixC_, typename MatrixD_,
          int TbM = 128, int TbN = 128, int TbK = 16, int WpM = 32, int WpN = 32, int WpK = 16,
          int Stages = 3, bool SplitKSerial = false, typename Swizzle = void>
class Gemm {
public:
    using MatrixA = MatrixA_;
    using MatrixB = MatrixB_;
    using MatrixC = MatrixC_;
    using MatrixD = MatrixD_;
    static_assert(MatrixC::layout == MatrixD::layout,
                  "lo_float::Gemm requires C and D to share the same layout");
    static_assert(std::is_same_v<typename MatrixA::scalar_type, float> &&
                      std::is_same_v<typename MatrixB::scalar_type, float> &&
                      std::is_same_v<typename MatrixC::scalar_type, float> &&
                      std::is_same_v<typename MatrixD::scalar_type, float>,
                  "the gfx950 GEMM works on FP32 matrices");

    enum class Status {
        kSuccess = 0,
        kErrorNotSupported = -1,
        kErrorAllocation = -2,
        kErrorInitialize = -3,
        kErrorExecution = -4
    };

    Gemm(int accum_mantissa_bits = 23, ProjSpec proj_spec = ProjSpec{})
        : accum_mantissa_bits_(accum_mantissa_bits), proj_spec_(proj_spec) {}

    void set_accum_mantissa_bits(int bits) { accum_mantissa_bits_ = bits; has_fmt_ = false; }
    void set_proj_spec(ProjSpec spec) { proj_spec_ = spec; }
    int get_accum_mantissa_bits() const { return accum_mantissa_bits_; }
    ProjSpec get_proj_spec() const { return proj_spec_; }
    template <class I, class N>
    void set_accum_format(FloatingPointParams<I, N> f) { fmt_ = detail::to_c(f); has_fmt_ = true; }
    void set_chunk(int chunk) { chunk_ = chunk; }

    Status can_implement(const MatrixA& A, const MatrixB& B, const MatrixC& C,
                         const MatrixD& D) const {
        if (A.n != B.m || C.m != A.m || C.n != B.n || D.m != A.m || D.n != B.n)
            return Status::kErrorNotSupported;
        if (A.m < 0 || A.n < 0 || B.n < 0) return Status::kErrorNotSupported;
        return Status::kSuccess;
    }

    static size_t get_workspace_size(const MatrixA&, const MatrixB&, const MatrixC&,
                                     const MatrixD&, int = 1) {
        return 0;
    }

    // D = (A * B) / (scale_a * scale_b); a zero product means 1 (cutlass_gemms.cuh:1777-1787)
    Status run_scaled(float scale_a, float scale_b, const MatrixA& A, const MatrixB& B,
                      const MatrixC& C, MatrixD& D, int split_k_slices = 1,
                      hipStream_t stream = nullptr) const {
        const float denom = scale_a * scale_b;
        const float alpha = denom == 0.0f ? 1.0f : 1.0f / denom;
        return (*this)(alpha, 0.0f, A, B, C, D, split_k_slices, stream);
    }

    Status operator()(float alpha, float beta, const MatrixA& A, const MatrixB& B,
                      const MatrixC& C, MatrixD& D, int split_k_slices = 1,
                      hipStream_t stream = nullptr) const {
        if (split_k_slices != 1 || can_implement(A, B, C, D) != Status::kSuccess)
            return Status::kErrorNotSupported;
        lof_gemm_desc d = make_desc(alpha, beta, A, B, C, D);
        const lof_status s = lof_gemm_f32(&d, -1, (lof_stream_t)stream);
        if (s == LOF_ERR_INVALID_ARGUMENT || s == LOF_ERR_UNSUPPORTED)
            return Status::kErrorNotSupported;
        return s == LOF_OK ? Status::kSuccess : Status::kErrorExecution;
    }

    // Average device time of `reps` runs in milliseconds (negative Status on error).
    float benchmark(float alpha, float beta, const MatrixA& A, const MatrixB& B,
                    const MatrixC& C, MatrixD& D, int split_k_slices = 1, int reps = 20,
                    hipStream_t stream = nullptr) const {
        Status st = (*this)(alpha, beta, A, B, C, D, split_k_slices, stream);
        if (st != Status::kSuccess) return (float)st;
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess) return (float)Status::kErrorExecution;
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            return (float)Status::kErrorExecution;
        }
        (void)hipEventRecord(e0, stream);
        for (int i = 0; i < reps && st == Status::kSuccess; ++i)
            st = (*this)(alpha, beta, A, B, C, D, split_k_slices, stream);
        (void)hipEventRecord(e1, stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return st == Status::kSuccess ? ms / (float)reps : (float)st;
    }

private:
    int accum_mantissa_bits_;
    ProjSpec proj_spec_;
    lof_format fmt_{};
    bool has_fmt_ = false;
    int chunk_ = 1;

    lof_accum accum() const {
        lof_accum a{};
        a.kind = has_fmt_ ? LOF_ACCUM_FORMAT : LOF_ACCUM_MANTISSA;
        a.mant_bits = accum_mantissa_bits_;
        a.fmt = fmt_;
        // the reference leaf rounds to nearest even whatever proj_spec says
        // (cutlass_gemms.cuh:122); a full-format accumulator honours the mode
        a.spec = lof_round_spec{has_fmt_ ? (int32_t)proj_spec_.rounding_mode : 0, 0, 0};
        a.chunk = chunk_;
        return a;
    }

    // A row-major D is computed directly; a column-major D as D^T = B^T * A^T, which has
    // the same per-output k order.
    lof_gemm_desc make_desc(float alpha, float beta, const MatrixA& A, const MatrixB& B,
                            const MatrixC& C, MatrixD& D) const {
        lof_gemm_desc d{};
        d.alpha = alpha;
        d.beta = beta;
        d.accum = accum();
        d.K = (int32_t)A.n;
        d.D = D.data;
        d.ldd = D.ld;
        d.C = C.data;
        d.ldc = C.ld;
        if constexpr (MatrixD::layout == RowMajor) {
            d.M = (int32_t)A.m;
            d.N = (int32_t)B.n;
            d.A = A.data; d.lda = A.ld; d.trans_a = MatrixA::layout == ColMajor;
            d.B = B.data; d.ldb = B.ld; d.trans_b = MatrixB::layout == ColMajor;
        } else {
            d.M = (int32_t)B.n;
            d.N = (int32_t)A.m;
            d.A = B.data; d.lda = B.ld; d.trans_a = MatrixB::layout == RowMajor;
            d.B = A.data; d.ldb = A.ld; d.trans_b = MatrixA::layout == RowMajor;
        }
        return d;
    }
};

// ---------------------------------------------------------------------------------------
// Conv2d forward (cutlass_conv.cuh:20-184).
//
// The problem description keeps the fields and constructor of CUTLASS's
// conv::Conv2dProblemSize that lo_float::Conv2d reads; replace
// `cutlass::conv::Conv2dProblemSize` / `cutlass::conv::Mode::kCrossCorrelation` by
// `lo_float::Conv2dProblemSize` / `lo_float::ConvMode::kCrossCorrelation`.
enum class ConvMode { kCrossCorrelation, kConvolution };

struct Conv2dProblemSize {
    int N, H, W, C, K, R, S, P, Q;
    int pad_h, pad_w, stride_h, stride_w, dilation_h, dilation_w;
    ConvMode mode;
    int split_k_slices;
    int groups;
    Conv2dProblemSize(int N_, int H_, int W_, int C_, int K_, int R_, int S_, int P_, int Q_,
                      int pad_h_, int pad_w_, int stride_h_, int stride_w_, int dilation_h_,
                      int dilation_w_, ConvMode mode_ = ConvMode::kCrossCorrelation,
                      int split_k_slices_ = 1, int groups_ = 1)
        : N(N_), H(H_), W(W_), C(C_), K(K_), R(R_), S(S_), P(P_), Q(Q_), pad_h(pad_h_),
          pad_w(pad_w_), stride_h(stride_h_), stride_w(stride_w_), dilation_h(dilation_h_),
          dilation_w(dilation_w_), mode(mode_), split_k_slices(split_k_slices_),
          groups(groups_) {}
};

// D = alpha * conv(A, B) + beta * C with A the NHWC activation (N, H, W, C), B the KRSC
// filter (K, R, S, C), C and D NHWC outputs (N, P, Q, K), cross-correlation, groups = 1,
// the accumulator re-rounded after every multiply-add.  Each output's reduction runs over
// (r, s, c) with c fastest, in ascending order.  The CUTLASS tile parameters are accepted
// for source compatibility; the tile is chosen by the library.  Extensions as in Gemm:
// set_accum_format(), set_chunk().
template <int TbM = 128, int TbN = 128, int TbK = 8, int WpM = 32, int WpN = 64, int WpK = 8,
          int Stages = 3>
class Conv2d {
public:
    using ElementA = float;
    using ElementB = float;
    using ElementC = float;
    using ElementCompute = float;

    enum class Status {
        kSuccess = 0,
        kErrorNotSupported = -1,
        kErrorAllocation = -2,
        kErrorInitialize = -3,
        kErrorExecution = -4
    };

    Conv2d(int accum_mantissa_bits = 23, ProjSpec proj_spec = ProjSpec{})
        : accum_mantissa_bits_(accum_mantissa_bits), proj_spec_(proj_spec) {}

    void set_accum_mantissa_bits(int bits) { accum_mantissa_bits_ = bits; has_fmt_ = false; }
    void set_proj_spec(ProjSpec spec) { proj_spec_ = spec; }
    template <class I, class N>
    void set_accum_format(FloatingPointParams<I, N> f) { fmt_ = detail::to_c(f); has_fmt_ = true; }
    void set_chunk(int chunk) { chunk_ = chunk; }

    Status operator()(const Conv2dProblemSize& ps, ElementA* dA, ElementB* dB, ElementC* dC,
                      ElementC* dD, ElementCompute alpha = ElementCompute(1),
                      ElementCompute beta = ElementCompute(0),
                      hipStream_t stream = nullptr) const {
        if (ps.groups != 1 || ps.split_k_slices != 1 || ps.mode != ConvMode::kCrossCorrelation)
            return Status::kErrorNotSupported;
        const auto out = [](int n, int pad, int dil, int k, int st) {
            return st > 0 ? (n + 2 * pad - dil * (k - 1) - 1) / st + 1 : -1;
        };
        if (ps.P != out(ps.H, ps.pad_h, ps.dilation_h, ps.R, ps.stride_h) ||
            ps.Q != out(ps.W, ps.pad_w, ps.dilation_w, ps.S, ps.stride_w))
            return Status::kErrorNotSupported;
        lof_conv2d_desc d{};
        d.N = ps.N; d.H = ps.H; d.W = ps.W; d.C = ps.C;
        d.K = ps.K; d.R = ps.R; d.S = ps.S;
        d.pad_h = ps.pad_h; d.pad_w = ps.pad_w;
        d.stride_h = ps.stride_h; d.stride_w = ps.stride_w;
        d.dil_h = ps.dilation_h; d.dil_w = ps.dilation_w;
        d.layout = LOF_CONV_NHWC;
        d.x = dA;
        d.w = dB;
        d.y = dD;
        d.alpha = alpha;
        d.accum.kind = has_fmt_ ? LOF_ACCUM_FORMAT : LOF_ACCUM_MANTISSA;
        d.accum.mant_bits = accum_mantissa_bits_;
        d.accum.fmt = fmt_;
        // the leaf rounds to nearest even whatever proj_spec says, as the GEMM's
        // (cutlass_gemms.cuh:122); a full-format accumulator honours the mode
        d.accum.spec = lof_round_spec{has_fmt_ ? (int32_t)proj_spec_.rounding_mode : 0, 0, 0};
        d.accum.chunk = chunk_;
        d.c = dC;
        d.beta = beta;
        const lof_status s = lof_conv2d_f32(&d, -1, (lof_stream_t)stream);
        if (s == LOF_ERR_INVALID_ARGUMENT || s == LOF_ERR_UNSUPPORTED)
            return Status::kErrorNotSupported;
        return s == LOF_OK ? Status::kSuccess : Status::kErrorExecution;
    }

private:
    int accum_mantissa_bits_;
    ProjSpec proj_spec_;
    lof_format fmt_{};
    bool has_fmt_ = false;
    int chunk_ = 1;
};

// ---------------------------------------------------------------------------------------
// n-bit pack codec (lo_pack.hpp:50, 392-531).
//
// The codec reads the Inf / NaN checker functors of FloatingPointParams.  A functor crosses
// the C-ABI as a lof_checker description, which it supplies through a static lof()
// member.  The checkers below are the reference's (lo_float.h:1411-1789,
// fp_tools.hpp:331-396) with that member; NoChecker matches nothing.  A user-defined
// checker passes the codec only once it adds lof().
enum class PackEndian { Forward, Reverse };

namespace detail {
constexpr lof_checker make_checker(int32_t kind, std::initializer_list<uint32_t> codes,
                                   uint32_t a, uint32_t b, uint32_t pat0, uint32_t pat1) {
    lof_checker c{kind, (int32_t)codes.size(), {0u, 0u, 0u}, a, b, pat0, pat1};
    int i = 0;
    for (uint32_t v : codes) c.codes[i++] = v;
    return c;
}
}  // namespace detail

struct Trivial_NaNChecker {
    bool operator()(uint32_t) const { return false; }
    uint32_t qNanBitPattern() const { return 0; }
    uint32_t sNanBitPattern() const { return 0; }
    static constexpr lof_checker lof() { return detail::make_checker(0, {}, 0, 0, 0, 0); }
};
struct Trivial_InfChecker {
    bool operator()(uint32_t) const { return false; }
    uint32_t minNegInf() const { return 0; }
    uint32_t minPosInf() const { return 0; }
    static constexpr lof_checker lof() { return detail::make_checker(0, {}, 0, 0, 0, 0); }
};
// e4m3fn / e5m2 (as the reference's param_float8_e5m2 uses it): NaN is 0xFF, no Inf.
struct OCP_F8E4M3_NaNChecker {
    bool operator()(uint32_t bits) const { return bits == 0xFFu; }
    uint32_t qNanBitPattern() const { return 0xFF; }
    uint32_t sNanBitPattern() const { return 0xFF; }
    static constexpr lof_checker lof() { return detail::make_checker(1, {0xFFu}, 0, 0, 0xFF, 0xFF); }
};
struct OCP_F8E4M3_InfChecker {
    bool operator()(uint32_t) const { return false; }
    uint32_t minNegInf() const { return 0; }
    uint32_t minPosInf() const { return 0; }
    static constexpr lof_checker lof() { return detail::make_checker(0, {}, 0, 0, 0, 0); }
};
struct OCP_F8E5M2_NaNChecker {
    bool operator()(uint32_t bits) const { return bits == 0xFFu || bits == 0xFEu || bits == 0xFDu; }
    uint32_t qNanBitPattern() const { return 0xFF; }
    uint32_t sNanBitPattern() const { return 0xFF; }
    static constexpr lof_checker lof() {
        return detail::make_checker(1, {0xFFu, 0xFEu, 0xFDu}, 0, 0, 0xFF, 0xFF);
    }
};
template <Signedness S = Signed>
struct IEEE_F8_NaNChecker {
    static constexpr uint32_t code = S == Signed ? 0x80u : 0xFFu;
    bool operator()(uint32_t bits) const { return bits == code; }
    uint32_t qNanBitPattern() const { return code; }
    uint32_t sNanBitPattern() const { return code; }
    static constexpr lof_checker lof() { return detail::make_checker(1, {code}, 0, 0, code, code); }
};
struct IEEE_F8_InfChecker {
    bool operator()(uint32_t bits) const { return bits == 0x7Fu || bits == 0xFFu; }
    uint32_t minNegInf() const { return 0xFF; }
    uint32_t minPosInf() const { return 0x7F; }
    static constexpr lof_checker lof() {
        return detail::make_checker(1, {0x7Fu, 0xFFu}, 0, 0, 0x7F, 0xFF);
    }
};
// P3109 binary<k>: NaN is 1 << (k-1) (signed) or all-ones (unsigned); with Extended
// behaviour, Inf is the code whose OR with 1 << (k-1) is the top code.
template <int k, Signedness S>
struct P_3109_NaNChecker {
    static constexpr uint32_t code = S == Signed ? (1u << (k - 1)) : (1u << k) - 1u;
    constexpr bool operator()(uint32_t bits) const { return bits == code; }
    uint32_t qNanBitPattern() const { return code; }
    uint32_t sNanBitPattern() const { return code; }
    static constexpr lof_checker lof() { return detail::make_checker(1, {code}, 0, 0, code, code); }
};
template <int k, Signedness S, Inf_Behaviors H>
struct P_3109_InfChecker {
    static constexpr uint32_t top = S == Signed ? (1u << k) - 1u : (1u << k) - 2u;
    constexpr bool operator()(uint32_t bits) const {
        return H != Saturating && (bits | (1u << (k - 1))) == top;
    }
    uint32_t minNegInf() const { return S == Signed ? (1u << k) - 1u : 0u; }
    uint32_t minPosInf() const { return S == Signed ? (1u << (k - 1)) - 1u : (1u << k) - 2u; }
    static constexpr lof_checker lof() {
        const uint32_t pos = S == Signed ? (1u << (k - 1)) - 1u : (1u << k) - 2u;
        const uint32_t neg = S == Signed ? (1u << k) - 1u : 0u;
        return H == Saturating ? detail::make_checker(0, {}, 0, 0, pos, neg)
                               : detail::make_checker(3, {top}, 1u << (k - 1), 0, pos, neg);
    }
};
struct OCP_E8M0_NaNChecker {
    bool operator()(uint32_t bits) const { return bits == 0xFFu; }
    uint32_t qNanBitPattern() const { return 0xFF; }
    uint32_t sNanBitPattern() const { return 0xFF; }
    static constexpr lof_checker lof() { return detail::make_checker(1, {0xFFu}, 0, 0, 0xFF, 0xFF); }
};
struct OCP_E8M0_InfChecker {
    bool operator()(uint32_t) const { return false; }
    uint32_t minNegInf() const { return 0; }
    uint32_t minPosInf() const { return 0; }
    static constexpr lof_checker lof() { return detail::make_checker(0, {}, 0, 0, 0, 0); }
};
struct FP16InfChecker {
    bool operator()(uint32_t bits) const { return (bits & 0x7C00u) == 0x7C00u && !(bits & 0x3FFu); }
    uint32_t minPosInf() const { return 0x7C00; }
    uint32_t minNegInf() const { return 0xFC00; }
    static constexpr lof_checker lof() {
        return detail::make_checker(2, {}, 0x7C00u, 0x3FFu, 0x7C00u, 0xFC00u);
    }
};
struct FP16NaNChecker {
    bool operator()(uint32_t bits) const { return (bits & 0x7C00u) == 0x7C00u && (bits & 0x3FFu); }
    uint32_t qNanBitPattern() const { return 0x7E00; }
    uint32_t sNanBitPattern() const { return 0x7C01; }
    static constexpr lof_checker lof() {
        return detail::make_checker(2, {}, 0x7C00u, 0x3FFu, 0x7E00u, 0x7C01u);
    }
};
struct BF16InfChecker {
    bool operator()(uint32_t bits) const { return (bits & 0x7F80u) == 0x7F80u && !(bits & 0x7Fu); }
    uint32_t minPosInf() const { return 0x7F80; }
    uint32_t minNegInf() const { return 0xFF80; }
    static constexpr lof_checker lof() {
        return detail::make_checker(2, {}, 0x7F80u, 0x7Fu, 0x7F80u, 0xFF80u);
    }
};
struct BF16NaNChecker {
    bool operator()(uint32_t bits) const { return (bits & 0x7F80u) == 0x7F80u && (bits & 0x7Fu); }
    uint32_t qNanBitPattern() const { return 0x7FC0; }
    uint32_t sNanBitPattern() const { return 0x7F81; }
    static constexpr lof_checker lof() {
        return detail::make_checker(2, {}, 0x7F80u, 0x7Fu, 0x7FC0u, 0x7F81u);
    }
};

namespace detail {
template <class C, class = void>
struct has_lof : std::false_type {};
template <class C>
struct has_lof<C, std::void_t<decltype(C::lof())>> : std::true_type {};

template <class C>
constexpr lof_checker checker_of() {
    if constexpr (std::is_same_v<C, NoChecker>) {
        return make_checker(0, {}, 0, 0, 0, 0);
    } else {
        static_assert(has_lof<C>::value,
                      "the pack codec needs a checker with a static lof() description");
        return C::lof();
    }
}
template <class I, class N>
inline lof_pack_format to_pack(const FloatingPointParams<I, N>& p) {
    return lof_pack_format{to_c(p), (int32_t)p.unsigned_behavior, checker_of<I>(),
                           checker_of<N>()};
}
}  // namespace detail

// Typed cast (lo_float.h:3979-3988, 4014-4022): Project over device arrays of reps, one per
// uint8_t (formats of <= 8 bits) or uint16_t (<= 16 bits) -- the storage of a
// Templated_Float<Fp> array -- with the scalar ConvertImpl semantics and ps's saturation
// mode.  The rep's low n bits are stored.
template <class I, class N, class Rep>
void Project(const float* d_x, Rep* d_y, int n, FloatingPointParams<I, N> params,
             ProjSpec ps = ProjSpec{}, hipStream_t stream = nullptr,
             const uint32_t* rand_bits = nullptr, uint64_t seed = 0) {
    static_assert(std::is_same_v<Rep, uint8_t> || std::is_same_v<Rep, uint16_t>,
                  "reps are stored in uint8_t or uint16_t");
    if (lof_rep_bytes(params.bitwidth) != (int32_t)sizeof(Rep))
        throw Error(LOF_ERR_INVALID_ARGUMENT, "Project: container size does not fit the format");
    const lof_pack_format f = detail::to_pack(params);
    detail::check(lof_cast_to_rep_f32(d_y, d_x, n, &f, detail::to_c(ps), rand_bits, seed,
                                      (lof_stream_t)stream),
                  "Project");
}

// The FP32 values of device reps (ConvertImpl<Templated_Float<Fp>, float>::run).
template <class I, class N, class Rep>
void Project(const Rep* d_x, float* d_y, int n, FloatingPointParams<I, N> params,
             hipStream_t stream = nullptr) {
    static_assert(std::is_same_v<Rep, uint8_t> || std::is_same_v<Rep, uint16_t>,
                  "reps are stored in uint8_t or uint16_t");
    if (lof_rep_bytes(params.bitwidth) != (int32_t)sizeof(Rep))
        throw Error(LOF_ERR_INVALID_ARGUMENT, "Project: container size does not fit the format");
    const lof_pack_format f = detail::to_pack(params);
    detail::check(lof_cast_from_rep_f32(d_y, d_x, n, &f, (lof_stream_t)stream), "Project");
}

// Bytes of a packed buffer of `count` n-bit elements (lo_pack.hpp:392-394).
inline size_t LoF_packed_size(int n, size_t count) {
    const int64_t v = lof_packed_size(n, (int64_t)count);
    if (v < 0) throw Error(LOF_ERR_INVALID_ARGUMENT, "LoF_packed_size: bitwidth outside [1, 16]");
    return (size_t)v;
}

// Round d_src[0..count) to params under ps and pack into d_dst (device pointers); returns
// the bytes written, LoF_packed_size(params.bitwidth, count) (lo_pack.hpp:504-517).
template <PackEndian E = PackEndian::Forward, class I, class N>
int LoF_encode_cuda(uint8_t* d_dst, const float* d_src, FloatingPointParams<I, N> params,
                    size_t count, ProjSpec ps = ProjSpec{}, hipStream_t stream = nullptr,
                    const uint32_t* rand_bits = nullptr, uint64_t seed = 0) {
    const lof_pack_format f = detail::to_pack(params);
    detail::check(lof_pack_encode_f32(d_dst, d_src, (int64_t)count, &f, detail::to_c(ps),
                                      E == PackEndian::Reverse ? LOF_PACK_REVERSE
                                                               : LOF_PACK_FORWARD,
                                      rand_bits, seed, (lof_stream_t)stream),
                  "LoF_encode_cuda");
    return (int)LoF_packed_size(params.bitwidth, count);
}

// Unpack `count` elements of params from d_src and widen them to float in d_dst; returns
// the bytes consumed (lo_pack.hpp:519-531).
template <PackEndian E = PackEndian::Forward, class I, class N>
int LoF_decode_cuda(float* d_dst, const uint8_t* d_src, FloatingPointParams<I, N> params,
                    size_t count, hipStream_t stream = nullptr) {
    const lof_pack_format f = detail::to_pack(params);
    detail::check(lof_pack_decode_f32(d_dst, d_src, (int64_t)count, &f,
                                      E == PackEndian::Reverse ? LOF_PACK_REVERSE
                                                               : LOF_PACK_FORWARD,
                                      (lof_stream_t)stream),
                  "LoF_decode_cuda");
    return (int)LoF_packed_size(params.bitwidth, count);
}

}  // namespace lo_float
