// test_api.cpp -- C++ API (include/lofloat/lo_float.hpp) against the oracle.
//
//   test_api cpu   argument checks that never touch a device
//   test_api gpu   round_fp_params / round_mantissa / run_virtual_mx_round and Gemm in
//                  every layout combination, bit-exact against liblof_oracle.so
// Built by __graft_entry__.build() with g++ (the header needs no HIP compiler).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "lofloat/lo_float.hpp"
#include "../../oracle/lof_oracle.h"

using namespace lo_float;

static int failures = 0;
#define EXPECT(c, ...)                       \
    do {                                     \
        if (!(c)) {                          \
            std::printf("FAIL: " __VA_ARGS__); \
            std::printf("\n");               \
            ++failures;                      \
        }                                    \
    } while (0)

static bool same_bits(float a, float b) {
    if (std::isnan(a) && std::isnan(b)) return true;
    uint32_t x, y;
    std::memcpy(&x, &a, 4);
    std::memcpy(&y, &b, 4);
    return x == y;
}

static int compare(const std::vector<float>& got, const std::vector<float>& want,
                   const char* what) {
    size_t bad = 0, first = 0;
    for (size_t i = 0; i < got.size(); ++i)
        if (!same_bits(got[i], want[i])) {
            if (!bad) first = i;
            ++bad;
        }
    EXPECT(bad == 0, "%s: %zu mismatches, first at %zu (%g vs %g)", what, bad, first,
           bad ? got[first] : 0.0, bad ? want[first] : 0.0);
    return (int)bad;
}

static orc_pack_format to_orc(const lof_pack_format& p) {
    static_assert(sizeof(orc_pack_format) == sizeof(lof_pack_format), "same layout");
    orc_pack_format o;
    std::memcpy(&o, &p, sizeof o);
    return o;
}

// The lof() description of each checker, as the oracle's decoder reads it, classifies every
// code of the format as the functor's own operator() does.
template <class I, class N>
static void checker_case(const char* name, FloatingPointParams<I, N> p) {
    const orc_pack_format of = to_orc(detail::to_pack(p));
    const uint32_t codes = 1u << p.bitwidth;
    std::vector<uint32_t> reps(codes);
    for (uint32_t r = 0; r < codes; ++r) reps[r] = r;
    std::vector<float> v(codes);
    orc_rep_to_float_array(reps.data(), v.data(), codes, &of);
    int bad = 0;
    for (uint32_t r = 0; r < codes; ++r) {
        const bool inf = p.OV_behavior != Saturating && p.IsInf(r);
        const bool nan = !inf && p.IsNaN(r);
        bad += std::isinf(v[r]) != inf || std::isnan(v[r]) != nan;
    }
    EXPECT(bad == 0, "checker description %s: %d codes classified differently", name, bad);
}

static void pack_cpu_checks() {
    EXPECT(LoF_packed_size(11, 17) == 44 && LoF_packed_size(3, 9) == 6 &&
               LoF_packed_size(16, 5) == 10 && LoF_packed_size(8, 0) == 0,
           "LoF_packed_size");
    try {
        (void)LoF_packed_size(17, 1);
        EXPECT(false, "bitwidth 17 must throw");
    } catch (const Error& e) {
        EXPECT(e.status == LOF_ERR_INVALID_ARGUMENT, "status for bitwidth 17");
    }
    checker_case("e4m3fn", FloatingPointParams(8, 3, 7, Saturating, _3109, Signed,
                                               OCP_F8E4M3_InfChecker{}, OCP_F8E4M3_NaNChecker{}));
    checker_case("e5m2", FloatingPointParams(8, 2, 15, Saturating, _3109, Signed,
                                             OCP_F8E4M3_InfChecker{}, OCP_F8E5M2_NaNChecker{}));
    checker_case("ieee_f8", FloatingPointParams(8, 3, 7, Extended, _3109, Signed,
                                                IEEE_F8_InfChecker{}, IEEE_F8_NaNChecker<Signed>{}));
    checker_case("ieee_f8u", FloatingPointParams(8, 3, 7, Extended, _3109, Unsigned,
                                                 IEEE_F8_InfChecker{}, IEEE_F8_NaNChecker<Unsigned>{}));
    checker_case("p3109<8,4>ext", FloatingPointParams(8, 3, 8, Extended, _3109, Signed,
                                                      P_3109_InfChecker<8, Signed, Extended>{},
                                                      P_3109_NaNChecker<8, Signed>{}));
    checker_case("p3109<6,3>u ext", FloatingPointParams(6, 2, 4, Extended, _3109, Unsigned,
                                                        P_3109_InfChecker<6, Unsigned, Extended>{},
                                                        P_3109_NaNChecker<6, Unsigned>{}));
    checker_case("p3109<11,5>", FloatingPointParams(11, 4, 32, Saturating, _3109, Signed,
                                                    P_3109_InfChecker<11, Signed, Saturating>{},
                                                    P_3109_NaNChecker<11, Signed>{}));
    checker_case("e8m0", FloatingPointParams(8, 0, 127, Saturating, _3109, Unsigned,
                                             OCP_E8M0_InfChecker{}, OCP_E8M0_NaNChecker{}));
    checker_case("half", FloatingPointParams(16, 10, 15, Extended, _3109, Signed, FP16InfChecker{},
                                             FP16NaNChecker{}));
    checker_case("bf16", FloatingPointParams(16, 7, 127, Extended, _3109, Signed, BF16InfChecker{},
                                             BF16NaNChecker{}));
    checker_case("trivial", FloatingPointParams(5, 2, 3, Saturating, _3109, Signed,
                                                Trivial_InfChecker{}, Trivial_NaNChecker{}));
}

static void cpu_checks() {
    ProjSpec ps;
    EXPECT(ps.rounding_mode == Rounding_Mode::RoundToNearestEven && ps.stoch_length == 0,
           "ProjSpec default");
    EXPECT((int)Rounding_Mode::RoundToOdd == 12 && (int)Rounding_Mode::StochasticRoundingD == 14,
           "Rounding_Mode values");
    FloatingPointParams<> bad(30, 24, 7, Saturating, _3109, Signed);
    try {
        round_fp_params<float>(nullptr, 0, bad, ProjSpec{}, 1.0f);
        EXPECT(false, "mantissa > 23 must throw");
    } catch (const Error& e) {
        EXPECT(e.status == LOF_ERR_UNSUPPORTED, "status for mantissa > 23");
    }
    using MatR = Matrix<float, int, RowMajor>;
    MatR A(nullptr, 4, 5, 5), B(nullptr, 6, 3, 3), C(nullptr, 4, 3, 3), D(nullptr, 4, 3, 3);
    using GemmR = Gemm<MatR, MatR, MatR, MatR>;
    GemmR g(10);
    EXPECT(g.can_implement(A, B, C, D) == GemmR::Status::kErrorNotSupported,
           "shape mismatch rejected");
    EXPECT(g.get_workspace_size(A, B, C, D) == 0, "no workspace");
    Conv2d<> conv(10);
    using CS = Conv2d<>::Status;
    EXPECT(conv(Conv2dProblemSize(1, 8, 8, 4, 6, 3, 3, 8, 8, 1, 1, 1, 1, 1, 1,
                                  ConvMode::kCrossCorrelation, 1, 2),
                nullptr, nullptr, nullptr, nullptr) == CS::kErrorNotSupported, "conv groups 2");
    EXPECT(conv(Conv2dProblemSize(1, 8, 8, 4, 6, 3, 3, 7, 8, 1, 1, 1, 1, 1, 1),
                nullptr, nullptr, nullptr, nullptr) == CS::kErrorNotSupported, "conv wrong P");
    EXPECT(conv(Conv2dProblemSize(1, 8, 8, 4, 6, 3, 3, 8, 8, 1, 1, 1, 1, 1, 1,
                                  ConvMode::kConvolution),
                nullptr, nullptr, nullptr, nullptr) == CS::kErrorNotSupported, "conv mode");
    pack_cpu_checks();
}

template <class T>
static T* dev_copy(const std::vector<T>& h) {
    T* d = nullptr;
    if (hipMalloc(&d, h.size() * sizeof(T) + 16) != hipSuccess) return nullptr;
    (void)hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    return d;
}
template <class T>
static std::vector<T> host_copy(const T* d, size_t n) {
    std::vector<T> h(n);
    (void)hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost);
    return h;
}

// A m x k matrix stored in `layout` with leading dimension ld, and its row-major image.
struct HostMat {
    std::vector<float> store, rowmajor;
    int ld;
};
static HostMat make_mat(int m, int k, Layout L, std::mt19937& rng) {
    std::normal_distribution<float> nd(0.f, 1.f);
    HostMat h;
    h.ld = (L == RowMajor ? k : m) + 3;
    h.store.assign((size_t)h.ld * (L == RowMajor ? m : k), 0.f);
    h.rowmajor.assign((size_t)m * k, 0.f);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < k; ++j) {
            const float v = nd(rng) / 8.f;
            h.rowmajor[(size_t)i * k + j] = v;
            h.store[L == RowMajor ? (size_t)i * h.ld + j : (size_t)j * h.ld + i] = v;
        }
    return h;
}

template <Layout LA, Layout LB, Layout LD>
static void gemm_case(std::mt19937& rng, float beta) {
    const int M = 67, N = 45, K = 93;
    HostMat a = make_mat(M, K, LA, rng), b = make_mat(K, N, LB, rng), c = make_mat(M, N, LD, rng);
    float *dA = dev_copy(a.store), *dB = dev_copy(b.store), *dC = dev_copy(c.store);
    std::vector<float> zero(c.store.size(), 0.f);
    float* dD = dev_copy(zero);
    using MA = Matrix<float, int, LA>;
    using MB = Matrix<float, int, LB>;
    using MD = Matrix<float, int, LD>;
    MA A(dA, M, K, a.ld);
    MB B(dB, K, N, b.ld);
    MD C(dC, M, N, c.ld), D(dD, M, N, c.ld);
    using G = Gemm<MA, MB, MD, MD>;
    G g(10);
    auto st = g(0.5f, beta, A, B, C, D);
    EXPECT(st == G::Status::kSuccess, "gemm status");
    (void)hipDeviceSynchronize();
    std::vector<float> out = host_copy(dD, c.store.size()), got((size_t)M * N);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j)
            got[(size_t)i * N + j] = out[LD == RowMajor ? (size_t)i * c.ld + j : (size_t)j * c.ld + i];
    orc_accum acc{};
    acc.kind = 0;
    acc.mant_bits = 10;
    acc.chunk = 1;
    std::vector<float> want((size_t)M * N);
    orc_gemm_f32(M, N, K, a.rowmajor.data(), K, 0, b.rowmajor.data(), N, 0, want.data(), N, 0.5f,
                 &acc, nullptr, 0, 0, c.rowmajor.data(), N, beta);
    char what[96];
    std::snprintf(what, sizeof(what), "Gemm layouts A%d B%d D%d beta %g", (int)LA, (int)LB,
                  (int)LD, beta);
    compare(got, want, what);
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC); (void)hipFree(dD);
}

// LoF_encode_cuda / LoF_decode_cuda against orc_pack_encode / orc_pack_decode: bytes and
// decoded values bit-exact.
template <PackEndian E, class I, class N>
static void pack_case(const char* name, FloatingPointParams<I, N> p, const std::vector<float>& x,
                      ProjSpec ps = ProjSpec{}, uint64_t seed = 0) {
    const size_t n = x.size(), nb = LoF_packed_size(p.bitwidth, n);
    float* dx = dev_copy(x);
    uint8_t* db = dev_copy(std::vector<uint8_t>(nb + 1, 0xA5));
    const int wrote = LoF_encode_cuda<E>(db, dx, p, n, ps, nullptr, nullptr, seed);
    float* dy = dev_copy(std::vector<float>(n, 0.f));
    const int read = LoF_decode_cuda<E>(dy, db, p, n);
    (void)hipDeviceSynchronize();
    EXPECT(wrote == (int)nb && read == (int)nb, "%s: byte counts %d %d vs %zu", name, wrote, read, nb);
    const orc_pack_format of = to_orc(detail::to_pack(p));
    const int endian = E == PackEndian::Reverse ? 1 : 0;
    std::vector<uint8_t> want(nb + 1, 0xA5);
    orc_pack_encode(want.data(), x.data(), (int64_t)n, &of,
                    orc_spec{(int32_t)ps.rounding_mode, (int32_t)ps.saturation_mode, ps.stoch_length},
                    endian, nullptr, seed);
    const std::vector<uint8_t> got = host_copy(db, nb + 1);
    size_t bad = 0;
    for (size_t i = 0; i <= nb; ++i) bad += got[i] != want[i];
    EXPECT(bad == 0, "%s: %zu packed bytes differ (incl. the guard byte)", name, bad);
    std::vector<float> dec(n);
    orc_pack_decode(dec.data(), want.data(), (int64_t)n, &of, endian);
    char what[96];
    std::snprintf(what, sizeof(what), "%s decode", name);
    compare(host_copy(dy, n), dec, what);
    (void)hipFree(dx); (void)hipFree(db); (void)hipFree(dy);
}

static void pack_gpu_checks(std::mt19937& rng) {
    std::vector<float> x(50001);
    std::normal_distribution<float> nd(0.f, 1.f);
    std::uniform_real_distribution<float> ud(-20.f, 20.f);
    for (auto& v : x) v = nd(rng) * std::exp2(ud(rng));
    x[0] = 0.f; x[1] = -0.f; x[2] = INFINITY; x[3] = -INFINITY; x[4] = NAN; x[5] = 1e-40f;
    const FloatingPointParams p11(11, 4, 32, Saturating, _3109, Signed,
                                  P_3109_InfChecker<11, Signed, Saturating>{},
                                  P_3109_NaNChecker<11, Signed>{});
    pack_case<PackEndian::Forward>("p3109<11,5> forward", p11, x);
    pack_case<PackEndian::Reverse>("p3109<11,5> reverse", p11, x);
    pack_case<PackEndian::Reverse>("e5m2 reverse", FloatingPointParams(
        8, 2, 15, Saturating, _3109, Signed, OCP_F8E4M3_InfChecker{}, OCP_F8E5M2_NaNChecker{}), x);
    pack_case<PackEndian::Forward>("half SatFinite", FloatingPointParams(
        16, 10, 15, Extended, _3109, Signed, FP16InfChecker{}, FP16NaNChecker{}), x,
        ProjSpec(Rounding_Mode::RoundToNearestEven, SatFinite));
    pack_case<PackEndian::Forward>("bf16 RTZ", FloatingPointParams(
        16, 7, 127, Extended, _3109, Signed, BF16InfChecker{}, BF16NaNChecker{}), x,
        ProjSpec(Rounding_Mode::RoundTowardsZero));
    pack_case<PackEndian::Reverse>("ieee_f8 SR-B", FloatingPointParams(
        8, 3, 7, Extended, _3109, Signed, IEEE_F8_InfChecker{}, IEEE_F8_NaNChecker<Signed>{}), x,
        ProjSpec(Rounding_Mode::StochasticRoundingB, OvfInf, 6), 77);
    pack_case<PackEndian::Forward>("e2m1 (no checkers)",
                                   FloatingPointParams<>(4, 1, 1, Saturating, _3109, Signed), x);
    pack_case<PackEndian::Forward>("p3109<6,3>u NegtoNaN", FloatingPointParams(
        6, 2, 4, Saturating, _3109, Unsigned, P_3109_InfChecker<6, Unsigned, Saturating>{},
        P_3109_NaNChecker<6, Unsigned>{}, NegtoNaN), x);
}

// Conv2d (NHWC activation, KRSC filter, NHWC output) against the oracle GEMM of the filter
// (K x RSC) with the unfolded input, reduction order (r, s, c); D = alpha * conv + beta * C.
static void conv_case(std::mt19937& rng) {
    const int N = 2, H = 11, W = 9, C = 5, K = 12, R = 3, S = 2, ph = 1, pw = 0, sh = 2, sw = 1,
              dh = 1, dw = 2;
    const int P = (H + 2 * ph - dh * (R - 1) - 1) / sh + 1, Q = (W + 2 * pw - dw * (S - 1) - 1) / sw + 1;
    std::normal_distribution<float> nd(0.f, 1.f);
    std::vector<float> x((size_t)N * H * W * C), w((size_t)K * R * S * C), c((size_t)N * P * Q * K);
    for (auto& v : x) v = nd(rng);
    for (auto& v : w) v = nd(rng);
    for (auto& v : c) v = nd(rng);
    const int RSC = R * S * C, NPQ = N * P * Q;
    std::vector<float> cols((size_t)RSC * NPQ, 0.f);
    for (int r = 0; r < R; ++r)
        for (int s = 0; s < S; ++s)
            for (int ch = 0; ch < C; ++ch)
                for (int b = 0; b < N; ++b)
                    for (int p = 0; p < P; ++p)
                        for (int q = 0; q < Q; ++q) {
                            const int ih = p * sh - ph + r * dh, iw = q * sw - pw + s * dw;
                            if (ih < 0 || ih >= H || iw < 0 || iw >= W) continue;
                            cols[(size_t)((r * S + s) * C + ch) * NPQ + (b * P + p) * Q + q] =
                                x[(((size_t)b * H + ih) * W + iw) * C + ch];
                        }
    // C (NHWC, i.e. NPQ x K) as the GEMM's K x NPQ operand
    std::vector<float> cT((size_t)K * NPQ);
    for (int n = 0; n < NPQ; ++n)
        for (int k = 0; k < K; ++k) cT[(size_t)k * NPQ + n] = c[(size_t)n * K + k];
    orc_accum acc{};
    acc.kind = 0;
    acc.mant_bits = 9;
    acc.chunk = 1;
    std::vector<float> D((size_t)K * NPQ);
    orc_gemm_f32(K, NPQ, RSC, w.data(), RSC, 0, cols.data(), NPQ, 0, D.data(), NPQ, 0.5f, &acc,
                 nullptr, 0, 0, cT.data(), NPQ, -2.0f);
    std::vector<float> want((size_t)NPQ * K);
    for (int n = 0; n < NPQ; ++n)
        for (int k = 0; k < K; ++k) want[(size_t)n * K + k] = D[(size_t)k * NPQ + n];
    float *dx = dev_copy(x), *dw_ = dev_copy(w), *dc = dev_copy(c);
    float* dy = dev_copy(std::vector<float>(want.size(), 0.f));
    Conv2d<> conv(9);
    const auto st = conv(Conv2dProblemSize(N, H, W, C, K, R, S, P, Q, ph, pw, sh, sw, dh, dw),
                         dx, dw_, dc, dy, 0.5f, -2.0f);
    (void)hipDeviceSynchronize();
    EXPECT(st == Conv2d<>::Status::kSuccess, "Conv2d status");
    compare(host_copy(dy, want.size()), want, "Conv2d NHWC/KRSC");
    (void)hipFree(dx); (void)hipFree(dw_); (void)hipFree(dc); (void)hipFree(dy);
}

// Project (typed cast) both ways against orc_float_to_rep_array / orc_rep_to_float_array.
static void project_case(std::mt19937& rng) {
    std::vector<float> x(10007);
    std::normal_distribution<float> nd(0.f, 1.f);
    for (auto& v : x) v = nd(rng) * 64.f;
    x[0] = INFINITY; x[1] = -INFINITY; x[2] = NAN; x[3] = 1e30f;
    const FloatingPointParams e5m2(8, 2, 15, Saturating, _3109, Signed, OCP_F8E4M3_InfChecker{},
                                   OCP_F8E5M2_NaNChecker{});
    const FloatingPointParams half(16, 10, 15, Extended, _3109, Signed, FP16InfChecker{},
                                   FP16NaNChecker{});
    float* dx = dev_copy(x);
    const auto run = [&](auto params, auto zero, ProjSpec ps, const char* what) {
        using Rep = decltype(zero);
        Rep* dr = dev_copy(std::vector<Rep>(x.size(), 0));
        Project(dx, dr, (int)x.size(), params, ps);
        float* dy = dev_copy(std::vector<float>(x.size(), 0.f));
        Project(dr, dy, (int)x.size(), params);
        (void)hipDeviceSynchronize();
        const orc_pack_format of = to_orc(detail::to_pack(params));
        std::vector<uint32_t> want(x.size());
        orc_float_to_rep_array(x.data(), want.data(), (int64_t)x.size(), &of,
                               orc_spec{(int32_t)ps.rounding_mode, (int32_t)ps.saturation_mode,
                                        ps.stoch_length}, nullptr, 0);
        const std::vector<Rep> got = host_copy(dr, x.size());
        size_t bad = 0;
        for (size_t i = 0; i < x.size(); ++i)
            bad += got[i] != (Rep)(want[i] & ((1u << params.bitwidth) - 1u));
        EXPECT(bad == 0, "%s: %zu reps differ", what, bad);
        std::vector<uint32_t> masked(x.size());
        for (size_t i = 0; i < x.size(); ++i) masked[i] = want[i] & ((1u << params.bitwidth) - 1u);
        std::vector<float> vals(x.size());
        orc_rep_to_float_array(masked.data(), vals.data(), (int64_t)x.size(), &of);
        compare(host_copy(dy, x.size()), vals, what);
        (void)hipFree(dr); (void)hipFree(dy);
    };
    run(e5m2, uint8_t{0}, ProjSpec(Rounding_Mode::RoundToNearestEven, SatFinite), "Project e5m2");
    run(half, uint16_t{0}, ProjSpec(Rounding_Mode::RoundTowardsZero, SatPropagate), "Project half");
    try {
        Project(dx, (uint16_t*)nullptr, 1, e5m2);
        EXPECT(false, "an 8-bit format in uint16_t containers must throw");
    } catch (const Error&) {
    }
    (void)hipFree(dx);
}

static uint32_t* g_words = nullptr;
static const uint32_t* dw_words(const std::vector<uint32_t>& w) {
    if (!g_words) g_words = dev_copy(w);
    return g_words;
}
static void free_words() {
    (void)hipFree(g_words);
    g_words = nullptr;
}

static void gpu_checks() {
    std::mt19937 rng(1234);
    const int n = 100003;
    std::vector<float> x(n);
    std::normal_distribution<float> nd(0.f, 1.f);
    std::uniform_real_distribution<float> ud(-12.f, 12.f);
    for (auto& v : x) v = nd(rng) * std::exp2(ud(rng));
    x[0] = 0.f; x[1] = -0.f; x[2] = INFINITY; x[3] = NAN; x[4] = 1e-40f;

    // round_fp_params: e4m3 (8,3,7) saturating, RNE, fused scale 2
    FloatingPointParams<> e4m3(8, 3, 7, Saturating, _3109, Signed);
    float* d = dev_copy(x);
    round_fp_params<float>(d, n, e4m3, ProjSpec{}, 2.0f);
    (void)hipDeviceSynchronize();
    std::vector<float> want(n);
    orc_format f{8, 3, 7, 1, 0, 0};
    orc_virtual_round_fmt_array(x.data(), want.data(), n, &f, orc_spec{0, 0, 0}, 2.0f, nullptr, 0, 0);
    compare(host_copy(d, n), want, "round_fp_params e4m3");

    // round_mantissa: 10 bits, SR-A with 5 random bits from explicit words
    std::vector<uint32_t> words(n);
    for (auto& w : words) w = rng();
    uint32_t* dw = dev_copy(words);
    float* dout = dev_copy(x);
    float* din = dev_copy(x);
    round_mantissa<float>(din, dout, n, 10, ProjSpec(Rounding_Mode::StochasticRoundingA, OvfInf, 5),
                          1.0f, nullptr, dw);
    (void)hipDeviceSynchronize();
    orc_virtual_round_mant_array(x.data(), want.data(), n, 10, orc_spec{3, 0, 5}, 1.0f,
                                 words.data(), 0, 0);
    compare(host_copy(dout, n), want, "round_mantissa SR-A");

    // run_virtual_mx_round: e2m1 elements, E8M0 scale, bs 32, in place
    FloatingPointParams<> e2m1(4, 1, 1, Saturating, _3109, Signed);
    FloatingPointParams<> e8m0(8, 0, 127, Extended, _3109, Unsigned);
    float* dm = dev_copy(x);
    run_virtual_mx_round(dm, n, 32, e8m0, e2m1, ProjSpec{}, ProjSpec{});
    (void)hipDeviceSynchronize();
    orc_format fe{4, 1, 1, 1, 0, 0}, fs{8, 0, 127, 0, 0, 1};
    orc_virtual_mx_round(x.data(), want.data(), n, 32, &fe, &fs, orc_spec{0, 0, 0}, 0, nullptr, 0, 0);
    compare(host_copy(dm, n), want, "run_virtual_mx_round e2m1");
    (void)hipFree(d); (void)hipFree(dw); (void)hipFree(dout); (void)hipFree(din); (void)hipFree(dm);

    // lo_float::virtual_round over arrays, format overload: e5m2 (8,2,15) saturating,
    // round-to-odd, out of place
    float* dv = dev_copy(x);
    float* dr = dev_copy(std::vector<float>(n, 0.f));
    FloatingPointParams<> e5m2(8, 2, 15, Saturating, _3109, Signed);
    virtual_round(dv, dr, n, e5m2, ProjSpec(Rounding_Mode::RoundToOdd));
    (void)hipDeviceSynchronize();
    orc_format f52{8, 2, 15, 1, 0, 0};
    orc_virtual_round_fmt_array(x.data(), want.data(), n, &f52, orc_spec{12, 0, 0}, 1.0f, nullptr,
                                0, 0);
    compare(host_copy(dr, n), want, "virtual_round(values, results, n, e5m2, RTO)");
    // mantissa overload: 7 bits, SR-C with 6 random bits from explicit words, in place
    float* di = dev_copy(x);
    virtual_round(di, di, 7, n, ProjSpec(Rounding_Mode::StochasticRoundingC, OvfInf, 6), nullptr,
                  dw_words(words));
    (void)hipDeviceSynchronize();
    orc_virtual_round_mant_array(x.data(), want.data(), n, 7, orc_spec{9, 0, 6}, 1.0f,
                                 words.data(), 0, 0);
    compare(host_copy(di, n), want, "virtual_round(values, results, 7, n, SR-C)");
    // lo_float::virtual_mx_round: e4m3 elements RTZ, E8M0 scale RNE, bs 16, n not a
    // multiple of the block (short last block)
    FloatingPointParams<> e4m3s(8, 3, 7, Saturating, _3109, Signed);
    float* dmx = dev_copy(x);
    virtual_mx_round(dmx, n, 16, e8m0, e4m3s, ProjSpec{}, ProjSpec(Rounding_Mode::RoundTowardsZero));
    (void)hipDeviceSynchronize();
    orc_format f43{8, 3, 7, 1, 0, 0};
    orc_virtual_mx_round(x.data(), want.data(), n, 16, &f43, &fs, orc_spec{1, 0, 0}, 0, nullptr, 0, 0);
    compare(host_copy(dmx, n), want, "virtual_mx_round(inout, n, 16, e8m0, e4m3, RNE, RTZ)");
    try {
        virtual_mx_round(dmx, n, 24, e8m0, e4m3s);
        EXPECT(false, "virtual_mx_round block size 24 must throw");
    } catch (const Error& e) {
        EXPECT(e.status == LOF_ERR_UNSUPPORTED, "status for block size 24");
    }
    virtual_mx_round(dmx, 0, 16, e8m0, e4m3s);   // n <= 0: nothing
    (void)hipFree(dv); (void)hipFree(dr); (void)hipFree(di); (void)hipFree(dmx);
    free_words();

    conv_case(rng);
    project_case(rng);
    pack_gpu_checks(rng);
    for (float beta : {0.0f, 0.75f}) {
        gemm_case<RowMajor, RowMajor, RowMajor>(rng, beta);
        gemm_case<ColMajor, RowMajor, RowMajor>(rng, beta);
        gemm_case<RowMajor, ColMajor, RowMajor>(rng, beta);
        gemm_case<ColMajor, ColMajor, RowMajor>(rng, beta);
        gemm_case<RowMajor, RowMajor, ColMajor>(rng, beta);
        gemm_case<ColMajor, RowMajor, ColMajor>(rng, beta);
        gemm_case<RowMajor, ColMajor, ColMajor>(rng, beta);
        gemm_case<ColMajor, ColMajor, ColMajor>(rng, beta);
    }
}

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    cpu_checks();
    if (gpu) gpu_checks();
    std::printf("%s checks: %d failure(s)\n", gpu ? "gpu" : "cpu", failures);
    return failures ? 1 : 0;
}
