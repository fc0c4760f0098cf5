// lof_gemm.hip -- emulated-accumulation GEMM launcher, tile-config table, autotuner and
// C-ABI (lof_gemm_f32, lof_gemm_num_configs, lof_gemm_config_valid, lof_gemm_config_name,
// lof_gemm_autotune).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>

#include "lof_gemm.hpp"
#include "lof_host.hpp"

namespace lofamd {

namespace {

template <int BM, int BN, int TM, int TN, int BK, bool SUBN>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    __shared__ float smem[gemm_smem_floats<BM, BN, BK>()];
    gemm_tile<BM, BN, TM, TN, BK, SUBN>(g, (int)blockIdx.y, (int)blockIdx.x, (int)threadIdx.x,
                                        smem);
}

// Occupancy: the 64x64 tile is held to 128 registers (4 waves per SIMD; chunk 1 spills a
// few), the 64x128 tile to 168 (3 waves): both are latency-bound, and a 4096^2 output's
// 4096 64x64 tiles then fill 1024 slots in exactly four rounds.
template <int WMB, int WNB, int CHUNK, bool SUBN, bool DRC = false>
__global__ __launch_bounds__(256, (WMB == 1 && WNB == 1) ? 4 : (WMB == 1 && WNB == 2) ? 3 : 1)
void k_gemm_mfma(GemmArgs g) {
    __shared__ __align__(16) float smem[MfmaTile<WMB, WNB>::SMEM_FLOATS];
    gemm_tile_mfma<WMB, WNB, CHUNK, SUBN, false, DRC>(g, (int)blockIdx.y, (int)blockIdx.x,
                                                      (int)threadIdx.x, smem);
}

using GemmKernel = void (*)(GemmArgs);

struct GemmConfig {
    const char* name;
    int bm, bn;
    GemmKernel kernel[2];          // VALU tile, any chunk: [subn]
    GemmKernel mfma[6][2];         // wave tile: [log2(chunk)][subn], chunk in {1, 2, ..., 32}
    // wave tile with directed-rounding chunk leaves, chunks 8-32: [log2(chunk) - 3] (nullptr
    // where the tile has none), for the chunk sums of mantissa-only accumulators of 20 or more
    // bits, which have no other fast leaf.  For narrower accumulators they measured no faster
    // than the TwoSum round-to-odd leaves (8e7m operands) and 30 % slower on exact sums.
    GemmKernel drc[3];
};

#define LOF_GEMM_CFG(BM, BN, TM, TN, BK)                                                    \
    {#BM "x" #BN "/" #TM "x" #TN "/k" #BK, BM, BN,                                           \
     {k_gemm<BM, BN, TM, TN, BK, false>, k_gemm<BM, BN, TM, TN, BK, true>}, {}, {}}
#define LOF_MFMA_PAIR(WMB, WNB, C) {k_gemm_mfma<WMB, WNB, C, false>, k_gemm_mfma<WMB, WNB, C, true>}
#define LOF_MFMA_TABLE(WMB, WNB)                                                            \
    {LOF_MFMA_PAIR(WMB, WNB, 1), LOF_MFMA_PAIR(WMB, WNB, 2), LOF_MFMA_PAIR(WMB, WNB, 4),      \
     LOF_MFMA_PAIR(WMB, WNB, 8), LOF_MFMA_PAIR(WMB, WNB, 16), LOF_MFMA_PAIR(WMB, WNB, 32)}
#define LOF_GEMM_MFMA_CFG(WMB, WNB)                                                         \
    {"mfma" #WMB #WNB, 64 * WMB, 64 * WNB, {nullptr, nullptr}, LOF_MFMA_TABLE(WMB, WNB), {}}
#define LOF_GEMM_MFMA_DRC_CFG(WMB, WNB)                                                     \
    {"mfma" #WMB #WNB, 64 * WMB, 64 * WNB, {nullptr, nullptr}, LOF_MFMA_TABLE(WMB, WNB),      \
     {k_gemm_mfma<WMB, WNB, 8, false, true>, k_gemm_mfma<WMB, WNB, 16, false, true>,          \
      k_gemm_mfma<WMB, WNB, 32, false, true>}}

const GemmConfig kConfigs[] = {
    LOF_GEMM_CFG(64, 64, 4, 4, 16),
    LOF_GEMM_CFG(128, 64, 8, 4, 16),
    LOF_GEMM_CFG(64, 128, 4, 8, 16),
    LOF_GEMM_CFG(32, 64, 2, 4, 32),
    LOF_GEMM_CFG(32, 32, 2, 2, 32),
    LOF_GEMM_MFMA_CFG(2, 2),
    LOF_GEMM_MFMA_CFG(2, 1),
    LOF_GEMM_MFMA_DRC_CFG(1, 1),
    LOF_GEMM_MFMA_DRC_CFG(1, 2),
};
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);
constexpr int kMfma64 = 7, kMfma64x128 = 8;

// chunks whose sums can run on MFMA
bool chunk_fits_mfma(int chunk) { return chunk >= 2 && chunk <= 32 && (32 % chunk) == 0; }
// chunks the wave tiles (MFMA configs) take: those, and chunk 1 on VALU
bool chunk_fits_wave_tile(int chunk) { return chunk == 1 || chunk_fits_mfma(chunk); }

bool is_mfma(int config) { return kConfigs[config].kernel[0] == nullptr; }

bool config_valid(int config, int chunk) {
    return config >= 0 && config < kNumConfigs &&
           (!is_mfma(config) || chunk_fits_wave_tile(chunk));
}

GemmKernel pick_kernel(const GemmConfig& c, const GemmArgs& g) {
    if (c.kernel[0]) return c.kernel[g.subn ? 1 : 0];
    const int l2 = 31 - __builtin_clz((unsigned)g.chunk);
    if (l2 >= 3 && c.drc[0] && g.dr_kind == 1 && g.range_f == 0u) return c.drc[l2 - 3];
    return c.mfma[l2][g.subn ? 1 : 0];
}

// autotuned choices, keyed by the problem's shape class and accumulator (kind, mantissa
// bits, format bias, chunk); not by the operands' values, which the tuner never inspects
using TuneKey = std::tuple<int, int, int, int, int, int, int, int>;
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tuned;

int log2_bucket(int v) {
    int b = 0;
    while ((1 << b) < v && b < 30) ++b;
    return b;
}

TuneKey tune_key(const lof_gemm_desc& d) {
    const bool fmt = d.accum.kind == LOF_ACCUM_FORMAT;
    return TuneKey{log2_bucket(d.M), log2_bucket(d.N), log2_bucket(d.K),
                   d.trans_a * 2 + d.trans_b, d.accum.kind,
                   fmt ? d.accum.fmt.mantissa_bits : d.accum.mant_bits, fmt ? d.accum.fmt.bias : 0,
                   d.accum.chunk};
}

// Untuned choice, from measurements at M = N = K = 4096 and on conv-shaped GEMMs
// (M = 64..512, e4m3 / 8e7m inputs): wave tiles for the chunks they take, 64x128 (each
// lane's A values serve two B columns) where that still gives enough tiles -- >= 384 (1.5
// per CU) for chunks 2..32 (+18-22 % over 64x64 at chunks 2 and 4), >= 768 for chunk 1,
// whose exact-sum leaves are LDS-bound (e4m3 +10-15 %, 8e7m on the RTO leaves -4-9 %) --
// else 64x64; the 32x64 VALU tile for other chunks; the 32x32 VALU tile when there are too
// few tiles to fill the chip.
int default_config(const lof_gemm_desc& d) {
    if (chunk_fits_wave_tile(d.accum.chunk)) {
        const int64_t tiles_64x128 = (int64_t)((d.M + 63) / 64) * ((d.N + 127) / 128);
        return tiles_64x128 >= (d.accum.chunk == 1 ? 768 : 384) ? kMfma64x128 : kMfma64;
    }
    if ((int64_t)((d.M + 31) / 32) * ((d.N + 63) / 64) >= 1024) return 3;
    return 4;
}

}  // namespace

void set_fast_path(GemmArgs* g, bool on, int m) {
    g->fast = on ? 1 : 0;
    if (on) {
        g->rh = 20 - m;
        g->half_m1 = (1u << (g->rh - 1)) - 1u;
        g->keep_mask = ~((1u << g->rh) - 1u);
        g->rf = 23 - m;
        g->half_m1f = (1u << (g->rf - 1)) - 1u;
        g->keep_maskf = ~((1u << g->rf) - 1u);
    } else {  // harmless shift values; range 0 sends every leaf to the general path
        g->rh = 1;
        g->half_m1 = 0u;
        g->keep_mask = ~1u;
        g->range = 0u;
        g->rf = 1;
        g->half_m1f = 0u;
        g->keep_maskf = ~1u;
        g->range_f = 0u;
        g->subn = 0;
    }
}

// Fast-path domain on |s|: [2^lo_e, 2^ehi) (double high words and float bits),
// lo_e = max(elo - m + 1, -126), which reaches into the format's subnormal range (rounding
// position 20 - m + (elo - E) <= 19 in the double's high word, 23 - m + (elo - E) <= 22 in
// the float's bits).
static void set_fast_domain(GemmArgs* g, int elo, int ehi, int m) {
    int lo_e = elo - m + 1;
    if (lo_e < -126) lo_e = -126;
    if (lo_e > elo) lo_e = elo;
    g->lo_bound = (uint32_t)(lo_e + 1023) << 20;
    g->range = ((uint32_t)(ehi + 1023) << 20) - g->lo_bound;
    g->lo_f = (uint32_t)(lo_e + 127) << 23;
    g->range_f = ((uint32_t)(ehi + 127) << 23) - g->lo_f;
    g->elo_b = elo + 127;
    g->elo_d = elo + 1023;
    g->subn = lo_e < elo ? 1 : 0;
    g->sum_ehi = ehi;
}

// Bounds for the MFMA tile's exact-sum leaves.  q: every value of the accumulator format is
// a multiple of 2^q (its smallest subnormal; -149 for a mantissa-only accumulator).
// sum_g = (1 + 2^-(m+1))^n for the n = 32 / chunk leaves of one K slice, rounded up to a
// float: one round-to-nearest step grows a magnitude by at most the factor 1 + 2^-(m+1).
static void set_exact_sum_bounds(GemmArgs* g, int m, int q) {
    if (q < -149) q = -149;
    g->sum_q = q;
    g->sum_lreq = q > -126 ? q : -126;
    g->sum_m = m;
    const int n = g->chunk >= 32 ? 1 : 32 / g->chunk;
    const double gr = std::pow(1.0 + std::ldexp(1.0, -(m + 1)), n);
    float gf = (float)gr;
    if ((double)gf < gr) gf = std::nextafter(gf, 1e30f);
    g->sum_g = gf;
    g->sum_n = (float)n;
}

lof_status prepare_gemm(const lof_gemm_desc* d, GemmArgs* g) {
    const char* what = "lof_gemm_f32";
    if (!d) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: desc is NULL", what);
    if (d->M < 0 || d->N < 0 || d->K < 0)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: negative dimension", what);
    const bool empty = d->M == 0 || d->N == 0;
    if (!empty) {
        if (!d->D) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: D is NULL", what);
        if (d->K > 0 && (!d->A || !d->B))
            return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: A or B is NULL", what);
        const int64_t min_lda = d->trans_a ? d->M : d->K;
        const int64_t min_ldb = d->trans_b ? d->K : d->N;
        if (d->K > 0 && (d->lda < min_lda || d->ldb < min_ldb))
            return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: lda/ldb too small", what);
        if (d->ldd < d->N) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: ldd < N", what);
        if (d->beta != 0.0f && (!d->C || d->ldc < d->N))
            return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: beta != 0 needs C with ldc >= N",
                             what);
    }
    const lof_accum& ac = d->accum;
    if (ac.chunk < 1) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: chunk < 1", what);
    if (!mode_valid(ac.spec.mode))
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad accumulator mode", what);
    if (is_stochastic(ac.spec.mode))
        return set_error(LOF_ERR_UNSUPPORTED,
                         "%s: stochastic accumulator rounding is not supported", what);
    *g = GemmArgs{};
    int m;
    if (ac.kind == LOF_ACCUM_MANTISSA) {
        if (ac.mant_bits < 0 || ac.mant_bits > 23)
            return set_error(LOF_ERR_UNSUPPORTED, "%s: accum mant_bits %d outside [0, 23]", what,
                             ac.mant_bits);
        m = ac.mant_bits;
        g->acc_kind = 0;
        g->acc_r = 23 - m;
        set_fast_domain(g, -126, 127, m);
    } else if (ac.kind == LOF_ACCUM_FORMAT) {
        lof_status s = validate_format(&ac.fmt, what);
        if (s != LOF_OK) return s;
        m = ac.fmt.mantissa_bits;
        g->acc_kind = 1;
        g->accF = make_fmt_dev(ac.fmt, lof_round_spec{ac.spec.mode, 0, 0});
        int elo = 1 - ac.fmt.bias, ehi = fmt_max_exp(ac.fmt);
        if (elo < -126) elo = -126;
        if (ehi > 127) ehi = 127;
        if (ehi < elo) ehi = elo;  // empty range: every leaf takes the general path
        set_fast_domain(g, elo, ehi, m);
    } else {
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad accumulator kind %d", what, ac.kind);
    }
    g->acc_mode = ac.spec.mode;
    g->M = d->M; g->N = d->N; g->K = d->K;
    g->A = d->A; g->lda = d->lda; g->ta = d->trans_a ? 1 : 0;
    g->B = d->B; g->ldb = d->ldb; g->tb = d->trans_b ? 1 : 0;
    g->D = d->D; g->ldd = d->ldd;
    g->alpha = d->alpha;
    g->beta = d->beta;
    g->C = d->beta != 0.0f ? d->C : nullptr;
    g->ldc = d->ldc;
    g->chunk = ac.chunk;
    g->vec = (((uintptr_t)d->A | (uintptr_t)d->B) & 15u) == 0 && (d->lda & 3) == 0 &&
             (d->ldb & 3) == 0;
    set_fast_path(g, ac.spec.mode == RNE && m <= 19, m);
    set_exact_sum_bounds(g, m, ac.kind == LOF_ACCUM_FORMAT ? 1 - ac.fmt.bias - m : -149);
    // a binary16-shaped format (10 mantissa bits, bias 15) rounds like the hardware's
    // binary16 conversion wherever |s| < 2^15, which the exact-sum test then requires
    if (ac.kind == LOF_ACCUM_FORMAT && m == 10 && ac.fmt.bias == 15) {
        g->sum_f16 = 1;
        if (g->sum_ehi > 15) g->sum_ehi = 15;
    }
    // the RTO leaves' rounding is exact for every float of these two kinds of accumulator
    g->sum_rto = (ac.kind == LOF_ACCUM_MANTISSA || g->sum_f16) ? 1 : 0;
    // a 10-bit mantissa-only accumulator rounds like binary16 on values scaled into its range
    g->sum_h10 = (ac.kind == LOF_ACCUM_MANTISSA && m == 10) ? 1 : 0;
    // directed-rounding chunk-1 leaves: round-to-nearest-even into a mantissa-only accumulator
    // of any width (a bitwise round of the RTO24 pattern) or a binary16-shaped format
    g->dr_kind = ac.spec.mode != RNE ? 0 : ac.kind == LOF_ACCUM_MANTISSA ? 1 : g->sum_f16 ? 2 : 0;
    g->dr_rf = 23 - m;
    g->dr_half = g->dr_rf > 0 ? (1u << (g->dr_rf - 1)) - 1u : 0u;
    g->dr_lsb = g->dr_rf > 0 ? 1u : 0u;
    g->dr_mask = ~((1u << g->dr_rf) - 1u);
    g->mfma_bf16 = 1;
    return LOF_OK;
}

lof_status launch_gemm(const GemmArgs& g, int config, hipStream_t st) {
    if (g.M == 0 || g.N == 0) return LOF_OK;
    if (!config_valid(config, g.chunk))
        return set_error(LOF_ERR_UNSUPPORTED,
                         "lof_gemm_f32: config %d needs a chunk in {1, 2, 4, 8, 16, 32} (got %d)",
                         config, g.chunk);
    const GemmConfig& c = kConfigs[config];
    dim3 grid((unsigned)((g.N + c.bn - 1) / c.bn), (unsigned)((g.M + c.bm - 1) / c.bm));
    if (grid.y > 65535u) return set_error(LOF_ERR_UNSUPPORTED, "lof_gemm_f32: M too large");
    hipLaunchKernelGGL(pick_kernel(c, g), grid, dim3(256), 0, st, g);
    return check_hip(hipGetLastError(), "lof_gemm_f32 launch");
}

int gemm_num_configs() { return kNumConfigs; }
bool gemm_chunk_fits_mfma(int chunk) { return chunk_fits_mfma(chunk); }
int gemm_config_bm(int c) { return kConfigs[c].bm; }
int gemm_config_bn(int c) { return kConfigs[c].bn; }

int resolve_gemm_config(const lof_gemm_desc& d, int config) {
    if (config >= 0) return config;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(tune_key(d));
    return it != g_tuned.end() ? it->second : default_config(d);
}

}  // namespace lofamd

using namespace lofamd;

extern "C" {

LOF_API lof_status lof_gemm_f32(const lof_gemm_desc* desc, int32_t config, lof_stream_t stream) {
    const int cfg_bits = config >= 0 ? (config & 0xFF) : -1;
    if (cfg_bits >= kNumConfigs)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_gemm_f32: config %d out of range",
                         config);
    GemmArgs g;
    lof_status s = prepare_gemm(desc, &g);
    if (s != LOF_OK) return s;
    // testing bits of config: 0x100 forces the general leaf path; 0x200 runs every chunk of
    // an MFMA config on the VALU chain; 0x400 runs every chunk on MFMA without the exactness
    // test (results then differ where a chunk is inexact); 0x800 turns the MFMA tile's
    // exact-sum leaves off, 0x1000 its RTO leaves, 0x2000 its binary16-held chunk-1 leaves,
    // 0x4000 only their packed binary16 form, 0x8000 the directed-rounding chunk-1 leaves,
    // 0x10000 the bf16 MFMA chunk sums (exact chunks of 8 to 32 then use the VALU chain)
    if (config >= 0 && (config & 0x800)) g.sum_off |= 1;
    if (config >= 0 && (config & 0x1000)) g.sum_off |= 2;
    if (config >= 0 && (config & 0x2000)) g.sum_off |= 4;
    if (config >= 0 && (config & 0x4000)) g.sum_off |= 8;
    if (config >= 0 && (config & 0x8000)) g.sum_off |= 16;
    if (config >= 0 && (config & 0x10000)) g.mfma_bf16 = 0;
    if (config >= 0 && (config & 0x20000)) g.sum_off |= 32;
    if (config >= 0 && (config & 0x100)) {
        set_fast_path(&g, false, 0);
        g.dr_kind = 0;
    }
    if (config >= 0 && (config & 0x200)) g.mfma_mode = 1;
    else if (config >= 0 && (config & 0x400)) g.mfma_mode = 2;
    return launch_gemm(g, resolve_gemm_config(*desc, cfg_bits), (hipStream_t)stream);
}

LOF_API int32_t lof_gemm_num_configs(void) { return kNumConfigs; }

LOF_API int32_t lof_gemm_config_valid(int32_t config, int32_t chunk) {
    return config_valid(config, chunk) ? 1 : 0;
}

LOF_API lof_status lof_gemm_config_name(int32_t config, char* buf, int32_t buflen) {
    if (config < 0 || config >= kNumConfigs || !buf || buflen <= 0)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_gemm_config_name: bad arguments");
    std::snprintf(buf, (size_t)buflen, "%s", kConfigs[config].name);
    return LOF_OK;
}

LOF_API lof_status lof_gemm_autotune(const lof_gemm_desc* desc, int32_t iters,
                                     lof_stream_t stream, int32_t* best_config,
                                     float* best_ms) {
    GemmArgs g;
    lof_status s = prepare_gemm(desc, &g);
    if (s != LOF_OK) return s;
    if (desc->beta != 0.0f && desc->C && desc->M > 0 && desc->N > 0) {
        // the timing launches write D; a source C inside D's rows would be lost
        const float* c0 = desc->C;
        const float* c1 = desc->C + (int64_t)(desc->M - 1) * desc->ldc + desc->N;
        const float* d0 = desc->D;
        const float* d1 = desc->D + (int64_t)(desc->M - 1) * desc->ldd + desc->N;
        if (c0 < d1 && d0 < c1)
            return set_error(LOF_ERR_INVALID_ARGUMENT,
                             "lof_gemm_autotune: C overlaps D, which the timing runs overwrite; "
                             "tune with a separate D");
    }
    g.beta = 0.0f;   // time the plain epilogue: C is never read
    g.C = nullptr;
    if (iters < 1) iters = 1;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if ((s = check_hip(hipEventCreate(&e0), "autotune event")) != LOF_OK) return s;
    if ((s = check_hip(hipEventCreate(&e1), "autotune event")) != LOF_OK) {
        (void)hipEventDestroy(e0);
        return s;
    }
    int best = -1;
    float best_t = 0.f;
    // candidates: the wave tiles for the chunks they take (the VALU tiles, which keep the
    // accumulator in double, measured 6-25x slower there), else the VALU tiles
    const bool wave = chunk_fits_wave_tile(g.chunk);
    for (int c = 0; c < kNumConfigs && s == LOF_OK; ++c) {
        if (!config_valid(c, g.chunk) || is_mfma(c) != wave) continue;
        s = launch_gemm(g, c, st);  // warm-up
        if (s != LOF_OK) break;
        (void)hipEventRecord(e0, st);
        for (int i = 0; i < iters && s == LOF_OK; ++i) s = launch_gemm(g, c, st);
        (void)hipEventRecord(e1, st);
        if (s != LOF_OK) break;
        s = check_hip(hipEventSynchronize(e1), "autotune sync");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= (float)iters;
        if (best < 0 || ms < best_t) { best = c; best_t = ms; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (s != LOF_OK) return s;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[tune_key(*desc)] = best;
    }
    if (best_config) *best_config = best;
    if (best_ms) *best_ms = best_t;
    return LOF_OK;
}

}  // extern "C"
