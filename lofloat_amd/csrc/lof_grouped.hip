// lof_grouped.hip -- one persistent launch over many rounding / MX / GEMM jobs.
//
// The plan uploads one GroupTask per job and the first tile index of each job (jobs are
// ordered GEMMs first, largest first, then rounding jobs, largest first).  GEMMs with
// k-chunk 8, 16 or 32 use the config's MFMA chunk-sum tile (64 x 64, 128 x 64 or 64 x 128),
// the others the config's VALU tile.  The kernel is held to two workgroups per CU (launch
// bounds), which the 128 x 64 MFMA tile would otherwise drop to one, and the configs with
// the 64 x 128 tile to three, as its single-GEMM kernel (measured 11.9 against 12.9 ms on
// the 64-job benchmark mix, although the bound costs spills).  A fixed grid of
// workgroups (as many as fit on the device at once) repeatedly claims the next tile with
// an atomic counter, finds its job by binary search over the start indices, and runs the
// job's tile routine: the same device routines as the single-job kernels, dispatched once
// per tile on the job's rounding mode.  The last workgroup to finish resets the counters,
// so a replayed launch (e.g. from a HIP graph) needs no memset and no host work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "lof_gemm.hpp"
#include "lof_host.hpp"
#include "lof_round_kernels.hpp"

namespace lofamd {

lof_status prepare_round(const float* in, float* out, int64_t n, const lof_format* fmt,
                         int mant_bits, bool mant, lof_round_spec spec, float scale,
                         const uint32_t* rand_bits, uint64_t seed, RoundArgs* a, int* rsrc);
lof_status prepare_mx(const float* in, float* out, int64_t n, int32_t bs,
                      const lof_format* elem_fmt, const lof_format* scale_fmt,
                      lof_round_spec elem_spec, int32_t scale_mode, const uint32_t* rand_bits,
                      uint64_t seed, MxArgs* a, int* rsrc);
lof_status launch_round(const RoundArgs& a, bool mant, int mode, int rsrc, hipStream_t st);
lof_status launch_mx(const MxArgs& a, int mode, int rsrc, hipStream_t st);

namespace {

struct GroupTask {
    int32_t kind;     // LOF_ITEM_*
    int32_t mode;     // element rounding mode
    int32_t rsrc;     // 0 no random words, 1 bits array, 2 generated words
    int32_t tiles_n;  // GEMM: tiles along N
    int32_t mfma;     // GEMM: 64 x 64 MFMA chunk-sum tiles (else the config's VALU tile)
    RoundArgs r;
    MxArgs x;
    GemmArgs g;
};

constexpr int kRVPT = 2;  // float4 slots per lane per sub-tile

static_assert(256 * kRVPT * 4 == kMxTileElems, "rounding and MX sub-tiles are 2048 floats");

template <int BM, int BN, int TM, int TN, int BK, int SUB, int WMB, int WNB>
struct GroupedCfg {
    static constexpr int bm = BM, bn = BN, tm = TM, tn = TN, bk = BK, sub = SUB;
    static constexpr int mfma_wmb = WMB;   // MFMA GEMM tile: (64 * WMB) x (64 * WNB)
    static constexpr int mfma_wnb = WNB;
    // floats per rounding / MX tile: SUB sub-tiles of 2048 floats
    static constexpr int64_t round_tile_elems = (int64_t)SUB * kMxTileElems;
};

// Each (mode, random source) tile routine is its own non-inlined function: one call per
// sub-tile, and the register allocator works on small functions instead of one kernel
// holding every variant.
template <int MODE, bool MANT, int RSRC>
__device__ __noinline__ void g_round_tile(const RoundArgs& a, int64_t st, int tid) {
    round_tile<MODE, MANT, RSRC, kRVPT>(a, st, tid);
}

template <int MODE, int RSRC>
__device__ __noinline__ void g_mx_tile(const MxArgs& a, int64_t st, int tid) {
    if (a.lanes_log2 == 0) mx_tile<MODE, RSRC, 4, false>(a, st, tid);  // block_size 4
    else mx_tile<MODE, RSRC, 8, false>(a, st, tid);
}

template <bool MANT>
__device__ __noinline__ void grouped_round(const RoundArgs& a, int mode, int rsrc,
                                           int64_t tile, int sub, int tid) {
    for (int s = 0; s < sub; ++s) {
        const int64_t st = tile * sub + s;
#define LOF_RC(M, R) \
    if (mode == M && rsrc == R) { g_round_tile<M, MANT, R>(a, st, tid); continue; }
        LOF_RC(RNE, 0) LOF_RC(RTZ, 0) LOF_RC(RAFZ, 0) LOF_RC(RNO, 0) LOF_RC(RDOWN, 0)
        LOF_RC(RUP, 0) LOF_RC(RTA, 0) LOF_RC(RTO, 0) LOF_RC(RTTZ, 0) LOF_RC(SRA, 0)
        LOF_RC(SRB, 0) LOF_RC(SRC, 0) LOF_RC(SRD, 0)
        LOF_RC(SRA, 1) LOF_RC(SRB, 1) LOF_RC(SRC, 1) LOF_RC(TRUESR, 1) LOF_RC(PROB, 1)
        LOF_RC(SRD, 1)
        LOF_RC(SRA, 2) LOF_RC(SRB, 2) LOF_RC(SRC, 2) LOF_RC(TRUESR, 2) LOF_RC(PROB, 2)
        LOF_RC(SRD, 2)
#undef LOF_RC
    }
}

__device__ __noinline__ void grouped_mx(const MxArgs& a, int mode, int rsrc, int64_t tile,
                                        int sub, int tid) {
    for (int s = 0; s < sub; ++s) {
        const int64_t st = tile * sub + s;
#define LOF_MC(M, R) \
    if (mode == M && rsrc == R) { g_mx_tile<M, R>(a, st, tid); continue; }
        LOF_MC(RNE, 0) LOF_MC(RTZ, 0) LOF_MC(RAFZ, 0) LOF_MC(RNO, 0) LOF_MC(RDOWN, 0)
        LOF_MC(RUP, 0) LOF_MC(RTA, 0) LOF_MC(RTO, 0) LOF_MC(RTTZ, 0) LOF_MC(SRA, 0)
        LOF_MC(SRB, 0) LOF_MC(SRC, 0) LOF_MC(SRD, 0)
        LOF_MC(SRA, 1) LOF_MC(SRB, 1) LOF_MC(SRC, 1) LOF_MC(TRUESR, 1) LOF_MC(PROB, 1)
        LOF_MC(SRD, 1)
        LOF_MC(SRA, 2) LOF_MC(SRB, 2) LOF_MC(SRC, 2) LOF_MC(TRUESR, 2) LOF_MC(PROB, 2)
        LOF_MC(SRD, 2)
#undef LOF_MC
    }
}

// GEMM tiles whose chunk sums run on MFMA (chunks 8, 16, 32): the 64 x 64 tile per
// (chunk, subnormal fast range) variant, inlined (as non-inlined calls they raise the kernel's register count
// enough to halve its occupancy).
template <int WMB, int WNB, int CHUNK, bool SUBN>
__device__ __forceinline__ void g_mfma_tile(const GemmArgs& g, int tm, int tn, int tid,
                                         float* smem) {
    gemm_tile_mfma<WMB, WNB, CHUNK, SUBN, false, false, true>(g, tm, tn, tid, smem);
}

template <int WMB, int WNB>
__device__ __forceinline__ void grouped_mfma(const GemmArgs& g, int tm, int tn, int tid,
                                          float* smem) {
#define LOF_GM(C)                                                             \
    if (g.chunk == C) {                                                       \
        if (g.subn) g_mfma_tile<WMB, WNB, C, true>(g, tm, tn, tid, smem);     \
        else g_mfma_tile<WMB, WNB, C, false>(g, tm, tn, tid, smem);           \
        return;                                                               \
    }
    LOF_GM(8) LOF_GM(16) LOF_GM(32)
#undef LOF_GM
}

template <class C>
__global__ __launch_bounds__(256, C::mfma_wnb == 2 ? 3 : 2) void k_grouped(const GroupTask* __restrict__ tasks,
                                                 const int64_t* __restrict__ starts, int ntasks,
                                                 int64_t total, unsigned long long* counters) {
    constexpr int kValuSmem = gemm_smem_floats<C::bm, C::bn, C::bk>();
    constexpr int kMfmaSmem = MfmaTile<C::mfma_wmb, C::mfma_wnb>::SMEM_FLOATS;
    __shared__ __align__(16) float smem[kValuSmem > kMfmaSmem ? kValuSmem : kMfmaSmem];
    __shared__ unsigned long long s_tile;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_tile = atomicAdd(&counters[0], 1ull);
        __syncthreads();
        const int64_t t = (int64_t)__builtin_amdgcn_readfirstlane((int)(s_tile & 0x7FFFFFFF));
        __syncthreads();
        if (t >= total) break;
        int lo = 0, hi = ntasks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (starts[mid] <= t) lo = mid; else hi = mid - 1;
        }
        const GroupTask& T = tasks[lo];
        const int64_t local = t - starts[lo];
        switch (T.kind) {
        case LOF_ITEM_ROUND:
            grouped_round<false>(T.r, T.mode, T.rsrc, local, C::sub, tid);
            break;
        case LOF_ITEM_ROUND_MANT:
            grouped_round<true>(T.r, T.mode, T.rsrc, local, C::sub, tid);
            break;
        case LOF_ITEM_MX:
            grouped_mx(T.x, T.mode, T.rsrc, local, C::sub, tid);
            break;
        default:
            if (T.mfma)
                grouped_mfma<C::mfma_wmb, C::mfma_wnb>(T.g, (int)(local / T.tiles_n), (int)(local % T.tiles_n),
                                          tid, smem);
            else if (T.g.subn)
                gemm_tile<C::bm, C::bn, C::tm, C::tn, C::bk, true>(
                    T.g, (int)(local / T.tiles_n), (int)(local % T.tiles_n), tid, smem);
            else
                gemm_tile<C::bm, C::bn, C::tm, C::tn, C::bk, false>(
                    T.g, (int)(local / T.tiles_n), (int)(local % T.tiles_n), tid, smem);
            __syncthreads();  // smem reuse by the next tile
            break;
        }
    }
    if (tid == 0) {
        __threadfence();
        const unsigned long long done = atomicAdd(&counters[1], 1ull);
        if (done == (unsigned long long)gridDim.x - 1) {  // last one out resets the queue
            atomicExch(&counters[0], 0ull);
            atomicExch(&counters[1], 0ull);
        }
    }
}

struct GroupedConfig {
    const char* name;
    int bm, bn;
    int mfma_bm, mfma_bn;   // MFMA GEMM tile (64 x 64, 128 x 64 or 64 x 128)
    int64_t round_tile_elems;
    void (*kernel)(const GroupTask*, const int64_t*, int, int64_t, unsigned long long*);
};

#define LOF_GCFG(BM, BN, TM, TN, BK, SUB, WMB, WNB)                                        \
    {#BM "x" #BN "/" #TM "x" #TN " r" #SUB " mfma" #WMB #WNB, BM, BN, 64 * WMB, 64 * WNB,   \
     GroupedCfg<BM, BN, TM, TN, BK, SUB, WMB, WNB>::round_tile_elems,                     \
     k_grouped<GroupedCfg<BM, BN, TM, TN, BK, SUB, WMB, WNB>>}

const GroupedConfig kGConfigs[] = {
    LOF_GCFG(64, 64, 4, 4, 16, 4, 1, 1),
    LOF_GCFG(64, 64, 4, 4, 16, 8, 1, 1),
    LOF_GCFG(32, 64, 2, 4, 32, 4, 1, 1),
    LOF_GCFG(32, 64, 2, 4, 32, 8, 1, 1),
    LOF_GCFG(64, 64, 4, 4, 16, 4, 2, 1),
    LOF_GCFG(64, 64, 4, 4, 16, 8, 2, 1),
    LOF_GCFG(64, 64, 4, 4, 16, 4, 1, 2),
    LOF_GCFG(64, 64, 4, 4, 16, 8, 1, 2),
};
constexpr int kNumGConfigs = sizeof(kGConfigs) / sizeof(kGConfigs[0]);
int g_default_gconfig = 4;   // 64 x 64 VALU, 128 x 64 MFMA, 4 rounding sub-tiles
std::mutex g_gmu;

// Fixed device workspace of lof_grouped_run_noalloc: the work-queue counters, the job
// start table and the task table, laid out as
//   [0, 256)                      two unsigned long long counters (then padding)
//   [kNoallocStartsOff, ...)      kNoallocMaxItems int64_t tile starts
//   [kNoallocTasksOff, ...)       kNoallocMaxItems GroupTask records
constexpr int kNoallocMaxItems = 256;
constexpr size_t kNoallocStartsOff = 256;
constexpr size_t kNoallocTasksOff =
    kNoallocStartsOff + ((kNoallocMaxItems * sizeof(int64_t) + 255) / 256) * 256;
constexpr size_t kNoallocBytes = kNoallocTasksOff + kNoallocMaxItems * sizeof(GroupTask);
static_assert(alignof(GroupTask) <= 256, "task table alignment");

}  // namespace

}  // namespace lofamd

// Device workspace of lof_grouped_run_noalloc (kNoallocBytes, about 170 KB per GPU, held
// until the process ends).  Host code never uses this name as an address: it takes each
// device's address from hipGetSymbolAddress once and passes that.
static __device__ __align__(256) unsigned char scratch[lofamd::kNoallocBytes];

using namespace lofamd;

struct lof_grouped {
    int device = 0;
    int config = 0;
    int grid = 0;
    int64_t total = 0;
    int ntasks = 0;
    GroupTask* d_tasks = nullptr;
    int64_t* d_starts = nullptr;
    unsigned long long* d_counters = nullptr;
    std::vector<GroupTask> tasks;   // host copy, in queue order
    std::vector<GroupTask> direct;  // device-filling GEMMs, each run as its own launch
};

struct lof_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

lof_status build_task(const lof_group_item& it, const GroupedConfig& cfg, GroupTask* t,
                      int64_t* tiles) {
    *t = GroupTask{};
    t->kind = it.kind;
    lof_status s;
    switch (it.kind) {
    case LOF_ITEM_ROUND:
    case LOF_ITEM_ROUND_MANT: {
        const bool mant = it.kind == LOF_ITEM_ROUND_MANT;
        s = prepare_round(it.in, it.out, it.n, mant ? nullptr : &it.fmt, it.mant_bits, mant,
                          it.spec, it.scale, it.rand_bits, it.seed, &t->r, &t->rsrc);
        if (s != LOF_OK) return s;
        t->mode = it.spec.mode;
        *tiles = (it.n + cfg.round_tile_elems - 1) / cfg.round_tile_elems;
        return LOF_OK;
    }
    case LOF_ITEM_MX:
        s = prepare_mx(it.in, it.out, it.n, it.block_size, &it.fmt, &it.scale_fmt, it.spec,
                       it.scale_mode, it.rand_bits, it.seed, &t->x, &t->rsrc);
        if (s != LOF_OK) return s;
        t->mode = it.spec.mode;
        *tiles = (it.n + cfg.round_tile_elems - 1) / cfg.round_tile_elems;
        return LOF_OK;
    case LOF_ITEM_GEMM:
        s = prepare_gemm(&it.gemm, &t->g);
        if (s != LOF_OK) return s;
        t->mfma = (it.gemm.accum.chunk == 8 || it.gemm.accum.chunk == 16 ||
                   it.gemm.accum.chunk == 32) ? 1 : 0;
        {
            const int bm = t->mfma ? cfg.mfma_bm : cfg.bm;
            const int bn = t->mfma ? cfg.mfma_bn : cfg.bn;
            t->tiles_n = (it.gemm.N + bn - 1) / bn;
            *tiles = (int64_t)t->tiles_n * ((it.gemm.M + bm - 1) / bm);
        }
        if (it.gemm.M == 0 || it.gemm.N == 0) *tiles = 0;
        return LOF_OK;
    default:
        return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_grouped_create: bad item kind %d",
                         it.kind);
    }
}

double task_cost(const GroupTask& t, int64_t tiles) {
    if (t.kind == LOF_ITEM_GEMM) return 1e12 + (double)t.g.M * t.g.N * (double)t.g.K;
    return (double)tiles;
}

lof_status free_plan(lof_grouped* p) {
    if (!p) return LOF_OK;
    if (p->d_tasks) (void)hipFree(p->d_tasks);
    if (p->d_starts) (void)hipFree(p->d_starts);
    if (p->d_counters) (void)hipFree(p->d_counters);
    delete p;
    return LOF_OK;
}


int resolve_gconfig(int config) {
    if (config >= 0) return config;
    std::lock_guard<std::mutex> lk(g_gmu);
    return g_default_gconfig;
}

// Validates every item and lays the jobs out in queue order: GEMMs first, largest first,
// then rounding jobs, largest first; jobs with no tiles are dropped.  A GEMM with at least
// `fill` tiles (the persistent grid's workgroups: it fills the device by itself) goes to
// `direct` instead, to run as its own launch of the single-GEMM kernel, whose tile outruns
// the combined kernel's (its registers are its own); the outputs are the same either way.
// work / tiles / order are caller-provided scratch of `count` entries; tasks / starts
// receive *ntasks entries, direct *ndirect.
lof_status order_tasks(const lof_group_item* items, int count, const GroupedConfig& cfg,
                       int64_t fill, GroupTask* work, int64_t* tiles, int* order,
                       GroupTask* tasks, int64_t* starts, int* ntasks, int64_t* total,
                       GroupTask* direct, int* ndirect) {
    for (int i = 0; i < count; ++i) {
        lof_status s = build_task(items[i], cfg, &work[i], &tiles[i]);
        if (s != LOF_OK) return s;
        order[i] = i;
    }
    std::stable_sort(order, order + count, [&](int a, int b) {
        return task_cost(work[a], tiles[a]) > task_cost(work[b], tiles[b]);
    });
    int nt = 0, nd = 0;
    int64_t tot = 0;
    for (int i = 0; i < count; ++i) {
        const int idx = order[i];
        if (tiles[idx] == 0) continue;
        if (work[idx].kind == LOF_ITEM_GEMM && tiles[idx] >= fill) {
            direct[nd++] = work[idx];
            continue;
        }
        starts[nt] = tot;
        tasks[nt] = work[idx];
        tot += tiles[idx];
        ++nt;
    }
    if (tot >= (int64_t)1 << 30)
        return set_error(LOF_ERR_UNSUPPORTED, "lof_grouped: more than 2^30 tiles");
    *ntasks = nt;
    *total = tot;
    *ndirect = nd;
    return LOF_OK;
}

// Workgroups of the config's kernel the device holds at once: the persistent grid is that
// many, at most one per tile.
lof_status resident_slots(const GroupedConfig& cfg, int device, int64_t* resident) {
    int cus = 0, per_cu = 0;
    lof_status s = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                                   device), "device attribute");
    if (s == LOF_OK)
        s = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &per_cu, reinterpret_cast<const void*>(cfg.kernel), 256, 0),
                      "occupancy");
    if (s == LOF_OK) *resident = (int64_t)cus * std::max(per_cu, 1);
    return s;
}

int persistent_grid(int64_t resident, int64_t total) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(resident, total));
}

// One job through its own single-job entry point.
lof_status launch_alone(const GroupTask& t, hipStream_t st) {
    switch (t.kind) {
    case LOF_ITEM_ROUND: return launch_round(t.r, false, t.mode, t.rsrc, st);
    case LOF_ITEM_ROUND_MANT: return launch_round(t.r, true, t.mode, t.rsrc, st);
    case LOF_ITEM_MX: return launch_mx(t.x, t.mode, t.rsrc, st);
    default: {
        lof_gemm_desc d{};
        d.M = t.g.M; d.N = t.g.N; d.K = t.g.K;
        d.trans_a = t.g.ta; d.trans_b = t.g.tb;
        d.accum.kind = t.g.acc_kind;
        d.accum.chunk = t.g.chunk;
        return launch_gemm(t.g, resolve_gemm_config(d, -1), st);
    }
    }
}

// Host state of lof_grouped_run_noalloc: each device's workspace address and one host image
// of the workspace with the planner's per-item scratch, all fixed-size.
constexpr int kNoallocMaxDevices = 64;
struct NoallocImage {
    alignas(256) unsigned char bytes[kNoallocBytes];
    GroupTask work[kNoallocMaxItems];
    GroupTask direct[kNoallocMaxItems];
    int64_t tiles[kNoallocMaxItems];
    int order[kNoallocMaxItems];
};
std::mutex g_noalloc_mu;
unsigned char* g_noalloc_ptr[kNoallocMaxDevices] = {};
NoallocImage g_noalloc_image;

}  // namespace

extern "C" {

LOF_API int32_t lof_grouped_num_configs(void) { return kNumGConfigs; }

LOF_API lof_status lof_grouped_config_name(int32_t config, char* buf, int32_t buflen) {
    if (config < 0 || config >= kNumGConfigs || !buf || buflen <= 0)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_grouped_config_name: bad arguments");
    std::snprintf(buf, (size_t)buflen, "%s", kGConfigs[config].name);
    return LOF_OK;
}

LOF_API lof_status lof_grouped_create(const lof_group_item* items, int32_t count, int32_t config,
                                      lof_grouped** plan) {
    if (!plan || count < 0 || (count > 0 && !items))
        return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_grouped_create: bad arguments");
    *plan = nullptr;
    if (config >= kNumGConfigs)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_grouped_create: config out of range");
    config = resolve_gconfig(config);
    const GroupedConfig& cfg = kGConfigs[config];
    auto* p = new lof_grouped();
    p->config = config;
    p->tasks.resize((size_t)count);
    p->direct.resize((size_t)count);
    std::vector<int64_t> starts((size_t)count);
    std::vector<GroupTask> work((size_t)count);
    std::vector<int64_t> tiles((size_t)count);
    std::vector<int> order((size_t)count);
    int64_t resident = 0;
    int ndirect = 0;
    lof_status s = check_hip(hipGetDevice(&p->device), "hipGetDevice");
    if (s == LOF_OK) s = resident_slots(cfg, p->device, &resident);
    if (s == LOF_OK)
        s = order_tasks(items, count, cfg, resident, work.data(), tiles.data(), order.data(),
                        p->tasks.data(), starts.data(), &p->ntasks, &p->total,
                        p->direct.data(), &ndirect);
    if (s == LOF_OK) {
        p->tasks.resize((size_t)p->ntasks);
        p->direct.resize((size_t)ndirect);
        starts.resize((size_t)p->ntasks);
        p->grid = persistent_grid(resident, p->total);
    }
    if (s == LOF_OK && p->ntasks > 0) {
        s = check_hip(hipMalloc(&p->d_tasks, sizeof(GroupTask) * p->tasks.size()), "hipMalloc");
        if (s == LOF_OK)
            s = check_hip(hipMalloc(&p->d_starts, sizeof(int64_t) * starts.size()), "hipMalloc");
        if (s == LOF_OK)
            s = check_hip(hipMalloc(&p->d_counters, 2 * sizeof(unsigned long long)), "hipMalloc");
        if (s == LOF_OK)
            s = check_hip(hipMemcpy(p->d_tasks, p->tasks.data(),
                                    sizeof(GroupTask) * p->tasks.size(), hipMemcpyHostToDevice),
                          "hipMemcpy");
        if (s == LOF_OK)
            s = check_hip(hipMemcpy(p->d_starts, starts.data(), sizeof(int64_t) * starts.size(),
                                    hipMemcpyHostToDevice), "hipMemcpy");
        if (s == LOF_OK)
            s = check_hip(hipMemset(p->d_counters, 0, 2 * sizeof(unsigned long long)),
                          "hipMemset");
        if (s == LOF_OK) s = check_hip(hipDeviceSynchronize(), "plan upload");
        if (s != LOF_OK) s = (s == LOF_ERR_HIP) ? LOF_ERR_ALLOC : s;
    }
    if (s != LOF_OK) {
        free_plan(p);
        return s;
    }
    *plan = p;
    return LOF_OK;
}

LOF_API lof_status lof_grouped_run(lof_grouped* p, lof_stream_t stream) {
    if (!p) return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_grouped_run: NULL plan");
    for (const GroupTask& t : p->direct) {
        const lof_status s = launch_alone(t, (hipStream_t)stream);
        if (s != LOF_OK) return s;
    }
    if (p->total == 0) return LOF_OK;
    const GroupedConfig& cfg = kGConfigs[p->config];
    hipLaunchKernelGGL(cfg.kernel, dim3((unsigned)p->grid), dim3(256), 0, (hipStream_t)stream,
                       p->d_tasks, p->d_starts, p->ntasks, p->total, p->d_counters);
    return check_hip(hipGetLastError(), "lof_grouped_run launch");
}

LOF_API lof_status lof_grouped_run_separate(lof_grouped* p, lof_stream_t stream) {
    if (!p) return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_grouped_run_separate: NULL plan");
    hipStream_t st = (hipStream_t)stream;
    for (const auto* list : {&p->direct, &p->tasks})
        for (const GroupTask& t : *list) {
            const lof_status s = launch_alone(t, st);
            if (s != LOF_OK) return s;
        }
    return LOF_OK;
}

LOF_API lof_status lof_grouped_destroy(lof_grouped* p) { return free_plan(p); }

LOF_API lof_status lof_grouped_info(const lof_grouped* p, int64_t* tiles, int32_t* grid) {
    if (!p) return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_grouped_info: NULL plan");
    if (tiles) *tiles = p->total;
    if (grid) *grid = p->grid;
    return LOF_OK;
}

LOF_API lof_status lof_grouped_autotune(const lof_group_item* items, int32_t count, int32_t iters,
                                        lof_stream_t stream, int32_t* best_config,
                                        float* best_ms) {
    if (iters < 1) iters = 1;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    lof_status s = check_hip(hipEventCreate(&e0), "event");
    if (s != LOF_OK) return s;
    if ((s = check_hip(hipEventCreate(&e1), "event")) != LOF_OK) {
        (void)hipEventDestroy(e0);
        return s;
    }
    int best = -1;
    float best_t = 0.f;
    for (int c = 0; c < kNumGConfigs && s == LOF_OK; ++c) {
        lof_grouped* p = nullptr;
        s = lof_grouped_create(items, count, c, &p);
        if (s != LOF_OK) break;
        s = lof_grouped_run(p, st);  // warm-up
        if (s == LOF_OK) s = check_hip(hipEventRecord(e0, st), "event record");
        for (int i = 0; i < iters && s == LOF_OK; ++i) s = lof_grouped_run(p, st);
        if (s == LOF_OK) s = check_hip(hipEventRecord(e1, st), "event record");
        if (s == LOF_OK) s = check_hip(hipEventSynchronize(e1), "event sync");
        float ms = 0.f;
        if (s == LOF_OK) (void)hipEventElapsedTime(&ms, e0, e1);
        free_plan(p);
        ms /= (float)iters;
        if (s == LOF_OK && (best < 0 || ms < best_t)) { best = c; best_t = ms; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (s != LOF_OK) return s;
    {
        std::lock_guard<std::mutex> lk(g_gmu);
        g_default_gconfig = best;
    }
    if (best_config) *best_config = best;
    if (best_ms) *best_ms = best_t;
    return LOF_OK;
}

LOF_API lof_status lof_grouped_run_noalloc(const lof_group_item* items, int32_t count,
                                           int32_t config, lof_stream_t stream) {
    if (count < 0 || (count > 0 && !items))
        return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_grouped_run_noalloc: bad arguments");
    if (count > kNoallocMaxItems)
        return set_error(LOF_ERR_UNSUPPORTED,
                         "lof_grouped_run_noalloc: %d items exceed the fixed workspace's %d "
                         "(use lof_grouped_create for larger lists)", count, kNoallocMaxItems);
    if (config >= kNumGConfigs)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_grouped_run_noalloc: config out of range");
    config = resolve_gconfig(config);
    const GroupedConfig& cfg = kGConfigs[config];
    int device = 0;
    lof_status s = check_hip(hipGetDevice(&device), "hipGetDevice");
    if (s != LOF_OK) return s;
    if (device < 0 || device >= kNoallocMaxDevices)
        return set_error(LOF_ERR_UNSUPPORTED, "lof_grouped_run_noalloc: device %d", device);
    // One host image of the workspace, built and copied under the lock: the copy from
    // pageable host memory has consumed the image when hipMemcpyAsync returns.
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    unsigned char* dptr = g_noalloc_ptr[device];
    if (!dptr) {   // first call on this device (which is current): its copy's address
        void* ptr = nullptr;
        s = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        if (s != LOF_OK) return s;
        dptr = g_noalloc_ptr[device] = static_cast<unsigned char*>(ptr);
    }
    NoallocImage& im = g_noalloc_image;
    int ntasks = 0, ndirect = 0;
    int64_t total = 0, resident = 0;
    s = resident_slots(cfg, device, &resident);
    if (s != LOF_OK) return s;
    s = order_tasks(items, count, cfg, resident, im.work, im.tiles, im.order,
                    reinterpret_cast<GroupTask*>(im.bytes + kNoallocTasksOff),
                    reinterpret_cast<int64_t*>(im.bytes + kNoallocStartsOff), &ntasks, &total,
                    im.direct, &ndirect);
    if (s != LOF_OK) return s;
    hipStream_t st = (hipStream_t)stream;
    for (int i = 0; i < ndirect; ++i) {   // launches only: nothing is allocated
        s = launch_alone(im.direct[i], st);
        if (s != LOF_OK) return s;
    }
    if (total == 0) return LOF_OK;
    const int grid = persistent_grid(resident, total);
    // counters start at zero (the kernel's last workgroup also leaves them at zero)
    std::memset(im.bytes, 0, kNoallocStartsOff);
    const size_t used = kNoallocTasksOff + (size_t)ntasks * sizeof(GroupTask);
    s = check_hip(hipMemcpyAsync(dptr, im.bytes, used, hipMemcpyHostToDevice, st),
                  "lof_grouped_run_noalloc upload");
    if (s != LOF_OK) return s;
    hipLaunchKernelGGL(cfg.kernel, dim3((unsigned)grid), dim3(256), 0, st,
                       reinterpret_cast<const GroupTask*>(dptr + kNoallocTasksOff),
                       reinterpret_cast<const int64_t*>(dptr + kNoallocStartsOff), ntasks, total,
                       reinterpret_cast<unsigned long long*>(dptr));
    return check_hip(hipGetLastError(), "lof_grouped_run_noalloc launch");
}

LOF_API lof_status lof_graph_capture_grouped(lof_grouped* p, int32_t separate,
                                             lof_stream_t stream, lof_graph** graph) {
    if (!p || !graph || !stream)
        return set_error(LOF_ERR_INVALID_ARGUMENT,
                         "lof_graph_capture_grouped: needs a plan and a non-default stream");
    hipStream_t st = (hipStream_t)stream;
    lof_status s = check_hip(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal),
                             "hipStreamBeginCapture");
    if (s != LOF_OK) return s;
    lof_status run = separate ? lof_grouped_run_separate(p, st) : lof_grouped_run(p, st);
    auto* g = new lof_graph();
    s = check_hip(hipStreamEndCapture(st, &g->graph), "hipStreamEndCapture");
    if (run != LOF_OK) s = run;
    if (s == LOF_OK)
        s = check_hip(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0),
                      "hipGraphInstantiate");
    if (s != LOF_OK) {
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return s;
    }
    *graph = g;
    return LOF_OK;
}

LOF_API lof_status lof_graph_launch(lof_graph* g, lof_stream_t stream) {
    if (!g) return set_error(LOF_ERR_INVALID_ARGUMENT, "lof_graph_launch: NULL graph");
    return check_hip(hipGraphLaunch(g->exec, (hipStream_t)stream), "hipGraphLaunch");
}

LOF_API lof_status lof_graph_destroy(lof_graph* g) {
    if (!g) return LOF_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return LOF_OK;
}

}  // extern "C"
