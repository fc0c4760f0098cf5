// lof_conv.hip -- Conv2d forward as an implicit GEMM on the wave tiles (C-ABI lof_conv2d_f32).
//
// Replaces lo_float::LoF_conv2d (Lof_kernel.cu:156-228) and lo_float::Conv2d
// (cutlass_conv.cuh:20-184), which run CUTLASS's implicit-GEMM fprop over the same
// accumulation leaf as the GEMM.  Here the convolution is the GEMM
//     D(K x N*P*Q) = W(K x C*R*S) * B(C*R*S x N*P*Q),   B(k, n) = x at tap k of position n,
// computed by the wave-tile kernel with CONV = true: each workgroup gathers its 32 x BN
// slice of B straight from x (zero in the padding) into LDS, where the plain GEMM would
// load a stored matrix, and writes D at the output tensor's strided positions.  The
// reduction, the leaf, the MFMA exactness test and the chain fallback are the GEMM's, so
// each output equals lof_gemm_f32 of W with the unfolded input.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lof_gemm.hpp"
#include "lof_host.hpp"

namespace lofamd {

namespace {

// the 64x64 tile at 4 waves per SIMD, as the GEMM's; the 128x64 tile held to 2 (256
// registers: at 260 it drops to one wave per SIMD and runs 1.5x longer)
template <int WMB, int WNB, int CHUNK, bool SUBN>
__global__ __launch_bounds__(256, WMB == 1 ? 4 : 2)
void k_conv_mfma(GemmArgs g) {
    __shared__ __align__(16) float smem[MfmaTile<WMB, WNB>::SMEM_FLOATS];
    gemm_tile_mfma<WMB, WNB, CHUNK, SUBN, true>(g, (int)blockIdx.y, (int)blockIdx.x,
                                                (int)threadIdx.x, smem);
}

using ConvKernel = void (*)(GemmArgs);

// [tile][log2 chunk][subn]: tile 0 = 64 x 64 (mfma11), tile 1 = 128 x 64 (mfma21), the two
// wave tiles the GEMM's default choice uses.
#define LOF_CONV_PAIR(WMB, C) {k_conv_mfma<WMB, 1, C, false>, k_conv_mfma<WMB, 1, C, true>}
#define LOF_CONV_TILE(WMB)                                                                 \
    {LOF_CONV_PAIR(WMB, 1), LOF_CONV_PAIR(WMB, 2), LOF_CONV_PAIR(WMB, 4),                  \
     LOF_CONV_PAIR(WMB, 8), LOF_CONV_PAIR(WMB, 16), LOF_CONV_PAIR(WMB, 32)}
const ConvKernel kConvKernels[2][6][2] = {LOF_CONV_TILE(1), LOF_CONV_TILE(2)};
#undef LOF_CONV_TILE
#undef LOF_CONV_PAIR

// GEMM config indices of the two tiles (lof_gemm_config_name: "mfma11", "mfma21")
constexpr int kCfgMfma21 = 6, kCfgMfma11 = 7;

int log2_chunk(int chunk) {
    switch (chunk) {
    case 1: return 0;
    case 2: return 1;
    case 4: return 2;
    case 8: return 3;
    case 16: return 4;
    case 32: return 5;
    default: return -1;
    }
}

}  // namespace

}  // namespace lofamd

using namespace lofamd;

extern "C" {

LOF_API lof_status lof_conv2d_f32(const lof_conv2d_desc* desc, int32_t config,
                                  lof_stream_t stream) {
    const char* what = "lof_conv2d_f32";
    if (!desc) return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: desc is NULL", what);
    const lof_conv2d_desc& c = *desc;
    if (c.N < 0 || c.H < 1 || c.W < 1 || c.C < 1 || c.K < 0 || c.R < 1 || c.S < 1 ||
        c.pad_h < 0 || c.pad_w < 0 || c.stride_h < 1 || c.stride_w < 1 || c.dil_h < 1 ||
        c.dil_w < 1)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad problem size", what);
    if (c.layout != LOF_CONV_NCHW && c.layout != LOF_CONV_NHWC)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: bad layout %d", what, c.layout);
    const int64_t P =
        ((int64_t)c.H + 2 * (int64_t)c.pad_h - (int64_t)c.dil_h * (c.R - 1) - 1) / c.stride_h + 1;
    const int64_t Q =
        ((int64_t)c.W + 2 * (int64_t)c.pad_w - (int64_t)c.dil_w * (c.S - 1) - 1) / c.stride_w + 1;
    if (P < 1 || Q < 1)
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: non-positive output size (P=%lld, Q=%lld)",
                         what, (long long)P, (long long)Q);
    const int64_t HW = (int64_t)c.H * c.W, x_elems = (int64_t)c.N * c.C * HW;
    const int64_t Ncols = (int64_t)c.N * P * Q, Kred = (int64_t)c.C * c.R * c.S;
    const int64_t lim = (int64_t)1 << 30;
    if (x_elems >= lim || Ncols >= lim || Kred >= lim || (int64_t)c.R * c.dil_h >= lim ||
        (int64_t)c.S * c.dil_w >= lim || (int64_t)c.H + c.pad_h >= lim ||
        (int64_t)c.W + c.pad_w >= lim || (int64_t)P * c.stride_h >= lim ||
        (int64_t)Q * c.stride_w >= lim)
        return set_error(LOF_ERR_UNSUPPORTED, "%s: sizes beyond 32-bit indexing", what);
    if (c.N > 0 && c.K > 0 && (!c.x || !c.w || !c.y))
        return set_error(LOF_ERR_INVALID_ARGUMENT, "%s: NULL tensor", what);

    // the GEMM this convolution is: D(K x N*P*Q) = W(K x CRS) * B(CRS x N*P*Q)
    lof_gemm_desc d{};
    d.M = c.K;
    d.N = (int32_t)Ncols;
    d.K = (int32_t)Kred;
    d.A = c.w;
    d.lda = Kred;
    d.B = c.x;
    d.ldb = Ncols;
    d.D = c.y;
    d.ldd = Ncols;
    d.alpha = c.alpha;
    d.accum = c.accum;
    d.C = c.c;
    d.ldc = Ncols;
    d.beta = c.beta;
    GemmArgs g;
    lof_status s = prepare_gemm(&d, &g);
    if (s != LOF_OK) return s;
    const int l2 = log2_chunk(g.chunk);
    if (l2 < 0)
        return set_error(LOF_ERR_UNSUPPORTED,
                         "%s: convolutions run on the wave tiles: chunk must be in "
                         "{1, 2, 4, 8, 16, 32} (got %d)", what, g.chunk);
    int tile;   // 0: 64 x 64, 1: 128 x 64
    const int cfg = config >= 0 ? (config & 0xFF) : -1;
    if (cfg < 0) {   // the GEMM's default rule: 128 x 64 for long chunks with >= 384 tiles
        const int64_t tiles_128x64 = ((int64_t)c.K + 127) / 128 * ((Ncols + 63) / 64);
        tile = (g.chunk >= 16 && tiles_128x64 >= 384) ? 1 : 0;
    } else if (cfg == kCfgMfma11) {
        tile = 0;
    } else if (cfg == kCfgMfma21) {
        tile = 1;
    } else {
        return set_error(LOF_ERR_UNSUPPORTED,
                         "%s: config %d: convolutions run on configs %d (mfma11) and %d (mfma21)",
                         what, cfg, kCfgMfma11, kCfgMfma21);
    }
    // testing bits as in lof_gemm_f32: 0x100 general leaf path, 0x200 every chunk on VALU,
    // 0x800 / 0x1000 / 0x2000 / 0x4000 / 0x8000 / 0x20000 the exact-sum / RTO /
    // binary16-held / packed binary16 / directed-rounding / mixed-precision FMA leaves off,
    // 0x10000 the bf16 MFMA chunk sums off
    if (config >= 0) {
        if (config & 0x800) g.sum_off |= 1;
        if (config & 0x1000) g.sum_off |= 2;
        if (config & 0x2000) g.sum_off |= 4;
        if (config & 0x4000) g.sum_off |= 8;
        if (config & 0x8000) g.sum_off |= 16;
        if (config & 0x10000) g.mfma_bf16 = 0;
        if (config & 0x20000) g.sum_off |= 32;
        if (config & 0x100) {
            set_fast_path(&g, false, 0);
            g.dr_kind = 0;
        }
        if (config & 0x200) g.mfma_mode = 1;
    }
    if (c.N == 0 || c.K == 0) return LOF_OK;

    g.tb = 0;
    g.vec = ((uintptr_t)c.w & 15u) == 0 && (Kred & 3) == 0;   // W's float4 loads
    ConvGeom& cv = g.cv;
    cv.conv = 1;
    cv.krsc = c.layout == LOF_CONV_NHWC ? 1 : 0;
    cv.H = c.H;
    cv.W = c.W;
    cv.P = (int32_t)P;
    cv.Q = (int32_t)Q;
    cv.ph = c.pad_h;
    cv.pw = c.pad_w;
    cv.sh = c.stride_h;
    cv.sw = c.stride_w;
    cv.dh = c.dil_h;
    cv.dw = c.dil_w;
    cv.PQ = (uint32_t)(P * Q);
    cv.D1 = cv.krsc ? (uint32_t)(c.S * c.C) : (uint32_t)(c.R * c.S);
    cv.D2 = cv.krsc ? c.C : c.S;
    cv.div_pq = make_fast_div(cv.PQ);
    cv.div_q = make_fast_div((uint32_t)Q);
    cv.div_d1 = make_fast_div(cv.D1);
    cv.div_d2 = make_fast_div((uint32_t)cv.D2);
    cv.sxb = (int64_t)c.C * HW;
    if (c.layout == LOF_CONV_NCHW) {
        cv.sxc = (int32_t)HW;
        cv.sxh = c.W;
        cv.sxw = 1;
        cv.sdb = (int64_t)c.K * P * Q;
        cv.sdm = P * Q;
        cv.sdpq = 1;
    } else {
        cv.sxc = 1;
        cv.sxh = c.W * c.C;
        cv.sxw = c.C;
        cv.sdb = (int64_t)P * Q * c.K;
        cv.sdm = 1;
        cv.sdpq = c.K;
    }
    const int bm = tile ? 128 : 64, bn = 64;
    const dim3 grid((unsigned)((Ncols + bn - 1) / bn), (unsigned)((c.K + bm - 1) / bm));
    if (grid.y > 65535u) return set_error(LOF_ERR_UNSUPPORTED, "%s: too many output channels", what);
    hipLaunchKernelGGL(kConvKernels[tile][l2][g.subn ? 1 : 0], grid, dim3(256), 0,
                       (hipStream_t)stream, g);
    return check_hip(hipGetLastError(), "lof_conv2d_f32 launch");
}

}  // extern "C"
