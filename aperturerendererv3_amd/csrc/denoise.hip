// Screen-space denoiser: Bilateral16x16<7> of the reference
// (core/denoise/denoise.cuh:352-444) with its G-buffer weight
// (ComputeSVGFWeight, :186-241) and Welch t-test gate (ComputeTStatisticFinal,
// :266-350), on the records the integrator keeps (denoise.hpp).
//
// Two kernels:
//   k_denoise_features  one thread per pixel: everything a tap needs that
//                       depends on one pixel only (log(1 + c), the normalised
//                       normal, the skewness-corrected Box-Cox mean theta, its
//                       variance s^2 / n, and Welch's per-side terms
//                       s^2 / n and (s^2 / n)^2 / (n - 1)), 29 floats, SoA.
//                       These are the same roundings the reference performs
//                       inside every tap, done once.
//   k_bilateral         16 x 16 pixels per 1024-thread block: the 30 x 30
//                       feature tile (7-pixel apron, clamped at the border)
//                       and the t quantile table are staged in LDS (106 KB,
//                       so one block per CU); four threads per pixel walk
//                       bands of 4, 4, 4 and 3 rows of its 15 x 15 window
//                       and their sums are added in row order, which gives
//                       each SIMD four waves to hide the latency of the LDS
//                       reads and of the IEEE divisions and square roots of
//                       the t statistic.
#include "ar_internal.hpp"
#include "denoise.hpp"

#include <cmath>

namespace ar {
namespace {

constexpr int kRadius = 7;
constexpr int kTile = 16;
constexpr int kSpan = kTile + 2 * kRadius;     // 30
constexpr int kSpanPix = kSpan * kSpan;        // 900
constexpr int kFeat = 29;
constexpr int kQuantiles = 1024;
#ifndef AR_DENOISE_SPLIT
#define AR_DENOISE_SPLIT 4
#endif
constexpr int kSplit = AR_DENOISE_SPLIT;        // threads per pixel, each taking a band of window rows
constexpr int kThreads = kSplit * kTile * kTile;
static_assert(kThreads <= 1024, "block too large");

// feature planes (each npix floats)
enum : int {
    F_LOG = 0,       // 3: log(c + 1)
    F_DEPTH = 3,
    F_ALB = 4,       // 3
    F_MET = 7,
    F_NRM = 8,       // 3, normalised
    F_ROUGH = 11,
    F_HIST = 12,     // 5
    F_TH = 17,       // 3: mean + M3 / n / 6 / s^2
    F_VT = 20,       // 3: s^2 / n (variance of theta)
    F_U = 23,        // 3: s^2 / (float)n
    F_L = 26,        // 3: u^2 / max(n - 1, 1)
};

__device__ const float kTQuantile[kQuantiles] = {
#include "t_quantile_table.inc"
};

// LookupTQuantile005(floorToInt(nu)) index: nu - 1 clamped to the table,
// NaN and nu < 1 giving 0 (floorToInt of NaN is 0 on the GPU), branch-free
__device__ __forceinline__ uint32_t quantile_index(float nu) {
    float f = floorf(nu);
    f = nu != nu ? 0.0f : f;
    f = fminf(fmaxf(f, 0.0f), (float)kQuantiles);
    const int k = (int)f - 1;
    return (uint32_t)(k < 0 ? 0 : k);
}

__global__ __launch_bounds__(256) void k_denoise_features(const float* __restrict__ hdr, const float* __restrict__ gbuf,
                                                          const float* __restrict__ stats, float* __restrict__ feat,
                                                          uint32_t npix) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    auto F = [&](int k) -> float& { return feat[(size_t)k * npix + i]; };
    const float* g = gbuf + (size_t)kGBufferFloats * i;
    const float* st = stats + (size_t)kStatsFloats * i;
    for (int c = 0; c < 3; ++c) F(F_LOG + c) = logf(hdr[3 * (size_t)i + c] + 1.0f);      // SeparateSample
    F(F_DEPTH) = g[6];
    for (int c = 0; c < 3; ++c) F(F_ALB + c) = g[c];
    F(F_MET) = g[7];
    float nx = g[3], ny = g[4], nz = g[5];
    const float l2 = nx * nx + ny * ny + nz * nz;
    if (l2 <= 1e-20f) {
        nx = ny = nz = 0.0f;
    } else {
        const float inv = 1.0f / sqrtf(l2);
        nx = nx * inv; ny = ny * inv; nz = nz * inv;
    }
    F(F_NRM) = nx; F(F_NRM + 1) = ny; F(F_NRM + 2) = nz;
    F(F_ROUGH) = g[8];
    for (int k = 0; k < 5; ++k) F(F_HIST + k) = g[9 + k];
    // ComputeTStatistic
    const uint32_t n = __float_as_uint(st[0]);
    const double num = (double)n;
    const uint32_t nm1 = n - 1u > 1u ? n - 1u : 1u;
    const double inv = 1.0 / (double)nm1;
    for (int c = 0; c < 3; ++c) {
        const float s2 = (float)(((double)st[4 + c] * num) * inv);
        F(F_VT + c) = (float)((double)s2 / num);
        F(F_TH + c) = st[1 + c] + st[7 + c] / 6.0f / s2;
        const float u = s2 / (float)n;                       // ComputeWelchNu's per-side terms
        F(F_U + c) = u;
        F(F_L + c) = u * u / (float)nm1;
    }
}

__global__ __launch_bounds__(kThreads) void k_bilateral(const float* __restrict__ feat, float* __restrict__ out,
                                                   int width, int height) {
    __shared__ float tile[kFeat][kSpanPix];
    __shared__ float tq[kQuantiles];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = ((int)threadIdx.z * kTile + ty) * kTile + tx;
    const int gx0 = blockIdx.x * kTile - kRadius, gy0 = blockIdx.y * kTile - kRadius;
    const size_t npix = (size_t)width * height;
    for (int k = tid; k < kQuantiles; k += kThreads) tq[k] = kTQuantile[k];
    for (int s = tid; s < kSpanPix; s += kThreads) {
        const int sy = s / kSpan, sx = s - sy * kSpan;
        const int gy = min(max(gy0 + sy, 0), height - 1);
        const int gx = min(max(gx0 + sx, 0), width - 1);
        const size_t gi = (size_t)gy * width + gx;
#pragma unroll
        for (int k = 0; k < kFeat; ++k) tile[k][s] = feat[(size_t)k * npix + gi];
    }
    __syncthreads();
    const int ox = blockIdx.x * kTile + tx, oy = blockIdx.y * kTile + ty;
    const bool inside = ox < width && oy < height;   // outside pixels still reach both barriers
    const int cs = (ty + kRadius) * kSpan + tx + kRadius;
    float c_alb[3], c_nrm[3], c_hist[5], c_th[3], c_vt[3], c_u[3], c_l[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        c_alb[c] = tile[F_ALB + c][cs]; c_nrm[c] = tile[F_NRM + c][cs];
        c_th[c] = tile[F_TH + c][cs]; c_vt[c] = tile[F_VT + c][cs];
        c_u[c] = tile[F_U + c][cs]; c_l[c] = tile[F_L + c][cs];
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) c_hist[k] = tile[F_HIST + k][cs];
    const float c_depth = tile[F_DEPTH][cs], c_met = tile[F_MET][cs], c_rough = tile[F_ROUGH][cs];
    float acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] = tile[F_LOG + c][cs] * 1.0e-6f;     // CENTER_PIXEL_EPS
    float aw = 1.0e-6f;
    constexpr float kInvModel = 1.0f / (2.0f * 0.1f * 0.1f);                 // FilterParams, in float
    // the window's rows are split between the two threads of the pixel
    // (threadIdx.z); the halves are added at the end
    const int band = (2 * kRadius + 1 + kSplit - 1) / kSplit;
    const int j0 = -kRadius + (int)threadIdx.z * band;
    const int j1 = min(j0 + band - 1, kRadius);
    if (threadIdx.z != 0) {
        acc[0] = acc[1] = acc[2] = 0.0f;
        aw = 0.0f;
    }
    for (int j = j0; j <= j1; ++j) {
#pragma unroll 3
        for (int i = -kRadius; i <= kRadius; ++i) {
            const int s = cs + j * kSpan + i;
            // ComputeSVGFWeight
            float w = 0.0f;
            {
                const float d0 = c_alb[0] - tile[F_ALB][s], d1 = c_alb[1] - tile[F_ALB + 1][s],
                            d2 = c_alb[2] - tile[F_ALB + 2][s];
                w = w + (d0 * d0 + d1 * d1 + d2 * d2) * 0.5f;
            }
            {
                const float d = 1.0f - (c_nrm[0] * tile[F_NRM][s] + c_nrm[1] * tile[F_NRM + 1][s] +
                                        c_nrm[2] * tile[F_NRM + 2][s]);
                w = w + d * d * 2.0f;
            }
            {
                const float bd = tile[F_DEPTH][s];
                const float den = fmaxf(fmaxf(c_depth, bd), 1e-6f);
                const float d = (c_depth - bd) / den;
                w = w + d * d * 8.0f;
            }
            {
                const float dm = c_met - tile[F_MET][s];
                w = w + dm * dm * 2.0f;
            }
            {
                const float dr = c_rough - tile[F_ROUGH][s];
                w = w + dr * dr * 2.0f;
            }
            {
                float l1 = 0.0f;
#pragma unroll
                for (int k = 0; k < 5; ++k) l1 = l1 + fabsf(c_hist[k] - tile[F_HIST + k][s]);
                const float tv = 0.5f * l1;
                w = w + tv * tv * kInvModel;
            }
            const float svgf = expf(-w);
            // ComputeTStatisticFinal
            bool pass = true;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float t = c_th[c] - tile[F_TH + c][s];
                const float t2 = t * t;
                const float vb = tile[F_VT + c][s];
                const float up = t2 + vb;
                const float lo = t2 + c_vt[c] + vb;
                const float asym = up / fmaxf(lo, 1.0e-6f);
                const float T = sqrtf(0.5f / (1.0f - asym) - 1.0f);
                const float uu = c_u[c] + tile[F_U + c][s];
                const float nu = uu * uu / (c_l[c] + tile[F_L + c][s]);
                pass = pass & (T < tq[quantile_index(nu)]);
            }
            const float wt = svgf * expf((float)(-(i * i + j * j)) * 0.125f) * (pass ? 1.0f : 0.0f);
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c] = acc[c] + tile[F_LOG + c][s] * wt;
            aw = aw + wt;
        }
    }
    __syncthreads();                       // the tile is no longer read; its first planes take the bands
    float* part = &tile[0][0];
    const int slot = ty * kTile + tx;
    if (threadIdx.z != 0) {
        float* q = part + 4 * ((threadIdx.z - 1) * kTile * kTile + slot);
        q[0] = acc[0]; q[1] = acc[1]; q[2] = acc[2]; q[3] = aw;
    }
    __syncthreads();
    if (threadIdx.z != 0 || !inside) return;
    for (int z = 1; z < kSplit; ++z) {      // bands added in row order
        const float* q = part + 4 * ((z - 1) * kTile * kTile + slot);
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] = acc[c] + q[c];
        aw = aw + q[3];
    }
    const size_t o = (size_t)oy * width + ox;
#pragma unroll
    for (int c = 0; c < 3; ++c) out[3 * o + c] = fmaxf(expf(acc[c] / aw) - 1.0f, 0.0f);   // CompositeSample
}

} // namespace
} // namespace ar

using namespace ar;

extern "C" int ar_denoise(ar_ctx* ctx, const float* hdr, const float* gbuffer, const float* stats, float* out,
                          uint32_t width, uint32_t height, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (!hdr || !gbuffer || !stats || !out) return set_error(ctx, AR_E_INVALID, "NULL buffer");
    if (width == 0 || height == 0 || width > 32768 || height > 32768)
        return set_error(ctx, AR_E_INVALID, "image size must be 1..32768 per side");
    if (out == hdr) return set_error(ctx, AR_E_INVALID, "out must not alias hdr");
    (void)hipSetDevice(ctx->device);
    const size_t npix = (size_t)width * height;
    if (ctx->denoise_floats < npix * kFeat) {
        if (ctx->denoise_feat) (void)hipFree(ctx->denoise_feat);
        ctx->denoise_feat = nullptr;
        ctx->denoise_floats = 0;
        const int rc = check_hip(ctx, hipMalloc((void**)&ctx->denoise_feat, npix * kFeat * sizeof(float)),
                                 "hipMalloc(denoise features)");
        if (rc != AR_OK) return rc;
        ctx->denoise_floats = npix * kFeat;
    }
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_denoise_features, dim3((uint32_t)((npix + 255) / 256)), dim3(256), 0, s, hdr, gbuffer, stats,
                       ctx->denoise_feat, (uint32_t)npix);
    hipLaunchKernelGGL(k_bilateral, dim3((width + kTile - 1) / kTile, (height + kTile - 1) / kTile),
                       dim3(kTile, kTile, kSplit), 0, s, ctx->denoise_feat, out, (int)width, (int)height);
    return check_hip(ctx, hipGetLastError(), "denoise launch");
}
