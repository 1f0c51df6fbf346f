// Output stage: display tonemap, odd/even luminance MSE and the BMP writer.
//
//   ar_tonemap          Tonemap kernel (core/scene.cu:22-133) with Gamma and
//                       ACES (core/tonemap.cuh:6-114); XYZ input goes through
//                       XYZ2SRGBLinearD65 (core/tonemap.cuh:116-126) first
//   ar_luminance_mse    DeviceMSELuminance (core/helper/measure.cuh:49-74), a
//                       two-stage reduction in double with a fixed order
//   ar_present          the output stage of SceneRenderer::Render (core/scene.cu:
//                       291-365): tonemap, denoise + tonemap, or a debug view
//   ar_render_host      the offline SceneRenderer::Render (core/scene.cu:215-237)
//   ar_write_bmp        SaveBMP (core/camera.cu:10-62): 32-bit BGRA, top-down
//
// One thread per pixel for the tonemap (HBM-bound: 12 B in, 4 B out).
#include "ar_internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace ar {
namespace {

constexpr uint32_t kOutBlock = 256;
constexpr uint32_t kReduceBlocks = 512;

__device__ __forceinline__ float sat(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

__device__ __forceinline__ void m3(const float* m, const float v[3], float o[3]) {
    for (int r = 0; r < 3; ++r) o[r] = m[3 * r] * v[0] + m[3 * r + 1] * v[1] + m[3 * r + 2] * v[2];
}

__device__ __forceinline__ float lerpf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

__constant__ float kSrgbAp0[9] = {0.4397010f, 0.3829780f, 0.1773350f, 0.0897923f, 0.8134230f,
                                  0.0967616f, 0.0175440f, 0.1115440f, 0.8707040f};
__constant__ float kAp0Ap1[9] = {1.4514393161f, -0.2365107469f, -0.2149285693f, -0.0765537734f, 1.1762296998f,
                                 -0.0996759264f, 0.0083161484f, -0.0060324498f, 0.9977163014f};
__constant__ float kAp1Xyz[9] = {0.6624541811f, 0.1340042065f, 0.1561876870f, 0.2722287168f, 0.6740817658f,
                                 0.0536895174f, -0.0055746495f, 0.0040607335f, 1.0103391003f};
__constant__ float kXyzAp1[9] = {1.6410233797f, -0.3248032942f, -0.2364246952f, -0.6636628587f, 1.6153315917f,
                                 0.0167563477f, 0.0117218943f, -0.0082844420f, 0.9883948585f};
__constant__ float kD60D65[9] = {0.98722400f, -0.00611327f, 0.0159533f, -0.00759836f, 1.00186000f,
                                 0.0053302f, 0.00307257f, -0.00509595f, 1.0816800f};
__constant__ float kXyz709[9] = {3.2409699419f, -1.5373831776f, -0.4986107603f, -0.9692436363f, 1.8759675015f,
                                 0.0415550574f, 0.0556300797f, -0.2039769589f, 1.0569715142f};

__device__ __forceinline__ void gamma3(float v[3]) {
    for (int k = 0; k < 3; ++k) v[k] = powf(v[k], 1.0f / 2.2f);
}

// ACES (core/tonemap.cuh:66-114)
__device__ void aces(const float in[3], float out[3]) {
    const float y2[3] = {0.272229f, 0.674082f, 0.0536895f};
    float a0[3], cg[3], post[3], xyz[3], lin[3], t[3];
    m3(kSrgbAp0, in, a0);
    m3(kAp0Ap1, a0, cg);
    float tmp = cg[0] * y2[0] + cg[1] * y2[1] + cg[2] * y2[2];
    for (int k = 0; k < 3; ++k) cg[k] = lerpf(tmp, cg[k], 0.96f);
    const float a = 278.5085f, b = 10.7772f, c = 293.6045f, d = 88.7122f, e = 80.6889f;
    for (int k = 0; k < 3; ++k) {
        const float x = cg[k];
        post[k] = (x * (x * a + b)) / (x * (x * c + d) + e);
    }
    m3(kAp1Xyz, post, xyz);
    const float div = fmaxf(xyz[0] + xyz[1] + xyz[2], 1.0e-4f);
    float xx = xyz[0] / div, yy = xyz[1] / div, Y = xyz[1];
    Y = fminf(fmaxf(Y, 0.0f), 65504.0f);
    Y = powf(Y, 0.9811f);
    const float mm = Y / fmaxf(yy, 1e-4f);
    t[0] = xx * mm;
    t[1] = Y;
    t[2] = (1.0f - xx - yy) * mm;
    m3(kXyzAp1, t, lin);
    tmp = lin[0] * y2[0] + lin[1] * y2[1] + lin[2] * y2[2];
    for (int k = 0; k < 3; ++k) lin[k] = lerpf(tmp, lin[k], 0.93f);
    m3(kAp1Xyz, lin, xyz);
    m3(kD60D65, xyz, t);
    m3(kXyz709, t, out);
    for (int k = 0; k < 3; ++k) out[k] = fmaxf(0.0f, out[k]);
    gamma3(out);
}

__global__ __launch_bounds__(kOutBlock) void k_tonemap(const float* __restrict__ hdr, uint32_t* __restrict__ ldr,
                                                       size_t npix, float exposure, uint32_t tone, uint32_t xyz) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    float v[3] = {hdr[3 * i], hdr[3 * i + 1], hdr[3 * i + 2]};
    if (xyz) {
        const float r = 3.2406f * v[0] + -1.5372f * v[1] + -0.4986f * v[2];
        const float g = -0.9689f * v[0] + 1.8758f * v[1] + 0.0415f * v[2];
        const float b = 0.0557f * v[0] + -0.204f * v[1] + 1.057f * v[2];
        v[0] = fmaxf(0.0f, r); v[1] = fmaxf(0.0f, g); v[2] = fmaxf(0.0f, b);
    }
    for (int k = 0; k < 3; ++k) v[k] = v[k] * exposure;
    if (tone == AR_TONE_GAMMA) {
        gamma3(v);
    } else if (tone == AR_TONE_ACES) {
        float o[3];
        aces(v, o);
        v[0] = o[0]; v[1] = o[1]; v[2] = o[2];
    }
    const uint32_t r = (uint32_t)(255.0f * sat(v[0]));
    const uint32_t g = (uint32_t)(255.0f * sat(v[1]));
    const uint32_t b = (uint32_t)(255.0f * sat(v[2]));
    ldr[i] = 0xff000000u | (r << 16) | (g << 8) | b;
}

// The Tonemap kernel's debug views (core/scene.cu:35-107) over the
// denoiser's per-pixel records (G-buffer: albedo[3] normal[3] depth metallic
// roughness ...; statistics: n mean[3] M2/n[3] M3/n[3]).
// HeatMapViridis (core/render/color.cuh:99-140): matplotlib's viridis map at
// t = 0, 1/4, 1/2, 3/4, 1, linear in between.  At t = 1 the reference's
// segment search lands on the last segment with weight 0 and so returns the
// 3/4 sample; kept so the views match.
__constant__ float kViridis[15] = {0.267004f, 0.004874f, 0.329415f, 0.229739f, 0.322361f, 0.545706f,
                                   0.127568f, 0.566949f, 0.550556f, 0.369214f, 0.788888f, 0.382914f,
                                   0.993248f, 0.906157f, 0.143936f};

__device__ __forceinline__ void viridis(float t, float o[3]) {
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const float a = t * 4.0f;
    const int i = (int)floorf(a);
    const float f = a - (float)i;
    const int s = i <= 0 ? 0 : (i >= 3 ? 3 : i);
    for (int k = 0; k < 3; ++k) o[k] = kViridis[3 * s + k] + (kViridis[3 * s + 3 + k] - kViridis[3 * s + k]) * f;
}

__global__ __launch_bounds__(kOutBlock) void k_view(const float* __restrict__ gbuf, const float* __restrict__ stats,
                                                    uint32_t* __restrict__ ldr, size_t npix, uint32_t view) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const float* g = gbuf ? gbuf + (size_t)AR_GBUFFER_FLOATS * i : nullptr;
    const float* st = stats ? stats + (size_t)AR_STATS_FLOATS * i : nullptr;
    float v[3] = {0.0f, 0.0f, 0.0f};
    switch (view) {
    case AR_VIEW_ALBEDO: v[0] = g[0]; v[1] = g[1]; v[2] = g[2]; break;
    case AR_VIEW_NORMAL:
        for (int k = 0; k < 3; ++k) v[k] = g[3 + k] * 0.5f + 0.5f;
        break;
    case AR_VIEW_DEPTH: viridis(1.0f / (g[6] + 1.0f), v); break;
    case AR_VIEW_METALLIC: viridis(g[7], v); break;
    case AR_VIEW_ROUGHNESS: viridis(g[8], v); break;
    case AR_VIEW_STAT_COUNT: v[0] = (float)log((double)__float_as_uint(st[0]) + 1.0) * 0.01f; break;
    default: {   // AR_VIEW_STAT_MEAN / _M2 / _M3: |moment| / 4
        const int at = 1 + 3 * (int)(view - AR_VIEW_STAT_MEAN);
        for (int k = 0; k < 3; ++k) v[k] = fabsf(st[at + k]) * 0.25f;
    }
    }
    const uint32_t r = (uint32_t)(255.0f * sat(v[0]));
    const uint32_t gg = (uint32_t)(255.0f * sat(v[1]));
    const uint32_t b = (uint32_t)(255.0f * sat(v[2]));
    ldr[i] = 0xff000000u | (r << 16) | (gg << 8) | b;
}

// first stage: each block sums a grid-stride share in a fixed order
__global__ __launch_bounds__(kOutBlock) void k_lum_mse_partial(const float* __restrict__ a, const float* __restrict__ b,
                                                               size_t npix, double* __restrict__ part) {
    __shared__ double red[kOutBlock];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x) {
        const float la = 0.2126f * a[3 * i] + 0.7152f * a[3 * i + 1] + 0.0722f * a[3 * i + 2];
        const float lb = 0.2126f * b[3 * i] + 0.7152f * b[3 * i + 1] + 0.0722f * b[3 * i + 2];
        const double d = (double)la - (double)lb;
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = kOutBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kOutBlock) void k_lum_mse_final(const double* __restrict__ part, uint32_t nparts,
                                                             size_t npix, double* __restrict__ out) {
    __shared__ double red[kOutBlock];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = kOutBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / (double)npix;
}

} // namespace
} // namespace ar

using namespace ar;

extern "C" {

int ar_tonemap(ar_ctx* ctx, const float* hdr, uint32_t* ldr, uint32_t width, uint32_t height, float exposure,
               uint32_t tone, uint32_t xyz_input, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (!hdr || !ldr) return set_error(ctx, AR_E_INVALID, "NULL image");
    if (tone > AR_TONE_ACES) return set_error(ctx, AR_E_INVALID, "unknown tone curve");
    const size_t npix = (size_t)width * height;
    if (npix == 0) return AR_OK;
    (void)hipSetDevice(ctx->device);
    hipLaunchKernelGGL(k_tonemap, dim3((uint32_t)((npix + kOutBlock - 1) / kOutBlock)), dim3(kOutBlock), 0,
                       (hipStream_t)stream, hdr, ldr, npix, exposure, tone, xyz_input ? 1u : 0u);
    return check_hip(ctx, hipGetLastError(), "tonemap launch");
}

int ar_present(ar_ctx* ctx, const float* hdr, const float* gbuffer, const float* stats, uint32_t* ldr, uint32_t width,
               uint32_t height, float exposure, uint32_t tone, uint32_t xyz_input, uint32_t view, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (!ldr) return set_error(ctx, AR_E_INVALID, "NULL ldr");
    if (tone > AR_TONE_ACES) return set_error(ctx, AR_E_INVALID, "unknown tone curve");
    if (view > AR_VIEW_STAT_M3) return set_error(ctx, AR_E_INVALID, "unknown view");
    if (view <= AR_VIEW_DENOISED && !hdr) return set_error(ctx, AR_E_INVALID, "NULL hdr");
    if ((view == AR_VIEW_DENOISED || (view >= AR_VIEW_ALBEDO && view <= AR_VIEW_ROUGHNESS)) && !gbuffer)
        return set_error(ctx, AR_E_INVALID, "this view needs the G-buffer");
    if ((view == AR_VIEW_DENOISED || view >= AR_VIEW_STAT_COUNT) && !stats)
        return set_error(ctx, AR_E_INVALID, "this view needs the sample statistics");
    const size_t npix = (size_t)width * height;
    if (npix == 0) return AR_OK;
    (void)hipSetDevice(ctx->device);
    const dim3 grid((uint32_t)((npix + kOutBlock - 1) / kOutBlock));
    if (view >= AR_VIEW_ALBEDO) {
        hipLaunchKernelGGL(k_view, grid, dim3(kOutBlock), 0, (hipStream_t)stream, gbuffer, stats, ldr, npix, view);
        return check_hip(ctx, hipGetLastError(), "view launch");
    }
    const float* src = hdr;
    if (view == AR_VIEW_DENOISED) {
        if (ctx->present_pixels < npix) {
            invalidate_graphs(ctx);
            if (ctx->present_buf) (void)hipFree(ctx->present_buf);
            ctx->present_buf = nullptr;
            ctx->present_pixels = 0;
            const int rc = check_hip(ctx, hipMalloc((void**)&ctx->present_buf, npix * 3 * sizeof(float)),
                                     "hipMalloc(denoised image)");
            if (rc != AR_OK) return rc;
            ctx->present_pixels = npix;
        }
        const int rc = ar_denoise(ctx, hdr, gbuffer, stats, ctx->present_buf, width, height, stream);
        if (rc != AR_OK) return rc;
        src = ctx->present_buf;
    }
    hipLaunchKernelGGL(k_tonemap, grid, dim3(kOutBlock), 0, (hipStream_t)stream, src, ldr, npix, exposure, tone,
                       xyz_input ? 1u : 0u);
    return check_hip(ctx, hipGetLastError(), "tonemap launch");
}

int ar_render_host(ar_ctx* ctx, const ar_camera* cam, const ar_render_opts* opts, float* hdr, uint32_t* ldr,
                   float exposure, uint32_t tone) {
    if (!ctx) return AR_E_INVALID;
    if (!cam || !opts || (!hdr && !ldr)) return set_error(ctx, AR_E_INVALID, "NULL camera, options or outputs");
    if (tone > AR_TONE_ACES) return set_error(ctx, AR_E_INVALID, "unknown tone curve");
    if (opts->width == 0 || opts->height == 0) return set_error(ctx, AR_E_INVALID, "empty image");
    const size_t npix = (size_t)opts->width * opts->height;
    (void)hipSetDevice(ctx->device);
    if (ctx->host_out_pixels < npix) {
        invalidate_graphs(ctx);
        if (ctx->host_out) (void)hipFree(ctx->host_out);
        ctx->host_out = nullptr;
        ctx->host_out_pixels = 0;
        const int rc = check_hip(ctx, hipMalloc((void**)&ctx->host_out, npix * 4 * sizeof(float)), "hipMalloc(render target)");
        if (rc != AR_OK) return rc;
        ctx->host_out_pixels = npix;
    }
    float* d_hdr = ctx->host_out;
    uint32_t* d_ldr = reinterpret_cast<uint32_t*>(ctx->host_out + 3 * npix);
    int rc = ar_render(ctx, cam, opts, d_hdr, ctx->stream);
    if (rc == AR_OK && ldr)
        rc = ar_tonemap(ctx, d_hdr, d_ldr, opts->width, opts->height, exposure, tone,
                        opts->spectral != AR_SPECTRAL_RGB && opts->output == AR_OUTPUT_XYZ, ctx->stream);
    if (rc == AR_OK) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "render");
    if (rc == AR_OK && hdr)
        rc = check_hip(ctx, hipMemcpy(hdr, d_hdr, npix * 3 * sizeof(float), hipMemcpyDeviceToHost), "copy render");
    if (rc == AR_OK && ldr)
        rc = check_hip(ctx, hipMemcpy(ldr, d_ldr, npix * sizeof(uint32_t), hipMemcpyDeviceToHost), "copy image");
    return rc;
}

int ar_luminance_mse(ar_ctx* ctx, const float* a, const float* b, size_t npix, double* out, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (!a || !b || !out) return set_error(ctx, AR_E_INVALID, "NULL argument");
    if (npix == 0) { *out = 0.0; return AR_OK; }
    (void)hipSetDevice(ctx->device);
    if (!ctx->reduce) {
        int rc = check_hip(ctx, hipMalloc((void**)&ctx->reduce, (kReduceBlocks + 1) * sizeof(double)), "hipMalloc(reduce)");
        if (rc != AR_OK) { ctx->reduce = nullptr; return rc; }
    }
    const hipStream_t s = (hipStream_t)stream;
    const uint32_t blocks = (uint32_t)std::min<size_t>(kReduceBlocks, (npix + kOutBlock - 1) / kOutBlock);
    hipLaunchKernelGGL(k_lum_mse_partial, dim3(blocks), dim3(kOutBlock), 0, s, a, b, npix, ctx->reduce);
    hipLaunchKernelGGL(k_lum_mse_final, dim3(1), dim3(kOutBlock), 0, s, ctx->reduce, blocks, npix,
                       ctx->reduce + kReduceBlocks);
    int rc = check_hip(ctx, hipGetLastError(), "luminance MSE launch");
    if (rc != AR_OK) return rc;
    rc = check_hip(ctx, hipMemcpyAsync(out, ctx->reduce + kReduceBlocks, sizeof(double), hipMemcpyDeviceToHost, s),
                   "luminance MSE copy");
    if (rc != AR_OK) return rc;
    return check_hip(ctx, hipStreamSynchronize(s), "luminance MSE sync");
}

int ar_write_bmp(const char* path, const uint32_t* bgra, uint32_t width, uint32_t height) {
    if (!path || !bgra || width == 0 || height == 0 || width > 0x7FFFFFFFu || height > 0x7FFFFFFFu)
        return AR_E_INVALID;
    const uint64_t data = (uint64_t)width * height * 4u;
    if (data + 54u > 0xFFFFFFFFull) return AR_E_CAPACITY;
    uint8_t h[54] = {0};
    auto put32 = [&](int at, uint32_t v) { for (int k = 0; k < 4; ++k) h[at + k] = (uint8_t)(v >> (8 * k)); };
    h[0] = 'B'; h[1] = 'M';
    put32(2, (uint32_t)(data + 54u));   // file size
    put32(10, 54u);                     // pixel data offset
    put32(14, 40u);                     // BITMAPINFOHEADER size
    put32(18, width);
    put32(22, (uint32_t)(-(int32_t)height));   // negative height: rows top-down
    h[26] = 1;                          // planes
    h[28] = 32;                         // bits per pixel (BI_RGB, no compression)
    put32(34, (uint32_t)data);          // image size
    FILE* f = std::fopen(path, "wb");
    if (!f) return AR_E_INVALID;
    bool ok = std::fwrite(h, 1, sizeof h, f) == sizeof h;
    // 0xAARRGGBB words are B, G, R, A bytes in memory: the BMP's 32-bit order
    ok = ok && std::fwrite(bgra, 4, (size_t)width * height, f) == (size_t)width * height;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? AR_OK : AR_E_INVALID;
}

} // extern "C"
