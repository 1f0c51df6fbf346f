// Screen-space denoiser records, updated per sample by the integrator's
// accumulate kernel and read by the bilateral filter (denoise.hip).
//
//   gbuffer_record   RecordScreenGBuffer + RecordShadingModel
//                    (core/denoise/denoise.cuh:82-163): running means of the
//                    first hit's albedo, normal, distance, metallic and
//                    roughness over the samples that hit, and of the
//                    shading-model one-hot over all samples
//   stats_record     RecordScreenStatisticBuffer (denoise_statistic.cuh:36-69):
//                    Welford mean, M2 / n and M3 / n of the Box-Cox (lambda
//                    1/2) transformed sample, M2 and M3 updated in double
//
// Records, floats per pixel (counts are uint32 bit patterns):
//   G-buffer (16): albedo[3] normal[3] depth metallic roughness
//                  model[5] (diffuse, light, glass, GGX, none) valid-frame count, frame count
//   statistics (12): n mean[3] M2/n[3] M3/n[3], 2 unused
// The reference keeps the G-buffer means in half precision; here they are
// float.  Box-Cox's pow(x, 0.5) is sqrtf (equal up to powf's last-ulp error,
// and bit-reproducible).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ar {

constexpr int kGBufferFloats = 16;
constexpr int kStatsFloats = 12;

namespace dn {

__device__ __forceinline__ void gbuffer_record(float* __restrict__ g, bool hit, const float alb[3], const float nrm[3],
                                               float depth, float metallic, float roughness, uint32_t model) {
    if (hit) {
        const uint32_t valid = __float_as_uint(g[14]);
        if (valid == 0) {
#pragma unroll
            for (int k = 0; k < 9; ++k) g[k] = 0.0f;
        }
        const float affix = 1.0f / (float)(valid + 1u);
        const float pre = (float)valid / (float)(valid + 1u);
#pragma unroll
        for (int k = 0; k < 9; ++k) g[k] = g[k] * pre;
#pragma unroll
        for (int k = 0; k < 3; ++k) g[k] = g[k] + affix * alb[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) g[3 + k] = g[3 + k] + affix * nrm[k];
        g[7] = g[7] + affix * metallic;
        g[8] = g[8] + affix * roughness;
        g[6] = g[6] + affix * depth;
        g[14] = __uint_as_float(valid + 1u);
    } else {
        model = 4u;
    }
    const uint32_t frame = __float_as_uint(g[15]);
    if (frame == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) g[9 + k] = 0.0f;
    }
    const float affix = 1.0f / (float)(frame + 1u);
    const float pre = (float)frame / (float)(frame + 1u);
#pragma unroll
    for (int k = 0; k < 5; ++k) g[9 + k] = g[9 + k] * pre;
    g[9 + model] = g[9 + model] + affix;
    g[15] = __uint_as_float(frame + 1u);
}

__device__ __forceinline__ void stats_record(float* __restrict__ st, const float x[3]) {
    const uint32_t n0 = __float_as_uint(st[0]);
    const uint32_t n1 = n0 + 1u;
    st[0] = __uint_as_float(n1);
    const float inv = 1.0f / (float)n1;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float xb = (sqrtf(fmaxf(x[c], 0.0f)) - 1.0f) * 2.0f;     // BoxCox05
        const float mean0 = st[1 + c];
        const float delta = xb - mean0;
        const float ddn = delta * inv;
        const float mean1 = mean0 + ddn;
        const double m2_0 = (double)st[4 + c] * (double)n0;
        const double m2_1 = m2_0 + (double)(delta * (delta - ddn));
        const double m3_0 = (double)st[7 + c] * (double)n0;
        const float d2 = delta * delta;
        const float ddn2 = ddn * ddn;
        const double m3_1 = m3_0 - 3.0 * ((double)ddn * m2_1) + (double)(delta * (d2 - ddn2));
        st[1 + c] = mean1;
        st[4 + c] = (float)(m2_1 / (double)n1);
        st[7 + c] = (float)(m3_1 / (double)n1);
    }
}

} // namespace dn
} // namespace ar
