// Uplift-table queries of the spectral basis network, shared by the query
// kernel (smlp.hip) and the integrator's material bases:
//   query_rgb      QueryRGBBasis (core/spectrum/spectrum_lut.cuh:124-181):
//                  trilinear over the [res^3][4] table (x fastest), coordinate
//                  c res - 0.5 clamped to [0, res - 1]
//   query_lambda   QueryLambda (:183-221): linear over the [res][4] table
//   lut_interact   Interact (:223-235): an RGB above 1 is divided by its
//                  largest channel and the clamped dot product scaled back
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace ar {
namespace sb {

// __linear_coord (spectrum_lut.cuh:93-107)
__device__ __forceinline__ void lin_coord(float gc, int res, int& i0, int& i1, float& t) {
    const float f = floorf(gc);
    i0 = (int)f;
    i1 = i0 + 1 >= res ? res - 1 : i0 + 1;
    t = i0 == i1 ? 0.0f : gc - f;
}

__device__ __forceinline__ float4 query_rgb(const float* __restrict__ lut, int res, float r, float g, float b) {
    const float mx = (float)res - 1.0f;
    const float gx = fminf(fmaxf(r * (float)res - 0.5f, 0.0f), mx);
    const float gy = fminf(fmaxf(g * (float)res - 0.5f, 0.0f), mx);
    const float gz = fminf(fmaxf(b * (float)res - 0.5f, 0.0f), mx);
    int x0, x1, y0, y1, z0, z1;
    float tx, ty, tz;
    lin_coord(gx, res, x0, x1, tx);
    lin_coord(gy, res, y0, y1, ty);
    lin_coord(gz, res, z0, z1, tz);
    const float wx0 = 1.0f - tx, wx1 = tx, wy0 = 1.0f - ty, wy1 = ty, wz0 = 1.0f - tz, wz1 = tz;
    const float w[8] = {wx0 * wy0 * wz0, wx1 * wy0 * wz0, wx0 * wy1 * wz0, wx1 * wy1 * wz0,
                        wx0 * wy0 * wz1, wx1 * wy0 * wz1, wx0 * wy1 * wz1, wx1 * wy1 * wz1};
    const size_t sy = (size_t)res, sz = (size_t)res * res;
    const size_t base[8] = {x0 + y0 * sy + z0 * sz, x1 + y0 * sy + z0 * sz, x0 + y1 * sy + z0 * sz,
                            x1 + y1 * sy + z0 * sz, x0 + y0 * sy + z1 * sz, x1 + y0 * sy + z1 * sz,
                            x0 + y1 * sy + z1 * sz, x1 + y1 * sy + z1 * sz};
    float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float4 v = *reinterpret_cast<const float4*>(lut + 4 * base[c]);
        o.x = fmaf(w[c], v.x, o.x);
        o.y = fmaf(w[c], v.y, o.y);
        o.z = fmaf(w[c], v.z, o.z);
        o.w = fmaf(w[c], v.w, o.w);
    }
    return o;
}

__device__ __forceinline__ float4 query_lambda(const float* __restrict__ lut, int res, float lam) {
    float u = (lam - 360.0f) / (780.0f - 360.0f);
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    float gc = u * (float)res - 0.5f;
    gc = fminf(fmaxf(gc, 0.0f), (float)res - 1.0f);
    int i0, i1;
    float t;
    lin_coord(gc, res, i0, i1, t);
    const float w0 = 1.0f - t, w1 = t;
    const float4 a = *reinterpret_cast<const float4*>(lut + 4 * (size_t)i0);
    const float4 b = *reinterpret_cast<const float4*>(lut + 4 * (size_t)i1);
    return make_float4(a.x * w0 + b.x * w1, a.y * w0 + b.y * w1, a.z * w0 + b.z * w1, a.w * w0 + b.w * w1);
}

// SpectralInteract (spectrum_basis.cuh:2462-2470): clamp of an FMA chain
__device__ __forceinline__ float dot_clamp(const float4& a, const float4& b) {
    float acc = 0.0f;
    acc = fmaf(a.x, b.x, acc);
    acc = fmaf(a.y, b.y, acc);
    acc = fmaf(a.z, b.z, acc);
    acc = fmaf(a.w, b.w, acc);
    return fminf(fmaxf(acc, 0.0f), 1.0f);
}

__device__ __forceinline__ float lut_interact(const float* __restrict__ rgb_lut, int rgb_res, const float4& kl, float r,
                                              float g, float b) {
    const float mx = fmaxf(fmaxf(r, g), b);
    if (mx > 1.0f) return dot_clamp(query_rgb(rgb_lut, rgb_res, r / mx, g / mx, b / mx), kl) * mx;
    return dot_clamp(query_rgb(rgb_lut, rgb_res, r, g, b), kl);
}

}  // namespace sb
}  // namespace ar
