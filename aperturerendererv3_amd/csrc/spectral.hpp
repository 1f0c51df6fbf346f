// Spectral support shared by host and device code.
//
//  * wavelength sampling: the reference's one-sample mixture (core/spectrum/
//    sample_ciexyz.cuh:288-326): 0.1 uniform on [360, 780] nm, 1 for a
//    two-Gaussian fit of xbar and 1 each for log-normal fits of ybar and zbar
//    (sample_ciexyz.cuh:159-285), inverted with Acklam's rational
//    approximation of the normal quantile plus one Newton step on erfc
//    (sample_ciexyz.cuh:28-156);
//  * CIE 1931 2-degree xbar/ybar/zbar and D65 at 1 nm with linear
//    interpolation (core/spectrum/ciexyz.cuh:67-85, 522-543); D65 normalised so
//    that integral D65 * ybar = 1 (S = 10567.0649781272);
//  * reflectance uplift: Jakob & Hanika 2019 sigmoid-polynomial spectra
//    s(x) = 1/2 + q / (2 sqrt(1 + q^2)), q = c0 x^2 + c1 x + c2 with
//    x = (lambda - 360) / 420, coefficients from a 3 x 64^3 table indexed by
//    the largest RGB channel (built in spectrum_lut.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace ar {
namespace spec {

constexpr float kLambdaMin = 360.0f;
constexpr float kLambdaMax = 780.0f;
constexpr int kTableCount = 421;          // 360..780 nm, 1 nm
constexpr float kD65Norm = 1.0f / 10567.065f;
constexpr int kLutRes = 64;               // per-axis resolution of the uplift table

#define AR_HD __host__ __device__ __forceinline__

AR_HD float clampf(float x, float a, float b) { return fminf(fmaxf(x, a), b); }

// Acklam's normal quantile with one Newton step (sample_ciexyz.cuh:28-156)
AR_HD float ndtri(float u) {
    u = clampf(u, 0.0f, 1.0f);
    const float a0 = -3.969683028665376e+01f, a1 = 2.209460984245205e+02f, a2 = -2.759285104469687e+02f,
                a3 = 1.383577518672690e+02f, a4 = -3.066479806614716e+01f, a5 = 2.506628277459239e+00f;
    const float b0 = -5.447609879822406e+01f, b1 = 1.615858368580409e+02f, b2 = -1.556989798598866e+02f,
                b3 = 6.680131188771972e+01f, b4 = -1.328068155288572e+01f;
    const float c0 = -7.784894002430293e-03f, c1 = -3.223964580411365e-01f, c2 = -2.400758277161838e+00f,
                c3 = -2.549732539343734e+00f, c4 = 4.374664141464968e+00f, c5 = 2.938163982698783e+00f;
    const float d0 = 7.784695709041462e-03f, d1 = 3.224671290700398e-01f, d2 = 2.445134137142996e+00f,
                d3 = 3.754408661907416e+00f;
    const float plow = 0.02425f, phigh = 1.0f - plow;
    float x;
    if (u < plow) {
        const float q = sqrtf(-2.0f * logf(u));
        float num = fmaf(fmaf(fmaf(fmaf(fmaf(c0, q, c1), q, c2), q, c3), q, c4), q, c5);
        float den = fmaf(fmaf(fmaf(fmaf(d0, q, d1), q, d2), q, d3), q, 1.0f);
        x = num / den;
    } else if (u > phigh) {
        const float q = sqrtf(-2.0f * log1pf(-u));
        float num = fmaf(fmaf(fmaf(fmaf(fmaf(c0, q, c1), q, c2), q, c3), q, c4), q, c5);
        float den = fmaf(fmaf(fmaf(fmaf(d0, q, d1), q, d2), q, d3), q, 1.0f);
        x = -num / den;
    } else {
        const float q = u - 0.5f, r = q * q;
        float num = fmaf(fmaf(fmaf(fmaf(fmaf(a0, r, a1), r, a2), r, a3), r, a4), r, a5) * q;
        float den = fmaf(fmaf(fmaf(fmaf(fmaf(b0, r, b1), r, b2), r, b3), r, b4), r, 1.0f);
        x = num / den;
    }
    const float inv_sqrt_2pi = 0.3989422804014327f, inv_sqrt2 = 0.7071067811865476f;
    const float pdf = expf(-0.5f * x * x) * inv_sqrt_2pi;
    const float cdf = 0.5f * erfcf(-x * inv_sqrt2);
    return x - (cdf - u) / pdf;
}

AR_HD float xbar_fit(float l) {
    const float k1 = (l - 595.8f) / 33.33f, k2 = (l - 446.8f) / 19.44f;
    return 1.065f * expf(-0.5f * k1 * k1) + 0.366f * expf(-0.5f * k2 * k2);
}
AR_HD float ybar_fit(float l) {
    const float k = (logf(l) - 6.3213077f) / 0.075f;
    return 1.014f * expf(-0.5f * k * k);
}
AR_HD float zbar_fit(float l) {
    const float k = (logf(l) - 6.1088030f) / 0.051f;
    return 1.839f * expf(-0.5f * k * k);
}

// density of the sampling mixture at l (sample_ciexyz.cuh:254-326)
AR_HD float lambda_pdf(float l) {
    const float total = 3.1f;
    const float pu = (0.1f / total) * (1.0f / (kLambdaMax - kLambdaMin));
    const float p = pu + (1.0f / total) * (xbar_fit(l) / 106.81109281f) + (1.0f / total) * (ybar_fit(l) / 106.34513640f) +
                    (1.0f / total) * (zbar_fit(l) / 105.88243587f);
    return fmaxf(p, 1e-16f);
}

AR_HD float sample_lambda(float u0, float u1, float u2) {
    const float total = 3.1f;
    const float ru = 0.1f / total, rx = 1.0f / total, ry = 1.0f / total;
    float l;
    if (u0 < ru) {
        l = u1 * (kLambdaMax - kLambdaMin) + kLambdaMin;
    } else if (u0 < ru + rx) {
        // two-Gaussian xbar fit (sample_ciexyz.cuh:159-203)
        if (u1 < 0.8330258723493694f) {
            const float t = 7.4884543e-13f + u2 * (0.9999999836707862f - 7.4884543e-13f);
            l = 595.8f + 33.33f * ndtri(t);
        } else {
            const float t = 0.0000040030528583f + u2 * (1.0f - 0.0000040030528583f);
            l = 446.8f + 19.44f * ndtri(t);
        }
    } else if (u0 < ru + rx + ry) {
        const float t = 0.0000000020798312f + u1 * (0.9999953206338883f - 0.0000000020798312f);
        l = expf(6.3269327170812479f + 0.075f * ndtri(t));
    } else {
        const float t = 0.0000049890553249f + u1 * (1.0f - 0.0000049890553249f);
        l = expf(6.1114040395252154f + 0.051f * ndtri(t));
    }
    return clampf(l, kLambdaMin, kLambdaMax);
}

// table lookups; tables hold 421 entries from 360 nm at 1 nm
AR_HD void cie_xyz(const float* __restrict__ xyz, float l, float& x, float& y, float& z) {
    if (l <= kLambdaMin) { x = xyz[0]; y = xyz[1]; z = xyz[2]; return; }
    if (l >= kLambdaMax) { x = xyz[3 * 420]; y = xyz[3 * 420 + 1]; z = xyz[3 * 420 + 2]; return; }
    const float u = l - kLambdaMin;
    const int i = (int)u;
    const float t = u - (float)i;
    const float* a = xyz + 3 * i;
    x = a[0] + (a[3] - a[0]) * t;
    y = a[1] + (a[4] - a[1]) * t;
    z = a[2] + (a[5] - a[2]) * t;
}

AR_HD float d65_norm(const float* __restrict__ d65, float l) {
    if (l <= kLambdaMin) return d65[0] * kD65Norm;
    if (l >= kLambdaMax) return d65[420] * kD65Norm;
    const float x = l - kLambdaMin;
    const int i = (int)floorf(x);
    const float t = x - (float)i;
    return (d65[i] + (d65[i + 1] - d65[i]) * t) * kD65Norm;
}

// Jakob-Hanika sigmoid polynomial at wavelength l
AR_HD float sigmoid_poly(float c0, float c1, float c2, float l) {
    const float x = (l - kLambdaMin) * (1.0f / (kLambdaMax - kLambdaMin));
    const float q = fmaf(fmaf(c0, x, c1), x, c2);
    if (!(fabsf(q) < 1e18f)) return q > 0.0f ? 1.0f : 0.0f;   // saturated (or inf)
    return 0.5f + q / (2.0f * sqrtf(1.0f + q * q));
}

// Largest-channel scale coordinate of the table (k -> smoothstep(smoothstep(k / (res-1))))
AR_HD float lut_scale(int k) {
    float t = (float)k / (float)(kLutRes - 1);
    t = t * t * (3.0f - 2.0f * t);
    return t * t * (3.0f - 2.0f * t);
}

#undef AR_HD

} // namespace spec
} // namespace ar
