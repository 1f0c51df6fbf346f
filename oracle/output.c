/*
 * Oracle of the output stage (test infrastructure only): the display
 * tonemap, the luminance MSE between the odd and even sample halves and the
 * progressive running mean.
 *
 *   orc_tonemap          Tonemap kernel (core/scene.cu:22-133): XYZ2SRGBLinearD65
 *                        for spectral input (core/tonemap.cuh:116-126), exposure,
 *                        Gamma / ACES (core/tonemap.cuh:6-114), 8-bit
 *                        0xff000000 | r << 16 | g << 8 | b
 *   orc_luminance_mse    DeviceMSELuminance (core/helper/measure.cuh:49-74)
 *   orc_running_mean     the lerp(target, result, 1 / (1 + frame)) running mean
 *                        of RenderCamera (core/scene.cu:176-189) folded over k
 *                        samples at once
 *   orc_viridis,         the Tonemap kernel's debug views (core/scene.cu:35-107)
 *   orc_view             with HeatMapViridis (core/render/color.cuh:99-140)
 *
 * Pins: tests/test_oracle_output.py (gamma closed form, ACES properties,
 * quantisation and packing, the MSE against numpy, the running mean against
 * the direct mean; the views: knots and midpoints of the heat map, t = 1
 * landing on the 3/4 sample as the reference's segment search does, each view's
 * channel formula on hand-made records).
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

static float sat(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

static void m3(const float m[9], const float v[3], float o[3])
{
    for (int r = 0; r < 3; ++r) o[r] = m[3 * r] * v[0] + m[3 * r + 1] * v[1] + m[3 * r + 2] * v[2];
}

static void gamma3(float v[3])
{
    for (int k = 0; k < 3; ++k) v[k] = powf(v[k], 1.0f / 2.2f);
}

static float lerpf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

/* ACES (core/tonemap.cuh:66-114): sRGB -> AP0 -> AP1, desaturate 0.96, the
 * (x (x a + b)) / (x (x c + d) + e) curve, dim surround (xyY with Y^0.9811),
 * desaturate 0.93, AP1 -> XYZ, D60 -> D65, XYZ -> Rec.709, clamp >= 0, gamma */
void orc_aces(const float in[3], float out[3])
{
    static const float srgb_ap0[9] = { 0.4397010f, 0.3829780f, 0.1773350f, 0.0897923f, 0.8134230f, 0.0967616f,
                                       0.0175440f, 0.1115440f, 0.8707040f };
    static const float ap0_ap1[9] = { 1.4514393161f, -0.2365107469f, -0.2149285693f, -0.0765537734f, 1.1762296998f,
                                      -0.0996759264f, 0.0083161484f, -0.0060324498f, 0.9977163014f };
    static const float ap1_xyz[9] = { 0.6624541811f, 0.1340042065f, 0.1561876870f, 0.2722287168f, 0.6740817658f,
                                      0.0536895174f, -0.0055746495f, 0.0040607335f, 1.0103391003f };
    static const float xyz_ap1[9] = { 1.6410233797f, -0.3248032942f, -0.2364246952f, -0.6636628587f, 1.6153315917f,
                                      0.0167563477f, 0.0117218943f, -0.0082844420f, 0.9883948585f };
    static const float d60_d65[9] = { 0.98722400f, -0.00611327f, 0.0159533f, -0.00759836f, 1.00186000f, 0.0053302f,
                                      0.00307257f, -0.00509595f, 1.0816800f };
    static const float xyz_709[9] = { 3.2409699419f, -1.5373831776f, -0.4986107603f, -0.9692436363f, 1.8759675015f,
                                      0.0415550574f, 0.0556300797f, -0.2039769589f, 1.0569715142f };
    const float y2[3] = { 0.272229f, 0.674082f, 0.0536895f };
    float aces[3], cg[3], post[3], xyz[3], lin[3], t[3];
    m3(srgb_ap0, in, aces);
    m3(ap0_ap1, aces, cg);
    float tmp = cg[0] * y2[0] + cg[1] * y2[1] + cg[2] * y2[2];
    for (int k = 0; k < 3; ++k) cg[k] = lerpf(tmp, cg[k], 0.96f);
    const float a = 278.5085f, b = 10.7772f, c = 293.6045f, d = 88.7122f, e = 80.6889f;
    for (int k = 0; k < 3; ++k) {
        const float x = cg[k];
        post[k] = (x * (x * a + b)) / (x * (x * c + d) + e);
    }
    /* darkSurround_to_dimSurround */
    m3(ap1_xyz, post, xyz);
    const float div = fmaxf(xyz[0] + xyz[1] + xyz[2], 1.0e-4f);
    float xyY[3] = { xyz[0] / div, xyz[1] / div, xyz[1] };
    xyY[2] = fminf(fmaxf(xyY[2], 0.0f), 65504.0f);
    xyY[2] = powf(xyY[2], 0.9811f);
    const float mm = xyY[2] / fmaxf(xyY[1], 1e-4f);
    t[0] = xyY[0] * mm;
    t[1] = xyY[2];
    t[2] = (1.0f - xyY[0] - xyY[1]) * mm;
    m3(xyz_ap1, t, lin);
    tmp = lin[0] * y2[0] + lin[1] * y2[1] + lin[2] * y2[2];
    for (int k = 0; k < 3; ++k) lin[k] = lerpf(tmp, lin[k], 0.93f);
    m3(ap1_xyz, lin, xyz);
    m3(d60_d65, xyz, t);
    m3(xyz_709, t, out);
    for (int k = 0; k < 3; ++k) out[k] = fmaxf(0.0f, out[k]);
    gamma3(out);
}

void orc_tonemap(const float* hdr, uint32_t* ldr, uint32_t npix, float exposure, int tone, int xyz_input)
{
    for (uint32_t i = 0; i < npix; ++i) {
        float v[3] = { hdr[3 * i], hdr[3 * i + 1], hdr[3 * i + 2] };
        if (xyz_input) {
            const float r = 3.2406f * v[0] + -1.5372f * v[1] + -0.4986f * v[2];
            const float g = -0.9689f * v[0] + 1.8758f * v[1] + 0.0415f * v[2];
            const float b = 0.0557f * v[0] + -0.204f * v[1] + 1.057f * v[2];
            v[0] = fmaxf(0.0f, r); v[1] = fmaxf(0.0f, g); v[2] = fmaxf(0.0f, b);
        }
        for (int k = 0; k < 3; ++k) v[k] = v[k] * exposure;
        if (tone == 1) gamma3(v);
        else if (tone == 2) { float o[3]; orc_aces(v, o); v[0] = o[0]; v[1] = o[1]; v[2] = o[2]; }
        const uint32_t r = (uint32_t)(255.0f * sat(v[0]));
        const uint32_t g = (uint32_t)(255.0f * sat(v[1]));
        const uint32_t b = (uint32_t)(255.0f * sat(v[2]));
        ldr[i] = 0xff000000u | (r << 16) | (g << 8) | b;
    }
}

double orc_luminance_mse(const float* a, const float* b, size_t npix)
{
    double sum = 0.0;
    for (size_t i = 0; i < npix; ++i) {
        const float la = 0.2126f * a[3 * i] + 0.7152f * a[3 * i + 1] + 0.0722f * a[3 * i + 2];
        const float lb = 0.2126f * b[3 * i] + 0.7152f * b[3 * i + 1] + 0.0722f * b[3 * i + 2];
        const double d = (double)la - (double)lb;
        sum += d * d;
    }
    return npix ? sum / (double)npix : 0.0;
}

/* acc holds the mean of n_before samples; fold in the mean of n_new more */
void orc_running_mean(float* acc, const float* mean_new, size_t nfloat, uint32_t n_before, uint32_t n_new)
{
    if (n_new == 0) return;
    const float total = (float)n_before + (float)n_new;
    const float wa = (float)n_before / total, wb = (float)n_new / total;
    for (size_t i = 0; i < nfloat; ++i) acc[i] = n_before == 0 ? mean_new[i] : acc[i] * wa + mean_new[i] * wb;
}

/* matplotlib's viridis at t = 0, 1/4, 1/2, 3/4, 1 */
static const float viridis_knots[15] = {
    0.267004f, 0.004874f, 0.329415f, 0.229739f, 0.322361f, 0.545706f, 0.127568f, 0.566949f, 0.550556f,
    0.369214f, 0.788888f, 0.382914f, 0.993248f, 0.906157f, 0.143936f,
};

/* Linear between the knots; segment floor(4t) clamped to 0..3 with the
 * weight 4t - floor(4t), so t = 1 gives the 3/4 knot (the reference's search
 * ends on the last segment with weight 0). */
void orc_viridis(float t, float out[3])
{
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const float a = t * 4.0f;
    const int i = (int)floorf(a);
    const float f = a - (float)i;
    const int s = i <= 0 ? 0 : (i >= 3 ? 3 : i);
    for (int k = 0; k < 3; ++k)
        out[k] = viridis_knots[3 * s + k] + (viridis_knots[3 * s + 3 + k] - viridis_knots[3 * s + k]) * f;
}

static uint32_t pack(const float v[3])
{
    const uint32_t r = (uint32_t)(255.0f * sat(v[0]));
    const uint32_t g = (uint32_t)(255.0f * sat(v[1]));
    const uint32_t b = (uint32_t)(255.0f * sat(v[2]));
    return 0xff000000u | (r << 16) | (g << 8) | b;
}

void orc_view(const float* gbuf, const float* stats, uint32_t* ldr, uint32_t npix, int view)
{
    for (uint32_t i = 0; i < npix; ++i) {
        const float* g = gbuf ? gbuf + 16 * (size_t)i : NULL;
        const float* st = stats ? stats + 12 * (size_t)i : NULL;
        float v[3] = { 0.0f, 0.0f, 0.0f };
        if (view == 2) {
            for (int k = 0; k < 3; ++k) v[k] = g[k];
        } else if (view == 3) {
            for (int k = 0; k < 3; ++k) v[k] = g[3 + k] * 0.5f + 0.5f;
        } else if (view == 4) {
            orc_viridis(1.0f / (g[6] + 1.0f), v);
        } else if (view == 5 || view == 6) {
            orc_viridis(g[view + 2], v);
        } else if (view == 7) {
            uint32_t n;
            memcpy(&n, st, 4);
            v[0] = (float)log((double)n + 1.0) * 0.01f;
        } else {
            const int at = 1 + 3 * (view - 8);
            for (int k = 0; k < 3; ++k) v[k] = fabsf(st[at + k]) * 0.25f;
        }
        ldr[i] = pack(v);
    }
}
