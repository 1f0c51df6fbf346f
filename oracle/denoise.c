/*
 * Oracle of the screen-space denoiser (test infrastructure only).
 *
 *   orc_gbuffer_record   RecordScreenGBuffer + RecordShadingModel
 *                        (core/denoise/denoise.cuh:82-163): running means of
 *                        the first hit's albedo, normal, distance, metallic and
 *                        roughness over the samples that hit, and of the
 *                        shading-model one-hot over all samples
 *   orc_stats_record     RecordScreenStatisticBuffer (denoise_statistic.cuh:
 *                        36-69): Welford mean, M2 / n and M3 / n of the
 *                        Box-Cox (lambda 1/2) transformed sample, M2 and M3
 *                        carried in double
 *   orc_svgf_weight      ComputeSVGFWeight (denoise.cuh:186-241)
 *   orc_tstat_gate       ComputeTStatisticFinal (denoise.cuh:266-350): Johnson's
 *                        skewness-corrected mean, the asymmetric t statistic,
 *                        Welch's degrees of freedom and the 0.5 % two-sided
 *                        t quantile
 *   orc_bilateral        Bilateral16x16<7> (denoise.cuh:352-444): 15 x 15 taps,
 *                        G-buffer weight x spatial Gaussian (sigma 2) x t-test
 *                        gate, averaged in log(1 + c)
 *
 * Records (floats per pixel; counts are uint32 bit patterns):
 *   G-buffer (16): albedo[3] normal[3] depth metallic roughness
 *                  model[5] (diffuse, light, glass, GGX, none) valid frames
 *   statistics (12): n mean[3] M2/n[3] M3/n[3] (2 unused)
 * The reference stores the G-buffer means in half precision; here they are
 * float (the library does the same).  Box-Cox's pow(x, 0.5) is sqrtf.
 *
 * Pins: tests/test_oracle_denoise.py (Welford against numpy moments, the
 * gate against scipy's Welch t-test, the filter against a direct numpy
 * evaluation of its weights, the quantile table against scipy and the
 * reference's table).
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

static const float kTQuantile[1024] = {
#include "../aperturerendererv3_amd/csrc/t_quantile_table.inc"
};

static uint32_t getu(const float* p) { uint32_t u; memcpy(&u, p, 4); return u; }
static void setu(float* p, uint32_t u) { memcpy(p, &u, 4); }

/* float -> int as the GPU converts: NaN -> 0, saturating at the int range */
static int floor_to_int(float x)
{
    if (x != x) return 0;
    const float f = floorf(x);
    if (f >= 2147483648.0f) return 2147483647;
    if (f < -2147483648.0f) return (-2147483647 - 1);
    return (int)f;
}

float orc_t_quantile(int nu)
{
    uint32_t idx = nu <= 0 ? 0u : (uint32_t)(nu - 1);
    if (idx > 1023u) idx = 1023u;
    return kTQuantile[idx];
}

void orc_gbuffer_record(float* g, int hit, const float albedo[3], const float normal[3], float depth, float metallic,
                        float roughness, uint32_t model)
{
    if (hit) {
        const uint32_t valid = getu(&g[14]);
        if (valid == 0) for (int k = 0; k < 9; ++k) g[k] = 0.0f;
        const float affix = 1.0f / (float)(valid + 1u);
        const float pre = (float)valid / (float)(valid + 1u);
        for (int k = 0; k < 9; ++k) g[k] = g[k] * pre;
        for (int k = 0; k < 3; ++k) g[k] = g[k] + affix * albedo[k];
        for (int k = 0; k < 3; ++k) g[3 + k] = g[3 + k] + affix * normal[k];
        g[7] = g[7] + affix * metallic;
        g[8] = g[8] + affix * roughness;
        g[6] = g[6] + affix * depth;
        setu(&g[14], valid + 1u);
    } else {
        model = 4u;
    }
    /* RecordShadingModel */
    const uint32_t frame = getu(&g[15]);
    if (frame == 0) for (int k = 0; k < 5; ++k) g[9 + k] = 0.0f;
    const float affix = 1.0f / (float)(frame + 1u);
    const float pre = (float)frame / (float)(frame + 1u);
    for (int k = 0; k < 5; ++k) g[9 + k] = g[9 + k] * pre;
    g[9 + model] = g[9 + model] + affix;
    setu(&g[15], frame + 1u);
}

void orc_stats_record(float* st, const float x[3])
{
    const uint32_t n0 = getu(&st[0]);
    const uint32_t n1 = n0 + 1u;
    setu(&st[0], n1);
    const float inv = 1.0f / (float)n1;
    for (int c = 0; c < 3; ++c) {
        const float xb = (sqrtf(fmaxf(x[c], 0.0f)) - 1.0f) * 2.0f;     /* BoxCox05 */
        const float mean0 = st[1 + c];
        const float delta = xb - mean0;
        const float ddn = delta * inv;
        const float mean1 = mean0 + ddn;
        const double m2_0 = (double)st[4 + c] * n0;
        const double m2_1 = m2_0 + (double)(delta * (delta - ddn));
        const double m3_0 = (double)st[7 + c] * n0;
        const float d2 = delta * delta;
        const float ddn2 = ddn * ddn;
        const double m3_1 = m3_0 - 3.0 * ((double)ddn * m2_1) + (double)(delta * (d2 - ddn2));
        st[1 + c] = mean1;
        st[4 + c] = (float)(m2_1 / n1);
        st[7 + c] = (float)(m3_1 / n1);
    }
}

static void normalize3f(float v[3])
{
    const float l2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (l2 <= 1e-20f) { v[0] = v[1] = v[2] = 0.0f; return; }
    const float inv = 1.0f / sqrtf(l2);
    v[0] = v[0] * inv; v[1] = v[1] * inv; v[2] = v[2] * inv;
}

float orc_svgf_weight(const float* a, const float* b)
{
    float w = 0.0f;
    {
        const float d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
        w = w + (d0 * d0 + d1 * d1 + d2 * d2) * 0.5f;              /* 1 / (2 * 1^2) */
    }
    {
        float na[3] = { a[3], a[4], a[5] }, nb[3] = { b[3], b[4], b[5] };
        normalize3f(na);
        normalize3f(nb);
        const float d = 1.0f - (na[0] * nb[0] + na[1] * nb[1] + na[2] * nb[2]);
        w = w + d * d * 2.0f;                                       /* 1 / (2 * 0.5^2) */
    }
    {
        const float den = fmaxf(fmaxf(a[6], b[6]), 1e-6f);
        const float d = (a[6] - b[6]) / den;
        w = w + d * d * 8.0f;                                       /* 1 / (2 * 0.25^2) */
    }
    {
        const float dm = a[7] - b[7];
        w = w + dm * dm * 2.0f;
    }
    {
        const float dr = a[8] - b[8];
        w = w + dr * dr * 2.0f;
    }
    {
        float l1 = 0.0f;
        for (int k = 0; k < 5; ++k) l1 = l1 + fabsf(a[9 + k] - b[9 + k]);
        const float tv = 0.5f * l1;
        w = w + tv * tv * (1.0f / (2.0f * 0.1f * 0.1f));           /* FilterParams in float: 49.999996 */
    }
    return expf(-w);
}

typedef struct {
    float s2[3];      /* Bessel-corrected variance */
    float vt[3];      /* variance of the estimator: s2 / n */
    float th[3];      /* mean + M3 / n / 6 / s2 */
} tstat;

static tstat make_tstat(const float* st)
{
    tstat r;
    const uint32_t n = getu(&st[0]);
    const double num = (double)n;
    const uint32_t nm1 = n - 1u > 1u ? n - 1u : 1u;
    const double inv = 1.0 / (double)nm1;
    for (int c = 0; c < 3; ++c) {
        const double m2 = (double)st[4 + c] * num;
        r.s2[c] = (float)(m2 * inv);
        r.vt[c] = (float)((double)r.s2[c] / num);
        r.th[c] = st[1 + c] + st[7 + c] / 6.0f / r.s2[c];
    }
    return r;
}

static float welch_nu(float sa, uint32_t na, float sb, uint32_t nb)
{
    const float ua = sa / (float)na, ub = sb / (float)nb;
    const float u = ua + ub;
    const float u2 = u * u;
    const float la = ua * ua / (float)(na - 1u > 1u ? na - 1u : 1u);
    const float lb = ub * ub / (float)(nb - 1u > 1u ? nb - 1u : 1u);
    return u2 / (la + lb);
}

float orc_tstat_gate(const float* sa, const float* sb)
{
    const tstat A = make_tstat(sa), B = make_tstat(sb);
    const uint32_t na = getu(&sa[0]), nb = getu(&sb[0]);
    int pass = 1;
    float T[3], q[3];
    for (int c = 0; c < 3; ++c) {
        const float t = A.th[c] - B.th[c];
        const float t2 = t * t;
        const float up = t2 + B.vt[c];
        const float lo = t2 + A.vt[c] + B.vt[c];
        const float aw = up / fmaxf(lo, 1.0e-6f);
        T[c] = sqrtf(0.5f / (1.0f - aw) - 1.0f);
        q[c] = orc_t_quantile(floor_to_int(welch_nu(A.s2[c], na, B.s2[c], nb)));
    }
    for (int c = 0; c < 3; ++c) pass = pass && T[c] < q[c];
    return pass ? 1.0f : 0.0f;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

void orc_bilateral(const float* hdr, const float* gbuf, const float* stats, float* out, uint32_t width,
                   uint32_t height)
{
    const int W = (int)width, H = (int)height, R = 7;
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int ci = y * W + x;
            const float* cg = &gbuf[16 * (size_t)ci];
            const float* cs = &stats[12 * (size_t)ci];
            float acc[3], aw = 1.0e-6f;
            for (int c = 0; c < 3; ++c) acc[c] = logf(hdr[3 * (size_t)ci + c] + 1.0f) * 1.0e-6f;
            for (int j = -R; j <= R; ++j) {
                const int ny = clampi(y + j, 0, H - 1);
                for (int i = -R; i <= R; ++i) {
                    const int nx = clampi(x + i, 0, W - 1);
                    const int ni = ny * W + nx;
                    const float w = orc_svgf_weight(cg, &gbuf[16 * (size_t)ni]) *
                                    expf((float)(-(i * i + j * j)) * 0.125f) *
                                    orc_tstat_gate(cs, &stats[12 * (size_t)ni]);
                    for (int c = 0; c < 3; ++c) acc[c] = acc[c] + logf(hdr[3 * (size_t)ni + c] + 1.0f) * w;
                    aw = aw + w;
                }
            }
            for (int c = 0; c < 3; ++c) out[3 * (size_t)ci + c] = fmaxf(expf(acc[c] / aw) - 1.0f, 0.0f);
        }
}
