/*
 * Oracle path integrator, written once for a scalar type R (float or double)
 * and included twice by render.c.  TEST INFRASTRUCTURE ONLY.
 *
 * Restates CalculateRadiance (core/render/render.cuh:460-998) for triangle
 * meshes in RGB mode, with the camera of RenderCamera (core/scene.cu:135-189):
 *   - Trace with tMin = INTERSECT_EPS(nonDeltaDepth) (core/config.cuh:3) and
 *     the hit-point offset eps*max(1,|p|)*(U+0.5) along the side of the next
 *     direction (render.cuh:75-163);
 *   - BSDF sample / eval / pdf for Lambert and GGX+Lambert
 *     (core/render/material.cuh:47-371, brdf.cuh:15-129, sample.cuh:63-248);
 *   - NEE taken with probability sqrt(sat(lum(throughput))) at non-delta
 *     depth < 10, shadow segment with tMin = 0 and tMax = dist - eps*max(1,
 *     |src|, |dst|), power-heuristic MIS in solid angle; light hits weighted
 *     against the previous vertex's NEE (render.cuh:205-343, 664-710, 766-827);
 *   - Russian roulette from non-delta depth 3 with q = clamp(0.5*lum +
 *     0.5*max, 0.05, 0.95) (render.cuh:838-859);
 *   - output clamp(checkNanInf(L), 0, 1e3) (render.cuh:966-997).
 * Triangle lights (emission > 0) replace the reference's MAT_LIGHT objects:
 * picked by the CDF of area * luminance(Le) (core/scene.cu:677-701), sampled
 * uniformly by area, emitting on the front face only, and ending the path
 * when hit (SampleBSDF MAT_LIGHT, material.cuh:196-203).
 * max_depth caps the bounce loop (the reference runs to MAX_TRACING_DEPTH);
 * NEE is skipped on the last iteration so that every path length up to
 * max_depth gets consistent MIS weights.
 *
 * Spectral modes restate the reference's wavelength path (render.cuh:466-474,
 * 657-700, 966-997): one wavelength from the CIE-fitted mixture with
 * density p (sample_ciexyz.cuh:15-326, Acklam's quantile + one Newton step),
 * or four hero wavelengths 100 nm apart in [380, 780); reflectance and
 * emission spectra from sigmoid-polynomial coefficients (Jakob & Hanika 2019)
 * looked up in the 3 x 64^3 table; emitters times D65 normalised to
 * integral D65 * ybar = 1 (ciexyz.cuh:67-85); output XYZ = xbar..zbar(lambda)
 * * L / p (linear interpolation of the 1 nm CIE table, ciexyz.cuh:522-543).
 *
 * Random numbers: Philox4x32-10 (Salmon et al., SC'11), key = seed, counter
 * = (pixel, sample, bounce, group); the GPU kernels draw the same numbers.
 */

#define RF(name) RF_CAT(name, SUF)
#define V3 RF(vec3)

typedef struct { R x, y, z; } V3;

static inline V3 RF(v3)(R x, R y, R z) { V3 v = { x, y, z }; return v; }
static inline V3 RF(add)(V3 a, V3 b) { return RF(v3)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 RF(sub)(V3 a, V3 b) { return RF(v3)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 RF(mul)(V3 a, V3 b) { return RF(v3)(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V3 RF(scale)(V3 a, R s) { return RF(v3)(a.x * s, a.y * s, a.z * s); }
static inline R RF(dot)(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 RF(cross)(V3 a, V3 b)
{
    return RF(v3)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline R RF(length)(V3 a) { return SQRT(RF(dot)(a, a)); }
/* normalize returns 0 below |v|^2 = 1e-20 (core/vector.cuh normalize) */
static inline V3 RF(normalize)(V3 v)
{
    const R l2 = RF(dot)(v, v);
    if (l2 <= (R)1e-20) return RF(v3)(0, 0, 0);
    const R inv = (R)1 / SQRT(l2);
    return RF(scale)(v, inv);
}
static inline R RF(maxr)(R a, R b) { return a > b ? a : b; }
static inline R RF(minr)(R a, R b) { return a < b ? a : b; }
static inline R RF(clampr)(R x, R a, R b) { return RF(minr)(RF(maxr)(x, a), b); }
static inline R RF(lum)(V3 c) { return c.x * (R)0.2126 + c.y * (R)0.7152 + c.z * (R)0.0722; }
static inline R RF(max3)(V3 c) { return RF(maxr)(c.x, RF(maxr)(c.y, c.z)); }
static inline V3 RF(reflect)(V3 v, V3 n) { return RF(sub)(v, RF(scale)(n, (R)2 * RF(dot)(v, n))); }
static inline R RF(fixnan)(R x) { return (isnan(x) || isinf(x)) ? (R)0 : x; }

static inline V3 RF(load3)(const float* p) { return RF(v3)((R)p[0], (R)p[1], (R)p[2]); }

/* per-path spectral quantity: RGB (channels 0..2), one wavelength (channel 0)
 * or four hero wavelengths */
#define S4 RF(s4t)
typedef struct { R v[4]; } S4;
static inline S4 RF(s4)(R a, R b, R c, R d) { S4 r; r.v[0] = a; r.v[1] = b; r.v[2] = c; r.v[3] = d; return r; }
static inline S4 RF(s4mul)(S4 a, S4 b) { return RF(s4)(a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]); }
static inline S4 RF(s4add)(S4 a, S4 b) { return RF(s4)(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]); }
static inline S4 RF(s4scale)(S4 a, R s) { return RF(s4)(a.v[0] * s, a.v[1] * s, a.v[2] * s, a.v[3] * s); }

/* ---------------------------------------------------------------------- */
/* Wavelengths (core/spectrum/sample_ciexyz.cuh, ciexyz.cuh)               */
/* ---------------------------------------------------------------------- */

static R RF(ndtri)(R u)
{
    u = u < 0 ? 0 : (u > 1 ? 1 : u);
    const R a0 = (R)-3.969683028665376e+01, a1 = (R)2.209460984245205e+02, a2 = (R)-2.759285104469687e+02,
            a3 = (R)1.383577518672690e+02, a4 = (R)-3.066479806614716e+01, a5 = (R)2.506628277459239e+00;
    const R b0 = (R)-5.447609879822406e+01, b1 = (R)1.615858368580409e+02, b2 = (R)-1.556989798598866e+02,
            b3 = (R)6.680131188771972e+01, b4 = (R)-1.328068155288572e+01;
    const R c0 = (R)-7.784894002430293e-03, c1 = (R)-3.223964580411365e-01, c2 = (R)-2.400758277161838e+00,
            c3 = (R)-2.549732539343734e+00, c4 = (R)4.374664141464968e+00, c5 = (R)2.938163982698783e+00;
    const R d0 = (R)7.784695709041462e-03, d1 = (R)3.224671290700398e-01, d2 = (R)2.445134137142996e+00,
            d3 = (R)3.754408661907416e+00;
    const R plow = (R)0.02425, phigh = (R)1 - plow;
    R x;
    if (u < plow) {
        const R q = SQRT((R)-2 * LOG(u));
        const R num = FMA(FMA(FMA(FMA(FMA(c0, q, c1), q, c2), q, c3), q, c4), q, c5);
        const R den = FMA(FMA(FMA(FMA(d0, q, d1), q, d2), q, d3), q, (R)1);
        x = num / den;
    } else if (u > phigh) {
        const R q = SQRT((R)-2 * LOG1P(-u));
        const R num = FMA(FMA(FMA(FMA(FMA(c0, q, c1), q, c2), q, c3), q, c4), q, c5);
        const R den = FMA(FMA(FMA(FMA(d0, q, d1), q, d2), q, d3), q, (R)1);
        x = -num / den;
    } else {
        const R q = u - (R)0.5, r = q * q;
        const R num = FMA(FMA(FMA(FMA(FMA(a0, r, a1), r, a2), r, a3), r, a4), r, a5) * q;
        const R den = FMA(FMA(FMA(FMA(FMA(b0, r, b1), r, b2), r, b3), r, b4), r, (R)1);
        x = num / den;
    }
    const R pdf = EXP((R)-0.5 * x * x) * (R)0.3989422804014327;
    const R cdf = (R)0.5 * ERFC(-x * (R)0.7071067811865476);
    return x - (cdf - u) / pdf;
}

static R RF(xfit)(R l) { const R a = (l - (R)595.8) / (R)33.33, b = (l - (R)446.8) / (R)19.44;
                         return (R)1.065 * EXP((R)-0.5 * a * a) + (R)0.366 * EXP((R)-0.5 * b * b); }
static R RF(yfit)(R l) { const R k = (LOG(l) - (R)6.3213077) / (R)0.075; return (R)1.014 * EXP((R)-0.5 * k * k); }
static R RF(zfit)(R l) { const R k = (LOG(l) - (R)6.1088030) / (R)0.051; return (R)1.839 * EXP((R)-0.5 * k * k); }

static R RF(lambda_pdf)(R l)
{
    const R total = (R)3.1;
    const R p = ((R)0.1 / total) * ((R)1 / (R)420) + ((R)1 / total) * (RF(xfit)(l) / (R)106.81109281) +
                ((R)1 / total) * (RF(yfit)(l) / (R)106.34513640) + ((R)1 / total) * (RF(zfit)(l) / (R)105.88243587);
    return p > (R)1e-16 ? p : (R)1e-16;
}

static R RF(sample_lambda)(R u0, R u1, R u2)
{
    const R total = (R)3.1, ru = (R)0.1 / total, rx = (R)1 / total, ry = (R)1 / total;
    R l;
    if (u0 < ru) l = u1 * (R)420 + (R)360;
    else if (u0 < ru + rx) {
        if (u1 < (R)0.8330258723493694) {
            const R t = (R)7.4884543e-13 + u2 * ((R)0.9999999836707862 - (R)7.4884543e-13);
            l = (R)595.8 + (R)33.33 * RF(ndtri)(t);
        } else {
            const R t = (R)0.0000040030528583 + u2 * ((R)1 - (R)0.0000040030528583);
            l = (R)446.8 + (R)19.44 * RF(ndtri)(t);
        }
    } else if (u0 < ru + rx + ry) {
        const R t = (R)0.0000000020798312 + u1 * ((R)0.9999953206338883 - (R)0.0000000020798312);
        l = EXP((R)6.3269327170812479 + (R)0.075 * RF(ndtri)(t));
    } else {
        const R t = (R)0.0000049890553249 + u1 * ((R)1 - (R)0.0000049890553249);
        l = EXP((R)6.1114040395252154 + (R)0.051 * RF(ndtri)(t));
    }
    return l < (R)360 ? (R)360 : (l > (R)780 ? (R)780 : l);
}

static V3 RF(cie)(const float* t, R l)
{
    if (l <= (R)360) return RF(load3)(t);
    if (l >= (R)780) return RF(load3)(t + 3 * 420);
    const R u = l - (R)360;
    const int i = (int)u;
    const R f = u - (R)i;
    const float* a = t + 3 * i;
    return RF(v3)((R)a[0] + ((R)a[3] - (R)a[0]) * f, (R)a[1] + ((R)a[4] - (R)a[1]) * f, (R)a[2] + ((R)a[5] - (R)a[2]) * f);
}

static R RF(d65n)(const float* t, R l)
{
    const R k = (R)1 / (R)10567.065f;
    if (l <= (R)360) return (R)t[0] * k;
    if (l >= (R)780) return (R)t[420] * k;
    const R x = l - (R)360;
    const int i = (int)FLOOR(x);
    const R f = x - (R)i;
    return ((R)t[i] + ((R)t[i + 1] - (R)t[i]) * f) * k;
}

static R RF(sigmoid)(const float* c, R l)
{
    const R x = (l - (R)360) * ((R)1 / (R)420);
    const R q = FMA(FMA((R)c[0], x, (R)c[1]), x, (R)c[2]);
    if (!(FABS(q) < (R)1e18)) return q > 0 ? (R)1 : (R)0;
    return (R)0.5 + q / ((R)2 * SQRT((R)1 + q * q));
}

typedef struct { int mode; R l[4]; } RF(lam);

static RF(lam) RF(make_lam)(int mode, R l0)
{
    RF(lam) L;
    L.mode = mode;
    L.l[0] = l0;
    for (int k = 1; k < 4; ++k) {
        R x = l0 - (R)380 + (R)100 * (R)k;
        if (x >= (R)400) x -= (R)400;
        L.l[k] = (R)380 + x;
    }
    return L;
}

static R RF(lum4)(S4 c, int mode)
{
    if (mode == 0) return c.v[0] * (R)0.2126 + c.v[1] * (R)0.7152 + c.v[2] * (R)0.0722;
    if (mode == 1) return c.v[0];
    return (c.v[0] + c.v[1] + c.v[2] + c.v[3]) * (R)0.25;
}

static R RF(max4)(S4 c, int mode)
{
    R a, b;
    if (mode == 0) { a = c.v[0] > c.v[1] ? c.v[0] : c.v[1]; return a > c.v[2] ? a : c.v[2]; }
    if (mode == 1) return c.v[0];
    a = c.v[0] > c.v[1] ? c.v[0] : c.v[1];
    b = c.v[2] > c.v[3] ? c.v[2] : c.v[3];
    return a > b ? a : b;
}

/* Russian-roulette survival probability of a path with throughput `mask`
 * before the current vertex's BSDF weight (render.cuh:838-852):
 * clamp(0.5 lum + 0.5 max, 0.05, 0.95), raised toward 1 by 0.99 inside a
 * medium (DELTA_RR_TERMINATE_REDUCE) */
static R RF(rr_q)(S4 mask, int mode, int in_medium)
{
    const R base = RF(clampr)(RF(lum4)(mask, mode) * (R)0.5 + RF(max4)(mask, mode) * (R)0.5, (R)0.05, (R)0.95);
    return in_medium ? base * ((R)1 - (R)0.99f) + (R)0.99f : base;
}

/* all active channels zero (allZero in render.cuh:364-407) */
static int RF(s4zero)(S4 c, int mode)
{
    if (mode == 0) return c.v[0] == 0 && c.v[1] == 0 && c.v[2] == 0;
    if (mode == 1) return c.v[0] == 0;
    return c.v[0] == 0 && c.v[1] == 0 && c.v[2] == 0 && c.v[3] == 0;
}

/* ---------------------------------------------------------------------- */
/* Sampling (core/render/sample.cuh)                                       */
/* ---------------------------------------------------------------------- */

/* BuildTangentBasisRandom (core/vector.cuh:284-294) */
static void RF(tangent_basis_random)(V3 n, R e, V3* t, V3* b)
{
    const V3 a = (FABS(n.z) < (R)0.999) ? RF(v3)(0, 0, 1) : RF(v3)(1, 0, 0);
    const V3 t0 = RF(normalize)(RF(cross)(a, n));
    const V3 b0 = RF(cross)(n, t0);
    const R ang = e * PI_R * (R)2;
    const R s = SIN(ang), c = COS(ang);
    *t = RF(add)(RF(scale)(t0, c), RF(scale)(b0, s));
    *b = RF(add)(RF(scale)(t0, -s), RF(scale)(b0, c));
}

/* ConcentricSampleDisk (sample.cuh:63-86) */
static void RF(concentric_disk)(R u0, R u1, R* dx, R* dy)
{
    const R sx = (R)2 * u0 - (R)1, sy = (R)2 * u1 - (R)1;
    if (sx == 0 && sy == 0) { *dx = 0; *dy = 0; return; }
    R r, phi;
    if (FABS(sx) > FABS(sy)) {
        r = sx;
        phi = (PI_R / (R)4) * (sy / RF(maxr)(FABS(sx), (R)1e-20));
    } else {
        r = sy;
        phi = (PI_R / (R)2) - (PI_R / (R)4) * (sx / RF(maxr)(FABS(sy), (R)1e-20));
    }
    *dx = r * COS(phi);
    *dy = r * SIN(phi);
}

/* SampleCosineWeightedHemisphere (sample.cuh:88-108) */
static V3 RF(sample_cosine)(R e0, R e1, R e2, V3 n, R* pdf)
{
    V3 t, b;
    RF(tangent_basis_random)(n, e2, &t, &b);
    R x, y;
    RF(concentric_disk)(e0, e1, &x, &y);
    const R z = SQRT(RF(maxr)((R)0, (R)1 - x * x - y * y));
    const V3 dir = RF(add)(RF(add)(RF(scale)(t, x), RF(scale)(b, y)), RF(scale)(n, z));
    *pdf = z / PI_R;
    return RF(normalize)(dir);
}

/* GGX_D (brdf.cuh:82-87) */
static R RF(ggx_d)(R alpha, R noh)
{
    const R a2 = alpha * alpha;
    const R d = noh * noh * (a2 - (R)1) + (R)1;
    return a2 / RF(maxr)(PI_R * d * d, (R)1e-12);
}

/* SampleGGXBoundedVNDF (sample.cuh:179-224; Eto & Tokuyoshi 2023) */
static V3 RF(sample_vndf)(V3 v, V3 n, R alpha, R r0, R r1, R r2, R* pdf)
{
    V3 X, Y;
    RF(tangent_basis_random)(n, r2, &X, &Y);
    const R xov = RF(dot)(X, v), yov = RF(dot)(Y, v), nov = RF(dot)(n, v);
    const R nov2 = RF(minr)((R)1, nov * nov);
    const R a2 = alpha * alpha;
    const R inv = (R)1 / SQRT(a2 + ((R)1 - a2) * nov2);
    const V3 vstd = RF(v3)(xov * alpha * inv, yov * alpha * inv, nov * inv);
    const R phi = (R)2 * PI_R * r0;
    const R s = (R)1 + SQRT((R)1 - nov2);
    const R s2 = s * s;
    const R k = ((R)1 - a2) * s2 / (s2 + a2 * nov2);
    const R b = nov > 0 ? k * vstd.z : vstd.z;
    const R z = ((R)1 - r1) * ((R)1 + b) - b;
    const R st = SQRT(RF(clampr)((R)1 - z * z, (R)0, (R)1));
    const V3 mstd = RF(v3)(vstd.x + st * COS(phi), vstd.y + st * SIN(phi), vstd.z + z);
    const V3 h = RF(normalize)(RF(v3)(mstd.x * alpha, mstd.y * alpha, mstd.z));
    const V3 lv = RF(v3)(xov, yov, nov);
    const R d2 = (R)2 * RF(dot)(lv, h);
    const V3 ll = RF(v3)(d2 * h.x - xov, d2 * h.y - yov, d2 * h.z - nov);
    const V3 dir = RF(add)(RF(add)(RF(scale)(X, ll.x), RF(scale)(Y, ll.y)), RF(scale)(n, ll.z));
    const R ndf = RF(ggx_d)(alpha, h.z);
    const R len2 = ((R)1 - nov2) * a2;
    const R t = SQRT(len2 + nov2);
    *pdf = ndf / ((R)2 * (k * nov + t));
    return dir;
}

/* GGXBoundedVNDF_PDF (sample.cuh:226-248) */
static R RF(vndf_pdf)(V3 v, V3 n, V3 l, R alpha)
{
    const R nov = RF(dot)(n, v), nol = RF(dot)(n, l);
    if (nov <= 0 || nol <= 0) return 0;
    const V3 h = RF(normalize)(RF(add)(v, l));
    const R nov2 = RF(minr)((R)1, nov * nov);
    const R a2 = alpha * alpha;
    const R s = (R)1 + SQRT((R)1 - nov2);
    const R s2 = s * s;
    const R k = ((R)1 - a2) * s2 / (s2 + a2 * nov2);
    const R noh = RF(maxr)((R)0, RF(dot)(n, h));
    const R ndf = RF(ggx_d)(alpha, noh);
    const R len2 = ((R)1 - nov2) * a2;
    const R t = SQRT(len2 + nov2);
    return RF(maxr)(ndf / ((R)2 * (k * nov + t)), (R)1e-12);
}

/* ---------------------------------------------------------------------- */
/* BSDF (core/render/material.cuh, brdf.cuh)                               */
/* ---------------------------------------------------------------------- */

typedef struct {
    V3 albedo, emission;
    R specular, metallic, roughness, escale, ascale;
    uint32_t model;
    float ca[4], ce[4];
    R ior, sigma_a, abbe, tilt;
} RF(mat);

static R RF(f0_ue)(R spec) { return (R)0.08 * RF(clampr)(spec, 0, 1); }

static R RF(g_correlated)(R alpha, R nol, R nov)
{
    const R a2 = alpha * alpha;
    const R gv = nol * SQRT(a2 + ((R)1 - a2) * nov * nov);
    const R gl = nov * SQRT(a2 + ((R)1 - a2) * nol * nol);
    const R den = gv + gl;
    return den > 0 ? ((R)2 * nol * nov) / den : 0;
}

static R RF(alpha_of)(const RF(mat)* m)
{
    const R r = RF(clampr)(m->roughness, (R)0.01, (R)1);
    return r * r;
}

/* QueryLambda through the scene's wavelength table (spectrum_lut.cuh:183-221) */
static void RF(lam_basis)(const orc_render_scene* sc, R l, float kl[4])
{
    float u = ((float)l - 360.0f) / (780.0f - 360.0f);
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    const int res = sc->lam_res;
    float g = u * (float)res - 0.5f;
    g = fminf(fmaxf(g, 0.0f), (float)res - 1.0f);
    const float f = floorf(g);
    const int i0 = (int)f, i1 = i0 + 1 >= res ? res - 1 : i0 + 1;
    const float t = i0 == i1 ? 0.0f : g - f;
    for (int k = 0; k < 4; ++k) kl[k] = sc->lam_tab[4 * i0 + k] * (1.0f - t) + sc->lam_tab[4 * i1 + k] * t;
}

/* reflectance of an uplift record at l: the sigmoid of the coefficients, or
 * SpectralInteract of the RGB basis with the wavelength basis (network) */
static R RF(uplift_at)(const orc_render_scene* sc, const float* c, R l)
{
    if (!sc->lam_tab) return RF(sigmoid)(c, l);
    float kl[4];
    RF(lam_basis)(sc, l, kl);
    return (R)orc_smlp_interact(c, kl);
}

static S4 RF(albedo_of)(const orc_render_scene* sc, const RF(mat)* m, const RF(lam)* L)
{
    if (L->mode == 0) return RF(s4)(m->albedo.x, m->albedo.y, m->albedo.z, 0);
    S4 r;
    for (int k = 0; k < 4; ++k)
        r.v[k] = (L->mode == 4 || k == 0) ? RF(uplift_at)(sc, m->ca, L->l[k]) * m->ascale : (R)0;
    return r;
}

static S4 RF(emission_of)(const orc_render_scene* sc, const RF(mat)* m, const RF(lam)* L, const float* d65)
{
    if (L->mode == 0) return RF(s4)(m->emission.x, m->emission.y, m->emission.z, 0);
    S4 r;
    for (int k = 0; k < 4; ++k)
        r.v[k] = (L->mode == 4 || k == 0) ? RF(uplift_at)(sc, m->ce, L->l[k]) * m->escale * RF(d65n)(d65, L->l[k])
                                          : (R)0;
    return r;
}

/* clamp(lum(F0), 0.1, 0.9) in RGB, clamp(F0(lambda_0), 0.1, 0.9) otherwise (material.cuh:257-264) */
static R RF(pspec_of)(const RF(mat)* m, S4 alb, int mode)
{
    const R f = RF(f0_ue)(m->specular);
    if (mode == 0) {
        const V3 f0 = RF(v3)(f + (m->albedo.x - f) * m->metallic, f + (m->albedo.y - f) * m->metallic,
                             f + (m->albedo.z - f) * m->metallic);
        return RF(clampr)(RF(lum)(f0), (R)0.1, (R)0.9);
    }
    return RF(clampr)(f + (alb.v[0] - f) * m->metallic, (R)0.1, (R)0.9);
}

/* EvaluateBSDFWithCos (material.cuh:47-165), rd toward the surface */
static S4 RF(eval_bsdf)(const RF(mat)* m, S4 alb, V3 rd, V3 l, V3 n)
{
    const R nol = RF(maxr)((R)0, RF(dot)(n, l));
    if (m->model == 0) return RF(s4scale)(alb, nol / PI_R);
    const V3 v = RF(normalize)(RF(scale)(rd, -1));
    const R nov = RF(maxr)((R)0, RF(dot)(n, v));
    if (nov <= 0 || nol <= 0) return RF(s4)(0, 0, 0, 0);
    const R alpha = RF(alpha_of)(m);
    const R f = RF(f0_ue)(m->specular);
    const R kd = (R)1 - m->metallic;
    const V3 h = RF(normalize)(RF(add)(v, l));
    const R noh = RF(maxr)((R)0, RF(dot)(n, h));
    const R loh = RF(maxr)((R)0, RF(dot)(l, h));
    const R c = RF(clampr)((R)1 - loh, 0, 1);
    const R c5 = c * c * c * c * c;
    const R D = RF(ggx_d)(alpha, noh);
    const R G = RF(g_correlated)(alpha, nol, nov);
    const R sden = RF(maxr)((R)4 * nov, (R)1e-7);
    const R dcos = nol / PI_R;
    S4 r;
    for (int k = 0; k < 4; ++k) {
        const R f0 = f + (alb.v[k] - f) * m->metallic;
        const R F = f0 + ((R)1 - f0) * c5;
        r.v[k] = F * D * G / sden + kd * ((R)1 - F) * alb.v[k] * dcos;
    }
    return r;
}

/* BSDF_PDF (material.cuh:316-371) */
static R RF(bsdf_pdf)(const RF(mat)* m, R pspec, V3 rd, V3 l, V3 n)
{
    const R nol = RF(maxr)((R)0, RF(dot)(n, l));
    if (m->model == 0) return nol > 0 ? nol * ((R)1 / PI_R) : 0;
    const V3 v = RF(scale)(rd, -1);
    const R nov = RF(maxr)((R)0, RF(dot)(n, v));
    if (nov <= 0 || nol <= 0) return 0;
    const R alpha = RF(alpha_of)(m);
    const R ps = RF(vndf_pdf)(v, n, l, alpha);
    const R pd = nol * ((R)1 / PI_R);
    return RF(maxr)((R)0, pspec * ps + ((R)1 - pspec) * pd);
}

/* SampleBSDF (material.cuh:167-314).  u[0] lobe, u[1..3] direction.
 * Returns 0 when the path terminates. */
static int RF(sample_bsdf)(const RF(mat)* m, R pspec, V3 rd, V3 n, const R* u, V3* next, R* pdf)
{
    if (m->model == 0) {
        *next = RF(sample_cosine)(u[1], u[2], u[3], n, pdf);
        return 1;
    }
    const V3 v = RF(scale)(rd, -1);
    const R nov = RF(maxr)((R)0, RF(dot)(n, v));
    if (nov <= 0) return 0;
    const R alpha = RF(alpha_of)(m);
    V3 l;
    R ps, pd;
    if (u[0] < pspec) {
        l = RF(sample_vndf)(v, n, alpha, u[1], u[2], u[3], &ps);
        const R nol = RF(maxr)((R)0, RF(dot)(n, l));
        if (nol <= 0) return 0;
        pd = nol / PI_R;
    } else {
        l = RF(sample_cosine)(u[1], u[2], u[3], n, &pd);
        ps = RF(vndf_pdf)(v, n, l, alpha);
    }
    *next = l;
    *pdf = pspec * ps + ((R)1 - pspec) * pd;
    return 1;
}

static R RF(power_heuristic)(R pa, R pb)
{
    const R a = pa * pa, b = pb * pb;
    return a / RF(maxr)((R)1e-24, a + b);
}

static R RF(intersect_eps)(int depth)
{
    return (R)1e-3 * EXP(RF(minr)((R)depth * (R)0.5, (R)4));
}

static V3 RF(sky)(V3 d)
{
    /* SkyLighting (core/render/render.cuh:58-73) */
    const V3 white = RF(v3)(1, 1, 1), blue = RF(v3)((R)0.05, (R)0.45, 1);
    const V3 down = RF(scale)(RF(v3)((R)0.2, (R)0.6, 1), (R)0.125);
    const R t1 = POW(RF(maxr)(d.y, 0), (R)0.75);
    const V3 a = RF(add)(blue, RF(scale)(RF(sub)(white, blue), t1));
    const R t2 = RF(minr)(RF(maxr)(-d.y * (R)0.5 + (R)0.5, 0), 1);
    const V3 e = RF(add)(a, RF(scale)(RF(sub)(down, a), t2));
    return RF(scale)(e, (R)0.5);
}

/* ---------------------------------------------------------------------- */
/* Scene view, traversal in R, lights                                      */
/* ---------------------------------------------------------------------- */

/* closest hit in R: prim (ORC_MISS on a miss), t, barycentric weights of v1,
 * v2, and for an object whether the ray started inside it */
static uint32_t RF(closest)(const orc_render_scene* sc, V3 o, V3 d, R tmin, R tmax, R* t, R* bu, R* bv,
                            int* inside);
static int RF(occluded)(const orc_render_scene* sc, V3 o, V3 d, R tmin, R tmax, R meps);

static void RF(tri)(const orc_render_scene* sc, uint32_t p, V3* a, V3* b, V3* c)
{
    *a = RF(load3)(&sc->verts[3 * sc->idx[3 * p + 0]]);
    *b = RF(load3)(&sc->verts[3 * sc->idx[3 * p + 1]]);
    *c = RF(load3)(&sc->verts[3 * sc->idx[3 * p + 2]]);
}

static RF(mat) RF(material)(const orc_render_scene* sc, uint32_t p)
{
    const int is_obj = p >= sc->ntri;
    const uint32_t mi = is_obj ? sc->nmats + (p - sc->ntri) : (sc->tri_mat ? sc->tri_mat[p] : 0);
    const orc_material* om = is_obj ? &sc->obj_mats[p - sc->ntri] : &sc->mats[mi];
    RF(mat) m;
    m.albedo = RF(load3)(om->albedo);
    m.emission = RF(load3)(om->emission);
    m.specular = om->specular; m.metallic = om->metallic; m.roughness = om->roughness;
    m.model = om->model;
    m.ior = om->ior; m.sigma_a = om->sigma_a; m.abbe = om->abbe; m.tilt = om->tilt;
    m.escale = 1;
    m.ascale = 1;
    for (int k = 0; k < 4; ++k) { m.ca[k] = 0.0f; m.ce[k] = 0.0f; }
    if (sc->mat_coef) {
        const float* mc = &sc->mat_coef[12 * mi];
        for (int k = 0; k < 4; ++k) { m.ca[k] = mc[k]; m.ce[k] = mc[4 + k]; }
        m.escale = (R)mc[8];
        m.ascale = (R)mc[9];
    }
    return m;
}

/* Light pick by the CDF walk of NEESampleLight (core/nee/nee.cuh:329-377);
 * cdf / prob are the FP32 tables built in orc_render_lights. */
static int RF(pick_light)(const orc_render_lights* L, R u)
{
    for (int i = 0; i < L->count; ++i)
        if (u < (R)L->cdf[i]) return i;
    return L->count - 1;
}

/* ---------------------------------------------------------------------- */
/* Analytic objects in R (the FP32 instance rounds exactly like trace.c's   */
/* orc_object_test / orc_object_surface and the GPU kernels)               */
/* ---------------------------------------------------------------------- */

static V3 RF(w2o_row)(const orc_object* ob, int r) { return RF(load3)(&ob->w2o[3 * r]); }

static void RF(obj_local)(const orc_object* ob, V3 o, V3 d, V3* lo, V3* ld)
{
    const V3 oc = RF(sub)(o, RF(load3)(ob->center));
    const V3 r0 = RF(w2o_row)(ob, 0), r1 = RF(w2o_row)(ob, 1), r2 = RF(w2o_row)(ob, 2);
    *lo = RF(v3)(RF(dot)(r0, oc), RF(dot)(r1, oc), RF(dot)(r2, oc));
    *ld = RF(normalize)(RF(v3)(RF(dot)(r0, d), RF(dot)(r1, d), RF(dot)(r2, d)));
}

static R RF(inv_safe_ref)(R x) { return x == 0 ? (x >= 0 ? (R)FLT_MAX : -(R)FLT_MAX) : (R)1 / x; }

static R RF(obj_radius)(const orc_object* ob)
{
    return (R)fmaxf(fmaxf(ob->extent[0], ob->extent[1]), ob->extent[2]);
}

/* SDF objects in R (oracle/sdf_impl.h) */
#define SR(name) RF(name)
#include "sdf_impl.h"
#undef SR

static int RF(is_sdf)(const orc_object* ob)
{
    return ob->type == ORC_OBJ_SDF_TORUS || ob->type == ORC_OBJ_SDF_VOLUME;
}

/* IntersectAABBObject / IntersectSphereObject / IntersectSDFObject (aabb.cuh:28-89,
 * sphere.cuh:11-68, sdf_base.cuh:118-167); meps = the SDF march tolerance */
static int RF(obj_test)(const orc_object* ob, V3 o, V3 d, R meps, R* t, int* inside)
{
    V3 lo, ld;
    RF(obj_local)(ob, o, d, &lo, &ld);
    *inside = 0;
    if (ob->type != ORC_OBJ_SPHERE) {
        const R lox[3] = { lo.x, lo.y, lo.z }, ldx[3] = { ld.x, ld.y, ld.z };
        R te = 0, tx = 0;
        for (int a = 0; a < 3; ++a) {
            const R inv = RF(inv_safe_ref)(ldx[a]);
            const R t1 = (-(R)ob->extent[a] - lox[a]) * inv;
            const R t2 = ((R)ob->extent[a] - lox[a]) * inv;
            const R mn = t1 < t2 ? t1 : t2, mx = t1 > t2 ? t1 : t2;
            te = a == 0 ? mn : (te > mn ? te : mn);
            tx = a == 0 ? mx : (tx < mx ? tx : mx);
        }
        if (te > tx) return 0;
        if (tx <= 0) return 0;
        if (RF(is_sdf)(ob)) {
            const R l3[3] = { lo.x, lo.y, lo.z }, d3[3] = { ld.x, ld.y, ld.z };
            return RF(sdf_march)(ob, l3, d3, te, tx, meps, t, inside);
        }
        *inside = te < 0;
        *t = *inside ? tx : te;
        return 1;
    }
    const R r = RF(obj_radius)(ob);
    const R b = RF(dot)(lo, ld);
    const R c = RF(dot)(lo, lo) - r * r;
    const R disc = b * b - c;
    if (disc < 0) return 0;
    const R sq = SQRT(disc);
    R tt = -b - sq;
    if (tt <= 0) { tt = -b + sq; *inside = 1; }
    if (tt <= 0) return 0;
    *t = tt;
    return 1;
}

/* hit point and world normal (box and sphere: flipped for a hit from inside;
 * SDF: the outward distance gradient) */
static void RF(obj_surface)(const orc_object* ob, V3 o, V3 d, R t, int inside, R meps, V3* pos, V3* n)
{
    V3 lo, ld;
    RF(obj_local)(ob, o, d, &lo, &ld);
    const V3 lp = RF(add)(lo, RF(scale)(ld, t));
    V3 ln = RF(v3)(0, 0, 0);
    if (RF(is_sdf)(ob)) {
        const R l3[3] = { lp.x, lp.y, lp.z };
        R n3[3];
        RF(sdf_normal)(ob, l3, meps, n3);
        ln = RF(v3)(n3[0], n3[1], n3[2]);
    } else if (ob->type == ORC_OBJ_BOX) {
        const R rx = FABS(lp.x) / (R)ob->extent[0], ry = FABS(lp.y) / (R)ob->extent[1],
                rz = FABS(lp.z) / (R)ob->extent[2];
        if (rx >= ry && rx >= rz) ln.x = lp.x > 0 ? (R)1 : (R)-1;
        else if (ry >= rz) ln.y = lp.y > 0 ? (R)1 : (R)-1;
        else ln.z = lp.z > 0 ? (R)1 : (R)-1;
        if (inside) ln = RF(scale)(ln, -1);
    } else {
        ln = RF(scale)(RF(normalize)(lp), inside ? (R)-1 : (R)1);
    }
    const V3 c0 = RF(v3)((R)ob->w2o[0], (R)ob->w2o[3], (R)ob->w2o[6]);
    const V3 c1 = RF(v3)((R)ob->w2o[1], (R)ob->w2o[4], (R)ob->w2o[7]);
    const V3 c2 = RF(v3)((R)ob->w2o[2], (R)ob->w2o[5], (R)ob->w2o[8]);
    *pos = RF(v3)(RF(dot)(c0, lp) + (R)ob->center[0], RF(dot)(c1, lp) + (R)ob->center[1],
                  RF(dot)(c2, lp) + (R)ob->center[2]);
    *n = RF(normalize)(RF(v3)(RF(dot)(c0, ln), RF(dot)(c1, ln), RF(dot)(c2, ln)));
}

/* OverlapAABBObject / OverlapSphereObject (aabb.cuh:91-105, sphere.cuh:70-84) */
static int RF(obj_overlap)(const orc_object* ob, V3 p)
{
    const V3 oc = RF(sub)(p, RF(load3)(ob->center));
    const V3 lo = RF(v3)(RF(dot)(RF(w2o_row)(ob, 0), oc), RF(dot)(RF(w2o_row)(ob, 1), oc),
                         RF(dot)(RF(w2o_row)(ob, 2), oc));
    if (RF(is_sdf)(ob)) {
        const R l3[3] = { lo.x, lo.y, lo.z };
        R u3[3];
        RF(sdf_to_unit)(ob, l3, u3);
        return RF(sdf_unit)(ob, u3[0], u3[1], u3[2]) <= 0;
    }
    if (ob->type == ORC_OBJ_BOX) {
        const R ex = ob->extent[0], ey = ob->extent[1], ez = ob->extent[2];
        return -ex < lo.x && -ey < lo.y && -ez < lo.z && ex > lo.x && ey > lo.y && ez > lo.z;
    }
    const R r = RF(obj_radius)(ob);
    return RF(dot)(lo, lo) < r * r;
}

/* The six facets of a box light: outward normal = +-normalised
 * worldToObject row (GetOBBAxes, nee.cuh:24-31), centre = c +- u e, tangent
 * and bitangent axes, half sides s1, s2 (nee.cuh:164-183). */
static void RF(box_facets)(const orc_object* ob, V3 nrm[6], V3 cen[6], V3 tan[6], V3 bit[6], R s1[6], R s2[6])
{
    const V3 ux = RF(normalize)(RF(w2o_row)(ob, 0)), uy = RF(normalize)(RF(w2o_row)(ob, 1)),
             uz = RF(normalize)(RF(w2o_row)(ob, 2));
    const V3 c = RF(load3)(ob->center);
    const R ex = ob->extent[0], ey = ob->extent[1], ez = ob->extent[2];
    const V3 ax[3] = { ux, uy, uz };
    const R e[3] = { ex, ey, ez };
    for (int f = 0; f < 6; ++f) {
        const int a = f >> 1;
        const R sg = (f & 1) ? (R)-1 : (R)1;
        nrm[f] = RF(scale)(ax[a], sg);
        cen[f] = (f & 1) ? RF(sub)(c, RF(scale)(ax[a], e[a])) : RF(add)(c, RF(scale)(ax[a], e[a]));
    }
    for (int f = 0; f < 2; ++f) { tan[f] = uy; bit[f] = uz; s1[f] = ey; s2[f] = ez; }
    for (int f = 2; f < 4; ++f) { tan[f] = ux; bit[f] = uz; s1[f] = ex; s2[f] = ez; }
    for (int f = 4; f < 6; ++f) { tan[f] = ux; bit[f] = uy; s1[f] = ex; s2[f] = ey; }
}

/* NEEFacingWeight_Box (nee.cuh:55-61) */
static R RF(facing_weight)(V3 n, V3 c, V3 x, R area)
{
    const V3 l2s = RF(normalize)(RF(sub)(x, c));
    return area * RF(maxr)((R)0, RF(dot)(n, l2s));
}

/* AreaToSolidAnglePDF (nee.cuh:35-46) */
static R RF(area_to_solid)(V3 x, V3 y, V3 ny, R pa_total)
{
    const V3 s2l = RF(sub)(y, x);
    const R dist2 = RF(maxr)(RF(dot)(s2l, s2l), (R)1e-20);
    const V3 dir = RF(scale)(s2l, (R)1 / SQRT(dist2));
    const R cosy = RF(maxr)((R)0, RF(dot)(ny, RF(scale)(dir, -1)));
    if (cosy <= 0) return 0;
    return pa_total * (dist2 / cosy);
}

/* NEEPDFOmegaFromHit_Box / _Sphere (nee.cuh:63-151, 253-271): solid-angle
 * density with which NEE from x would have sampled light point y (normal ny)
 * of the object, select probability included */
static R RF(obj_light_pdf)(const orc_object* ob, R select, V3 x, V3 y, V3 ny)
{
    if (ob->type == ORC_OBJ_BOX) {
        V3 nrm[6], cen[6], tan[6], bit[6];
        R s1[6], s2[6];
        RF(box_facets)(ob, nrm, cen, tan, bit, s1, s2);
        int hit = -1;
        for (int f = 0; f < 6; ++f)
            if (RF(dot)(nrm[f], ny) > (R)0.999) { hit = f; break; }
        if (hit < 0) {
            R best = (R)1e30;
            for (int f = 0; f < 6; ++f) {
                const R dd = FABS(RF(dot)(nrm[f], RF(sub)(y, cen[f])));
                if (dd < best) { best = dd; hit = f; }
            }
        }
        R sum = 0, whit = 0, ahit = 0;
        for (int f = 0; f < 6; ++f) {
            const R area = (R)4 * s1[f] * s2[f];
            const R w = RF(facing_weight)(nrm[f], cen[f], x, area);
            sum += w;
            if (f == hit) { whit = w; ahit = area; }
        }
        if (sum <= 0 || whit <= 0) return 0;
        const R pa = select * ((R)1 / ahit) * (whit / sum);
        return RF(area_to_solid)(x, y, ny, pa);
    }
    const V3 c = RF(load3)(ob->center);
    const R r = RF(obj_radius)(ob);
    const R dist = RF(length)(RF(sub)(x, c));
    if (dist <= r) return 0;
    const R cosc = RF(clampr)(r / dist, 0, 1);
    const R cap = (R)2 * PI_R * r * r * ((R)1 - cosc);
    if (cap <= (R)1e-20) return 0;
    return RF(area_to_solid)(x, y, ny, select / cap);
}

/* BuildTangentBasis (core/vector.cuh:277-282) */
static void RF(tangent_basis)(V3 n, V3* t, V3* b)
{
    const V3 a = (FABS(n.z) < (R)0.999) ? RF(v3)(0, 0, 1) : RF(v3)(1, 0, 0);
    *t = RF(normalize)(RF(cross)(a, n));
    *b = RF(cross)(n, *t);
}

/* NEESampleBoxAreaFacing / NEESampleSphereCapArea (nee.cuh:153-251,
 * 273-310): a light point y with normal ny seen from x and its area density
 * pa (facet or cap, without the select probability).  ub: facet pick (box),
 * u0, u1: the point.  Returns 0 when nothing is sampleable from x. */
static int RF(obj_light_sample)(const orc_object* ob, V3 x, R ub, R u0, R u1, V3* y, V3* ny, R* pa)
{
    if (ob->type == ORC_OBJ_BOX) {
        V3 nrm[6], cen[6], tan[6], bit[6];
        R s1[6], s2[6];
        RF(box_facets)(ob, nrm, cen, tan, bit, s1, s2);
        int idx[6], cnt = 0;
        R w[6], sum = 0;
        for (int f = 0; f < 6; ++f) {
            const R area = (R)4 * s1[f] * s2[f];
            const R wf = RF(facing_weight)(nrm[f], cen[f], x, area);
            if (wf > 0) { idx[cnt] = f; w[cnt] = wf; sum += wf; ++cnt; }
        }
        if (cnt == 0) return 0;
        int slot = 0;
        const R r = ub * sum;
        R acc = 0;
        for (int k = 0; k < cnt; ++k) {
            acc += w[k];
            if (r <= acc) { slot = k; break; }
        }
        const int f = idx[slot];
        const R a1 = (R)2 * u0 - (R)1, a2 = (R)2 * u1 - (R)1;
        *y = RF(add)(RF(add)(cen[f], RF(scale)(tan[f], a1 * s1[f])), RF(scale)(bit[f], a2 * s2[f]));
        *ny = nrm[f];
        *pa = (w[slot] / sum) * ((R)1 / ((R)4 * s1[f] * s2[f]));
        return 1;
    }
    const V3 c = RF(load3)(ob->center);
    const R r = RF(obj_radius)(ob);
    const V3 pc = RF(sub)(x, c);
    const R dist = RF(length)(pc);
    if (dist <= r) return 0;
    const V3 axis = RF(scale)(pc, (R)1 / dist);
    const R cosc = RF(clampr)(r / dist, 0, 1);
    const R cap = (R)2 * PI_R * r * r * ((R)1 - cosc);
    const R cost = cosc * ((R)1 - u0) + u0;
    const R sint = SQRT(RF(maxr)((R)0, (R)1 - cost * cost));
    const R phi = (R)2 * PI_R * u1;
    V3 t, b;
    RF(tangent_basis)(axis, &t, &b);
    const V3 nn = RF(add)(RF(add)(RF(scale)(t, COS(phi) * sint), RF(scale)(b, SIN(phi) * sint)),
                          RF(scale)(axis, cost));
    *y = RF(add)(c, RF(scale)(nn, r));
    *ny = nn;
    *pa = (R)1 / cap;
    return 1;
}

/* ---------------------------------------------------------------------- */
/* Glass (core/render/material.cuh:116-161, 206-257; brdf.cuh:15-33;        */
/* core/spectrum/spectrum.cuh:16-54)                                        */
/* ---------------------------------------------------------------------- */

/* IORFromAbbeNM: n(lambda) = n_d + K (1/lambda - 1/lambda_d),
 * K = (n_d - 1) / (V_d (1/lambda_F - 1/lambda_C)); K = 0 for V_d <= 0 */
static R RF(ior_at)(R nd, R vd, R lambda)
{
    if (!(lambda > 0)) return nd;
    const R invf = (R)1 / (R)486.1327f, invc = (R)1 / (R)656.2725f, invd = (R)1 / (R)587.5618f;
    const R k = (vd > 0) ? (nd - (R)1) / (vd * (invf - invc)) : (R)0;
    return nd + k * ((R)1 / lambda - invd);
}

/* SigmaAFromTiltNM: sigma_a (1 + tilt (1/lambda - 1/lambda_d) / (1/lambda_F - 1/lambda_C)), >= 0 */
static R RF(sigma_at)(R sa, R lambda, R tilt)
{
    const R invf = (R)1 / (R)486.1327f, invc = (R)1 / (R)656.2725f, invd = (R)1 / (R)587.5618f;
    const R u = (R)1 / RF(maxr)(lambda, (R)1e-6);
    const R du = (u - invd) / (invf - invc);
    return RF(maxr)((R)0, sa * ((R)1 + tilt * du));
}

/* FresnelDielectricExact (brdf.cuh:25-31), eta = n_in / n_out */
static R RF(fresnel)(R cosi, R cost, R eta)
{
    const R inve = (R)1 / eta;
    const R rpar = (inve * cosi - cost) / (inve * cosi + cost);
    const R rperp = (cosi - inve * cost) / (cosi + inve * cost);
    return (R)0.5 * (rpar * rpar + rperp * rperp);
}

/* refract (core/vector.cuh:264-272) */
static V3 RF(refract)(V3 in, V3 normal, R eta)
{
    const V3 n = RF(dot)(normal, in) > 0 ? RF(scale)(normal, -1) : normal;
    const R cosi = -RF(dot)(n, in);
    const R k = (R)1 - (eta * eta) * RF(maxr)((R)0, (R)1 - cosi * cosi);
    const R cost = SQRT(RF(maxr)((R)0, k));
    return RF(add)(RF(scale)(in, eta), RF(scale)(n, eta * cosi - cost));
}

/* ---------------------------------------------------------------------- */
/* One path (CalculateRadiance, core/render/render.cuh:460-998)            */
/* ---------------------------------------------------------------------- */

/* guiding: slot radiance over the hero/one wavelength(s) as XYZ, or RGB */
static V3 RF(slot_xyz)(const orc_render_opts* op, const RF(lam)* LM, S4 r)
{
    if (LM->mode == 0) return RF(v3)(r.v[0], r.v[1], r.v[2]);
    if (LM->mode == 1) {
        const V3 c = RF(cie)(op->cie, LM->l[0]);
        const R pdf = RF(lambda_pdf)(LM->l[0]);
        return RF(v3)(c.x * r.v[0] / pdf, c.y * r.v[0] / pdf, c.z * r.v[0] / pdf);
    }
    R X = 0, Y = 0, Z = 0;
    for (int k = 0; k < 4; ++k) {
        const V3 c = RF(cie)(op->cie, LM->l[k]);
        X = X + c.x * r.v[k]; Y = Y + c.y * r.v[k]; Z = Z + c.z * r.v[k];
    }
    const R psum = (R)4 / (R)400;
    return RF(v3)(X / psum, Y / psum, Z / psum);
}

/* maskForProbeQuery (render.cuh:893-900): max(w, 1e-6) / (sum + 1e-6), with
 * w = throughput (RGB) or throughput times xbar..zbar of the wavelength(s) */
static void RF(probe_query)(const orc_render_opts* op, const RF(lam)* LM, S4 mask, float q[3])
{
    R w[3];
    if (LM->mode == 0) { w[0] = mask.v[0]; w[1] = mask.v[1]; w[2] = mask.v[2]; }
    else if (LM->mode == 1) {
        const V3 c = RF(cie)(op->cie, LM->l[0]);
        w[0] = mask.v[0] * c.x; w[1] = mask.v[0] * c.y; w[2] = mask.v[0] * c.z;
    } else {
        w[0] = w[1] = w[2] = 0;
        for (int k = 0; k < 4; ++k) {
            const V3 c = RF(cie)(op->cie, LM->l[k]);
            w[0] = w[0] + mask.v[k] * c.x; w[1] = w[1] + mask.v[k] * c.y; w[2] = w[2] + mask.v[k] * c.z;
        }
    }
    for (int k = 0; k < 3; ++k) q[k] = (float)RF(maxr)(w[k], (R)1e-6f);
    const float sum = (float)((double)(q[0] + q[1] + q[2]) + 1e-6);
    for (int k = 0; k < 3; ++k) q[k] = q[k] / sum;
}

static inline V3 RF(reflectv)(V3 v, V3 n) { return RF(sub)(v, RF(scale)(n, (R)2 * RF(dot)(v, n))); }

typedef struct { uint64_t key; float dir[3], rad[3]; } RF(gsample);

/* CollectRadiance with an endpoint emission: every open slot started before
 * this vertex receives throughput * Le (render.cuh:364-376) */
static void RF(g_endpoint)(S4* gm, S4* gr, int* gval, S4 le, int non_delta, int mode)
{
    if (RF(s4zero)(le, mode)) return;
    const int hi = non_delta < 5 ? non_delta : 5;
    for (int i = 0; i < hi; ++i)
        if (!RF(s4zero)(gm[i], mode)) { gr[i] = RF(s4add)(gr[i], RF(s4mul)(gm[i], le)); gval[i] = 1; }
}

/* closest hit of the path's next ray: the whole scene, or only the medium
 * object while inside one (Trace's solo trace, render.cuh:89-92) */
static uint32_t RF(next_hit)(const orc_render_scene* sc, int medium, V3 o, V3 d, R tmin, R* t, R* bu, R* bv,
                             int* inside)
{
    *inside = 0;
    *bu = 0; *bv = 0;
    if (medium >= 0) {
        R tt;
        int ins;
        if (RF(obj_test)(&sc->objs[medium], o, d, tmin * (R)0.5, &tt, &ins) && tt >= tmin) {
            *t = tt;
            *inside = ins;
            return sc->ntri + (uint32_t)medium;
        }
        return ORC_MISS;
    }
    return RF(closest)(sc, o, d, tmin, (R)INFINITY, t, bu, bv, inside);
}

/* The first camera ray starts inside the first object containing the camera
 * (Overlap, intersection.cuh:312-327; CalculateRadiance:536-538). */
static int RF(camera_medium)(const orc_render_scene* sc, V3 o)
{
    for (uint32_t k = 0; k < sc->nobj; ++k)
        if (RF(obj_overlap)(&sc->objs[k], o)) return (int)k;
    return -1;
}

/* fh (optional, 11 floats): the first vertex for the denoiser's G-buffer
 * (RecordScreenGBuffer at depth 0, render.cuh:640-643): hit, albedo[3],
 * normal[3] (the intersector's for objects, the one facing the ray for
 * triangles), distance, metallic, roughness, shading model (0 diffuse,
 * 1 light, 2 glass, 3 GGX, 4 none) */
static V3 RF(trace_path)(const orc_render_scene* sc, const orc_render_lights* L, const orc_render_opts* op,
                         uint32_t pix, uint32_t smp, V3 o, V3 d, R lambda0,
                         const orc_guide* G, float alpha, RF(gsample)* gout, int* ngout, float* fh)
{
    if (fh) { for (int k = 0; k < 11; ++k) fh[k] = 0.0f; fh[10] = 4.0f; }
    const int mode = op->spectral;
    const RF(lam) LM = RF(make_lam)(mode, lambda0);
    S4 rad = RF(s4)(0, 0, 0, 0), mask = RF(s4)(1, 1, 1, 1);
    int non_delta = 0;
    int prev_nee = 0;
    R last_pdf = 0;
    V3 last_pos = RF(v3)(0, 0, 0);
    /* media state (MediaContext, material.cuh:12-18): the object the path is
     * inside (-1 none) and the vertex where it entered; hero wavelengths
     * dropped at a dispersive interface */
    int medium = RF(camera_medium)(sc, o);
    V3 mpos = o;
    int dropped = 0;
    /* guiding slots: throughput since the slot's vertex, collected radiance,
     * sampled direction and probe key (CollectRadiance, render.cuh:345-410) */
    S4 gm[5], gr[5];
    V3 gdir[5];
    uint64_t gkey[5];
    int gval[5];
    for (int i = 0; i < 5; ++i) {
        gm[i] = RF(s4)(0, 0, 0, 0); gr[i] = gm[i]; gdir[i] = RF(v3)(0, 0, 0); gkey[i] = ORC_NO_KEY; gval[i] = 0;
    }
    float gq[3] = { 1.0f, 1.0f, 1.0f };
    const int guided = G != NULL && orc_guide_count(G) > 0;
    if (ngout) *ngout = 0;
    d = RF(normalize)(d);
    for (int depth = 0; depth < op->max_depth; ++depth) {
        float uf[16];
        orc_philox_uniforms(op->seed, pix, smp, (uint32_t)depth, guided ? 4 : 3, uf);
        R u[12];
        for (int k = 0; k < 12; ++k) u[k] = (R)uf[k];
        const R eps = RF(intersect_eps)(non_delta);
        R t, bu, bv;
        int inside = 0;
        const uint32_t prim = RF(next_hit)(sc, medium, o, d, eps, &t, &bu, &bv, &inside);
        if (prim == ORC_MISS && medium >= 0) {
            /* inside a medium with no exit found (bMissThinSurfaceInsideMedia_,
             * render.cuh:107-124): leave the medium and continue straight from
             * the same origin as a delta vertex of weight 1 */
            medium = -1;
            R qrr = 1;
            if (non_delta >= op->rr_min_depth) {
                qrr = RF(rr_q)(mask, mode, 0);
                if (u[9] > qrr) break;
            }
            if (gout) {
                const R invq = (R)(1.0 / (double)qrr);
                const int hi = non_delta + 1 < 5 ? non_delta + 1 : 5;
                for (int i = 0; i < hi; ++i)
                    if (!RF(s4zero)(gm[i], mode)) gm[i] = RF(s4scale)(gm[i], invq);
            }
            mask = RF(s4scale)(mask, (R)1 / qrr);
            RF(probe_query)(op, &LM, mask, gq);
            continue;
        }
        if (prim == ORC_MISS) {
            S4 e = RF(s4)(0, 0, 0, 0);
            if (op->sky == 1) {
                const V3 sk = RF(sky)(d);
                e = RF(s4)(sk.x, sk.y, sk.z, 0);
                rad = RF(s4add)(rad, RF(s4mul)(mask, e));
            } else if (op->sky == 2) {
                if (mode == 0) e = RF(s4)((R)op->env[0], (R)op->env[1], (R)op->env[2], 0);
                else {
                    RF(mat) em;
                    em.escale = (R)sc->env_coef[4];
                    for (int k = 0; k < 4; ++k) em.ce[k] = sc->env_coef[k];
                    e = RF(emission_of)(sc, &em, &LM, op->d65);
                }
                rad = RF(s4add)(rad, RF(s4mul)(mask, e));
            }
            RF(g_endpoint)(gm, gr, gval, e, non_delta, mode);
            break;
        }
        const int is_obj = prim >= sc->ntri;
        const orc_object* ob = is_obj ? &sc->objs[prim - sc->ntri] : NULL;
        V3 p, ng, n;
        R ndd;
        if (is_obj) {
            /* object hits shade with the intersector's normal: outward, or
             * inward for a hit from inside (aabb.cuh:60-73, sphere.cuh:52-53) */
            RF(obj_surface)(ob, o, d, t, inside, eps * (R)0.5, &p, &ng);
            ndd = RF(dot)(ng, d);
            n = ng;
        } else {
            V3 a, b, c;
            RF(tri)(sc, prim, &a, &b, &c);
            const R bw = (R)1 - bu - bv;
            p = RF(add)(RF(add)(RF(scale)(a, bw), RF(scale)(b, bu)), RF(scale)(c, bv));
            ng = RF(normalize)(RF(cross)(RF(sub)(b, a), RF(sub)(c, a)));
            ndd = RF(dot)(ng, d);
            n = ndd < 0 ? ng : RF(scale)(ng, -1);
        }
        const RF(mat) m = RF(material)(sc, prim);
        const int is_light = sc->light_index[prim] >= 0;
        if (fh && depth == 0) {
            fh[0] = 1.0f;
            fh[1] = (float)m.albedo.x; fh[2] = (float)m.albedo.y; fh[3] = (float)m.albedo.z;
            fh[4] = (float)n.x; fh[5] = (float)n.y; fh[6] = (float)n.z;
            fh[7] = (float)t;
            fh[8] = (float)m.metallic;
            fh[9] = (float)m.roughness;
            fh[10] = is_light ? 1.0f : (m.model == 2 ? 2.0f : (m.model == 1 ? 3.0f : 0.0f));
        }
        if (is_light) {
            /* triangle lights emit from the front face; box and sphere lights
             * from every hit (MAT_LIGHT objects, render.cuh:664-710) */
            if (is_obj || ndd < 0) {
                R w = 1;
                if (op->nee && depth > 0 && prev_nee) {
                    const int li = sc->light_index[prim];
                    R pl;
                    if (is_obj) {
                        pl = RF(obj_light_pdf)(ob, (R)L->prob[li], last_pos, p, ng);
                    } else {
                        const V3 tl = RF(sub)(p, last_pos);
                        const R dist2 = RF(maxr)(RF(dot)(tl, tl), (R)1e-20);
                        const V3 dir = RF(scale)(tl, (R)1 / SQRT(dist2));
                        const R cosy = RF(maxr)((R)0, -RF(dot)(ng, dir));
                        const R pa = (R)L->prob[li] / (R)L->area[li];
                        pl = cosy <= 0 ? (R)0 : pa * (dist2 / cosy);
                    }
                    if (pl >= 0 && last_pdf > 0) w = RF(power_heuristic)(last_pdf, pl);
                }
                const S4 le = RF(emission_of)(sc, &m, &LM, op->d65);
                rad = RF(s4add)(rad, RF(s4scale)(RF(s4mul)(mask, le), w));
                RF(g_endpoint)(gm, gr, gval, RF(s4scale)(le, w), non_delta, mode);
            }
            break;
        }
        if (m.model == 2) {
            /* glass: Fresnel reflection or refraction (SampleBSDF MAT_GLASS,
             * material.cuh:206-257) weighted by EvaluateBSDFWithCos
             * (material.cuh:116-161); Beer-Lambert over the path inside */
            const int ins = medium >= 0;
            R ior = m.ior;
            if (mode != 0) ior = RF(ior_at)(m.ior, m.abbe, LM.l[0]);
            if (mode == 4 && m.abbe > 0 && !dropped) {
                /* dispersion: the hero wavelength continues alone with the
                 * weight of all four (the secondary wavelengths' estimates
                 * end here) */
                mask = RF(s4)(mask.v[0] * (R)4, 0, 0, 0);
                dropped = 1;
            }
            S4 tr = RF(s4)(1, 1, 1, 1);
            if (ins) {
                const R dist = RF(length)(RF(sub)(mpos, p));
                const S4 alb = RF(albedo_of)(sc, &m, &LM);
                for (int k = 0; k < 4; ++k) {
                    const R sa = mode == 0 ? m.sigma_a : RF(sigma_at)(m.sigma_a, LM.l[k], m.tilt);
                    const R a = RF(minr)(RF(maxr)(alb.v[k], (R)1e-12f), (R)1);
                    tr.v[k] = EXP(dist * sa * LOG(a));
                }
            }
            const V3 nin = RF(dot)(ng, d) > 0 ? RF(scale)(ng, -1) : ng;
            const R eta = ins ? ior : (R)1 / ior;
            const R cosi = -RF(dot)(nin, d);
            const R sin2i = RF(maxr)((R)0, (R)1 - cosi * cosi);
            const R k = (R)1 - sin2i * (eta * eta);
            V3 next;
            R pdf;
            S4 fcos;
            int next_medium = medium;
            if (k <= 0) {
                next = RF(reflect)(d, nin);
                pdf = 1;
                fcos = RF(s4scale)(tr, (R)0.99f);
            } else {
                const R cost = SQRT(k);
                const R fr = RF(fresnel)(cosi, cost, eta);
                const V3 rdir = RF(reflect)(d, nin);
                const V3 tdir = RF(refract)(d, nin, eta);
                if (u[0] < fr) {
                    next = rdir;
                    pdf = fr;
                } else {
                    next = tdir;
                    pdf = (R)1 - fr;
                    next_medium = ins ? -1 : (int)(prim - sc->ntri);
                }
                if (RF(dot)(next, rdir) >= RF(dot)(next, tdir)) {
                    fcos = RF(s4scale)(tr, fr);
                } else {
                    const R wgt = (eta * eta) * (FABS(cost) / RF(maxr)((R)1e-6f, cosi));
                    fcos = RF(s4scale)(tr, ((R)1 - fr) * wgt);
                }
            }
            mpos = p;
            medium = next_medium;
            const R off = eps * RF(maxr)((R)1, RF(length)(p)) * (u[4] + (R)0.5);
            const R sgn = RF(dot)(ng, next) > 0 ? (R)1 : (RF(dot)(ng, next) < 0 ? (R)-1 : (R)0);
            p = RF(add)(p, RF(scale)(ng, sgn * off));
            prev_nee = 0;
            last_pdf = pdf;
            last_pos = p;
            const S4 contrib = RF(s4scale)(fcos, (R)1 / RF(maxr)(pdf, (R)1e-12));
            R qrr = 1;
            if (non_delta >= op->rr_min_depth) {
                qrr = RF(rr_q)(mask, mode, medium >= 0);
                if (u[9] > qrr) break;
            }
            if (gout) {
                const R invq = (R)(1.0 / (double)qrr);
                const int hi = non_delta + 1 < 5 ? non_delta + 1 : 5;
                for (int i = 0; i < hi; ++i)
                    if (!RF(s4zero)(gm[i], mode)) gm[i] = RF(s4mul)(gm[i], RF(s4scale)(contrib, invq));
            }
            mask = RF(s4scale)(RF(s4mul)(mask, contrib), (R)1 / qrr);
            RF(probe_query)(op, &LM, mask, gq);
            o = p;
            d = next;
            continue;
        }
        const S4 alb = RF(albedo_of)(sc, &m, &LM);
        const R pspec = m.model == 0 ? (R)0 : RF(pspec_of)(&m, alb, mode);
        V3 next;
        R pdf;
        if (!RF(sample_bsdf)(&m, pspec, d, n, u, &next, &pdf)) break;
        S4 fcos = RF(eval_bsdf)(&m, alb, d, next, n);
        mpos = p;
        /* offset along the side of the next direction (render.cuh:155-160) */
        const R off = eps * RF(maxr)((R)1, RF(length)(p)) * (u[4] + (R)0.5);
        const R sgn = RF(dot)(n, next) > 0 ? (R)1 : (RF(dot)(n, next) < 0 ? (R)-1 : (R)0);
        p = RF(add)(p, RF(scale)(n, sgn * off));
        /* guide/BSDF one-sample mixture (render.cuh:553-638) */
        const float* ptex = NULL;
        const V3 rvec = RF(reflectv)(d, n);
        const float nf[3] = { (float)n.x, (float)n.y, (float)n.z };
        if (guided) {
            const float pf[3] = { (float)p.x, (float)p.y, (float)p.z };
            const float rf[3] = { (float)rvec.x, (float)rvec.y, (float)rvec.z };
            const int pi = orc_guide_find(G, orc_voxel_key(pf, rf));
            if (pi >= 0) {
                ptex = orc_guide_texels(G, pi);
                const float xf[3] = { (float)next.x, (float)next.y, (float)next.z };
                const R pb = pdf;
                const R pp = (R)orc_probe_hemi_pdf(ptex, xf, nf, gq);
                if (pb > (R)1e-12f && pp > (R)1e-12f) {
                    if (u[10] < (R)alpha) {
                        const float ug[5] = { uf[11], uf[12], uf[13], uf[14], uf[15] };
                        float gd[3], gp;
                        orc_probe_hemi_sample(ptex, nf, gq, ug, gd, &gp);
                        const V3 gv = RF(v3)((R)gd[0], (R)gd[1], (R)gd[2]);
                        if (RF(dot)(gv, n) > (R)1e-12f) {
                            const R pbs = RF(bsdf_pdf)(&m, pspec, d, gv, n);
                            next = gv;
                            fcos = RF(eval_bsdf)(&m, alb, d, gv, n);
                            pdf = ((R)1 - (R)alpha) * RF(maxr)(pbs, (R)1e-12f) + (R)alpha * (R)gp;
                        } else {
                            ptex = NULL;
                        }
                    } else {
                        pdf = ((R)1 - (R)alpha) * pb + (R)alpha * pp;
                    }
                } else {
                    ptex = NULL;
                }
            }
        }
        int did_nee = 0;
        if (op->nee && L->count > 0 && depth < op->max_depth - 1 && non_delta < 10 && medium < 0 &&
            u[5] < SQRT(RF(clampr)(RF(lum4)(mask, mode), 0, 1))) {
            const int li = RF(pick_light)(L, u[6]);
            const uint32_t lp = (uint32_t)L->tri[li];
            V3 q, ln;
            R pa_total;
            int valid = 1;
            if (lp < sc->ntri) {
                V3 la, lb, lc;
                RF(tri)(sc, lp, &la, &lb, &lc);
                const R su = SQRT(u[7]);
                q = RF(add)(RF(add)(RF(scale)(la, (R)1 - su), RF(scale)(lb, su * ((R)1 - u[8]))),
                            RF(scale)(lc, su * u[8]));
                ln = RF(normalize)(RF(cross)(RF(sub)(lb, la), RF(sub)(lc, la)));
                pa_total = RF(maxr)((R)1e-20, (R)L->prob[li] / (R)L->area[li]);
            } else {
                /* box facet pick from a fifth group of uniforms */
                float all[20];
                orc_philox_uniforms(op->seed, pix, smp, (uint32_t)depth, 5, all);
                R pa;
                valid = RF(obj_light_sample)(&sc->objs[lp - sc->ntri], p, (R)all[16], u[7], u[8], &q, &ln, &pa);
                pa_total = valid ? RF(maxr)((R)1e-20, pa * (R)L->prob[li]) : (R)1;
            }
            if (valid) {
                const RF(mat) lm = RF(material)(sc, lp);
                const V3 lu = RF(sub)(q, p);
                const R dist2 = RF(maxr)(RF(dot)(lu, lu), (R)1e-12);
                const V3 ldir = RF(scale)(lu, (R)1 / SQRT(dist2));
                /* AnyHitSegmentBVH (intersection.cuh:285-310) */
                const R dist = SQRT(RF(dot)(lu, lu));
                int occl = 0;
                if (dist > 0) {
                    const V3 sdir = RF(scale)(lu, (R)1 / dist);
                    const R seps = RF(intersect_eps)(depth) * RF(maxr)((R)1, RF(maxr)(RF(length)(p), RF(length)(q)));
                    const R tmx = RF(maxr)((R)0, dist - seps);
                    if (tmx > 0) occl = RF(occluded)(sc, p, sdir, (R)0, tmx, RF(intersect_eps)(depth) * (R)0.5);
                }
                if (occl) {
                    did_nee = 1;
                } else {
                    const R cosx = RF(maxr)((R)0, RF(dot)(n, ldir));
                    const R cosy = RF(maxr)((R)0, -RF(dot)(ln, ldir));
                    if (cosx > 0 && cosy > 0) {
                        const S4 f = RF(eval_bsdf)(&m, alb, d, ldir, n);
                        const R g = cosy / dist2;
                        const S4 contrib = RF(s4scale)(RF(s4mul)(f, RF(emission_of)(sc, &lm, &LM, op->d65)), g / pa_total);
                        const R pl = pa_total * (dist2 / cosy);
                        R pb = RF(maxr)(RF(bsdf_pdf)(&m, pspec, d, ldir, n), (R)0);
                        if (ptex) {
                            const float lf[3] = { (float)ldir.x, (float)ldir.y, (float)ldir.z };
                            pb = ((R)1 - (R)alpha) * pb + (R)alpha * (R)orc_probe_hemi_pdf(ptex, lf, nf, gq);
                        }
                        const R w = RF(power_heuristic)(pl, pb);
                        rad = RF(s4add)(rad, RF(s4mul)(mask, RF(s4scale)(contrib, w)));
                        did_nee = 1;
                    }
                }
            }
        }
        prev_nee = did_nee;
        last_pdf = pdf;
        last_pos = p;
        const R inv_pdf = (R)1 / RF(maxr)(pdf, (R)1e-12);
        const S4 contrib = RF(s4scale)(fcos, inv_pdf);
        R qrr = 1;
        if (non_delta >= op->rr_min_depth) {
            qrr = RF(rr_q)(mask, mode, medium >= 0);
            if (u[9] > qrr) break;
        }
        if (gout) {
            if (non_delta < 5 && medium < 0 && RF(s4zero)(gm[non_delta], mode)) {
                const float pf[3] = { (float)p.x, (float)p.y, (float)p.z };
                const float rf[3] = { (float)rvec.x, (float)rvec.y, (float)rvec.z };
                gdir[non_delta] = next;
                gkey[non_delta] = orc_voxel_key(pf, rf);
                gm[non_delta] = RF(s4)(1, 1, 1, 1);
            }
            const R invq = (R)(1.0 / (double)qrr);
            const int hi = non_delta + 1 < 5 ? non_delta + 1 : 5;
            for (int i = 0; i < hi; ++i)
                if (!RF(s4zero)(gm[i], mode)) gm[i] = RF(s4mul)(gm[i], RF(s4scale)(contrib, invq));
        }
        mask = RF(s4scale)(RF(s4mul)(mask, contrib), (R)1 / qrr);
        RF(probe_query)(op, &LM, mask, gq);
        o = p;
        d = next;
        non_delta++;
    }
    /* guiding samples of this path (render.cuh:909-964) */
    if (gout) {
        int nw = 0;
        for (int i = 0; i < 5; ++i) {
            if (!gval[i]) continue;
            V3 x = RF(slot_xyz)(op, &LM, gr[i]);
            R strength = RF(max3)(x);
            if (mode != 0) strength = RF(minr)(strength, (R)1.0e5f);
            if (strength < (R)1.0e-5f) continue;
            RF(gsample)* g = &gout[nw++];
            g->key = gkey[i];
            g->dir[0] = (float)gdir[i].x; g->dir[1] = (float)gdir[i].y; g->dir[2] = (float)gdir[i].z;
            g->rad[0] = (float)RF(maxr)(x.x, 0); g->rad[1] = (float)RF(maxr)(x.y, 0); g->rad[2] = (float)RF(maxr)(x.z, 0);
        }
        *ngout = nw;
    }
    /* pixel value (render.cuh:966-997) */
    if (mode == 0) {
        V3 out = RF(v3)(RF(fixnan)(rad.v[0]), RF(fixnan)(rad.v[1]), RF(fixnan)(rad.v[2]));
        return RF(v3)(RF(clampr)(out.x, 0, (R)1e3), RF(clampr)(out.y, 0, (R)1e3), RF(clampr)(out.z, 0, (R)1e3));
    }
    R X = 0, Y = 0, Z = 0;
    if (mode == 1) {
        const V3 c = RF(cie)(op->cie, LM.l[0]);
        const R pdf = RF(lambda_pdf)(LM.l[0]);
        const R v = RF(fixnan)(rad.v[0]);
        X = c.x * v / pdf; Y = c.y * v / pdf; Z = c.z * v / pdf;
    } else {
        const R psum = (R)4 / (R)400;
        for (int k = 0; k < 4; ++k) {
            const V3 c = RF(cie)(op->cie, LM.l[k]);
            const R v = RF(fixnan)(rad.v[k]);
            X = X + c.x * v; Y = Y + c.y * v; Z = Z + c.z * v;
        }
        X = X / psum; Y = Y / psum; Z = Z / psum;
    }
    return RF(v3)(RF(clampr)(X, 0, (R)1e3), RF(clampr)(Y, 0, (R)1e3), RF(clampr)(Z, 0, (R)1e3));
}

/* RenderCamera (core/scene.cu:135-189): jittered pinhole, vertical fov. */
static V3 RF(camera_dir)(const orc_render_opts* op, uint32_t i, uint32_t j, R u0, R u1)
{
    const R u = ((R)i + u0) / (R)op->width;
    const R v = ((R)j + u1) / (R)op->height;
    const R ndx = u * (R)2 - (R)1, ndy = ((R)1 - v) * (R)2 - (R)1;
    const R fov = (R)op->vfov_deg * (R)3.1415926 / (R)180;
    const R aspect = (R)op->width / RF(maxr)((R)1, (R)op->height);
    const R thy = TAN((R)0.5 * fov);
    const R thx = thy * aspect;
    const V3 fwd = RF(load3)(op->cam_dir);
    const V3 right = RF(normalize)(RF(cross)(fwd, RF(v3)(0, 1, 0)));
    const V3 up = RF(normalize)(RF(cross)(fwd, right));
    return RF(normalize)(RF(add)(RF(add)(fwd, RF(scale)(right, ndx * thx)), RF(scale)(up, ndy * thy)));
}

/* one camera sample of pixel px (RenderCamera, core/scene.cu:135-189) */
static V3 RF(sample_pixel)(const orc_render_scene* sc, const orc_render_lights* L, const orc_render_opts* op,
                           long long px, uint32_t smp, const orc_guide* G, float alpha, RF(gsample)* gout,
                           int* ngout, float* fh)
{
    if (op->crop_w) px = (long long)(op->crop_y0 + px / op->crop_w) * op->width + op->crop_x0 + px % op->crop_w;
    const uint32_t i = (uint32_t)(px % op->width), j = (uint32_t)(px / op->width);
    float uf[8];
    orc_philox_uniforms(op->seed, (uint32_t)px, smp, ORC_CAMERA_BOUNCE, 1, uf);
    const V3 d = RF(camera_dir)(op, i, j, (R)uf[0], (R)uf[1]);
    R l0 = 0;
    if (op->spectral != 0) {
        float w[4];
        orc_philox_uniforms(op->seed, (uint32_t)px, smp, ORC_CAMERA_BOUNCE, 2, uf);
        for (int k = 0; k < 4; ++k) w[k] = uf[4 + k];
        l0 = op->spectral == 1 ? RF(sample_lambda)((R)w[0], (R)w[1], (R)w[2]) : (R)380 + (R)400 * (R)w[0];
    }
    return RF(trace_path)(sc, L, op, (uint32_t)px, smp, RF(load3)(op->cam_pos), d, l0, G, alpha, gout, ngout, fh);
}

static void RF(finish_pixel)(const orc_render_opts* op, V3 acc, R* o)
{
    R x = acc.x / (R)op->spp, y = acc.y / (R)op->spp, z = acc.z / (R)op->spp;
    if (op->spectral != 0 && op->output == 1) {
        /* XYZ2SRGBLinearD65 (core/tonemap.cuh:116-126) */
        const R r = (R)3.2406 * x + (R)-1.5372 * y + (R)-0.4986 * z;
        const R g = (R)-0.9689 * x + (R)1.8758 * y + (R)0.0415 * z;
        const R b = (R)0.0557 * x + (R)-0.204 * y + (R)1.057 * z;
        x = RF(maxr)((R)0, r); y = RF(maxr)((R)0, g); z = RF(maxr)((R)0, b);
    }
    o[0] = x;
    o[1] = y;
    o[2] = z;
}

static void RF(render)(const orc_render_scene* sc, const orc_render_lights* L, const orc_render_opts* op,
                       R* out, int threads)
{
    const long long npix = op->crop_w ? (long long)op->crop_w * op->crop_h : (long long)op->width * op->height;
    if (!op->guiding || !op->guide) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
        for (long long px = 0; px < npix; ++px) {
            V3 acc = RF(v3)(0, 0, 0);
            for (uint32_t s = 0; s < (uint32_t)op->spp; ++s) {
                float fh[11];
                const int aux = op->gbuffer || op->stats;
                const V3 v = RF(sample_pixel)(sc, L, op, px, s + op->sample_offset, NULL, 0.0f, NULL, NULL,
                                              aux ? fh : NULL);
                acc = RF(add)(acc, v);
                /* the denoiser's per-pixel records, folded in sample order */
                if (op->gbuffer)
                    orc_gbuffer_record(&op->gbuffer[16 * px], fh[0] != 0.0f, &fh[1], &fh[4], fh[7], fh[8], fh[9],
                                       (uint32_t)fh[10]);
                if (op->stats) {
                    const float x[3] = { (float)v.x, (float)v.y, (float)v.z };
                    orc_stats_record(&op->stats[12 * px], x);
                }
            }
            RF(finish_pixel)(op, acc, &out[3 * px]);
        }
        return;
    }
    /* voxel guiding: frames of one sample per pixel, probes updated after
     * each frame with its samples in pixel order (SceneRenderer::Render +
     * UpdatePathGuidingSample, core/scene.cu:249-270, 409-570) */
    V3* acc = malloc(sizeof(V3) * (size_t)npix);
    RF(gsample)* gs = malloc(sizeof(RF(gsample)) * 5 * (size_t)npix);
    int* gn = malloc(sizeof(int) * (size_t)npix);
    uint64_t* keys = malloc(sizeof(uint64_t) * 5 * (size_t)npix);
    float* dirs = malloc(sizeof(float) * 15 * (size_t)npix);
    float* rads = malloc(sizeof(float) * 15 * (size_t)npix);
    for (long long px = 0; px < npix; ++px) acc[px] = RF(v3)(0, 0, 0);
    for (uint32_t s = 0; s < (uint32_t)op->spp; ++s) {
        const uint32_t smp = s + op->sample_offset;
        const float alpha = orc_guide_alpha(smp);
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
        for (long long px = 0; px < npix; ++px)
            acc[px] = RF(add)(acc[px], RF(sample_pixel)(sc, L, op, px, smp, op->guide, alpha, &gs[5 * px], &gn[px],
                                                        NULL));
        int ns = 0;
        for (long long px = 0; px < npix; ++px)
            for (int k = 0; k < gn[px]; ++k, ++ns) {
                const RF(gsample)* g = &gs[5 * px + k];
                keys[ns] = g->key;
                for (int c = 0; c < 3; ++c) { dirs[3 * ns + c] = g->dir[c]; rads[3 * ns + c] = g->rad[c]; }
            }
        orc_guide_update(op->guide, ns, keys, dirs, rads);
    }
    for (long long px = 0; px < npix; ++px) RF(finish_pixel)(op, acc[px], &out[3 * px]);
    free(acc); free(gs); free(gn); free(keys); free(dirs); free(rads);
}

#undef V3
#undef S4
#undef RF
#undef R
#undef SUF
#undef SQRT
#undef SIN
#undef COS
#undef TAN
#undef EXP
#undef POW
#undef FABS
#undef PI_R
#undef LOG
#undef LOG1P
#undef FMA
#undef ERFC
#undef FLOOR
