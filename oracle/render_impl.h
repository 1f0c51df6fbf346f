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
    R specular, metallic, roughness;
    uint32_t model;
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

static V3 RF(f0_of)(const RF(mat)* m)
{
    const R f = RF(f0_ue)(m->specular);
    return RF(v3)(f + (m->albedo.x - f) * m->metallic, f + (m->albedo.y - f) * m->metallic,
                  f + (m->albedo.z - f) * m->metallic);
}

static R RF(alpha_of)(const RF(mat)* m)
{
    const R r = RF(clampr)(m->roughness, (R)0.01, (R)1);
    return r * r;
}

/* EvaluateBSDFWithCos (material.cuh:47-165), rd toward the surface */
static V3 RF(eval_bsdf)(const RF(mat)* m, V3 rd, V3 l, V3 n)
{
    const R nol = RF(maxr)((R)0, RF(dot)(n, l));
    if (m->model == 0) return RF(scale)(m->albedo, nol / PI_R);
    const V3 v = RF(normalize)(RF(scale)(rd, -1));
    const R nov = RF(maxr)((R)0, RF(dot)(n, v));
    if (nov <= 0 || nol <= 0) return RF(v3)(0, 0, 0);
    const R alpha = RF(alpha_of)(m);
    const V3 f0 = RF(f0_of)(m);
    const R kd = (R)1 - m->metallic;
    const V3 h = RF(normalize)(RF(add)(v, l));
    const R noh = RF(maxr)((R)0, RF(dot)(n, h));
    const R loh = RF(maxr)((R)0, RF(dot)(l, h));
    const R c = RF(clampr)((R)1 - loh, 0, 1);
    const R c5 = c * c * c * c * c;
    const V3 F = RF(v3)(f0.x + ((R)1 - f0.x) * c5, f0.y + ((R)1 - f0.y) * c5, f0.z + ((R)1 - f0.z) * c5);
    const R D = RF(ggx_d)(alpha, noh);
    const R G = RF(g_correlated)(alpha, nol, nov);
    const R sden = RF(maxr)((R)4 * nov, (R)1e-7);
    const R dcos = nol / PI_R;
    return RF(v3)(F.x * D * G / sden + kd * ((R)1 - F.x) * m->albedo.x * dcos,
                  F.y * D * G / sden + kd * ((R)1 - F.y) * m->albedo.y * dcos,
                  F.z * D * G / sden + kd * ((R)1 - F.z) * m->albedo.z * dcos);
}

/* BSDF_PDF (material.cuh:316-371) */
static R RF(bsdf_pdf)(const RF(mat)* m, V3 rd, V3 l, V3 n)
{
    const R nol = RF(maxr)((R)0, RF(dot)(n, l));
    if (m->model == 0) return nol > 0 ? nol * ((R)1 / PI_R) : 0;
    const V3 v = RF(scale)(rd, -1);
    const R nov = RF(maxr)((R)0, RF(dot)(n, v));
    if (nov <= 0 || nol <= 0) return 0;
    const R alpha = RF(alpha_of)(m);
    const R pspec = RF(clampr)(RF(lum)(RF(f0_of)(m)), (R)0.1, (R)0.9);
    const R ps = RF(vndf_pdf)(v, n, l, alpha);
    const R pd = nol * ((R)1 / PI_R);
    return RF(maxr)((R)0, pspec * ps + ((R)1 - pspec) * pd);
}

/* SampleBSDF (material.cuh:167-314).  u[0] lobe, u[1..3] direction.
 * Returns 0 when the path terminates. */
static int RF(sample_bsdf)(const RF(mat)* m, V3 rd, V3 n, const R* u, V3* next, R* pdf)
{
    if (m->model == 0) {
        *next = RF(sample_cosine)(u[1], u[2], u[3], n, pdf);
        return 1;
    }
    const V3 v = RF(scale)(rd, -1);
    const R nov = RF(maxr)((R)0, RF(dot)(n, v));
    if (nov <= 0) return 0;
    const R alpha = RF(alpha_of)(m);
    const R pspec = RF(clampr)(RF(lum)(RF(f0_of)(m)), (R)0.1, (R)0.9);
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

/* closest hit in R: prim (ORC_MISS on a miss), t, barycentric weights of v1, v2 */
static uint32_t RF(closest)(const orc_render_scene* sc, V3 o, V3 d, R tmin, R tmax, R* t, R* bu, R* bv);
static int RF(occluded)(const orc_render_scene* sc, V3 o, V3 d, R tmin, R tmax);

static void RF(tri)(const orc_render_scene* sc, uint32_t p, V3* a, V3* b, V3* c)
{
    *a = RF(load3)(&sc->verts[3 * sc->idx[3 * p + 0]]);
    *b = RF(load3)(&sc->verts[3 * sc->idx[3 * p + 1]]);
    *c = RF(load3)(&sc->verts[3 * sc->idx[3 * p + 2]]);
}

static RF(mat) RF(material)(const orc_render_scene* sc, uint32_t p)
{
    const orc_material* om = &sc->mats[sc->tri_mat ? sc->tri_mat[p] : 0];
    RF(mat) m;
    m.albedo = RF(load3)(om->albedo);
    m.emission = RF(load3)(om->emission);
    m.specular = om->specular; m.metallic = om->metallic; m.roughness = om->roughness;
    m.model = om->model;
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
/* One path (CalculateRadiance, core/render/render.cuh:460-998)            */
/* ---------------------------------------------------------------------- */

static V3 RF(trace_path)(const orc_render_scene* sc, const orc_render_lights* L, const orc_render_opts* op,
                         uint32_t pix, uint32_t smp, V3 o, V3 d)
{
    V3 rad = RF(v3)(0, 0, 0), mask = RF(v3)(1, 1, 1);
    int non_delta = 0;
    int prev_nee = 0;
    R last_pdf = 0;
    V3 last_pos = RF(v3)(0, 0, 0);
    d = RF(normalize)(d);
    for (int depth = 0; depth < op->max_depth; ++depth) {
        float uf[12];
        orc_philox_uniforms(op->seed, pix, smp, (uint32_t)depth, 3, uf);
        R u[12];
        for (int k = 0; k < 12; ++k) u[k] = (R)uf[k];
        const R eps = RF(intersect_eps)(non_delta);
        R t, bu, bv;
        const uint32_t prim = RF(closest)(sc, o, d, eps, (R)INFINITY, &t, &bu, &bv);
        if (prim == ORC_MISS) {
            if (op->sky == 1) rad = RF(add)(rad, RF(mul)(mask, RF(sky)(d)));
            else if (op->sky == 2) rad = RF(add)(rad, RF(mul)(mask, RF(load3)(op->env)));
            break;
        }
        V3 a, b, c;
        RF(tri)(sc, prim, &a, &b, &c);
        const R bw = (R)1 - bu - bv;
        V3 p = RF(add)(RF(add)(RF(scale)(a, bw), RF(scale)(b, bu)), RF(scale)(c, bv));
        const V3 ng = RF(normalize)(RF(cross)(RF(sub)(b, a), RF(sub)(c, a)));
        const R ndd = RF(dot)(ng, d);
        const V3 n = ndd < 0 ? ng : RF(scale)(ng, -1);
        const RF(mat) m = RF(material)(sc, prim);
        const int is_light = sc->light_index[prim] >= 0;
        if (is_light) {
            if (ndd < 0) {
                R w = 1;
                if (op->nee && depth > 0 && prev_nee) {
                    const int li = sc->light_index[prim];
                    const V3 tl = RF(sub)(p, last_pos);
                    const R dist2 = RF(maxr)(RF(dot)(tl, tl), (R)1e-20);
                    const V3 dir = RF(scale)(tl, (R)1 / SQRT(dist2));
                    const R cosy = RF(maxr)((R)0, -RF(dot)(ng, dir));
                    const R pa = (R)L->prob[li] / (R)L->area[li];
                    const R pl = cosy <= 0 ? (R)0 : pa * (dist2 / cosy);
                    if (pl >= 0 && last_pdf > 0) w = RF(power_heuristic)(last_pdf, pl);
                }
                rad = RF(add)(rad, RF(scale)(RF(mul)(mask, m.emission), w));
            }
            break;
        }
        V3 next;
        R pdf;
        if (!RF(sample_bsdf)(&m, d, n, u, &next, &pdf)) break;
        const V3 fcos = RF(eval_bsdf)(&m, d, next, n);
        /* offset along the side of the next direction (render.cuh:155-160) */
        const R off = eps * RF(maxr)((R)1, RF(length)(p)) * (u[4] + (R)0.5);
        const R sgn = RF(dot)(n, next) > 0 ? (R)1 : (RF(dot)(n, next) < 0 ? (R)-1 : (R)0);
        p = RF(add)(p, RF(scale)(n, sgn * off));
        int did_nee = 0;
        if (op->nee && L->count > 0 && depth < op->max_depth - 1 && non_delta < 10 &&
            u[5] < SQRT(RF(clampr)(RF(lum)(mask), 0, 1))) {
            const int li = RF(pick_light)(L, u[6]);
            const uint32_t lp = (uint32_t)L->tri[li];
            V3 la, lb, lc;
            RF(tri)(sc, lp, &la, &lb, &lc);
            const R su = SQRT(u[7]);
            const V3 q = RF(add)(RF(add)(RF(scale)(la, (R)1 - su), RF(scale)(lb, su * ((R)1 - u[8]))),
                                 RF(scale)(lc, su * u[8]));
            const V3 ln = RF(normalize)(RF(cross)(RF(sub)(lb, la), RF(sub)(lc, la)));
            const RF(mat) lm = RF(material)(sc, lp);
            const R pa_total = RF(maxr)((R)1e-20, (R)L->prob[li] / (R)L->area[li]);
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
                if (tmx > 0) occl = RF(occluded)(sc, p, sdir, (R)0, tmx);
            }
            if (occl) {
                did_nee = 1;
            } else {
                const R cosx = RF(maxr)((R)0, RF(dot)(n, ldir));
                const R cosy = RF(maxr)((R)0, -RF(dot)(ln, ldir));
                if (cosx > 0 && cosy > 0) {
                    const V3 f = RF(eval_bsdf)(&m, d, ldir, n);
                    const R g = cosy / dist2;
                    const V3 contrib = RF(scale)(RF(mul)(f, lm.emission), g / pa_total);
                    const R pl = pa_total * (dist2 / cosy);
                    const R pb = RF(maxr)(RF(bsdf_pdf)(&m, d, ldir, n), (R)0);
                    const R w = RF(power_heuristic)(pl, pb);
                    rad = RF(add)(rad, RF(mul)(mask, RF(scale)(contrib, w)));
                    did_nee = 1;
                }
            }
        }
        prev_nee = did_nee;
        last_pdf = pdf;
        last_pos = p;
        const R inv_pdf = (R)1 / RF(maxr)(pdf, (R)1e-12);
        const V3 contrib = RF(scale)(fcos, inv_pdf);
        R q = 1;
        if (non_delta >= op->rr_min_depth) {
            q = RF(clampr)(RF(lum)(mask) * (R)0.5 + RF(max3)(mask) * (R)0.5, (R)0.05, (R)0.95);
            if (u[9] > q) break;
        }
        mask = RF(scale)(RF(mul)(mask, contrib), (R)1 / q);
        o = p;
        d = next;
        non_delta++;
    }
    V3 out = RF(v3)(RF(fixnan)(rad.x), RF(fixnan)(rad.y), RF(fixnan)(rad.z));
    out = RF(v3)(RF(clampr)(out.x, 0, (R)1e3), RF(clampr)(out.y, 0, (R)1e3), RF(clampr)(out.z, 0, (R)1e3));
    return out;
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

static void RF(render)(const orc_render_scene* sc, const orc_render_lights* L, const orc_render_opts* op,
                       R* out, int threads)
{
    const long long npix = (long long)op->width * op->height;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (long long px = 0; px < npix; ++px) {
        const uint32_t i = (uint32_t)(px % op->width), j = (uint32_t)(px / op->width);
        V3 acc = RF(v3)(0, 0, 0);
        for (uint32_t s = 0; s < (uint32_t)op->spp; ++s) {
            float uf[4];
            orc_philox_uniforms(op->seed, (uint32_t)px, s + op->sample_offset, ORC_CAMERA_BOUNCE, 1, uf);
            const V3 d = RF(camera_dir)(op, i, j, (R)uf[0], (R)uf[1]);
            const V3 r = RF(trace_path)(sc, L, op, (uint32_t)px, s + op->sample_offset, RF(load3)(op->cam_pos), d);
            acc = RF(add)(acc, r);
        }
        out[3 * px + 0] = acc.x / (R)op->spp;
        out[3 * px + 1] = acc.y / (R)op->spp;
        out[3 * px + 2] = acc.z / (R)op->spp;
    }
}

#undef V3
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
