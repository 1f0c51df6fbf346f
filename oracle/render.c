/*
 * Oracle integrator (test infrastructure only): Philox RNG, light table,
 * FP32 / FP64 traversal glue and the two instantiations of render_impl.h.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------- Philox4x32-10 ---------------- */

static void mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo)
{
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n1 = lo1;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        const uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_philox_uniforms(uint64_t seed, uint32_t pix, uint32_t smp, uint32_t bounce, int ngroups, float* out)
{
    const uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    for (int g = 0; g < ngroups; ++g) {
        const uint32_t ctr[4] = { pix, smp, bounce, (uint32_t)g };
        uint32_t x[4];
        orc_philox4x32_10(ctr, key, x);
        for (int k = 0; k < 4; ++k) out[4 * g + k] = (float)(x[k] >> 8) * 0x1p-24f;
    }
}

/* ---------------- instantiations ---------------- */

#define RF_CAT(a, b) RF_CAT2(a, b)
#define RF_CAT2(a, b) a##_##b

#define R float
#define SUF f32
#define SQRT sqrtf
#define SIN sinf
#define COS cosf
#define TAN tanf
#define EXP expf
#define POW powf
#define FABS fabsf
#define PI_R 3.14159265358979323846f
#define LOG logf
#define LOG1P log1pf
#define FMA fmaf
#define ERFC erfcf
#define FLOOR floorf
#include "render_impl.h"

#define R double
#define SUF f64
#define SQRT sqrt
#define SIN sin
#define COS cos
#define TAN tan
#define EXP exp
#define POW pow
#define FABS fabs
#define PI_R 3.14159265358979323846
#define LOG log
#define LOG1P log1p
#define FMA fma
#define ERFC erfc
#define FLOOR floor
#include "render_impl.h"

/* ---------------- FP64 traversal (watertight test in double) ---------------- */

static int tri_test_d(const double o[3], const double d[3], const float* v0, const float* v1, const float* v2,
                      double* t, double* u, double* v)
{
    const double ax = fabs(d[0]), ay = fabs(d[1]), az = fabs(d[2]);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > fabs(d[kz])) kz = 2;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (d[kz] < 0) { const int tmp = kx; kx = ky; ky = tmp; }
    if (d[kz] == 0) return 0;
    const double sx = d[kx] / d[kz], sy = d[ky] / d[kz], sz = 1.0 / d[kz];
    const double A[3] = { v0[0] - o[0], v0[1] - o[1], v0[2] - o[2] };
    const double B[3] = { v1[0] - o[0], v1[1] - o[1], v1[2] - o[2] };
    const double C[3] = { v2[0] - o[0], v2[1] - o[1], v2[2] - o[2] };
    const double Ax = A[kx] - sx * A[kz], Ay = A[ky] - sy * A[kz];
    const double Bx = B[kx] - sx * B[kz], By = B[ky] - sy * B[kz];
    const double Cx = C[kx] - sx * C[kz], Cy = C[ky] - sy * C[kz];
    const double U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return 0;
    const double det = U + V + W;
    if (det == 0) return 0;
    const double T = U * (sz * A[kz]) + V * (sz * B[kz]) + W * (sz * C[kz]);
    *t = T / det;
    *u = V / det;
    *v = W / det;
    return 1;
}

static int box_hit_d(const double o[3], const double inv[3], const orc_node* nd, double tmin, double tmax)
{
    double m = fmax(fabs(o[0]), fmax(fabs(o[1]), fabs(o[2])));
    for (int a = 0; a < 3; ++a) m = fmax(m, fmax(fabs(nd->bmin[a]), fabs(nd->bmax[a])));
    const double pad = m * 0x1p-16 + 1e-30;
    double te = -INFINITY, tx = INFINITY;
    for (int a = 0; a < 3; ++a) {
        const double t0 = (nd->bmin[a] - pad - o[a]) * inv[a], t1 = (nd->bmax[a] + pad - o[a]) * inv[a];
        te = fmax(te, fmin(t0, t1));
        tx = fmin(tx, fmax(t0, t1));
    }
    return tx >= fmax(0.0, tmin) && te <= tmax && te <= tx;
}

/* any = 1: first accepted hit.  Canonical rule (min t, then min id); objects
 * by the FP64 instance of the object intersector. */
static uint32_t trace_d(const orc_render_scene* sc, const double o[3], const double d[3], double tmin, double tmax,
                        double meps, int any, double* tout, double* uout, double* vout, int* inside)
{
    *inside = 0;
    if (sc->node_count == 0 || !(tmax >= tmin)) return ORC_MISS;
    double inv[3];
    for (int a = 0; a < 3; ++a) inv[a] = d[a] == 0 ? (signbit(d[a]) ? -DBL_MAX : DBL_MAX) : 1.0 / d[a];
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    uint32_t best = ORC_MISS;
    double bt = tmax, bu = 0, bv = 0;
    while (sp) {
        const orc_node* nd = &sc->nodes[stack[--sp]];
        if (!box_hit_d(o, inv, nd, tmin, bt)) continue;
        if (nd->count > 0) {
            for (int k = 0; k < nd->count; ++k) {
                const uint32_t p = (uint32_t)sc->indices[nd->first + k];
                double t, u = 0, v = 0;
                int ins = 0;
                if (p < sc->ntri) {
                    const float* a = &sc->verts[3 * sc->idx[3 * p]];
                    const float* b = &sc->verts[3 * sc->idx[3 * p + 1]];
                    const float* c = &sc->verts[3 * sc->idx[3 * p + 2]];
                    if (!tri_test_d(o, d, a, b, c, &t, &u, &v)) continue;
                } else {
                    const vec3_f64 oo = { o[0], o[1], o[2] }, dd = { d[0], d[1], d[2] };
                    if (!obj_test_f64(&sc->objs[p - sc->ntri], oo, dd, meps, &t, &ins)) continue;
                }
                const int acc = t >= tmin && (best != ORC_MISS ? (t < bt || (t == bt && p < best)) : t < tmax);
                if (acc) {
                    best = p; bt = t; bu = u; bv = v; *inside = ins;
                    if (any) { *tout = bt; return best; }
                }
            }
        } else {
            if (sp + 2 > 128) abort();
            stack[sp++] = nd->right;
            stack[sp++] = nd->left;
        }
    }
    *tout = bt; *uout = bu; *vout = bv;
    return best;
}

static orc_geom scene_geom(const orc_render_scene* sc)
{
    orc_geom g = { sc->verts, sc->idx, sc->ntri, sc->objs, sc->nobj };
    return g;
}

static uint32_t closest_f32(const orc_render_scene* sc, vec3_f32 o, vec3_f32 d, float tmin, float tmax,
                            float* t, float* bu, float* bv, int* inside)
{
    const float oo[3] = { o.x, o.y, o.z }, dd[3] = { d.x, d.y, d.z };
    const orc_geom g = scene_geom(sc);
    uint32_t p, f = 0;
    *bu = *bv = 0.0f;
    orc_closest1_ge(sc->nodes, sc->node_count, sc->indices, &g, oo, dd, tmin, tmax, 0.5f * tmin, &p, t, bu, bv, &f);
    *inside = (f & ORC_HIT_INSIDE) != 0;
    return p;
}

static int occluded_f32(const orc_render_scene* sc, vec3_f32 o, vec3_f32 d, float tmin, float tmax, float meps)
{
    const float oo[3] = { o.x, o.y, o.z }, dd[3] = { d.x, d.y, d.z };
    const orc_geom g = scene_geom(sc);
    return orc_any1_ge(sc->nodes, sc->node_count, sc->indices, &g, oo, dd, tmin, tmax, meps);
}

static uint32_t closest_f64(const orc_render_scene* sc, vec3_f64 o, vec3_f64 d, double tmin, double tmax,
                            double* t, double* bu, double* bv, int* inside)
{
    const double oo[3] = { o.x, o.y, o.z }, dd[3] = { d.x, d.y, d.z };
    *bu = *bv = 0.0;
    return trace_d(sc, oo, dd, tmin, tmax, 0.5 * tmin, 0, t, bu, bv, inside);
}

static int occluded_f64(const orc_render_scene* sc, vec3_f64 o, vec3_f64 d, double tmin, double tmax, double meps)
{
    const double oo[3] = { o.x, o.y, o.z }, dd[3] = { d.x, d.y, d.z };
    double t, u, v;
    int ins;
    return trace_d(sc, oo, dd, tmin, tmax, meps, 1, &t, &u, &v, &ins) != ORC_MISS;
}

/* ---------------- RGB uplift lookup ---------------- */

static float lut_scale(int k)
{
    float t = (float)k / 63.0f;
    t = t * t * (3.0f - 2.0f * t);
    return t * t * (3.0f - 2.0f * t);
}

void orc_uplift(const float* lut, const float rgb_in[3], float c[3])
{
    float rgb[3];
    for (int k = 0; k < 3; ++k) rgb[k] = fminf(fmaxf(rgb_in[k], 0.0f), 1.0f);
    int l = 0;
    if (rgb[1] > rgb[l]) l = 1;
    if (rgb[2] > rgb[l]) l = 2;
    const float z = rgb[l];
    if (!(z > 0.0f)) { c[0] = 0.0f; c[1] = 0.0f; c[2] = -1e30f; return; }
    const float x = rgb[(l + 1) % 3] / z, y = rgb[(l + 2) % 3] / z;
    int k = 0;
    while (k < 62 && lut_scale(k + 1) <= z) ++k;
    const float z0 = lut_scale(k), z1 = lut_scale(k + 1);
    float tz = (z - z0) / (z1 - z0);
    tz = fminf(fmaxf(tz, 0.0f), 1.0f);
    const float fx = x * 63.0f, fy = y * 63.0f;
    int i = (int)fx, j = (int)fy;
    if (i > 62) i = 62;
    if (j > 62) j = 62;
    const float tx = fx - (float)i, ty = fy - (float)j;
    for (int m = 0; m < 3; ++m) {
        float acc = 0.0f;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const float w = (dx ? tx : 1.0f - tx) * (dy ? ty : 1.0f - ty) * (dz ? tz : 1.0f - tz);
                    acc += w * lut[((((size_t)l * 64 + (size_t)(k + dz)) * 64 + (size_t)(j + dy)) * 64 + (size_t)(i + dx)) * 3 + (size_t)m];
                }
        c[m] = acc;
    }
}

/* ---------------- light table (core/scene.cu:677-701) ---------------- */

static float tri_area_f(const float* a, const float* b, const float* c)
{
    const float e1[3] = { b[0] - a[0], b[1] - a[1], b[2] - a[2] };
    const float e2[3] = { c[0] - a[0], c[1] - a[1], c[2] - a[2] };
    const float cx = e1[1] * e2[2] - e1[2] * e2[1];
    const float cy = e1[2] * e2[0] - e1[0] * e2[2];
    const float cz = e1[0] * e2[1] - e1[1] * e2[0];
    return 0.5f * sqrtf(cx * cx + cy * cy + cz * cz);
}

int orc_light_table(const float* verts, const uint32_t* idx, uint32_t ntri, const uint32_t* tri_mat,
                    const orc_material* mats, const orc_object* objects, uint32_t nobj, const orc_material* obj_mats,
                    int32_t* light_index, int32_t* ltri, float* larea, float* lprob, float* lcdf)
{
    const uint32_t np = ntri + nobj;
    int nl = 0;
    float total = 0.0f;
    for (uint32_t p = 0; p < np; ++p) {
        const orc_material* m = p < ntri ? &mats[tri_mat ? tri_mat[p] : 0] : &obj_mats[p - ntri];
        light_index[p] = -1;
        if (m->emission[0] > 0.0f || m->emission[1] > 0.0f || m->emission[2] > 0.0f) {
            const float area = p < ntri ? tri_area_f(&verts[3 * idx[3 * p]], &verts[3 * idx[3 * p + 1]],
                                                     &verts[3 * idx[3 * p + 2]])
                                        : orc_object_area(&objects[p - ntri]);
            const float lum = m->emission[0] * 0.2126f + m->emission[1] * 0.7152f + m->emission[2] * 0.0722f;
            const float w = fmaxf(area * lum, 1.0e-20f);
            light_index[p] = nl;
            ltri[nl] = (int32_t)p;
            larea[nl] = area;
            lprob[nl] = w;
            total += w;
            ++nl;
        }
    }
    float acc = 0.0f;
    for (int i = 0; i < nl; ++i) {
        lprob[i] = lprob[i] / total;
        acc += lprob[i];
        lcdf[i] = acc;
    }
    return nl;
}

int orc_render(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t ntri, const uint32_t* tri_mat,
               const orc_material* mats, uint32_t nmats, const orc_render_opts* opts, int precision,
               void* out, int threads)
{
    if (opts->crop_w && (opts->guiding || opts->gbuffer || opts->stats || opts->crop_h == 0 ||
                         opts->crop_x0 + opts->crop_w > opts->width || opts->crop_y0 + opts->crop_h > opts->height))
        return -1;
    (void)nv;
    orc_render_scene sc;
    memset(&sc, 0, sizeof sc);
    const uint32_t nobj = opts->objects ? opts->nobj : 0;
    const uint32_t np = ntri + nobj;
    const size_t cap = np ? np : 1;
    sc.verts = verts; sc.idx = idx; sc.ntri = ntri; sc.tri_mat = tri_mat; sc.mats = mats;
    sc.objs = opts->objects; sc.nobj = nobj; sc.obj_mats = opts->obj_mats; sc.nmats = nmats;
    float* pmin = malloc(sizeof(float) * 3 * cap);
    float* pmax = malloc(sizeof(float) * 3 * cap);
    float* pcen = malloc(sizeof(float) * 3 * cap);
    orc_node* nodes = malloc(sizeof(orc_node) * 2 * cap);
    int32_t* indices = malloc(sizeof(int32_t) * cap);
    int32_t* light_index = malloc(sizeof(int32_t) * cap);
    int32_t* ltri = malloc(sizeof(int32_t) * cap);
    float* larea = malloc(sizeof(float) * cap);
    float* lprob = malloc(sizeof(float) * cap);
    float* lcdf = malloc(sizeof(float) * cap);
    const orc_geom g = { verts, idx, ntri, opts->objects, nobj };
    orc_geom_proxies(&g, pmin, pmax, pcen);
    sc.node_count = orc_bvh_build(pmin, pmax, pcen, (int)np, 8, 12, nodes, indices);
    sc.nodes = nodes;
    sc.indices = indices;
    const int nl = orc_light_table(verts, idx, ntri, tri_mat, mats, opts->objects, nobj, opts->obj_mats, light_index,
                                   ltri, larea, lprob, lcdf);
    sc.light_index = light_index;
    float* mat_coef = NULL;
    float* lam_tab = NULL;
    sc.lam_tab = NULL;
    sc.lam_res = 0;
    if (opts->spectral != 0) {
        const uint32_t nm = nmats + nobj;
        mat_coef = calloc((size_t)12 * (nm ? nm : 1), sizeof(float));
        const int net = opts->smlp != NULL;
        if (net) {
            lam_tab = malloc((size_t)4 * opts->smlp_lambda_res * sizeof(float));
            for (int i = 0; i < opts->smlp_lambda_res; ++i)
                orc_smlp_lambda_entry(opts->smlp, opts->smlp_lambda_res, i, &lam_tab[4 * i]);
            sc.lam_tab = lam_tab;
            sc.lam_res = opts->smlp_lambda_res;
        }
        for (uint32_t i = 0; i < nm; ++i) {
            const orc_material* mi = i < nmats ? &mats[i] : &opts->obj_mats[i - nmats];
            float* mc = &mat_coef[12 * i];
            float a[3] = { mi->albedo[0], mi->albedo[1], mi->albedo[2] };
            const float ma = fmaxf(a[0], fmaxf(a[1], a[2]));
            mc[9] = 1.0f;
            float e[3] = { mi->emission[0], mi->emission[1], mi->emission[2] };
            const float mx = fmaxf(e[0], fmaxf(e[1], e[2]));
            float scale = 1.0f;
            if (mx > 1.0f) { for (int k = 0; k < 3; ++k) e[k] /= mx; scale = mx; }
            if (net) {
                /* Interact: an RGB above 1 is queried divided by its largest channel */
                if (ma > 1.0f) { for (int k = 0; k < 3; ++k) a[k] /= ma; mc[9] = ma; }
                orc_smlp_query_rgb(opts->smlp, opts->smlp_rgb_res, a, mc);
                orc_smlp_query_rgb(opts->smlp, opts->smlp_rgb_res, e, mc + 4);
            } else {
                orc_uplift(opts->lut, mi->albedo, mc);
                orc_uplift(opts->lut, e, mc + 4);
            }
            mc[8] = scale;
        }
        float e[3] = { opts->env[0], opts->env[1], opts->env[2] };
        const float mx = fmaxf(e[0], fmaxf(e[1], e[2]));
        sc.env_coef[4] = 1.0f;
        if (mx > 1.0f) { for (int k = 0; k < 3; ++k) e[k] /= mx; sc.env_coef[4] = mx; }
        if (net) orc_smlp_query_rgb(opts->smlp, opts->smlp_rgb_res, e, sc.env_coef);
        else orc_uplift(opts->lut, e, sc.env_coef);
    }
    sc.mat_coef = mat_coef;
    orc_render_lights L = { nl, ltri, larea, lprob, lcdf };
    int nt = threads;
#ifdef _OPENMP
    if (nt <= 0) nt = omp_get_max_threads();
#else
    nt = 1;
#endif
    if (precision == 64) render_f64(&sc, &L, opts, (double*)out, nt);
    else render_f32(&sc, &L, opts, (float*)out, nt);
    free(pmin); free(pmax); free(pcen); free(nodes); free(indices); free(light_index);
    free(ltri); free(larea); free(lprob); free(lcdf); free(mat_coef); free(lam_tab);
    return 0;
}

/* Wavelength helpers exposed for the tests (FP32 instantiation). */
float orc_ndtri_f(float u) { return ndtri_f32(u); }
float orc_lambda_pdf_f(float l) { return lambda_pdf_f32(l); }
float orc_sample_lambda_f(float u0, float u1, float u2) { return sample_lambda_f32(u0, u1, u2); }
float orc_sigmoid_f(const float c[3], float l) { return sigmoid_f32(c, l); }

/* Object light sampler and its density, FP64 instance (for the tests). */
int orc_obj_light_sample_d(const orc_object* ob, const double x[3], double ub, double u0, double u1, double y[3],
                           double ny[3], double* pa)
{
    const vec3_f64 xv = { x[0], x[1], x[2] };
    vec3_f64 yv, nv;
    const int ok = obj_light_sample_f64(ob, xv, ub, u0, u1, &yv, &nv, pa);
    y[0] = yv.x; y[1] = yv.y; y[2] = yv.z;
    ny[0] = nv.x; ny[1] = nv.y; ny[2] = nv.z;
    return ok;
}

double orc_obj_light_pdf_d(const orc_object* ob, double select, const double x[3], const double y[3],
                           const double ny[3])
{
    const vec3_f64 xv = { x[0], x[1], x[2] }, yv = { y[0], y[1], y[2] }, nv = { ny[0], ny[1], ny[2] };
    return obj_light_pdf_f64(ob, select, xv, yv, nv);
}

/* BSDF pieces and the light pick, FP64 instances (for the tests).  A
 * material is evaluated in RGB mode, as a triangle material. */
static mat_f64 test_material(const orc_material* om)
{
    orc_render_scene sc;
    memset(&sc, 0, sizeof sc);
    sc.mats = om;
    sc.ntri = 1;
    return material_f64(&sc, 0);
}

static vec3_f64 v3d(const double* p) { return v3_f64(p[0], p[1], p[2]); }
static void put3(vec3_f64 v, double* p) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

void orc_ggx_d_d(double alpha, size_t count, const double* noh, double* out)
{
    for (size_t i = 0; i < count; ++i) out[i] = ggx_d_f64(alpha, noh[i]);
}

void orc_g_correlated_d(double alpha, size_t count, const double* nol, const double* nov, double* out)
{
    for (size_t i = 0; i < count; ++i) out[i] = g_correlated_f64(alpha, nol[i], nov[i]);
}

void orc_sample_vndf_d(const double v[3], const double n[3], double alpha, size_t count, const double* r,
                       double* l, double* pdf)
{
    for (size_t i = 0; i < count; ++i)
        put3(sample_vndf_f64(v3d(v), v3d(n), alpha, r[3 * i], r[3 * i + 1], r[3 * i + 2], &pdf[i]), &l[3 * i]);
}

void orc_vndf_pdf_d(const double v[3], const double n[3], double alpha, size_t count, const double* l, double* out)
{
    for (size_t i = 0; i < count; ++i) out[i] = vndf_pdf_f64(v3d(v), v3d(n), v3d(&l[3 * i]), alpha);
}

void orc_sample_cosine_d(const double n[3], size_t count, const double* e, double* l, double* pdf)
{
    for (size_t i = 0; i < count; ++i)
        put3(sample_cosine_f64(e[3 * i], e[3 * i + 1], e[3 * i + 2], v3d(n), &pdf[i]), &l[3 * i]);
}

double orc_pspec_d(const orc_material* om)
{
    const mat_f64 m = test_material(om);
    return pspec_of_f64(&m, s4_f64(0, 0, 0, 0), 0);
}

void orc_eval_bsdf_d(const orc_material* om, const double rd[3], const double n[3], size_t count, const double* l,
                     double* out)
{
    const mat_f64 m = test_material(om);
    const s4t_f64 alb = s4_f64(m.albedo.x, m.albedo.y, m.albedo.z, 0);
    for (size_t i = 0; i < count; ++i) {
        const s4t_f64 f = eval_bsdf_f64(&m, alb, v3d(rd), v3d(&l[3 * i]), v3d(n));
        out[3 * i] = f.v[0]; out[3 * i + 1] = f.v[1]; out[3 * i + 2] = f.v[2];
    }
}

void orc_bsdf_pdf_d(const orc_material* om, const double rd[3], const double n[3], size_t count, const double* l,
                    double* out)
{
    const mat_f64 m = test_material(om);
    const double ps = orc_pspec_d(om);
    for (size_t i = 0; i < count; ++i) out[i] = bsdf_pdf_f64(&m, ps, v3d(rd), v3d(&l[3 * i]), v3d(n));
}

void orc_sample_bsdf_d(const orc_material* om, const double rd[3], const double n[3], size_t count, const double* u,
                       double* l, double* pdf, int32_t* ok)
{
    const mat_f64 m = test_material(om);
    const double ps = orc_pspec_d(om);
    for (size_t i = 0; i < count; ++i) {
        vec3_f64 next = v3_f64(0, 0, 0);
        pdf[i] = 0;
        ok[i] = sample_bsdf_f64(&m, ps, v3d(rd), v3d(n), &u[4 * i], &next, &pdf[i]);
        put3(next, &l[3 * i]);
    }
}

void orc_pick_light_d(const float* cdf, int nlights, size_t count, const double* u, int32_t* out)
{
    orc_render_lights L;
    memset(&L, 0, sizeof L);
    L.count = nlights;
    L.cdf = cdf;
    for (size_t i = 0; i < count; ++i) out[i] = pick_light_f64(&L, u[i]);
}

double orc_rr_survival_d(const double mask[3])
{
    return rr_q_f64(s4_f64(mask[0], mask[1], mask[2], 0), 0, 0);
}
