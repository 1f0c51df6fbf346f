/*
 * Oracle voxel path guiding.  TEST INFRASTRUCTURE ONLY: tests/, smoke() and
 * bench.py's CPU baseline may call it; the product library never does.
 *
 * Restates core/guiding (FP32, as the reference computes it):
 *   - equal-area octahedral map of Clarberg 2008 (octahedron.cuh:13-134),
 *     Jacobian 4*pi (octahedron.cuh:179-187);
 *   - icosahedral 20-face direction binning and the 0.05 voxel grid
 *     (guiding_bin.cuh:40-113, config.cuh:43-48);
 *   - probe texture of 8x8 leaf texels plus 4x4, 2x2, 1x1 mips of float4
 *     (Texture<4>, octahedron.cuh:203-525): GenerateMipmaps with the 1e-5
 *     leaf floor, SampleOctahedron (3-level descent by dot(texel, query), then
 *     uniform in the leaf), OctahedronPDF, and the hemisphere fold;
 *   - the per-frame update (guiding.cuh:336-517, scene.cu:409-570): samples
 *     accumulate radiance * 4 pi into the leaf texel of their direction and
 *     count per probe; every touched probe merges old * n_o/(n_o+n_n) +
 *     sum/(n_o+n_n) in FP64, count capped at 1e6, then rebuilds its mips.
 * Probes are keyed by (voxel x, y, z, face) packed into 64 bits; the reference
 * indexes a dense voxel volume instead, which gives the same probe per key.
 * Samples are accumulated sequentially in the order given (pixel order, then
 * slot order), which is one of the orders the reference's atomics may take.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_F 3.141592654f

static float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

void orc_uv_to_dir(float u01, float v01, float d[3])
{
    const float u = 2.0f * u01 - 1.0f, v = 2.0f * v01 - 1.0f;
    const float a = v + u, b = v - u;
    float r, phi4, zs;
    if (v >= 0.0f) {
        if (u >= 0.0f) {
            if (a <= 1.0f) { r = a; zs = 1.0f; phi4 = r == 0.0f ? 0.0f : v / r; }
            else { r = 2.0f - a; zs = -1.0f; phi4 = (1.0f - u) / r; }
        } else {
            if (b <= 1.0f) { r = b; zs = 1.0f; phi4 = 1.0f - u / r; }
            else { r = 2.0f - b; zs = -1.0f; phi4 = 1.0f + (1.0f - v) / r; }
        }
    } else {
        if (u < 0.0f) {
            if (a >= -1.0f) { r = -a; zs = 1.0f; phi4 = 2.0f - v / r; }
            else { r = 2.0f + a; zs = -1.0f; phi4 = 2.0f + (1.0f + u) / r; }
        } else {
            if (b >= -1.0f) { r = -b; zs = 1.0f; phi4 = 3.0f + u / r; }
            else { r = 2.0f + b; zs = -1.0f; phi4 = 3.0f + (1.0f + v) / r; }
        }
    }
    if (r == 0.0f) phi4 = 0.0f;
    const float r2 = r * r;
    const float phi = phi4 * (0.5f * PI_F);
    const float st = r * sqrtf(fmaxf(0.0f, 2.0f - r2));
    const float ct = zs * (1.0f - r2);
    d[0] = st * cosf(phi);
    d[1] = st * sinf(phi);
    d[2] = ct;
}

void orc_dir_to_uv(const float d[3], float uv[2])
{
    const float p = atan2f(d[1], d[0]) * (2.0f / PI_F);
    float u, v;
    if (d[2] < 0.0f) {
        const float r = sqrtf(fmaxf(0.0f, 1.0f + d[2]));
        if (p >= 0.0f) {
            if (p <= 1.0f) { u = 1.0f - r * p; v = 2.0f - r - u; }
            else { u = r * (2.0f - p) - 1.0f; v = 2.0f - r + u; }
        } else {
            if (p >= -1.0f) { u = r * p + 1.0f; v = r - 2.0f + u; }
            else { u = r * (2.0f + p) - 1.0f; v = r - 2.0f - u; }
        }
    } else {
        const float r = sqrtf(fmaxf(0.0f, 1.0f - d[2]));
        if (p >= 0.0f) {
            if (p < 1.0f) { v = r * p; u = r - v; }
            else { v = r * (2.0f - p); u = v - r; }
        } else {
            if (p > -1.0f) { v = r * p; u = r + v; }
            else { v = -r * (2.0f + p); u = -(r + v); }
        }
    }
    uv[0] = 0.5f * (u + 1.0f);
    uv[1] = 0.5f * (v + 1.0f);
}

/* face normals of the regular icosahedron, in the reference's index order
 * (guiding_bin.cuh:40-66) */
static const float kFace20[20][3] = {
    {-0.577350269f, 0.577350269f, 0.577350269f}, {0.000000000f, 0.934172359f, 0.356822090f},
    {0.000000000f, 0.934172359f, -0.356822090f}, {-0.577350269f, 0.577350269f, -0.577350269f},
    {-0.934172359f, 0.356822090f, 0.000000000f}, {0.577350269f, 0.577350269f, 0.577350269f},
    {-0.356822090f, 0.000000000f, 0.934172359f}, {-0.934172359f, -0.356822090f, 0.000000000f},
    {-0.356822090f, 0.000000000f, -0.934172359f}, {0.577350269f, 0.577350269f, -0.577350269f},
    {0.577350269f, -0.577350269f, 0.577350269f}, {0.000000000f, -0.934172359f, 0.356822090f},
    {0.000000000f, -0.934172359f, -0.356822090f}, {0.577350269f, -0.577350269f, -0.577350269f},
    {0.934172359f, -0.356822090f, 0.000000000f}, {0.356822090f, 0.000000000f, 0.934172359f},
    {-0.577350269f, -0.577350269f, 0.577350269f}, {-0.577350269f, -0.577350269f, -0.577350269f},
    {0.356822090f, 0.000000000f, -0.934172359f}, {0.934172359f, 0.356822090f, 0.000000000f},
};

const float* orc_face20_normals(void) { return &kFace20[0][0]; }

static void normalize3f(const float v[3], float o[3])
{
    const float l2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (l2 <= 1e-20f) { o[0] = o[1] = o[2] = 0.0f; return; }
    const float inv = 1.0f / sqrtf(l2);
    o[0] = v[0] * inv; o[1] = v[1] * inv; o[2] = v[2] * inv;
}

int orc_face20(const float n[3])
{
    float d[3];
    normalize3f(n, d);
    int best = 0;
    float bd = -1e30f;
    for (int f = 0; f < 20; ++f) {
        const float v = d[0] * kFace20[f][0] + d[1] * kFace20[f][1] + d[2] * kFace20[f][2];
        if (v > bd) { bd = v; best = f; }
    }
    return best;
}

uint64_t orc_voxel_key(const float p[3], const float dir[3])
{
    int64_t c[3];
    for (int k = 0; k < 3; ++k) {
        const float s = floorf(p[k] * 20.0f);
        if (!(s >= -262144.0f && s < 262144.0f)) return ORC_NO_KEY;
        c[k] = (int64_t)s;
    }
    const uint64_t f = (uint64_t)orc_face20(dir);
    return ((uint64_t)(c[0] + 262144) & 0x7FFFFull) | (((uint64_t)(c[1] + 262144) & 0x7FFFFull) << 19) |
           (((uint64_t)(c[2] + 262144) & 0x7FFFFull) << 38) | (f << 57);
}

int orc_leaf_pixel(const float dir[3])
{
    float d[3], uv[2];
    normalize3f(dir, d);
    orc_dir_to_uv(d, uv);
    int ix = (int)floorf(uv[0] * 8.0f), iy = (int)floorf(uv[1] * 8.0f);
    ix = ix < 0 ? 0 : (ix > 7 ? 7 : ix);
    iy = iy < 0 ? 0 : (iy > 7 ? 7 : iy);
    return iy * 8 + ix;
}

/* texel (x, y) of level l (0 = 8x8 leaves ... 3 = 1x1) */
static int texel_index(int x, int y, int level)
{
    static const int off[4] = { 0, 64, 80, 84 };
    return off[level] + y * (8 >> level) + x;
}

void orc_probe_mips(float* tex)
{
    for (int i = 0; i < 64; ++i)
        for (int c = 0; c < 4; ++c) tex[4 * i + c] = fmaxf(tex[4 * i + c], 1.0e-5f);
    for (int level = 0; level < 3; ++level) {
        const int dr = 8 >> (level + 1);
        for (int y = 0; y < dr; ++y)
            for (int x = 0; x < dr; ++x) {
                const float* A = &tex[4 * texel_index(2 * x, 2 * y, level)];
                const float* B = &tex[4 * texel_index(2 * x + 1, 2 * y, level)];
                const float* C = &tex[4 * texel_index(2 * x, 2 * y + 1, level)];
                const float* D = &tex[4 * texel_index(2 * x + 1, 2 * y + 1, level)];
                float* o = &tex[4 * texel_index(x, y, level + 1)];
                for (int c = 0; c < 4; ++c) o[c] = (A[c] + B[c] + C[c] + D[c]) * 0.25f;
            }
    }
    for (int c = 0; c < 4; ++c) tex[4 * 85 + c] = 0.0f;
}

static float wdot(const float* t, const float q[3]) { return t[0] * q[0] + t[1] * q[1] + t[2] * q[2]; }

float orc_probe_pdf(const float* tex, const float dir[3], const float q[3])
{
    float uv[2];
    orc_dir_to_uv(dir, uv);
    int ix = (int)floorf(uv[0] * 8.0f), iy = (int)floorf(uv[1] * 8.0f);
    ix = ix < 0 ? 0 : (ix > 7 ? 7 : ix);
    iy = iy < 0 ? 0 : (iy > 7 ? 7 : iy);
    const float wl = wdot(&tex[4 * texel_index(ix, iy, 0)], q);
    const float wa = wdot(&tex[4 * texel_index(0, 0, 3)], q);
    const float inv_g = (float)(1.0 / (double)(4.0f * PI_F));
    if (!(wa > 0.0f)) return inv_g;
    return fmaxf((wl / wa) * inv_g, 1.0e-12f);
}

/* u[0..2]: the level picks (top level first), u[3..4]: position in the leaf */
void orc_probe_sample(const float* tex, const float q[3], const float u[5], float dir[3], float* pdf)
{
    int cx = 0, cy = 0;
    float pt = 1.0f;
    for (int level = 3, k = 0; level > 0; --level, ++k) {
        const float ws = 4.0f * wdot(&tex[4 * texel_index(cx, cy, level)], q);
        const int nx = cx << 1, ny = cy << 1;
        const float r = u[k];
        if (ws <= 0.0f) {
            int pick = (int)(r * 4.0f);
            if (pick > 3) pick = 3;
            cx = nx + (pick & 1);
            cy = ny + (pick >> 1);
            pt *= 0.25f;
        } else {
            float cdf = 0.0f;
            int done = 0;
            for (int x = 0; x < 2 && !done; ++x)
                for (int y = 0; y < 2 && !done; ++y) {
                    const float w = wdot(&tex[4 * texel_index(nx + x, ny + y, level - 1)], q) / ws;
                    cdf += w;
                    if (r <= cdf || (x == 1 && y == 1)) {
                        cx = nx + x; cy = ny + y; pt *= w; done = 1;
                    }
                }
        }
    }
    const float su = ((float)cx + (1.0f - u[3])) / 8.0f, sv = ((float)cy + (1.0f - u[4])) / 8.0f;
    orc_uv_to_dir(clamp01(su), clamp01(sv), dir);
    *pdf = fmaxf(pt * 64.0f / (4.0f * PI_F), 1.0e-12f);
}

/* ReflectHemisphereNormal (octahedron.cuh:270-273) */
static void mirror_of(const float w[3], const float n[3], float m[3])
{
    float wn[3], nn[3];
    normalize3f(w, wn);
    normalize3f(n, nn);
    const float d = nn[0] * wn[0] + nn[1] * wn[1] + nn[2] * wn[2];
    for (int k = 0; k < 3; ++k) m[k] = wn[k] - 2.0f * d * nn[k];
}

float orc_probe_hemi_pdf(const float* tex, const float dir[3], const float n[3], const float q[3])
{
    if (dir[0] * n[0] + dir[1] * n[1] + dir[2] * n[2] <= 0.0f) return 0.0f;
    float m[3];
    mirror_of(dir, n, m);
    return orc_probe_pdf(tex, dir, q) + orc_probe_pdf(tex, m, q);
}

void orc_probe_hemi_sample(const float* tex, const float n[3], const float q[3], const float u[5], float dir[3],
                           float* pdf)
{
    orc_probe_sample(tex, q, u, dir, pdf);
    float m[3];
    mirror_of(dir, n, m);
    *pdf += orc_probe_pdf(tex, m, q);
    if (dir[0] * n[0] + dir[1] * n[1] + dir[2] * n[2] < 0.0f) { dir[0] = m[0]; dir[1] = m[1]; dir[2] = m[2]; }
}

/* ------------------------------------------------------------------------ */
/* probe store                                                              */
/* ------------------------------------------------------------------------ */

struct orc_guide {
    int cap, count;
    uint64_t* key;
    float* tex;          /* cap * 86 * 4 */
    uint64_t* n;
    float* acc;          /* cap * 64 * 4, this frame's sums */
    uint64_t* accn;
    uint32_t tcap;
    uint64_t* tkey;
    int32_t* tval;
};

static uint64_t mix64(uint64_t x)
{
    x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27; x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

orc_guide* orc_guide_create(int cap)
{
    orc_guide* g = calloc(1, sizeof *g);
    g->cap = cap;
    g->key = calloc((size_t)cap, sizeof(uint64_t));
    g->tex = calloc((size_t)cap * 86 * 4, sizeof(float));
    g->n = calloc((size_t)cap, sizeof(uint64_t));
    g->acc = calloc((size_t)cap * 64 * 4, sizeof(float));
    g->accn = calloc((size_t)cap, sizeof(uint64_t));
    g->tcap = 1;
    while (g->tcap < 2u * (uint32_t)cap) g->tcap <<= 1;
    g->tkey = malloc(sizeof(uint64_t) * g->tcap);
    g->tval = malloc(sizeof(int32_t) * g->tcap);
    for (uint32_t i = 0; i < g->tcap; ++i) { g->tkey[i] = ORC_NO_KEY; g->tval[i] = -1; }
    return g;
}

void orc_guide_destroy(orc_guide* g)
{
    if (!g) return;
    free(g->key); free(g->tex); free(g->n); free(g->acc); free(g->accn); free(g->tkey); free(g->tval);
    free(g);
}

int orc_guide_count(const orc_guide* g) { return g->count; }

int orc_guide_find(const orc_guide* g, uint64_t key)
{
    if (!g || g->count == 0 || key == ORC_NO_KEY) return -1;
    uint32_t h = (uint32_t)mix64(key) & (g->tcap - 1);
    for (uint32_t k = 0; k < g->tcap; ++k, h = (h + 1) & (g->tcap - 1)) {
        if (g->tkey[h] == key) return g->tval[h];
        if (g->tkey[h] == ORC_NO_KEY) return -1;
    }
    return -1;
}

static int guide_insert(orc_guide* g, uint64_t key)
{
    uint32_t h = (uint32_t)mix64(key) & (g->tcap - 1);
    for (uint32_t k = 0; k < g->tcap; ++k, h = (h + 1) & (g->tcap - 1)) {
        if (g->tkey[h] == key) return g->tval[h];
        if (g->tkey[h] == ORC_NO_KEY) {
            if (g->count >= g->cap) return -1;
            const int i = g->count++;
            g->tkey[h] = key;
            g->tval[h] = i;
            g->key[i] = key;
            memset(&g->tex[(size_t)i * 86 * 4], 0, 86 * 4 * sizeof(float));
            g->n[i] = 0;
            return i;
        }
    }
    return -1;
}

const float* orc_guide_texels(const orc_guide* g, int i) { return &g->tex[(size_t)i * 86 * 4]; }

void orc_guide_export(const orc_guide* g, uint64_t* keys, float* tex, uint64_t* counts)
{
    for (int i = 0; i < g->count; ++i) {
        if (keys) keys[i] = g->key[i];
        if (tex) memcpy(&tex[(size_t)i * 85 * 4], &g->tex[(size_t)i * 86 * 4], 85 * 4 * sizeof(float));
        if (counts) counts[i] = g->n[i];
    }
}

/* One frame's update: samples (key, direction, radiance) in order. */
void orc_guide_update(orc_guide* g, int ns, const uint64_t* keys, const float* dirs, const float* rad)
{
    int* touched = malloc(sizeof(int) * (size_t)(ns > 0 ? ns : 1));
    int nt = 0;
    const float J = 4.0f * PI_F;
    for (int s = 0; s < ns; ++s) {
        if (keys[s] == ORC_NO_KEY) continue;
        const int i = guide_insert(g, keys[s]);
        if (i < 0) continue;
        const int leaf = orc_leaf_pixel(&dirs[3 * s]);
        float* a = &g->acc[((size_t)i * 64 + (size_t)leaf) * 4];
        a[0] += rad[3 * s + 0] * J;
        a[1] += rad[3 * s + 1] * J;
        a[2] += rad[3 * s + 2] * J;
        if (g->accn[i]++ == 0) touched[nt++] = i;
    }
    for (int t = 0; t < nt; ++t) {
        const int i = touched[t];
        const uint64_t nn = g->accn[i], no = g->n[i];
        const double inv = 1.0 / (double)(no + nn);
        const double alpha = (double)no * inv;
        float* tex = &g->tex[(size_t)i * 86 * 4];
        float* a = &g->acc[(size_t)i * 64 * 4];
        for (int k = 0; k < 64; ++k)
            for (int c = 0; c < 3; ++c)
                tex[4 * k + c] = (float)((double)tex[4 * k + c] * alpha + (double)a[4 * k + c] * inv);
        g->n[i] = no + nn < 1000000ull ? no + nn : 1000000ull;
        orc_probe_mips(tex);
        memset(a, 0, 64 * 4 * sizeof(float));
        g->accn[i] = 0;
    }
    free(touched);
}

/* GetPathGuidingProb (render.cuh:195-200): lerp(0, 0.5, min(1, f/64)^0.5) */
float orc_guide_alpha(uint32_t frame)
{
    const float t = powf(fminf(1.0f, (float)frame / 64.0f), 0.5f);
    const float a = 0.0f * (1.0f - t) + 0.5f * t;
    return a < 0.0f ? 0.0f : (a > 1.0f ? 1.0f : a);
}
