/*
 * Oracle (test infrastructure only): BVH build, watertight triangle test and
 * BVH traversal, restated in plain C from the reference and the literature.
 * Header oracle.h lists what pins each function.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013).      */
/* ------------------------------------------------------------------------ */

typedef struct {
    int kx, ky, kz;
    float sx, sy, sz;
    int valid;
} ray_pre;

static ray_pre ray_precompute(const float d[3])
{
    ray_pre p;
    const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > fabsf(d[kz])) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) { int tmp = kx; kx = ky; ky = tmp; }
    p.kx = kx; p.ky = ky; p.kz = kz;
    p.valid = (d[kz] != 0.0f);
    p.sx = d[kx] / d[kz];
    p.sy = d[ky] / d[kz];
    p.sz = 1.0f / d[kz];
    return p;
}

/* Returns 1 when the ray's line crosses the triangle (edge functions of one
 * sign, det != 0) and writes t = T/det, the weights of v1 and v2 and flags:
 * bit0/1/2 = weight of v0/v1/v2 is exactly zero (edge/vertex hit),
 * bit3 = det < 0 (the triangle's winding is clockwise seen from the ray). */
static int tri_test_pre(const float o[3], const ray_pre* p,
                        const float v0[3], const float v1[3], const float v2[3],
                        float* t_out, float* u_out, float* v_out, uint32_t* f_out)
{
    if (!p->valid) return 0;
    const float A[3] = { v0[0] - o[0], v0[1] - o[1], v0[2] - o[2] };
    const float B[3] = { v1[0] - o[0], v1[1] - o[1], v1[2] - o[2] };
    const float C[3] = { v2[0] - o[0], v2[1] - o[1], v2[2] - o[2] };
    const int kx = p->kx, ky = p->ky, kz = p->kz;
    const float Ax = A[kx] - p->sx * A[kz];
    const float Ay = A[ky] - p->sy * A[kz];
    const float Bx = B[kx] - p->sx * B[kz];
    const float By = B[ky] - p->sy * B[kz];
    const float Cx = C[kx] - p->sx * C[kz];
    const float Cy = C[ky] - p->sy * C[kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        /* exact-zero edge function: redo the three products in double */
        const double CxBy = (double)Cx * (double)By;
        const double CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        const double AxCy = (double)Ax * (double)Cy;
        const double AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        const double BxAy = (double)Bx * (double)Ay;
        const double ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return 0;
    const float det = U + V + W;
    if (det == 0.0f) return 0;
    const float Az = p->sz * A[kz];
    const float Bz = p->sz * B[kz];
    const float Cz = p->sz * C[kz];
    const float T = U * Az + V * Bz + W * Cz;
    *t_out = T / det;
    if (u_out) *u_out = V / det;
    if (v_out) *v_out = W / det;
    if (f_out) {
        uint32_t f = 0;
        if (U == 0.0f) f |= 1u;
        if (V == 0.0f) f |= 2u;
        if (W == 0.0f) f |= 4u;
        if (det < 0.0f) f |= 8u;
        *f_out = f;
    }
    return 1;
}

int orc_tri_test(const float org[3], const float dir[3],
                 const float v0[3], const float v1[3], const float v2[3],
                 float* t, float* u, float* v, uint32_t* flags)
{
    const ray_pre p = ray_precompute(dir);
    return tri_test_pre(org, &p, v0, v1, v2, t, u, v, flags);
}

/* ------------------------------------------------------------------------ */
/* BVH build: binned SAH, bvh.cuh:142-422                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    const float *pmin, *pmax, *pcen;
    int32_t* idx;
    int32_t* tmp;
    orc_node* nodes;
    int node_count;
    int leaf_size, bin_count;
} build_state;

static float box_area(const float mn[3], const float mx[3])
{
    /* AABBArea, bvh.cuh:94-98 */
    const float dx = fmaxf(0.0f, mx[0] - mn[0]);
    const float dy = fmaxf(0.0f, mx[1] - mn[1]);
    const float dz = fmaxf(0.0f, mx[2] - mn[2]);
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

static int bin_of(float key, float cmin, float invw, int nb)
{
    int b = invw > 0.0f ? (int)((key - cmin) * invw) : 0;
    return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
}

static const float* g_sort_cen;
static int g_sort_axis;
static int cmp_key_then_index(const void* a, const void* b)
{
    const int32_t ia = *(const int32_t*)a, ib = *(const int32_t*)b;
    const float ka = g_sort_cen[3 * ia + g_sort_axis];
    const float kb = g_sort_cen[3 * ib + g_sort_axis];
    if (ka < kb) return -1;
    if (ka > kb) return 1;
    return (ia > ib) - (ia < ib);
}

static void make_leaf(orc_node* node, int start, int count)
{
    node->left = -1; node->right = -1;
    node->first = start; node->count = count;
}

static int build_rec(build_state* s, int start, int end)
{
    const int count = end - start;
    const int node_idx = s->node_count++;
    orc_node* node = &s->nodes[node_idx];

    float bmin[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, bmax[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
    float cmin[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, cmax[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
    for (int i = start; i < end; ++i) {
        const int p = s->idx[i];
        for (int a = 0; a < 3; ++a) {
            bmin[a] = fminf(bmin[a], s->pmin[3 * p + a]);
            bmax[a] = fmaxf(bmax[a], s->pmax[3 * p + a]);
            cmin[a] = fminf(cmin[a], s->pcen[3 * p + a]);
            cmax[a] = fmaxf(cmax[a], s->pcen[3 * p + a]);
        }
    }
    memcpy(node->bmin, bmin, sizeof bmin);
    memcpy(node->bmax, bmax, sizeof bmax);

    if (count <= s->leaf_size) { make_leaf(node, start, count); return node_idx; }

    /* longest centroid axis, bvh.cuh:227-240 */
    const float ext[3] = { cmax[0] - cmin[0], cmax[1] - cmin[1], cmax[2] - cmin[2] };
    int axis = 0;
    if (ext[1] > ext[0] && ext[1] >= ext[2]) axis = 1;
    else if (ext[2] > ext[0] && ext[2] >= ext[1]) axis = 2;
    if (ext[0] <= 1e-6f && ext[1] <= 1e-6f && ext[2] <= 1e-6f) {
        make_leaf(node, start, count);
        return node_idx;
    }

    const int nb = s->bin_count;
    const float c0 = cmin[axis], c1 = cmax[axis];
    const float invw = (c1 > c0) ? (float)nb / (c1 - c0) : 0.0f;

    float bin_min[64][3], bin_max[64][3];
    int bin_cnt[64];
    for (int b = 0; b < nb; ++b) {
        for (int a = 0; a < 3; ++a) { bin_min[b][a] = FLT_MAX; bin_max[b][a] = -FLT_MAX; }
        bin_cnt[b] = 0;
    }
    for (int i = start; i < end; ++i) {
        const int p = s->idx[i];
        const int b = bin_of(s->pcen[3 * p + axis], c0, invw, nb);
        for (int a = 0; a < 3; ++a) {
            bin_min[b][a] = fminf(bin_min[b][a], s->pmin[3 * p + a]);
            bin_max[b][a] = fmaxf(bin_max[b][a], s->pmax[3 * p + a]);
        }
        bin_cnt[b]++;
    }
    float left_area[64], right_area[64];
    int left_cnt[64], right_cnt[64];
    {
        float rmn[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, rmx[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
        int rc = 0;
        for (int b = 0; b < nb; ++b) {
            if (bin_cnt[b] > 0)
                for (int a = 0; a < 3; ++a) { rmn[a] = fminf(rmn[a], bin_min[b][a]); rmx[a] = fmaxf(rmx[a], bin_max[b][a]); }
            rc += bin_cnt[b];
            left_area[b] = rc > 0 ? box_area(rmn, rmx) : 0.0f;
            left_cnt[b] = rc;
        }
    }
    {
        float rmn[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, rmx[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
        int rc = 0;
        for (int b = nb - 1; b >= 0; --b) {
            if (bin_cnt[b] > 0)
                for (int a = 0; a < 3; ++a) { rmn[a] = fminf(rmn[a], bin_min[b][a]); rmx[a] = fmaxf(rmx[a], bin_max[b][a]); }
            rc += bin_cnt[b];
            right_area[b] = rc > 0 ? box_area(rmn, rmx) : 0.0f;
            right_cnt[b] = rc;
        }
    }
    /* SAH with costI = costT = 1, bvh.cuh:314-333 */
    const float parent_area = box_area(bmin, bmax);
    float best_cost = FLT_MAX;
    int best_split = -1;
    for (int b = 0; b < nb - 1; ++b) {
        const int nl = left_cnt[b], nr = right_cnt[b + 1];
        if (nl == 0 || nr == 0) continue;
        const float sah = 1.0f + (left_area[b] * (float)nl + right_area[b + 1] * (float)nr)
                                     / fmaxf(parent_area, 1e-20f) * 1.0f;
        if (sah < best_cost) { best_cost = sah; best_split = b; }
    }

    int mid = -1;
    if (best_split >= 0) {
        /* partition by bin id (bvh.cuh:336-351); stable here, so the leaf
         * order is deterministic (std::partition's is implementation-defined) */
        int nl = 0, nr = 0;
        for (int i = start; i < end; ++i) {
            const int p = s->idx[i];
            if (bin_of(s->pcen[3 * p + axis], c0, invw, nb) <= best_split) s->idx[start + nl++] = p;
            else s->tmp[nr++] = p;
        }
        memcpy(&s->idx[start + nl], s->tmp, (size_t)nr * sizeof(int32_t));
        mid = start + nl;
        if (mid == start || mid == end) best_split = -1;
    }
    if (best_split < 0) {
        /* median split on the centroid key (bvh.cuh:356-370); ties broken by
         * primitive index so the result is deterministic */
        const int m = start + count / 2;
        g_sort_cen = s->pcen; g_sort_axis = axis;
        qsort(&s->idx[start], (size_t)count, sizeof(int32_t), cmp_key_then_index);
        mid = m;
        if (mid == start || mid == end) { make_leaf(node, start, count); return node_idx; }
    }
    const int l = build_rec(s, start, mid);
    const int r = build_rec(s, mid, end);
    node = &s->nodes[node_idx];
    node->left = l; node->right = r; node->first = -1; node->count = 0;
    return node_idx;
}

int orc_bvh_build(const float* pmin, const float* pmax, const float* pcen,
                  int n, int leaf_size, int bin_count,
                  orc_node* nodes_out, int32_t* indices_out)
{
    if (n <= 0) return 0;
    if (bin_count < 4) bin_count = 4;
    if (bin_count > 64) bin_count = 64;
    if (leaf_size < 1) leaf_size = 1;
    if (leaf_size > 16) leaf_size = 16;
    build_state s;
    s.pmin = pmin; s.pmax = pmax; s.pcen = pcen;
    s.idx = indices_out;
    s.tmp = (int32_t*)malloc((size_t)n * sizeof(int32_t));
    s.nodes = nodes_out;
    s.node_count = 0;
    s.leaf_size = leaf_size; s.bin_count = bin_count;
    for (int i = 0; i < n; ++i) s.idx[i] = i;
    build_rec(&s, 0, n);
    free(s.tmp);
    return s.node_count;
}

void orc_tri_proxies(const float* verts, const uint32_t* idx, int ntri,
                     float* pmin, float* pmax, float* pcen)
{
    for (int i = 0; i < ntri; ++i) {
        const float* a = &verts[3 * idx[3 * i + 0]];
        const float* b = &verts[3 * idx[3 * i + 1]];
        const float* c = &verts[3 * idx[3 * i + 2]];
        for (int k = 0; k < 3; ++k) {
            const float mn = fminf(a[k], fminf(b[k], c[k]));
            const float mx = fmaxf(a[k], fmaxf(b[k], c[k]));
            pmin[3 * i + k] = mn;
            pmax[3 * i + k] = mx;
            pcen[3 * i + k] = 0.5f * (mn + mx);   /* AABBCenter, bvh.cuh:100-103 */
        }
    }
}

static int depth_rec(const orc_node* nodes, int i)
{
    if (nodes[i].count > 0 || nodes[i].left < 0) return 0;
    const int a = depth_rec(nodes, nodes[i].left), b = depth_rec(nodes, nodes[i].right);
    return 1 + (a > b ? a : b);
}

int orc_bvh_depth(const orc_node* nodes, int node_count)
{
    return node_count > 0 ? depth_rec(nodes, 0) : 0;
}

/* ------------------------------------------------------------------------ */
/* Traversal                                                                */
/* ------------------------------------------------------------------------ */

static float inv_safe(float x)
{
    /* invSafe, core/vector.cuh:151-154 */
    return (x == 0.0f) ? (x >= 0.0f ? FLT_MAX : -FLT_MAX) : 1.0f / x;
}

/* FastAABBHitWithEnter (bvh.cuh:446-470).  The reference accepts a box when
 * tExit >= max(0, tMin) and tEnter <= tMax, without requiring
 * tEnter <= tExit, so a ray that passes beside a box still "hits" it until a
 * closer hit shrinks tMax (mode 1 keeps that, mode 0 adds the check).  With
 * pad != 0 the box is grown by pad on every side first; the canonical
 * traversal uses that to make the test conservative with respect to the
 * triangle test's rounding. */
static int box_hit(const float o[3], const float inv[3], const float mn[3], const float mx[3],
                   float tmin, float tmax, float pad, int strict, float* tenter)
{
    const float lx = mn[0] - pad, hx = mx[0] + pad;
    const float ly = mn[1] - pad, hy = mx[1] + pad;
    const float lz = mn[2] - pad, hz = mx[2] + pad;
    const float tx0 = (lx - o[0]) * inv[0], tx1 = (hx - o[0]) * inv[0];
    float te = fminf(tx0, tx1), tx = fmaxf(tx0, tx1);
    const float ty0 = (ly - o[1]) * inv[1], ty1 = (hy - o[1]) * inv[1];
    te = fmaxf(te, fminf(ty0, ty1)); tx = fminf(tx, fmaxf(ty0, ty1));
    const float tz0 = (lz - o[2]) * inv[2], tz1 = (hz - o[2]) * inv[2];
    te = fmaxf(te, fminf(tz0, tz1)); tx = fminf(tx, fmaxf(tz0, tz1));
    if (tx >= fmaxf(0.0f, tmin) && te <= tmax && (!strict || te <= tx)) { *tenter = te; return 1; }
    return 0;
}

/* Padding used by the canonical box test: 2^-16 of the largest coordinate
 * magnitude of the box and of the ray origin.  The triangle test's rounding
 * is relative to |v - o|, so a hit point it reports may lie a few ulps of
 * that magnitude outside the exact triangle; the padded box still holds it. */
static float box_pad(const float mn[3], const float mx[3], float omax)
{
    float m = omax;
    for (int a = 0; a < 3; ++a) { m = fmaxf(m, fabsf(mn[a])); m = fmaxf(m, fabsf(mx[a])); }
    return m * 0x1p-16f + 1e-30f;
}

/* default SDF march tolerance of a ray: SDF_EPS(0) = INTERSECT_EPS(0) / 2 for
 * tmin <= 1e-3, tmin / 2 above (integrator rays carry tmin = INTERSECT_EPS(depth)) */
static float default_meps(float tmin) { return 0.5f * fmaxf(tmin, 1e-3f); }

static void load_ray(orc_rays r, size_t i, float o[3], float d[3], float* tmin, float* tmax, float* meps)
{
    o[0] = r.ox[i]; o[1] = r.oy[i]; o[2] = r.oz[i];
    d[0] = r.dx[i]; d[1] = r.dy[i]; d[2] = r.dz[i];
    *tmin = r.tmin[i]; *tmax = r.tmax[i];
    *meps = r.meps ? r.meps[i] : default_meps(*tmin);
}

static void tri_verts(const float* verts, const uint32_t* idx, int p,
                      const float** a, const float** b, const float** c)
{
    *a = &verts[3 * idx[3 * p + 0]];
    *b = &verts[3 * idx[3 * p + 1]];
    *c = &verts[3 * idx[3 * p + 2]];
}

/* ------------------------------------------------------------------------ */
/* Analytic objects: oriented box and sphere (core/object/aabb.cuh:9-106,   */
/* core/object/sphere.cuh:10-85, SceneObject in core/render/object.cuh).    */
/* ------------------------------------------------------------------------ */

static float dot3(const float a[3], const float b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* C = A * B (float3x3::mul, core/vector.cuh: rows of A dotted with columns of B) */
static void mat3_mul(const float A[9], const float B[9], float C[9])
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            C[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
}

/* normalize (core/vector.cuh:229-238): zero below |v|^2 = 1e-20; the
 * reciprocal square root is taken as 1 / sqrtf (IEEE) */
static void normalize3(float v[3])
{
    const float l2 = dot3(v, v);
    if (l2 <= 1e-20f) { v[0] = v[1] = v[2] = 0.0f; return; }
    const float inv = 1.0f / sqrtf(l2);
    v[0] = v[0] * inv; v[1] = v[1] * inv; v[2] = v[2] * inv;
}

static float max3f(const float v[3]) { return fmaxf(fmaxf(v[0], v[1]), v[2]); }

/* float instance of the SDF march (oracle/sdf_impl.h) */
#define R float
#define SR(name) name##_f32
#define SQRT sqrtf
#define FLOOR floorf
#define FABS fabsf
#include "sdf_impl.h"
#undef R
#undef SR
#undef SQRT
#undef FLOOR
#undef FABS

static int is_sdf(const orc_object* ob) { return ob->type == ORC_OBJ_SDF_TORUS || ob->type == ORC_OBJ_SDF_VOLUME; }

/* SceneObject constructor (object.cuh:286-305): objectToWorld = Rz (Ry Rx)
 * from Euler angles in degrees (float3x3::rotX/Y/Z, vector.cuh:394-425),
 * worldToObject = its transpose; extents are taken by absolute value. */
void orc_object_init(orc_object* ob, uint32_t type, const float center[3], const float extent[3],
                     const float rot_deg[3])
{
    const float k = 3.141592654f / 180.0f;
    const float ax = rot_deg[0] * k, ay = rot_deg[1] * k, az = rot_deg[2] * k;
    const float cx = cosf(ax), sx = sinf(ax), cy = cosf(ay), sy = sinf(ay), cz = cosf(az), sz = sinf(az);
    const float Rx[9] = { 1, 0, 0, 0, cx, -sx, 0, sx, cx };
    const float Ry[9] = { cy, 0, sy, 0, 1, 0, -sy, 0, cy };
    const float Rz[9] = { cz, -sz, 0, sz, cz, 0, 0, 0, 1 };
    float RyRx[9], o2w[9];
    mat3_mul(Ry, Rx, RyRx);
    mat3_mul(Rz, RyRx, o2w);
    memset(ob, 0, sizeof *ob);
    ob->type = type;
    for (int a = 0; a < 3; ++a) { ob->center[a] = center[a]; ob->extent[a] = fabsf(extent[a]); }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) ob->w2o[3 * r + c] = o2w[3 * c + r];
}

/* UpdateProxy (object.cuh:259-284): box and SDF = centre +- sum |axis_k| *
 * extent_k, sphere = centre +- max extent; the proxy centre is the object centre. */
void orc_object_proxy(const orc_object* ob, float pmin[3], float pmax[3], float pcen[3])
{
    float half[3];
    if (ob->type != ORC_OBJ_SPHERE) {
        const float* ax = &ob->w2o[0];   /* objectToWorld * e_x = row 0 of worldToObject */
        const float* ay = &ob->w2o[3];
        const float* az = &ob->w2o[6];
        for (int a = 0; a < 3; ++a)
            half[a] = fabsf(ax[a]) * ob->extent[0] + fabsf(ay[a]) * ob->extent[1] + fabsf(az[a]) * ob->extent[2];
    } else {
        const float r = max3f(ob->extent);
        half[0] = half[1] = half[2] = r;
    }
    for (int a = 0; a < 3; ++a) {
        pmin[a] = ob->center[a] - half[a];
        pmax[a] = ob->center[a] + half[a];
        pcen[a] = ob->center[a];
    }
}

/* GetArea (object.cuh:239-252): the box's (and an SDF object's box's) 8 (ex ey
 * + ex ez + ey ez); the sphere's 4 pi extent.x^2 (extent.x, not the
 * intersected max extent) */
float orc_object_area(const orc_object* ob)
{
    const float ex = ob->extent[0], ey = ob->extent[1], ez = ob->extent[2];
    if (ob->type != ORC_OBJ_SPHERE) return 8.0f * (ex * ey + ex * ez + ey * ez);
    return 4.0f * 3.141592654f * ex * ex;
}

static void obj_local_ray(const orc_object* ob, const float o[3], const float d[3], float lo[3], float ld[3])
{
    const float oc[3] = { o[0] - ob->center[0], o[1] - ob->center[1], o[2] - ob->center[2] };
    for (int r = 0; r < 3; ++r) {
        lo[r] = dot3(&ob->w2o[3 * r], oc);
        ld[r] = dot3(&ob->w2o[3 * r], d);
    }
    normalize3(ld);
}

/* IntersectAABBObject / IntersectSphereObject / IntersectSDFObject
 * (sdf_base.cuh:118-167): the first surface crossing along the normalised
 * local direction with t > 0 (t is measured along normalize(d)); inside = the
 * ray starts inside the object.  An SDF object is marched between the
 * entry and exit of its box. */
int orc_object_test_eps(const orc_object* ob, const float o[3], const float d[3], float meps, float* t, int* inside)
{
    float lo[3], ld[3];
    obj_local_ray(ob, o, d, lo, ld);
    *inside = 0;
    if (ob->type != ORC_OBJ_SPHERE) {
        float te = -INFINITY, tx = INFINITY;
        for (int a = 0; a < 3; ++a) {
            const float inv = inv_safe(ld[a]);
            const float t1 = (-ob->extent[a] - lo[a]) * inv;
            const float t2 = (ob->extent[a] - lo[a]) * inv;
            const float mn = fminf(t1, t2), mx = fmaxf(t1, t2);
            te = a == 0 ? mn : fmaxf(te, mn);
            tx = a == 0 ? mx : fminf(tx, mx);
        }
        if (te > tx) return 0;
        if (tx <= 0.0f) return 0;
        if (is_sdf(ob)) return sdf_march_f32(ob, lo, ld, te, tx, meps, t, inside);
        *inside = te < 0.0f;
        *t = *inside ? tx : te;
        return 1;
    }
    const float r = max3f(ob->extent);
    const float b = dot3(lo, ld);
    const float c = dot3(lo, lo) - r * r;
    const float disc = b * b - c;
    if (disc < 0.0f) return 0;
    const float s = sqrtf(disc);
    float tt = -b - s;
    if (tt <= 0.0f) { tt = -b + s; *inside = 1; }
    if (tt <= 0.0f) return 0;
    *t = tt;
    return 1;
}

int orc_object_test(const orc_object* ob, const float o[3], const float d[3], float* t, int* inside)
{
    return orc_object_test_eps(ob, o, d, 5e-4f, t, inside);
}

/* Hit point and normal of an object hit (the tail of IntersectAABBObject /
 * IntersectSphereObject / IntersectSDFObject): local point lo + ld t; box
 * normal along the axis of largest |p_k| / extent_k, sphere normal p / |p|,
 * both flipped for a hit from inside; SDF normal the distance gradient
 * (outward, never flipped); rotated to world space and normalised. */
void orc_object_surface_eps(const orc_object* ob, const float o[3], const float d[3], float t, int inside,
                            float meps, float pos[3], float n[3])
{
    float lo[3], ld[3], lp[3], ln[3];
    obj_local_ray(ob, o, d, lo, ld);
    for (int a = 0; a < 3; ++a) lp[a] = lo[a] + ld[a] * t;
    if (is_sdf(ob)) {
        sdf_normal_f32(ob, lp, meps, ln);
    } else if (ob->type == ORC_OBJ_BOX) {
        const float rx = fabsf(lp[0]) / ob->extent[0];
        const float ry = fabsf(lp[1]) / ob->extent[1];
        const float rz = fabsf(lp[2]) / ob->extent[2];
        ln[0] = ln[1] = ln[2] = 0.0f;
        if (rx >= ry && rx >= rz) ln[0] = lp[0] > 0.0f ? 1.0f : -1.0f;
        else if (ry >= rz) ln[1] = lp[1] > 0.0f ? 1.0f : -1.0f;
        else ln[2] = lp[2] > 0.0f ? 1.0f : -1.0f;
        if (inside) { ln[0] = -ln[0]; ln[1] = -ln[1]; ln[2] = -ln[2]; }
    } else {
        ln[0] = lp[0]; ln[1] = lp[1]; ln[2] = lp[2];
        normalize3(ln);
        const float sg = inside ? -1.0f : 1.0f;
        ln[0] = ln[0] * sg; ln[1] = ln[1] * sg; ln[2] = ln[2] * sg;
    }
    for (int r = 0; r < 3; ++r) {
        /* objectToWorld row r = column r of worldToObject */
        const float col[3] = { ob->w2o[r], ob->w2o[3 + r], ob->w2o[6 + r] };
        pos[r] = dot3(col, lp) + ob->center[r];
        n[r] = dot3(col, ln);
    }
    normalize3(n);
}

void orc_object_surface(const orc_object* ob, const float o[3], const float d[3], float t, int inside,
                        float pos[3], float n[3])
{
    orc_object_surface_eps(ob, o, d, t, inside, 5e-4f, pos, n);
}

/* OverlapAABBObject / OverlapSphereObject: strict containment of a point;
 * OverlapSDFObject (sdf_base.cuh:169-180): distance <= 0 */
int orc_object_overlap(const orc_object* ob, const float p[3])
{
    const float oc[3] = { p[0] - ob->center[0], p[1] - ob->center[1], p[2] - ob->center[2] };
    float lo[3];
    for (int r = 0; r < 3; ++r) lo[r] = dot3(&ob->w2o[3 * r], oc);
    if (is_sdf(ob)) {
        float u[3];
        sdf_to_unit_f32(ob, lo, u);
        return sdf_unit_f32(ob, u[0], u[1], u[2]) <= 0.0f;
    }
    if (ob->type == ORC_OBJ_BOX)
        return -ob->extent[0] < lo[0] && -ob->extent[1] < lo[1] && -ob->extent[2] < lo[2] &&
               ob->extent[0] > lo[0] && ob->extent[1] > lo[1] && ob->extent[2] > lo[2];
    const float r = max3f(ob->extent);
    return dot3(lo, lo) < r * r;
}

void orc_object_set_sdf(orc_object* ob, const float params[4], const float* vol, const uint32_t dims[3])
{
    for (int k = 0; k < 4; ++k) ob->sdf[k] = params ? params[k] : 0.0f;
    ob->vol = vol;
    for (int k = 0; k < 3; ++k) ob->dims[k] = dims ? dims[k] : 0u;
    if (dims) {
        const float ix = 1.0f / fmaxf((float)dims[0], 1.0f), iy = 1.0f / fmaxf((float)dims[1], 1.0f),
                    iz = 1.0f / fmaxf((float)dims[2], 1.0f);
        ob->vmax = sqrtf(ix * ix + iy * iy + iz * iz) * 2.5f;
    }
}

float orc_sdf_value(const orc_object* ob, const float u[3])
{
    return sdf_unit_f32(ob, u[0], u[1], u[2]);
}

/* PlaneSDF (sdf_diamond.cuh:15-19) */
static float plane_sdf(const float p[3], float nx, float ny, float nz, float d)
{
    const float len = fmaxf(sqrtf(nx * nx + ny * ny + nz * nz), 1e-8f);
    return (p[0] * nx + p[1] * ny + p[2] * nz - d) / len;
}

/* SdfDiamondUnit (sdf_diamond.cuh:21-66): a brilliant cut as the maximum of
 * its facet planes, in the sector of the azimuth of p */
static float sdf_diamond(const float u[3], const float prm[4])
{
    const float dd = prm[0] != 0.0f ? prm[0] : 0.94f;
    const float bias = prm[1] != 0.0f ? prm[1] : 0.5f;
    const float table = prm[2] != 0.0f ? prm[2] : 0.3f;
    const float scale = prm[3] != 0.0f ? prm[3] : 1.0f;
    const float si = 1.0f / scale;
    const float p[3] = { u[0] * si, u[1] * si, u[2] * si };
    const float af2 = 4.0f / 3.14159265358979f;
    const float s = atan2f(p[1], p[0]);
    const float sf = floorf(s * af2 + bias) / af2;
    const float sf2 = floorf(s * af2) / af2;
    float m = -1e20f;
    m = fmaxf(m, plane_sdf(p, cosf(sf), sinf(sf), 1.444f, dd));
    m = fmaxf(m, plane_sdf(p, cosf(sf), sinf(sf), -1.072f, dd));
    m = fmaxf(m, plane_sdf(p, 0.0f, 0.0f, 1.0f, table));
    m = fmaxf(m, plane_sdf(p, 0.0f, 0.0f, -1.0f, 0.865f));
    m = fmaxf(m, plane_sdf(p, cosf(s), sinf(s), 0.0f, 0.911f));
    m = fmaxf(m, plane_sdf(p, cosf(sf + 0.21f), sinf(sf + 0.21f), -1.02f, 0.9193f));
    m = fmaxf(m, plane_sdf(p, cosf(sf - 0.21f), sinf(sf - 0.21f), -1.02f, 0.9193f));
    m = fmaxf(m, plane_sdf(p, cosf(sf - 0.21f), sinf(sf - 0.21f), 1.03f, 0.912f));
    m = fmaxf(m, plane_sdf(p, cosf(sf + 0.21f), sinf(sf + 0.21f), 1.03f, 0.912f));
    m = fmaxf(m, plane_sdf(p, cosf(sf2 + 0.393f), sinf(sf2 + 0.393f), 2.21f, 1.131f));
    return m * scale;
}

void orc_sdf_bake(uint32_t shape, const float params[4], const uint32_t dims[3], float* out)
{
    orc_object tor;
    memset(&tor, 0, sizeof tor);
    tor.type = ORC_OBJ_SDF_TORUS;
    tor.sdf[0] = params[0]; tor.sdf[1] = params[1];
    for (uint32_t z = 0; z < dims[2]; ++z)
        for (uint32_t y = 0; y < dims[1]; ++y)
            for (uint32_t x = 0; x < dims[0]; ++x) {
                const float p01[3] = { dims[0] > 1 ? (float)x / (float)(dims[0] - 1) : 0.0f,
                                       dims[1] > 1 ? (float)y / (float)(dims[1] - 1) : 0.0f,
                                       dims[2] > 1 ? (float)z / (float)(dims[2] - 1) : 0.0f };
                const float u[3] = { p01[0] - 0.5f, p01[1] - 0.5f, p01[2] - 0.5f };
                out[((size_t)z * dims[1] + y) * dims[0] + x] =
                    shape == 0 ? sdf_torus_f32(&tor, u[0], u[1], u[2]) : sdf_diamond(u, params);
            }
}

void orc_geom_proxies(const orc_geom* g, float* pmin, float* pmax, float* pcen)
{
    if (g->ntri) orc_tri_proxies(g->verts, g->idx, (int)g->ntri, pmin, pmax, pcen);
    for (uint32_t k = 0; k < g->nobj; ++k) {
        const size_t i = 3 * ((size_t)g->ntri + k);
        orc_object_proxy(&g->objs[k], &pmin[i], &pmax[i], &pcen[i]);
    }
}

/* Primitive p of the geometry (triangles first, then objects): 1 when the
 * ray's line (triangle) or the forward ray (object) crosses it, with t. */
static int prim_test(const orc_geom* g, uint32_t p, const float o[3], const float d[3], const ray_pre* pre,
                     float meps, float* t)
{
    if (p < g->ntri) {
        const float *a, *b, *c;
        tri_verts(g->verts, g->idx, (int)p, &a, &b, &c);
        return tri_test_pre(o, pre, a, b, c, t, NULL, NULL, NULL);
    }
    int inside;
    return orc_object_test_eps(&g->objs[p - g->ntri], o, d, meps, t, &inside);
}

#define ORC_STACK 128

/* One ray.  any = 1: stop at the first accepted hit.  Returns 1 on a hit. */
static int traverse_one(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
                        const float o[3], const float d[3], float tmin, float tmax, float meps,
                        int mode, int any,
                        uint32_t* prim_out, float* t_out, uint32_t* nbox, uint32_t* ntri)
{
    uint32_t nb = 0, nt = 0;
    int have = 0;
    uint32_t best_p = ORC_MISS;
    float best_t = tmax;
    if (node_count > 0 && tmax >= tmin) {
        const ray_pre pre = ray_precompute(d);
        const float inv[3] = { inv_safe(d[0]), inv_safe(d[1]), inv_safe(d[2]) };
        const float omax = fmaxf(fabsf(o[0]), fmaxf(fabsf(o[1]), fabsf(o[2])));
        int stack[ORC_STACK];
        int sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const int ni = stack[--sp];
            const orc_node* nd = &nodes[ni];
            float te;
            nb++;
            const float pad = mode == 0 ? box_pad(nd->bmin, nd->bmax, omax) : 0.0f;
            if (!box_hit(o, inv, nd->bmin, nd->bmax, tmin, best_t, pad, mode == 0, &te)) continue;
            if (nd->count > 0) {
                for (int k = 0; k < nd->count; ++k) {
                    const uint32_t p = (uint32_t)indices[nd->first + k];
                    float t;
                    nt++;
                    if (!prim_test(g, p, o, d, &pre, meps, &t)) continue;
                    int accept;
                    if (mode == 0)
                        accept = t >= tmin && (have ? (t < best_t || (t == best_t && p < best_p)) : t < tmax);
                    else
                        accept = t >= tmin && t < best_t;     /* intersection.cuh:73 */
                    if (accept) {
                        have = 1; best_t = t; best_p = p;
                        if (any) goto done;
                    }
                }
            } else {
                const orc_node* L = &nodes[nd->left];
                const orc_node* R = &nodes[nd->right];
                float tl = 0.0f, tr = 0.0f;
                nb += 2;
                const float pl = mode == 0 ? box_pad(L->bmin, L->bmax, omax) : 0.0f;
                const float pr = mode == 0 ? box_pad(R->bmin, R->bmax, omax) : 0.0f;
                const int hl = box_hit(o, inv, L->bmin, L->bmax, tmin, best_t, pl, mode == 0, &tl);
                const int hr = box_hit(o, inv, R->bmin, R->bmax, tmin, best_t, pr, mode == 0, &tr);
                /* sorted push, intersection.cuh:218-243 (near child popped first) */
                if (hl && hr) {
                    if (sp + 2 > ORC_STACK) abort();
                    if (tl > tr) { stack[sp++] = nd->left; stack[sp++] = nd->right; }
                    else { stack[sp++] = nd->right; stack[sp++] = nd->left; }
                } else if (hl) {
                    if (sp + 1 > ORC_STACK) abort();
                    stack[sp++] = nd->left;
                } else if (hr) {
                    if (sp + 1 > ORC_STACK) abort();
                    stack[sp++] = nd->right;
                }
            }
        }
    }
done:
    if (nbox) *nbox = nb;
    if (ntri) *ntri = nt;
    *prim_out = have ? best_p : ORC_MISS;
    *t_out = best_t;
    return have;
}

/* Full hit record: triangles give barycentrics and edge flags, objects
 * u = v = 0 and ORC_HIT_INSIDE for a hit from inside. */
static void hit_record(const orc_geom* g, const float o[3], const float d[3], float meps, uint32_t p,
                       float* t, float* u, float* v, uint32_t* flags)
{
    *u = 0.0f; *v = 0.0f; *flags = 0u;
    if (p < g->ntri) {
        const float *a, *b, *c;
        tri_verts(g->verts, g->idx, (int)p, &a, &b, &c);
        const ray_pre pre = ray_precompute(d);
        tri_test_pre(o, &pre, a, b, c, t, u, v, flags);
    } else {
        int inside = 0;
        orc_object_test_eps(&g->objs[p - g->ntri], o, d, meps, t, &inside);
        if (inside) *flags = ORC_HIT_INSIDE;
    }
}

static void finish_hit(const orc_geom* g, const float o[3], const float d[3], float meps, uint32_t p,
                       orc_hits hits, size_t i)
{
    hits.prim[i] = p;
    if (p == ORC_MISS) {
        hits.t[i] = INFINITY; hits.u[i] = 0.0f; hits.v[i] = 0.0f; hits.flags[i] = 0u;
        return;
    }
    float t = 0.0f, u = 0.0f, v = 0.0f;
    uint32_t f = 0;
    hit_record(g, o, d, meps, p, &t, &u, &v, &f);
    hits.t[i] = t; hits.u[i] = u; hits.v[i] = v; hits.flags[i] = f;
}

static int nthreads(int threads)
{
#ifdef _OPENMP
    return threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
    return 1;
#endif
}

static orc_geom tri_geom(const float* verts, const uint32_t* idx, int ntri)
{
    orc_geom g = { verts, idx, (uint32_t)(ntri > 0 ? ntri : 0), NULL, 0 };
    return g;
}

void orc_trace_closest_g(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
                         orc_rays rays, size_t n, orc_hits hits, int mode,
                         uint32_t* boxes, uint32_t* tris, int threads)
{
    const int nt = nthreads(threads);
    (void)nt;
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nt)
    for (long long ii = 0; ii < (long long)n; ++ii) {
        const size_t i = (size_t)ii;
        float o[3], d[3], tmin, tmax, meps, t;
        uint32_t p, nb, ntr;
        load_ray(rays, i, o, d, &tmin, &tmax, &meps);
        traverse_one(nodes, node_count, indices, g, o, d, tmin, tmax, meps, mode, 0, &p, &t, &nb, &ntr);
        if (boxes) boxes[i] = nb;
        if (tris) tris[i] = ntr;
        finish_hit(g, o, d, meps, p, hits, i);
    }
}

void orc_trace_any_g(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
                     orc_rays rays, size_t n, uint32_t* occluded_bits, int mode,
                     uint32_t* boxes, uint32_t* tris, int threads)
{
    const int nt = nthreads(threads);
    (void)nt;
    const long long nw = (long long)((n + 31) / 32);
#pragma omp parallel for schedule(dynamic, 32) num_threads(nt)
    for (long long w = 0; w < nw; ++w) {
        uint32_t word = 0;
        for (int b = 0; b < 32; ++b) {
            const size_t i = (size_t)w * 32 + (size_t)b;
            if (i >= n) break;
            float o[3], d[3], tmin, tmax, meps, t;
            uint32_t p, nb, ntr;
            load_ray(rays, i, o, d, &tmin, &tmax, &meps);
            if (traverse_one(nodes, node_count, indices, g, o, d, tmin, tmax, meps, mode, 1, &p, &t, &nb, &ntr))
                word |= 1u << b;
            if (boxes) boxes[i] = nb;
            if (tris) tris[i] = ntr;
        }
        occluded_bits[w] = word;
    }
}

void orc_trace_brute_g(const orc_geom* g, orc_rays rays, size_t n, orc_hits hits, int threads)
{
    const int nt = nthreads(threads);
    (void)nt;
    const uint32_t np = g->ntri + g->nobj;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (long long ii = 0; ii < (long long)n; ++ii) {
        const size_t i = (size_t)ii;
        float o[3], d[3], tmin, tmax, meps;
        load_ray(rays, i, o, d, &tmin, &tmax, &meps);
        const ray_pre pre = ray_precompute(d);
        int have = 0;
        uint32_t best_p = ORC_MISS;
        float best_t = tmax;
        for (uint32_t p = 0; p < np; ++p) {
            float t;
            if (!prim_test(g, p, o, d, &pre, meps, &t)) continue;
            if (t >= tmin && (have ? t < best_t : t < tmax)) {   /* p ascends: ties keep the smaller id */
                have = 1; best_t = t; best_p = p;
            }
        }
        finish_hit(g, o, d, meps, have ? best_p : ORC_MISS, hits, i);
    }
}

void orc_trace_any_brute_g(const orc_geom* g, orc_rays rays, size_t n, uint32_t* occluded_bits, int threads)
{
    const int nt = nthreads(threads);
    (void)nt;
    const long long nw = (long long)((n + 31) / 32);
    const uint32_t np = g->ntri + g->nobj;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
    for (long long w = 0; w < nw; ++w) {
        uint32_t word = 0;
        for (int b = 0; b < 32; ++b) {
            const size_t i = (size_t)w * 32 + (size_t)b;
            if (i >= n) break;
            float o[3], d[3], tmin, tmax, meps;
            load_ray(rays, i, o, d, &tmin, &tmax, &meps);
            const ray_pre pre = ray_precompute(d);
            for (uint32_t p = 0; p < np; ++p) {
                float t;
                if (prim_test(g, p, o, d, &pre, meps, &t) && t >= tmin && t < tmax) {
                    word |= 1u << b;
                    break;
                }
            }
        }
        occluded_bits[w] = word;
    }
}

void orc_trace_closest(const orc_node* nodes, int node_count, const int32_t* indices,
                       const float* verts, const uint32_t* idx,
                       orc_rays rays, size_t n, orc_hits hits, int mode,
                       uint32_t* boxes, uint32_t* tris, int threads)
{
    /* triangle count is not needed by the traversal: every index is a triangle */
    const orc_geom g = tri_geom(verts, idx, 0x7FFFFFFF);
    orc_trace_closest_g(nodes, node_count, indices, &g, rays, n, hits, mode, boxes, tris, threads);
}

void orc_trace_any(const orc_node* nodes, int node_count, const int32_t* indices,
                   const float* verts, const uint32_t* idx,
                   orc_rays rays, size_t n, uint32_t* occluded_bits, int mode,
                   uint32_t* boxes, uint32_t* tris, int threads)
{
    const orc_geom g = tri_geom(verts, idx, 0x7FFFFFFF);
    orc_trace_any_g(nodes, node_count, indices, &g, rays, n, occluded_bits, mode, boxes, tris, threads);
}

void orc_trace_brute(const float* verts, const uint32_t* idx, int ntri,
                     orc_rays rays, size_t n, orc_hits hits, int threads)
{
    const orc_geom g = tri_geom(verts, idx, ntri);
    orc_trace_brute_g(&g, rays, n, hits, threads);
}

void orc_trace_any_brute(const float* verts, const uint32_t* idx, int ntri,
                         orc_rays rays, size_t n, uint32_t* occluded_bits, int threads)
{
    const orc_geom g = tri_geom(verts, idx, ntri);
    orc_trace_any_brute_g(&g, rays, n, occluded_bits, threads);
}

/* Single-ray entry points used by the oracle integrator (render.c). */
int orc_closest1_ge(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
                    const float o[3], const float d[3], float tmin, float tmax, float meps,
                    uint32_t* prim, float* t, float* u, float* v, uint32_t* flags)
{
    float tt;
    uint32_t p;
    const int hit = traverse_one(nodes, node_count, indices, g, o, d, tmin, tmax, meps, 0, 0, &p, &tt, NULL, NULL);
    *prim = p;
    *t = tt;
    *flags = 0u;
    if (hit) hit_record(g, o, d, meps, p, t, u, v, flags);
    return hit;
}

int orc_any1_ge(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
                const float o[3], const float d[3], float tmin, float tmax, float meps)
{
    float tt;
    uint32_t p;
    return traverse_one(nodes, node_count, indices, g, o, d, tmin, tmax, meps, 0, 1, &p, &tt, NULL, NULL);
}

int orc_closest1_g(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
                   const float o[3], const float d[3], float tmin, float tmax,
                   uint32_t* prim, float* t, float* u, float* v, uint32_t* flags)
{
    return orc_closest1_ge(nodes, node_count, indices, g, o, d, tmin, tmax, default_meps(tmin), prim, t, u, v,
                           flags);
}

int orc_any1_g(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
               const float o[3], const float d[3], float tmin, float tmax)
{
    return orc_any1_ge(nodes, node_count, indices, g, o, d, tmin, tmax, default_meps(tmin));
}

int orc_closest1(const orc_node* nodes, int node_count, const int32_t* indices, const float* verts,
                 const uint32_t* idx, const float o[3], const float d[3], float tmin, float tmax,
                 uint32_t* prim, float* t, float* u, float* v)
{
    const orc_geom g = tri_geom(verts, idx, 0x7FFFFFFF);
    uint32_t f;
    return orc_closest1_g(nodes, node_count, indices, &g, o, d, tmin, tmax, prim, t, u, v, &f);
}

int orc_any1(const orc_node* nodes, int node_count, const int32_t* indices, const float* verts,
             const uint32_t* idx, const float o[3], const float d[3], float tmin, float tmax)
{
    const orc_geom g = tri_geom(verts, idx, 0x7FFFFFFF);
    return orc_any1_g(nodes, node_count, indices, &g, o, d, tmin, tmax);
}
