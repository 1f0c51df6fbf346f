/*
 * libar_cpu.so: the oracle behind the C ABI of include/ar.h.  Test
 * infrastructure only: a program written against ar.h (or the SceneRenderer
 * wrapper, include/ar/scene_renderer.hpp) runs on the CPU restatement when it
 * links this library instead of libar_hip.so, so the tests can swap backends
 * under one program.
 *
 * Every pointer ar.h calls a device pointer is a host pointer here; streams
 * are ignored and every call has finished when it returns.
 *
 * Implemented (same arguments, results and error codes as libar_hip.so):
 *   context, scene (meshes, analytic objects, SDF volumes), ar_bvh_build (the
 *   oracle's binned-SAH BVH2: widths 2, 4 and 8 are accepted, the tree is binary),
 *   ar_trace_closest / _counted / _noalloc and ar_trace_any (canonical hit
 *   rule: the same hits as the GPU, bit for bit), ar_render, ar_render_host
 *   and ar_render_progressive in RGB (the FP32 oracle integrator; voxel
 *   guiding with one sample per frame), batches (each scene rendered alone,
 *   which is what a GPU batch reproduces), ar_tonemap, ar_present,
 *   ar_denoise, ar_luminance_mse, ar_write_bmp, and the voxel-guiding
 *   configure / reset / info / snapshot calls.
 * Everything else returns AR_E_UNSUPPORTED: spectral modes (their uplift
 * tables are built by libar_hip.so), learned guiding, guided frames of more
 * than one sample, graphs and autotuning, the spectral network, shadow-ray
 * generation and probe restore.
 */
#include "../include/ar.h"
#include "oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
    float* data;
    uint32_t dims[3];
} cpu_volume;

struct ar_ctx {
    char err[512];
    /* scene */
    float* verts;
    uint32_t nv;
    uint32_t* idx;
    uint32_t ntri;
    uint32_t* tri_mat;
    orc_material* mats;
    uint32_t nmats;
    int scene_set;
    ar_object* aobj;
    orc_object* obj;
    orc_material* omat;
    uint32_t nobj;
    cpu_volume* vols;
    uint32_t nvol;
    /* BVH */
    orc_node* nodes;
    int32_t* bidx;
    int node_count;
    int bvh_ready;
    ar_bvh_info info;
    /* voxel guiding */
    orc_guide* guide;
    uint64_t guide_cap;
    /* ar_present's denoised image */
    float* den;
    size_t den_px;
    /* batch: the scene contexts */
    ar_ctx** batch;
    uint32_t nbatch;
};

/* ar_material and orc_material share one layout */
typedef char material_layout_check[sizeof(ar_material) == sizeof(orc_material) ? 1 : -1];

static int set_error(ar_ctx* c, int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof c->err, fmt, ap);
    va_end(ap);
    return code;
}

static int finite3(const float v[3]) { return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]); }

/* every float of a material is finite (the HIP library refuses the rest) */
static int material_finite(const ar_material* m)
{
    return finite3(m->albedo) && finite3(m->emission) && isfinite(m->specular) && isfinite(m->metallic) &&
           isfinite(m->roughness) && isfinite(m->ior) && isfinite(m->sigma_a) && isfinite(m->abbe) && isfinite(m->tilt);
}

/* ------------------------------------------------------------------ context */

int ar_abi_version(void) { return AR_ABI_VERSION; }

int ar_create(ar_ctx** out, const ar_device_opts* opts)
{
    (void)opts;
    if (!out) return AR_E_INVALID;
    ar_ctx* c = calloc(1, sizeof *c);
    if (!c) return AR_E_OOM;
    c->guide_cap = 1u << 20;
    *out = c;
    return AR_OK;
}

static void free_bvh(ar_ctx* c)
{
    free(c->nodes);
    free(c->bidx);
    c->nodes = NULL;
    c->bidx = NULL;
    c->node_count = 0;
    c->bvh_ready = 0;
}

static void drop_guide(ar_ctx* c)
{
    if (c->guide) orc_guide_destroy(c->guide);
    c->guide = NULL;
}

int ar_destroy(ar_ctx* c)
{
    if (!c) return AR_E_INVALID;
    free_bvh(c);
    drop_guide(c);
    free(c->verts); free(c->idx); free(c->tri_mat); free(c->mats);
    free(c->aobj); free(c->obj); free(c->omat);
    for (uint32_t k = 0; k < c->nvol; ++k) free(c->vols[k].data);
    free(c->vols);
    free(c->den);
    free(c->batch);
    free(c);
    return AR_OK;
}

const char* ar_last_error(const ar_ctx* c) { return c ? c->err : "null context"; }

/* ------------------------------------------------------------------ scene */

int ar_scene_set(ar_ctx* c, const ar_mesh* meshes, size_t nmesh)
{
    if (!c) return AR_E_INVALID;
    if (nmesh && !meshes) return set_error(c, AR_E_INVALID, "meshes is NULL");
    size_t nv = 0, nt = 0, nm = 0;
    for (size_t m = 0; m < nmesh; ++m) {
        const ar_mesh* me = &meshes[m];
        if (me->ntri && (!me->vertices || !me->indices))
            return set_error(c, AR_E_INVALID, "mesh %zu: NULL vertices or indices", m);
        for (size_t k = 0; k < 3 * (size_t)me->nv; ++k)
            if (!isfinite(me->vertices[k])) return set_error(c, AR_E_INVALID, "mesh %zu: non-finite vertex", m);
        for (size_t t = 0; t < me->ntri; ++t) {
            for (int k = 0; k < 3; ++k)
                if (me->indices[3 * t + k] >= me->nv)
                    return set_error(c, AR_E_INVALID, "mesh %zu: index out of range", m);
            const uint32_t mi = me->tri_material ? me->tri_material[t] : 0;
            if (me->nmaterials && mi >= me->nmaterials)
                return set_error(c, AR_E_INVALID, "mesh %zu: material out of range", m);
        }
        if (me->nmaterials && !me->materials) return set_error(c, AR_E_INVALID, "materials is NULL");
        for (uint32_t k = 0; k < me->nmaterials; ++k) {
            if (!material_finite(&me->materials[k]))
                return set_error(c, AR_E_INVALID, "mesh %zu: material %u has a non-finite value", m, k);
            if (me->materials[k].model > AR_MODEL_GGX)
                return set_error(c, me->materials[k].model == AR_MODEL_GLASS ? AR_E_UNSUPPORTED : AR_E_INVALID,
                                 "mesh %zu: triangle materials are Lambert or GGX (glass is for analytic objects)", m);
        }
        nv += me->nv;
        nt += me->ntri;
        nm += me->nmaterials;
    }
    if (nt > 0x7FFFFFFu) return set_error(c, AR_E_CAPACITY, "more than 2^27-1 triangles");
    float* verts = malloc(sizeof(float) * 3 * (nv ? nv : 1));
    uint32_t* idx = malloc(sizeof(uint32_t) * 3 * (nt ? nt : 1));
    uint32_t* tm = malloc(sizeof(uint32_t) * (nt ? nt : 1));
    orc_material* mats = calloc(nm ? nm : 1, sizeof(orc_material));
    if (!verts || !idx || !tm || !mats) {
        free(verts); free(idx); free(tm); free(mats);
        return set_error(c, AR_E_OOM, "host allocation failed");
    }
    size_t vb = 0, tb = 0, mb = 0;
    for (size_t m = 0; m < nmesh; ++m) {
        const ar_mesh* me = &meshes[m];
        if (me->nv) memcpy(verts + 3 * vb, me->vertices, sizeof(float) * 3 * me->nv);
        for (size_t t = 0; t < me->ntri; ++t) {
            for (int k = 0; k < 3; ++k) idx[3 * (tb + t) + k] = (uint32_t)vb + me->indices[3 * t + k];
            const uint32_t mi = me->tri_material ? me->tri_material[t] : 0;
            tm[tb + t] = me->nmaterials ? (uint32_t)mb + mi : 0;
        }
        if (me->nmaterials) memcpy(mats + mb, me->materials, sizeof(orc_material) * me->nmaterials);
        vb += me->nv;
        tb += me->ntri;
        mb += me->nmaterials;
    }
    if (nm == 0) {   /* the default material: grey Lambert */
        mats[0].albedo[0] = mats[0].albedo[1] = mats[0].albedo[2] = 0.75f;
        mats[0].roughness = 1.0f;
        nm = 1;
    }
    free(c->verts); free(c->idx); free(c->tri_mat); free(c->mats);
    c->verts = verts; c->nv = (uint32_t)nv;
    c->idx = idx; c->ntri = (uint32_t)nt;
    c->tri_mat = tm;
    c->mats = mats; c->nmats = (uint32_t)nm;
    c->scene_set = 1;
    free_bvh(c);
    drop_guide(c);   /* probes describe the previous scene's light field */
    return AR_OK;
}

/* orc_object of every ar_object (volumes looked up by index) */
static void bind_objects(ar_ctx* c)
{
    for (uint32_t k = 0; k < c->nobj; ++k) {
        const ar_object* a = &c->aobj[k];
        orc_object_init(&c->obj[k], a->type, a->center, a->extent, a->rotation);
        if (a->type == AR_OBJ_SDF_TORUS) {
            orc_object_set_sdf(&c->obj[k], a->sdf, NULL, NULL);
        } else if (a->type == AR_OBJ_SDF_VOLUME) {
            const cpu_volume* v = &c->vols[a->sdf_volume];
            orc_object_set_sdf(&c->obj[k], a->sdf, v->data, v->dims);
        }
        memcpy(&c->omat[k], &a->material, sizeof(orc_material));
    }
}

int ar_scene_set_objects(ar_ctx* c, const ar_object* objects, size_t nobj)
{
    if (!c) return AR_E_INVALID;
    if (nobj && !objects) return set_error(c, AR_E_INVALID, "objects is NULL");
    if (nobj > 0xFFFFFFu) return set_error(c, AR_E_CAPACITY, "too many objects");
    for (size_t k = 0; k < nobj; ++k) {
        const ar_object* o = &objects[k];
        if (o->type > AR_OBJ_SDF_VOLUME) return set_error(c, AR_E_INVALID, "object %zu: unknown type", k);
        if (!finite3(o->center) || !finite3(o->extent) || !finite3(o->rotation))
            return set_error(c, AR_E_INVALID, "object %zu: non-finite center, extent or rotation", k);
        const float ex = fabsf(o->extent[0]), ey = fabsf(o->extent[1]), ez = fabsf(o->extent[2]);
        const int empty = o->type != AR_OBJ_SPHERE ? !(ex > 0.0f && ey > 0.0f && ez > 0.0f)
                                                   : !(fmaxf(ex, fmaxf(ey, ez)) > 0.0f);
        if (empty) return set_error(c, AR_E_INVALID, "object %zu: empty extent", k);
        if (o->type == AR_OBJ_SDF_TORUS && !(isfinite(o->sdf[0]) && isfinite(o->sdf[1])))
            return set_error(c, AR_E_INVALID, "object %zu: non-finite torus radii", k);
        if (o->type == AR_OBJ_SDF_VOLUME && (o->sdf_volume < 0 || (uint32_t)o->sdf_volume >= c->nvol))
            return set_error(c, AR_E_INVALID, "object %zu: no such SDF volume", k);
        const float* e = o->material.emission;
        if ((o->type == AR_OBJ_SDF_TORUS || o->type == AR_OBJ_SDF_VOLUME) && (e[0] > 0.0f || e[1] > 0.0f || e[2] > 0.0f))
            return set_error(c, AR_E_UNSUPPORTED, "object %zu: SDF objects cannot be lights", k);
        if (!material_finite(&o->material)) return set_error(c, AR_E_INVALID, "object %zu: non-finite material value", k);
        if (o->material.model > AR_MODEL_GLASS) return set_error(c, AR_E_INVALID, "object %zu: unknown material model", k);
        if (o->material.model == AR_MODEL_GLASS && !(o->material.ior > 0.0f))
            return set_error(c, AR_E_INVALID, "object %zu: glass needs ior > 0", k);
    }
    free(c->aobj); free(c->obj); free(c->omat);
    c->aobj = malloc(sizeof(ar_object) * (nobj ? nobj : 1));
    c->obj = calloc(nobj ? nobj : 1, sizeof(orc_object));
    c->omat = calloc(nobj ? nobj : 1, sizeof(orc_material));
    if (!c->aobj || !c->obj || !c->omat) return set_error(c, AR_E_OOM, "host allocation failed");
    if (nobj) memcpy(c->aobj, objects, sizeof(ar_object) * nobj);
    c->nobj = (uint32_t)nobj;
    bind_objects(c);
    free_bvh(c);
    drop_guide(c);
    return AR_OK;
}

/* ------------------------------------------------------------------ SDF volumes */

static int check_dims(ar_ctx* c, const uint32_t dims[3], size_t* n)
{
    if (!dims) return set_error(c, AR_E_INVALID, "dims is NULL");
    for (int k = 0; k < 3; ++k)
        if (dims[k] < 1 || dims[k] > 4096) return set_error(c, AR_E_INVALID, "volume dims must be 1..4096");
    *n = (size_t)dims[0] * dims[1] * dims[2];
    return AR_OK;
}

static int append_volume(ar_ctx* c, float* data, const uint32_t dims[3], uint32_t* index)
{
    cpu_volume* v = realloc(c->vols, sizeof(cpu_volume) * (c->nvol + 1));
    if (!v) { free(data); return set_error(c, AR_E_OOM, "host allocation failed"); }
    c->vols = v;
    v[c->nvol].data = data;
    memcpy(v[c->nvol].dims, dims, sizeof(uint32_t) * 3);
    *index = c->nvol++;
    if (c->nobj) bind_objects(c);   /* the volume array moved */
    return AR_OK;
}

int ar_sdf_volume_bake(ar_ctx* c, uint32_t shape, const float params[4], const uint32_t dims[3], uint32_t* index)
{
    if (!c || !index) return AR_E_INVALID;
    if (shape != AR_SDF_SHAPE_TORUS && shape != AR_SDF_SHAPE_DIAMOND)
        return set_error(c, AR_E_INVALID, "unknown SDF shape");
    size_t n = 0;
    const int rc = check_dims(c, dims, &n);
    if (rc != AR_OK) return rc;
    float* d = malloc(sizeof(float) * n);
    if (!d) return set_error(c, AR_E_OOM, "host allocation failed");
    const float zero[4] = { 0, 0, 0, 0 };
    orc_sdf_bake(shape, params ? params : zero, dims, d);
    return append_volume(c, d, dims, index);
}

int ar_sdf_volume_add(ar_ctx* c, const float* voxels, const uint32_t dims[3], uint32_t* index)
{
    if (!c || !index) return AR_E_INVALID;
    if (!voxels) return set_error(c, AR_E_INVALID, "voxels is NULL");
    size_t n = 0;
    const int rc = check_dims(c, dims, &n);
    if (rc != AR_OK) return rc;
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(voxels[i])) return set_error(c, AR_E_INVALID, "non-finite voxel");
    float* d = malloc(sizeof(float) * n);
    if (!d) return set_error(c, AR_E_OOM, "host allocation failed");
    memcpy(d, voxels, sizeof(float) * n);
    return append_volume(c, d, dims, index);
}

int ar_sdf_volume_read(ar_ctx* c, uint32_t index, float* voxels, uint32_t dims[3])
{
    if (!c) return AR_E_INVALID;
    if (index >= c->nvol) return set_error(c, AR_E_INVALID, "no such volume");
    const cpu_volume* v = &c->vols[index];
    if (dims) memcpy(dims, v->dims, sizeof(uint32_t) * 3);
    if (voxels) memcpy(voxels, v->data, sizeof(float) * v->dims[0] * v->dims[1] * v->dims[2]);
    return AR_OK;
}

int ar_sdf_volume_clear(ar_ctx* c)
{
    if (!c) return AR_E_INVALID;
    for (uint32_t k = 0; k < c->nobj; ++k)
        if (c->aobj[k].type == AR_OBJ_SDF_VOLUME)
            return set_error(c, AR_E_INVALID, "the scene's objects refer to a volume: set the objects first");
    for (uint32_t k = 0; k < c->nvol; ++k) free(c->vols[k].data);
    free(c->vols);
    c->vols = NULL;
    c->nvol = 0;
    return AR_OK;
}

/* ------------------------------------------------------------------ BVH and traversal */

static orc_geom geom_of(const ar_ctx* c)
{
    const orc_geom g = { c->verts, c->idx, c->ntri, c->obj, c->nobj };
    return g;
}

int ar_bvh_build(ar_ctx* c, const ar_bvh_opts* opts)
{
    if (!c) return AR_E_INVALID;
    if (!c->scene_set) return set_error(c, AR_E_NOSCENE, "no scene: call ar_scene_set first");
    ar_bvh_opts o = { 4, 2, 16, 64 };
    if (opts) o = *opts;
    if (o.width != 2 && o.width != 4 && o.width != 8) return set_error(c, AR_E_UNSUPPORTED, "BVH width must be 2, 4 or 8");
    if (o.leaf_size < 1 || o.leaf_size > 15) return set_error(c, AR_E_INVALID, "leaf_size must be 1..15");
    if (o.bin_count < 4 || o.bin_count > 64) return set_error(c, AR_E_INVALID, "bin_count must be 4..64");
    free_bvh(c);
    const clock_t t0 = clock();
    const size_t np = (size_t)c->ntri + c->nobj, cap = np ? np : 1;
    float* pmin = malloc(sizeof(float) * 3 * cap);
    float* pmax = malloc(sizeof(float) * 3 * cap);
    float* pcen = malloc(sizeof(float) * 3 * cap);
    c->nodes = malloc(sizeof(orc_node) * 2 * cap);
    c->bidx = malloc(sizeof(int32_t) * cap);
    if (!pmin || !pmax || !pcen || !c->nodes || !c->bidx) {
        free(pmin); free(pmax); free(pcen);
        free_bvh(c);
        return set_error(c, AR_E_OOM, "host allocation failed");
    }
    const orc_geom g = geom_of(c);
    orc_geom_proxies(&g, pmin, pmax, pcen);
    c->node_count = np ? orc_bvh_build(pmin, pmax, pcen, (int)np, (int)o.leaf_size, (int)o.bin_count, c->nodes,
                                       c->bidx)
                       : 0;
    free(pmin); free(pmax); free(pcen);
    memset(&c->info, 0, sizeof c->info);
    c->info.node_count = (uint64_t)c->node_count;
    for (int i = 0; i < c->node_count; ++i) c->info.leaf_count += c->nodes[i].count > 0;
    c->info.depth = c->node_count ? (uint32_t)orc_bvh_depth(c->nodes, c->node_count) : 0;
    c->info.width = 2;
    c->info.build_seconds = (double)(clock() - t0) / CLOCKS_PER_SEC;
    c->bvh_ready = 1;
    return AR_OK;
}

int ar_bvh_info_get(const ar_ctx* c, ar_bvh_info* info)
{
    if (!c || !info) return AR_E_INVALID;
    if (!c->bvh_ready) return AR_E_NOSCENE;
    *info = c->info;
    return AR_OK;
}

static orc_rays rays_of(const ar_rays* r)
{
    const orc_rays o = { r->ox, r->oy, r->oz, r->tmin, r->dx, r->dy, r->dz, r->tmax, r->meps };
    return o;
}

static int trace_check(ar_ctx* c, const ar_rays* rays, size_t n)
{
    if (!rays) return set_error(c, AR_E_INVALID, "NULL rays");
    if (!c->bvh_ready) return set_error(c, AR_E_NOSCENE, "no BVH: call ar_scene_set and ar_bvh_build first");
    if (n > ((size_t)1 << 31)) return set_error(c, AR_E_CAPACITY, "more than 2^31 rays");
    return AR_OK;
}

static void miss_all(const ar_hits* h, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        h->prim[i] = AR_MISS;
        h->t[i] = INFINITY;
        h->u[i] = h->v[i] = 0.0f;
        h->flags[i] = 0;
    }
}

int ar_trace_closest_counted(ar_ctx* c, const ar_rays* rays, size_t n, const ar_hits* hits, uint32_t* box_tests,
                             uint32_t* tri_tests, void* stream)
{
    (void)stream;
    if (!c) return AR_E_INVALID;
    int rc = trace_check(c, rays, n);
    if (rc != AR_OK) return rc;
    if (!hits) return set_error(c, AR_E_INVALID, "NULL hits");
    if (c->node_count == 0) {
        miss_all(hits, n);
        if (box_tests) memset(box_tests, 0, sizeof(uint32_t) * n);
        if (tri_tests) memset(tri_tests, 0, sizeof(uint32_t) * n);
        return AR_OK;
    }
    const orc_geom g = geom_of(c);
    const orc_hits h = { hits->prim, hits->t, hits->u, hits->v, hits->flags };
    orc_trace_closest_g(c->nodes, c->node_count, c->bidx, &g, rays_of(rays), n, h, 0, box_tests, tri_tests, 0);
    return AR_OK;
}

int ar_trace_closest(ar_ctx* c, const ar_rays* rays, size_t n, const ar_hits* hits, void* stream)
{
    return ar_trace_closest_counted(c, rays, n, hits, NULL, NULL, stream);
}

int ar_trace_closest_noalloc(ar_ctx* c, const ar_rays* rays, size_t n, const ar_hits* hits, void* stream)
{
    return ar_trace_closest_counted(c, rays, n, hits, NULL, NULL, stream);
}

int ar_trace_any(ar_ctx* c, const ar_rays* rays, size_t n, uint32_t* occluded, void* stream)
{
    (void)stream;
    if (!c) return AR_E_INVALID;
    const int rc = trace_check(c, rays, n);
    if (rc != AR_OK) return rc;
    if (!occluded) return set_error(c, AR_E_INVALID, "NULL occluded");
    if (c->node_count == 0) {
        memset(occluded, 0, sizeof(uint32_t) * ((n + 31) / 32));
        return AR_OK;
    }
    const orc_geom g = geom_of(c);
    orc_trace_any_g(c->nodes, c->node_count, c->bidx, &g, rays_of(rays), n, occluded, 0, NULL, NULL, 0);
    return AR_OK;
}

/* ------------------------------------------------------------------ rendering */

static int render_opts(ar_ctx* c, const ar_camera* cam, const ar_render_opts* o, orc_render_opts* r)
{
    if (!cam || !o) return set_error(c, AR_E_INVALID, "NULL camera or options");
    if (o->width == 0 || o->height == 0 || o->spp == 0) return set_error(c, AR_E_INVALID, "empty image or spp = 0");
    if (!c->bvh_ready) return set_error(c, AR_E_NOSCENE, "no BVH: call ar_scene_set and ar_bvh_build first");
    if (o->spectral != AR_SPECTRAL_RGB)
        return set_error(c, AR_E_UNSUPPORTED, "the CPU backend renders RGB only");
    if (o->guiding == AR_GUIDING_MLP) return set_error(c, AR_E_UNSUPPORTED, "the CPU backend has no learned guiding");
    if (o->guiding > AR_GUIDING_MLP) return set_error(c, AR_E_INVALID, "unknown guiding mode");
    if (o->guiding == AR_GUIDING_VOXEL && (o->frame_spp > 1 || o->frame_growth > 1))
        return set_error(c, AR_E_UNSUPPORTED, "the CPU backend guides frames of one sample");
    if (o->sky > AR_SKY_CONSTANT) return set_error(c, AR_E_INVALID, "unknown sky");
    memset(r, 0, sizeof *r);
    r->width = o->width; r->height = o->height; r->spp = o->spp; r->sample_offset = o->sample_offset;
    r->seed = o->seed;
    r->max_depth = (int)o->max_depth;
    r->rr_min_depth = (int)o->rr_min_depth;
    r->nee = o->nee ? 1 : 0;
    r->sky = (int)o->sky;
    memcpy(r->env, o->env, sizeof r->env);
    memcpy(r->cam_pos, cam->pos, sizeof r->cam_pos);
    memcpy(r->cam_dir, cam->dir, sizeof r->cam_dir);
    r->vfov_deg = cam->vfov_deg;
    if (o->guiding == AR_GUIDING_VOXEL) {
        if (!c->guide) c->guide = orc_guide_create((int)c->guide_cap);
        if (!c->guide) return set_error(c, AR_E_OOM, "host allocation failed");
        r->guiding = 1;
        r->guide = c->guide;
    }
    r->objects = c->nobj ? c->obj : NULL;
    r->nobj = c->nobj;
    r->obj_mats = c->omat;
    return AR_OK;
}

static int run(ar_ctx* c, const orc_render_opts* r, float* out)
{
    if (orc_render(c->verts, c->nv, c->idx, c->ntri, c->tri_mat, c->mats, c->nmats, r, 32, out, 0) != 0)
        return set_error(c, AR_E_INVALID, "render options rejected");
    return AR_OK;
}

int ar_render(ar_ctx* c, const ar_camera* cam, const ar_render_opts* opts, float* out_rgb, void* stream)
{
    (void)stream;
    if (!c) return AR_E_INVALID;
    if (c->batch) return set_error(c, AR_E_INVALID, "a batch context renders with ar_render_batch");
    if (!out_rgb) return set_error(c, AR_E_INVALID, "NULL output");
    orc_render_opts r;
    const int rc = render_opts(c, cam, opts, &r);
    return rc != AR_OK ? rc : run(c, &r, out_rgb);
}

int ar_render_host(ar_ctx* c, const ar_camera* cam, const ar_render_opts* opts, float* hdr, uint32_t* ldr,
                   float exposure, uint32_t tone)
{
    if (!c) return AR_E_INVALID;
    if (!cam || !opts || (!hdr && !ldr)) return set_error(c, AR_E_INVALID, "NULL camera, options or outputs");
    if (tone > AR_TONE_ACES) return set_error(c, AR_E_INVALID, "unknown tone curve");
    if (opts->width == 0 || opts->height == 0) return set_error(c, AR_E_INVALID, "empty image");
    const size_t npix = (size_t)opts->width * opts->height;
    float* img = hdr ? hdr : malloc(sizeof(float) * 3 * npix);
    if (!img) return set_error(c, AR_E_OOM, "host allocation failed");
    int rc = ar_render(c, cam, opts, img, NULL);
    if (rc == AR_OK && ldr) orc_tonemap(img, ldr, (uint32_t)npix, exposure, (int)tone, 0);
    if (img != hdr) free(img);
    return rc;
}

/* running mean of the GPU's accumulate pass: sum / n, then
 * dst * before / (before + n) + mean * n / (before + n) */
static void fold(float* dst, const float* sum, size_t nf, uint32_t n, uint32_t before)
{
    for (size_t i = 0; i < nf; ++i) {
        float x = sum[i] / (float)n;
        if (before > 0) {
            const float total = (float)before + (float)n;
            const float wa = (float)before / total, wb = (float)n / total;
            x = dst[i] * wa + x * wb;
        }
        dst[i] = x;
    }
}

int ar_render_progressive(ar_ctx* c, const ar_camera* cam, const ar_render_opts* opts, uint32_t frame, float* accum,
                          float* odd, float* even, float* gbuffer, float* stats, void* stream)
{
    (void)stream;
    if (!c) return AR_E_INVALID;
    if (!cam || !opts || !accum) return set_error(c, AR_E_INVALID, "NULL camera, options or accumulation buffer");
    if ((odd == NULL) != (even == NULL)) return set_error(c, AR_E_INVALID, "odd and even go together");
    orc_render_opts r;
    int rc = render_opts(c, cam, opts, &r);
    if (rc != AR_OK) return rc;
    const size_t npix = (size_t)opts->width * opts->height, nf = 3 * npix;
    float* one = malloc(sizeof(float) * nf);
    float* sum = calloc(3 * nf, sizeof(float));   /* all, odd, even */
    if (!one || !sum) { free(one); free(sum); return set_error(c, AR_E_OOM, "host allocation failed"); }
    if (frame == 0) {
        if (gbuffer) memset(gbuffer, 0, sizeof(float) * AR_GBUFFER_FLOATS * npix);
        if (stats) memset(stats, 0, sizeof(float) * AR_STATS_FLOATS * npix);
    }
    r.spp = 1;
    r.gbuffer = gbuffer;
    r.stats = stats;
    /* one sample at a time, summed in sample order as the GPU's accumulate pass sums */
    for (uint32_t s = 0; s < opts->spp && rc == AR_OK; ++s) {
        r.sample_offset = frame + s;
        rc = run(c, &r, one);
        float* half = sum + ((frame + s) & 1u ? nf : 2 * nf);
        for (size_t i = 0; i < nf && rc == AR_OK; ++i) {
            sum[i] = sum[i] + one[i];
            half[i] = half[i] + one[i];
        }
    }
    if (rc == AR_OK) {
        fold(accum, sum, nf, opts->spp, frame);
        if (odd) {
            const uint32_t f = frame, e = f + opts->spp;
            if (e / 2 - f / 2) fold(odd, sum + nf, nf, e / 2 - f / 2, f / 2);
            if ((e + 1) / 2 - (f + 1) / 2) fold(even, sum + 2 * nf, nf, (e + 1) / 2 - (f + 1) / 2, (f + 1) / 2);
        }
    }
    free(one);
    free(sum);
    return rc;
}

int ar_batch_create(ar_ctx* const* scenes, uint32_t n, ar_ctx** out)
{
    if (!scenes || !out || n == 0) return AR_E_INVALID;
    for (uint32_t k = 0; k < n; ++k) {
        if (!scenes[k]) return AR_E_INVALID;
        if (!scenes[k]->bvh_ready) return set_error(scenes[k], AR_E_NOSCENE, "scene %u has no BVH", k);
        if (scenes[k]->batch) return set_error(scenes[k], AR_E_INVALID, "scene %u is a batch", k);
    }
    ar_ctx* b = NULL;
    int rc = ar_create(&b, NULL);
    if (rc != AR_OK) return rc;
    b->batch = malloc(sizeof(ar_ctx*) * n);
    if (!b->batch) { ar_destroy(b); return AR_E_OOM; }
    memcpy(b->batch, scenes, sizeof(ar_ctx*) * n);
    b->nbatch = n;
    *out = b;
    return AR_OK;
}

int ar_batch_size(const ar_ctx* c) { return c ? (int)c->nbatch : AR_E_INVALID; }

int ar_render_batch(ar_ctx* c, uint32_t n, const ar_camera* cams, const ar_render_opts* opts, float* const* outs,
                    void* stream)
{
    (void)stream;
    if (!c) return AR_E_INVALID;
    if (!c->batch) return set_error(c, AR_E_INVALID, "not a batch context");
    if (n != c->nbatch || !cams || !opts || !outs) return set_error(c, AR_E_INVALID, "one camera, option set and output per scene");
    for (uint32_t k = 0; k < n; ++k) {
        if (opts[k].guiding != AR_GUIDING_OFF)
            return set_error(c, AR_E_UNSUPPORTED, "the CPU backend renders batches unguided");
        const int rc = ar_render(c->batch[k], &cams[k], &opts[k], outs[k], NULL);
        if (rc != AR_OK) return set_error(c, rc, "scene %u: %s", k, c->batch[k]->err);
    }
    return AR_OK;
}

int ar_set_tile(ar_ctx* c, uint32_t threads)
{
    if (!c) return AR_E_INVALID;
    if (threads != 64 && threads != 128 && threads != 256 && threads != 512)
        return set_error(c, AR_E_INVALID, "tile must be 64, 128, 256 or 512");
    return AR_OK;
}

/* ------------------------------------------------------------------ output */

int ar_tonemap(ar_ctx* c, const float* hdr, uint32_t* ldr, uint32_t width, uint32_t height, float exposure,
               uint32_t tone, uint32_t xyz_input, void* stream)
{
    (void)stream;
    if (!c) return AR_E_INVALID;
    if (!hdr || !ldr) return set_error(c, AR_E_INVALID, "NULL image");
    if (tone > AR_TONE_ACES) return set_error(c, AR_E_INVALID, "unknown tone curve");
    orc_tonemap(hdr, ldr, width * height, exposure, (int)tone, xyz_input ? 1 : 0);
    return AR_OK;
}

int ar_denoise(ar_ctx* c, const float* hdr, const float* gbuffer, const float* stats, float* out, uint32_t width,
               uint32_t height, void* stream)
{
    (void)stream;
    if (!c) return AR_E_INVALID;
    if (!hdr || !gbuffer || !stats || !out) return set_error(c, AR_E_INVALID, "NULL buffer");
    if (width == 0 || height == 0 || width > 32768 || height > 32768)
        return set_error(c, AR_E_INVALID, "image size must be 1..32768 per side");
    if (out == hdr) return set_error(c, AR_E_INVALID, "out must not alias hdr");
    orc_bilateral(hdr, gbuffer, stats, out, width, height);
    return AR_OK;
}

int ar_present(ar_ctx* c, const float* hdr, const float* gbuffer, const float* stats, uint32_t* ldr, uint32_t width,
               uint32_t height, float exposure, uint32_t tone, uint32_t xyz_input, uint32_t view, void* stream)
{
    if (!c) return AR_E_INVALID;
    if (!ldr) return set_error(c, AR_E_INVALID, "NULL ldr");
    if (tone > AR_TONE_ACES) return set_error(c, AR_E_INVALID, "unknown tone curve");
    if (view > AR_VIEW_STAT_M3) return set_error(c, AR_E_INVALID, "unknown view");
    if (view <= AR_VIEW_DENOISED && !hdr) return set_error(c, AR_E_INVALID, "NULL hdr");
    if ((view == AR_VIEW_DENOISED || (view >= AR_VIEW_ALBEDO && view <= AR_VIEW_ROUGHNESS)) && !gbuffer)
        return set_error(c, AR_E_INVALID, "this view needs the G-buffer");
    if ((view == AR_VIEW_DENOISED || view >= AR_VIEW_STAT_COUNT) && !stats)
        return set_error(c, AR_E_INVALID, "this view needs the sample statistics");
    const size_t npix = (size_t)width * height;
    if (npix == 0) return AR_OK;
    if (view >= AR_VIEW_ALBEDO) {
        orc_view(gbuffer, stats, ldr, (uint32_t)npix, (int)view);
        return AR_OK;
    }
    const float* src = hdr;
    if (view == AR_VIEW_DENOISED) {
        if (c->den_px < npix) {
            free(c->den);
            c->den = malloc(sizeof(float) * 3 * npix);
            c->den_px = c->den ? npix : 0;
            if (!c->den) return set_error(c, AR_E_OOM, "host allocation failed");
        }
        const int rc = ar_denoise(c, hdr, gbuffer, stats, c->den, width, height, stream);
        if (rc != AR_OK) return rc;
        src = c->den;
    }
    orc_tonemap(src, ldr, (uint32_t)npix, exposure, (int)tone, xyz_input ? 1 : 0);
    return AR_OK;
}

int ar_luminance_mse(ar_ctx* c, const float* a, const float* b, size_t npix, double* out, void* stream)
{
    (void)stream;
    if (!c) return AR_E_INVALID;
    if (!a || !b || !out) return set_error(c, AR_E_INVALID, "NULL argument");
    *out = orc_luminance_mse(a, b, npix);
    return AR_OK;
}

int ar_write_bmp(const char* path, const uint32_t* bgra, uint32_t width, uint32_t height)
{
    if (!path || !bgra || width == 0 || height == 0 || width > 0x7FFFFFFFu || height > 0x7FFFFFFFu)
        return AR_E_INVALID;
    const uint64_t data = (uint64_t)width * height * 4u;
    if (data + 54u > 0xFFFFFFFFull) return AR_E_CAPACITY;
    uint8_t h[54] = { 0 };
    const uint32_t fields[5][2] = { { 2, (uint32_t)(data + 54u) }, { 10, 54u }, { 14, 40u }, { 18, width },
                                    { 22, (uint32_t)(-(int32_t)height) } };
    for (int f = 0; f < 5; ++f)
        for (int k = 0; k < 4; ++k) h[fields[f][0] + k] = (uint8_t)(fields[f][1] >> (8 * k));
    h[0] = 'B'; h[1] = 'M';
    h[26] = 1;    /* planes */
    h[28] = 32;   /* bits per pixel */
    for (int k = 0; k < 4; ++k) h[34 + k] = (uint8_t)((uint32_t)data >> (8 * k));
    FILE* f = fopen(path, "wb");
    if (!f) return AR_E_INVALID;
    int ok = fwrite(h, 1, sizeof h, f) == sizeof h;
    ok = ok && fwrite(bgra, 4, (size_t)width * height, f) == (size_t)width * height;
    ok = (fclose(f) == 0) && ok;
    return ok ? AR_OK : AR_E_INVALID;
}

/* ------------------------------------------------------------------ voxel guiding state */

int ar_guiding_configure(ar_ctx* c, uint64_t max_probes)
{
    if (!c) return AR_E_INVALID;
    if (max_probes == 0 || max_probes > (1ull << 30)) return set_error(c, AR_E_INVALID, "max_probes must be in [1, 2^30]");
    if (max_probes > (1u << 24)) return set_error(c, AR_E_UNSUPPORTED, "the CPU backend holds at most 2^24 probes");
    drop_guide(c);
    c->guide_cap = max_probes;
    return AR_OK;
}

int ar_guiding_reset(ar_ctx* c)
{
    if (!c) return AR_E_INVALID;
    drop_guide(c);
    return AR_OK;
}

int ar_guiding_info_get(ar_ctx* c, ar_guiding_info* out)
{
    if (!c || !out) return AR_E_INVALID;
    memset(out, 0, sizeof *out);
    out->probes = c->guide ? (uint64_t)orc_guide_count(c->guide) : 0;
    out->capacity = c->guide_cap;
    return AR_OK;
}

int ar_guiding_snapshot(ar_ctx* c, uint64_t* keys, float* texels, uint64_t* counts, uint64_t cap, uint64_t* n)
{
    if (!c) return AR_E_INVALID;
    const uint64_t count = c->guide ? (uint64_t)orc_guide_count(c->guide) : 0;
    if (n) *n = count;
    if (count > cap && (keys || texels || counts)) return set_error(c, AR_E_CAPACITY, "snapshot larger than cap");
    if (c->guide && (keys || texels || counts)) orc_guide_export(c->guide, keys, texels, counts);
    return AR_OK;
}

/* ------------------------------------------------------------------ not in the CPU backend */

#define UNSUPPORTED(c) ((c) ? set_error((c), AR_E_UNSUPPORTED, "not in the CPU backend") : AR_E_INVALID)

int ar_make_shadow_rays(ar_ctx* c, const ar_rays* rays, const ar_hits* hits, size_t n, const float target[3],
                        float tmin, float tmax_scale, const ar_rays* out, void* stream)
{
    (void)rays; (void)hits; (void)n; (void)target; (void)tmin; (void)tmax_scale; (void)out; (void)stream;
    return UNSUPPORTED(c);
}

int ar_graph_capture(ar_ctx* c, uint32_t n, const ar_camera* cams, const ar_render_opts* opts, float* const* outs,
                     ar_graph** out)
{
    (void)n; (void)cams; (void)opts; (void)outs; (void)out;
    return UNSUPPORTED(c);
}

int ar_graph_launch(ar_graph* graph, void* stream) { (void)graph; (void)stream; return AR_E_UNSUPPORTED; }
int ar_graph_destroy(ar_graph* graph) { return graph ? AR_E_UNSUPPORTED : AR_OK; }

int ar_autotune_batch(ar_ctx* const* scenes, uint32_t n, const ar_camera* cams, const ar_render_opts* opts,
                      uint32_t probe_spp, ar_tune_result* out)
{
    (void)n; (void)cams; (void)opts; (void)probe_spp; (void)out;
    return scenes && scenes[0] ? UNSUPPORTED(scenes[0]) : AR_E_INVALID;
}

int ar_guiding_restore(ar_ctx* c, const uint64_t* keys, const float* texels, const uint64_t* counts, uint64_t n)
{
    (void)keys; (void)texels; (void)counts; (void)n;
    return UNSUPPORTED(c);
}

int ar_render_counters(ar_ctx* c, int enable) { (void)enable; return UNSUPPORTED(c); }
int ar_render_counters_read(ar_ctx* c, uint64_t out[5]) { (void)out; return UNSUPPORTED(c); }

int ar_mlp_configure(ar_ctx* c, const ar_mlp_opts* opts) { (void)opts; return UNSUPPORTED(c); }
int ar_mlp_params_get(ar_ctx* c, float* host_out, size_t count) { (void)host_out; (void)count; return UNSUPPORTED(c); }
int ar_mlp_params_set(ar_ctx* c, const float* host_in, size_t count) { (void)host_in; (void)count; return UNSUPPORTED(c); }

int ar_mlp_records_get(ar_ctx* c, float* x, float* dir, float* weight, size_t cap, uint32_t* count,
                       uint32_t* total)
{
    (void)x; (void)dir; (void)weight; (void)cap; (void)count; (void)total;
    return UNSUPPORTED(c);
}

int ar_mlp_forward(ar_ctx* c, const float* x5, size_t n, float* out40, void* stream)
{
    (void)x5; (void)n; (void)out40; (void)stream;
    return UNSUPPORTED(c);
}

int ar_mlp_gradient(ar_ctx* c, const float* x5, const float* dir, const float* weight, size_t n, float* grad,
                    float* loss, void* stream)
{
    (void)x5; (void)dir; (void)weight; (void)n; (void)grad; (void)loss; (void)stream;
    return UNSUPPORTED(c);
}

int ar_mlp_train(ar_ctx* c, const float* x5, const float* dir, const float* weight, size_t n, uint32_t steps,
                 float* loss, void* stream)
{
    (void)x5; (void)dir; (void)weight; (void)n; (void)steps; (void)loss; (void)stream;
    return UNSUPPORTED(c);
}

int ar_spectrum_lut_build(float* out, size_t count) { (void)out; (void)count; return AR_E_UNSUPPORTED; }
int ar_spectrum_uplift(const float* rgb, size_t n, float* coeffs) { (void)rgb; (void)n; (void)coeffs; return AR_E_UNSUPPORTED; }
int ar_spectral_set_basis(ar_ctx* c, uint32_t basis) { (void)basis; return UNSUPPORTED(c); }
int ar_spectral_mlp_set(ar_ctx* c, const float* params) { (void)params; return UNSUPPORTED(c); }
int ar_spectral_mlp_get(ar_ctx* c, float* params, int ema) { (void)params; (void)ema; return UNSUPPORTED(c); }

int ar_spectral_mlp_eval(ar_ctx* c, const float* x, float* out, size_t n, void* stream)
{
    (void)x; (void)out; (void)n; (void)stream;
    return UNSUPPORTED(c);
}

int ar_spectral_mlp_build_tables(ar_ctx* c, uint32_t rgb_res, uint32_t lambda_res, void* stream)
{
    (void)rgb_res; (void)lambda_res; (void)stream;
    return UNSUPPORTED(c);
}

int ar_spectral_mlp_read_tables(ar_ctx* c, float* rgb_table, float* lambda_table)
{
    (void)rgb_table; (void)lambda_table;
    return UNSUPPORTED(c);
}

int ar_spectral_mlp_query(ar_ctx* c, const float* x, float* out, size_t n, void* stream)
{
    (void)x; (void)out; (void)n; (void)stream;
    return UNSUPPORTED(c);
}

int ar_spectral_mlp_train(ar_ctx* c, const float* x, const float* y, size_t n, const ar_smlp_train_opts* opts,
                          float* loss_history, void* stream)
{
    (void)x; (void)y; (void)n; (void)opts; (void)loss_history; (void)stream;
    return UNSUPPORTED(c);
}

int ar_spectral_mlp_gradient(ar_ctx* c, const float* x, const float* y, size_t n, uint32_t batch, uint32_t step,
                             uint32_t seed, int noise, float* grad, float* loss, void* stream)
{
    (void)x; (void)y; (void)n; (void)batch; (void)step; (void)seed; (void)noise; (void)grad; (void)loss; (void)stream;
    return UNSUPPORTED(c);
}

int ar_spectral_uplift_dataset(ar_ctx* c, size_t n, uint32_t seed, float* x, float* y, void* stream)
{
    (void)n; (void)seed; (void)x; (void)y; (void)stream;
    return UNSUPPORTED(c);
}
