// Batched rendering: many scenes in one wavefront, hipGraph capture of a
// render, and the tile/BVH-width autotuner.
//
// ar_batch_create merges the device arrays of scene contexts (BVH nodes,
// leaf triangles, per-triangle shading data, materials, lights, analytic
// objects and their signed-distance voxels) into one batch context,
// rebasing every reference: node and leaf-slot references of the BVH, prim
// ids of leaf triangles and object ids of object leaf slots, material and
// light ids of the shading data and of the object records, the object
// records' own leaf slots and voxel offsets, and the prim ids of the light
// table (triangles first: objects follow all triangles of the batch, as in
// one scene).  A batch render then traces the paths of all scenes through
// the same queues; each ray carries its scene's BVH root and each path its
// scene's camera, light range and starting medium.
#include "ar_internal.hpp"
#include "objects.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

namespace ar {
namespace {

__device__ __forceinline__ uint32_t rebase_ref(uint32_t r, uint32_t node_off, uint32_t slot_off) {
    if (r == kRefEmpty) return r;
    if (r & kLeafBit) return r + slot_off;   // first slot in bits 0..26 (checked on the host)
    return r + node_off;
}

__global__ void k_rebase_nodes2(GpuNode2* n, uint32_t count, uint32_t node_off, uint32_t slot_off) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    n[i].d[0] = rebase_ref(n[i].d[0], node_off, slot_off);
    n[i].d[1] = rebase_ref(n[i].d[1], node_off, slot_off);
}

__global__ void k_rebase_nodes4(GpuNode4* n, uint32_t count, uint32_t node_off, uint32_t slot_off) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) n[i].ref[k] = rebase_ref(n[i].ref[k], node_off, slot_off);
}

__global__ void k_rebase_nodes8(GpuNode8* n, uint32_t count, uint32_t node_off, uint32_t slot_off) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
#pragma unroll
    for (int k = 0; k < 8; ++k) n[i].ref[k] = rebase_ref(n[i].ref[k], node_off, slot_off);
}

// leaf slots: triangles (v0, prim id bits), (v1, 0), (v2, 0); objects
// (0, kObjSlot | object id bits), (0), (0)
__global__ void k_rebase_tris(float4* t, uint32_t slots, uint32_t tri_off, uint32_t obj_off) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    const uint32_t w = __float_as_uint(t[3 * (size_t)i].w);
    t[3 * (size_t)i].w = __uint_as_float(w + ((w & kObjSlot) ? obj_off : tri_off));
}

// light table prim ids: the scene's triangles (< ntri) move by tri_off, its
// objects (ntri + k) to obj_prim + k
__global__ void k_rebase_lights(int32_t* v, uint32_t n, uint32_t ntri, uint32_t tri_off, uint32_t obj_prim) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t l = (uint32_t)v[i];
    v[i] = (int32_t)(l < ntri ? l + tri_off : l - ntri + obj_prim);
}

// shading triangles: (v0, material bits), (v1, light index bits or -1), (v2, 0)
__global__ void k_rebase_prims(float4* p, uint32_t tris, uint32_t mat_off, uint32_t light_off) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tris) return;
    p[3 * (size_t)i].w = __uint_as_float(__float_as_uint(p[3 * (size_t)i].w) + mat_off);
    const int l = __float_as_int(p[3 * (size_t)i + 1].w);
    if (l >= 0) p[3 * (size_t)i + 1].w = __int_as_float(l + (int)light_off);
}

uint32_t blocks(size_t n) { return (uint32_t)std::max<size_t>(1, (n + 255) / 256); }

template <typename T>
int alloc(ar_ctx* ctx, T** p, size_t n) {
    const hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(ctx, AR_E_OOM, "batch: device allocation failed");
    }
    return AR_OK;
}

} // namespace

// an object record's own leaf slot (a one-slot leaf reference) in the merged slots
uint32_t rebase_leaf(uint32_t r, uint32_t slot_off) { return r == kRefEmpty ? r : r + slot_off; }

int batch_merge(ar_ctx* out, ar_ctx* const* scenes, uint32_t n) {
    const uint32_t width = scenes[0]->bvh.width;
    const size_t nb = width == 8 ? sizeof(GpuNode8) : width == 4 ? sizeof(GpuNode4) : sizeof(GpuNode2);
    // Each scene's nodes and leaf slots start where its own arrays did
    // relative to the 128-byte lines (the builder places leaves to touch the
    // fewest lines): node offsets in whole lines, slot offsets in multiples
    // of 8 slots (8 x 48 B = 3 lines).  The gaps are never referenced.
    const size_t node_align = 128 / nb, slot_align = 8;
    auto up = [](size_t x, size_t a) { return (x + a - 1) / a * a; };
    size_t nodes = 0, slots = 0, tris = 0, mats = 0, lights = 0, nobjs = 0, voxels = 0;
    uint32_t stack = 0, nsdf = 0;
    float olim = INFINITY;
    for (uint32_t i = 0; i < n; ++i) {
        const ar_ctx* c = scenes[i];
        if (!c->bvh_ready) return set_error(out, AR_E_NOSCENE, "batch: a scene has no BVH");
        if (c->bvh.width != width) return set_error(out, AR_E_INVALID, "batch: scenes must share the BVH width");
        if (!c->batch.empty()) return set_error(out, AR_E_INVALID, "batch: a batch cannot contain a batch");
        nodes = up(nodes, node_align) + c->bvh.node_count;
        slots = up(slots, slot_align) + c->bvh.tri_slots;
        tris += c->scene.ntri;
        mats += c->scene.host_mats.size();
        lights += c->scene.nlights;
        nobjs += c->scene.nobj;
        nsdf += c->scene.nsdf;
        voxels += c->sdf.host.size();
        stack = std::max(stack, c->bvh.stack_entries);
        if (c->bvh.root != kRefEmpty) olim = std::min(olim, c->bvh.origin_limit);
    }
    if (slots >= (1u << 27) || nodes >= (1u << 31) || tris + nobjs >= (1u << 31))
        return set_error(out, AR_E_CAPACITY, "batch: more than 2^27 triangle slots");
    if (nobjs >= kObjSlot || voxels > (size_t)UINT32_MAX)
        return set_error(out, AR_E_CAPACITY, "batch: too many objects or SDF voxels");
    char* dnodes = nullptr;
    float4* dtris = nullptr;
    float4* dprims = nullptr;
    GpuMaterial* dmats = nullptr;
    int32_t* ltri = nullptr;
    float *larea = nullptr, *lprob = nullptr, *lcdf = nullptr;
    float4* dobjs = nullptr;
    int rc = alloc(out, &dnodes, nodes * nb);
    if (rc == AR_OK) rc = alloc(out, &dtris, 3 * slots);
    if (rc == AR_OK) rc = alloc(out, &dprims, 3 * tris);
    if (rc == AR_OK) rc = alloc(out, &dmats, mats);
    if (rc == AR_OK) rc = alloc(out, &ltri, lights);
    if (rc == AR_OK) rc = alloc(out, &larea, lights);
    if (rc == AR_OK) rc = alloc(out, &lprob, lights);
    if (rc == AR_OK) rc = alloc(out, &lcdf, lights);
    if (rc == AR_OK && nobjs) rc = alloc(out, &dobjs, (size_t)obj::kRecord * nobjs);
    auto free_all = [&]() {
        for (void* p : {(void*)dnodes, (void*)dtris, (void*)dprims, (void*)dmats, (void*)ltri, (void*)larea,
                        (void*)lprob, (void*)lcdf, (void*)dobjs})
            if (p) (void)hipFree(p);
    };
    if (rc != AR_OK) {
        free_all();
        return rc;
    }
    // the alignment gaps hold zeros
    if (hipMemset(dnodes, 0, nodes * nb) != hipSuccess || hipMemset(dtris, 0, 3 * slots * sizeof(float4)) != hipSuccess) {
        free_all();
        return set_error(out, AR_E_DEVICE, "batch: hipMemset failed");
    }
    std::vector<GpuMaterial> hm;
    hm.reserve(mats);
    std::vector<float4> ho;      // merged object records (host copy)
    ho.reserve((size_t)obj::kRecord * nobjs);
    SdfPool pool;                // merged voxels
    std::vector<BatchEntry> entries;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    size_t on = 0, os = 0, ot = 0, om = 0, ol = 0, oo = 0;
    hipError_t e = hipSuccess;
    for (uint32_t i = 0; i < n && e == hipSuccess; ++i) {
        const ar_ctx* c = scenes[i];
        on = up(on, node_align);
        os = up(os, slot_align);
        const uint32_t cn = c->bvh.node_count, cs = c->bvh.tri_slots, ct = c->scene.ntri, cl = c->scene.nlights;
        const uint32_t co = c->scene.nobj;
        const uint32_t vox_off = (uint32_t)pool.host.size();
        for (uint32_t k = 0; k < co; ++k) {
            float4 r[obj::kRecord];
            std::memcpy(r, &c->scene.host_objs[(size_t)obj::kRecord * k], sizeof r);
            uint32_t type, mat, solo;
            int32_t light;
            std::memcpy(&type, &r[0].w, 4);
            std::memcpy(&mat, &r[2].w, 4);
            std::memcpy(&light, &r[3].w, 4);
            std::memcpy(&solo, &r[4].w, 4);
            mat += (uint32_t)om;
            if (light >= 0) light += (int32_t)ol;
            solo = rebase_leaf(solo, (uint32_t)os);
            std::memcpy(&r[2].w, &mat, 4);
            std::memcpy(&r[3].w, &light, 4);
            std::memcpy(&r[4].w, &solo, 4);
            if (type == AR_OBJ_SDF_VOLUME) {
                uint32_t first;
                std::memcpy(&first, &r[5].x, 4);
                first += vox_off;
                std::memcpy(&r[5].x, &first, 4);
            }
            ho.insert(ho.end(), r, r + obj::kRecord);
        }
        pool.host.insert(pool.host.end(), c->sdf.host.begin(), c->sdf.host.end());
        for (uint32_t v : c->sdf.off) pool.off.push_back(v + vox_off);
        pool.dims.insert(pool.dims.end(), c->sdf.dims.begin(), c->sdf.dims.end());
        if (cn) e = hipMemcpy(dnodes + on * nb, c->bvh.nodes, cn * nb, hipMemcpyDeviceToDevice);
        if (e == hipSuccess && cs) e = hipMemcpy(dtris + 3 * os, c->bvh.tris, 3 * cs * sizeof(float4), hipMemcpyDeviceToDevice);
        if (e == hipSuccess && ct) e = hipMemcpy(dprims + 3 * ot, c->scene.prims, 3 * ct * sizeof(float4), hipMemcpyDeviceToDevice);
        if (e == hipSuccess && cl) {
            e = hipMemcpy(ltri + ol, c->scene.light_tri, cl * 4, hipMemcpyDeviceToDevice);
            if (e == hipSuccess) e = hipMemcpy(larea + ol, c->scene.light_area, cl * 4, hipMemcpyDeviceToDevice);
            if (e == hipSuccess) e = hipMemcpy(lprob + ol, c->scene.light_prob, cl * 4, hipMemcpyDeviceToDevice);
            if (e == hipSuccess) e = hipMemcpy(lcdf + ol, c->scene.light_cdf, cl * 4, hipMemcpyDeviceToDevice);
        }
        if (e != hipSuccess) break;
        if (cn) {
            if (width == 8)
                hipLaunchKernelGGL(k_rebase_nodes8, dim3(blocks(cn)), dim3(256), 0, nullptr,
                                   (GpuNode8*)(dnodes + on * nb), cn, (uint32_t)on, (uint32_t)os);
            else if (width == 4)
                hipLaunchKernelGGL(k_rebase_nodes4, dim3(blocks(cn)), dim3(256), 0, nullptr,
                                   (GpuNode4*)(dnodes + on * nb), cn, (uint32_t)on, (uint32_t)os);
            else
                hipLaunchKernelGGL(k_rebase_nodes2, dim3(blocks(cn)), dim3(256), 0, nullptr,
                                   (GpuNode2*)(dnodes + on * nb), cn, (uint32_t)on, (uint32_t)os);
        }
        if (cs)
            hipLaunchKernelGGL(k_rebase_tris, dim3(blocks(cs)), dim3(256), 0, nullptr, dtris + 3 * os, cs, (uint32_t)ot,
                               (uint32_t)oo);
        if (ct)
            hipLaunchKernelGGL(k_rebase_prims, dim3(blocks(ct)), dim3(256), 0, nullptr, dprims + 3 * ot, ct,
                               (uint32_t)om, (uint32_t)ol);
        if (cl)
            hipLaunchKernelGGL(k_rebase_lights, dim3(blocks(cl)), dim3(256), 0, nullptr, ltri + ol, cl, ct,
                               (uint32_t)ot, (uint32_t)(tris + oo));
        BatchEntry be{};
        be.root = c->bvh.root == kRefEmpty ? kRefEmpty : (uint32_t)(c->bvh.root + on);
        be.light_off = (uint32_t)ol;
        be.light_cnt = cl;
        be.tri_off = (uint32_t)ot;
        be.obj_off = (uint32_t)oo;
        be.obj_cnt = co;
        for (int a = 0; a < 3; ++a) { be.box_lo[a] = c->scene.box_lo[a]; be.box_inv[a] = c->scene.box_inv[a]; }
        entries.push_back(be);
        hm.insert(hm.end(), c->scene.host_mats.begin(), c->scene.host_mats.end());
        for (int a = 0; a < 3; ++a) {
            if (ct) {
                lo[a] = std::min(lo[a], c->scene.box_lo[a]);
                hi[a] = std::max(hi[a], c->scene.box_lo[a] + 1.0f / c->scene.box_inv[a]);
            }
        }
        on += cn; os += cs; ot += ct; om += c->scene.host_mats.size(); ol += cl; oo += co;
    }
    if (e == hipSuccess && !hm.empty()) e = hipMemcpy(dmats, hm.data(), hm.size() * sizeof(GpuMaterial), hipMemcpyHostToDevice);
    if (e == hipSuccess && nobjs) e = hipMemcpy(dobjs, ho.data(), ho.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        free_all();
        return check_hip(out, e, "batch merge");
    }
    out->sdf.host.swap(pool.host);
    out->sdf.off.swap(pool.off);
    out->sdf.dims.swap(pool.dims);
    out->sdf.dirty = !out->sdf.host.empty();
    rc = sdf_upload(out);
    if (rc != AR_OK) {
        free_all();
        return rc;
    }
    out->bvh.width = width;
    out->bvh.nodes = (GpuNode2*)dnodes;
    out->bvh.tris = (float*)dtris;
    out->bvh.node_count = (uint32_t)nodes;
    out->bvh.tri_slots = (uint32_t)slots;
    out->bvh.root = entries.empty() ? kRefEmpty : entries[0].root;
    out->bvh.origin_limit = std::isfinite(olim) ? olim : 0.0f;
    out->bvh.stack_entries = std::max(stack, 2u);
    out->bvh_ready = true;
    out->scene_set = true;
    out->info.width = width;
    out->info.node_count = nodes;
    SceneDev& s = out->scene;
    s.prims = dprims;
    s.mats = dmats;
    s.light_tri = ltri;
    s.light_area = larea;
    s.light_prob = lprob;
    s.light_cdf = lcdf;
    s.nlights = (uint32_t)lights;
    s.ntri = (uint32_t)tris;
    s.objs = dobjs;
    s.nobj = (uint32_t)nobjs;
    s.nsdf = nsdf;
    s.host_objs.swap(ho);
    s.host_mats = hm;
    s.spectral_ready = false;
    for (int a = 0; a < 3; ++a) {
        s.box_lo[a] = std::isfinite(lo[a]) ? lo[a] : 0.0f;
        const float ext = hi[a] - lo[a];
        s.box_inv[a] = std::isfinite(ext) && ext > 0.0f ? 1.0f / ext : 1.0f;
    }
    out->batch = entries;
    return AR_OK;
}

} // namespace ar

using namespace ar;

extern "C" {

int ar_batch_create(ar_ctx* const* scenes, uint32_t n, ar_ctx** out) {
    if (!scenes || !out || n == 0) return AR_E_INVALID;
    *out = nullptr;
    ar_device_opts dev{scenes[0]->device, 0};
    ar_ctx* b = nullptr;
    int rc = ar_create(&b, &dev);
    if (rc != AR_OK) return rc;
    rc = batch_merge(b, scenes, n);
    if (rc != AR_OK) {
        // keep the message for the caller on the first scene
        set_error(scenes[0], rc, ar_last_error(b));
        ar_destroy(b);
        return rc;
    }
    *out = b;
    return AR_OK;
}

int ar_batch_size(const ar_ctx* ctx) { return ctx ? (int)ctx->batch.size() : 0; }

int ar_render_batch(ar_ctx* ctx, uint32_t n, const ar_camera* cams, const ar_render_opts* opts, float* const* outs,
                    void* stream) {
    if (!ctx || !cams || !opts || !outs || n == 0) return ctx ? set_error(ctx, AR_E_INVALID, "NULL argument") : AR_E_INVALID;
    if (ctx->batch.empty() ? n != 1 : n != ctx->batch.size())
        return set_error(ctx, AR_E_INVALID, "n must equal the number of scenes of the batch context");
    if (!ctx->bvh_ready) return set_error(ctx, AR_E_NOSCENE, "no BVH");
    for (uint32_t i = 0; i < n; ++i)
        if (opts[i].sky > AR_SKY_CONSTANT) return set_error(ctx, AR_E_INVALID, "unknown sky mode");
    const int rc = order_stream(ctx, (hipStream_t)stream);
    if (rc != AR_OK) return rc;
    return render_batch(ctx, n, cams, opts, outs, (hipStream_t)stream);
}

int ar_set_tile(ar_ctx* ctx, uint32_t threads) {
    if (!ctx) return AR_E_INVALID;
    if (threads != 64 && threads != 128 && threads != 256 && threads != 512)
        return set_error(ctx, AR_E_INVALID, "tile must be 64, 128, 256 or 512 threads");
    ctx->tile = threads;
    return AR_OK;
}

int ar_graph_capture(ar_ctx* ctx, uint32_t n, const ar_camera* cams, const ar_render_opts* opts, float* const* outs,
                     ar_graph** out) {
    if (!ctx || !cams || !opts || !outs || !out || n == 0) return ctx ? set_error(ctx, AR_E_INVALID, "NULL argument") : AR_E_INVALID;
    *out = nullptr;
    if (ctx->batch.empty() ? n != 1 : n != ctx->batch.size())
        return set_error(ctx, AR_E_INVALID, "n must equal the number of scenes of the context");
    if (!ctx->bvh_ready) return set_error(ctx, AR_E_NOSCENE, "no BVH");
    GraphRec* g = new GraphRec();
    g->ctx = ctx;
    const int rc = graph_capture(ctx, n, cams, opts, outs, g);
    if (rc != AR_OK) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        free_job(g->job);
        delete g;
        return rc;
    }
    ctx->graphs.push_back(g);
    *out = reinterpret_cast<ar_graph*>(g);
    return AR_OK;
}

int ar_graph_launch(ar_graph* graph, void* stream) {
    GraphRec* g = reinterpret_cast<GraphRec*>(graph);
    if (!g) return AR_E_INVALID;
    if (!g->ctx) return AR_E_INVALID;   // its context was destroyed
    if (g->generation != g->ctx->buffer_generation)
        return set_error(g->ctx, AR_E_INVALID,
                         "graph outdated: device memory it uses was freed or reallocated (scene, BVH, render, "
                         "guiding or table buffers)");
    const int rc = order_stream(g->ctx, (hipStream_t)stream);
    if (rc != AR_OK) return rc;
    return check_hip(g->ctx, hipGraphLaunch(g->exec, (hipStream_t)stream), "hipGraphLaunch");
}

int ar_graph_destroy(ar_graph* graph) {
    GraphRec* g = reinterpret_cast<GraphRec*>(graph);
    if (!g) return AR_E_INVALID;
    (void)hipDeviceSynchronize();
    if (g->ctx) {
        std::vector<GraphRec*>& v = g->ctx->graphs;
        v.erase(std::remove(v.begin(), v.end(), g), v.end());
    }
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    free_job(g->job);
    delete g;
    return AR_OK;
}

int ar_autotune_batch(ar_ctx* const* scenes, uint32_t n, const ar_camera* cams, const ar_render_opts* opts,
                      uint32_t probe_spp, ar_tune_result* out) {
    if (!scenes || !cams || !opts || !out || n == 0) return AR_E_INVALID;
    std::memset(out, 0, sizeof *out);
    ar_ctx* c0 = scenes[0];
    std::vector<ar_render_opts> po(opts, opts + n);
    for (ar_render_opts& o : po) o.spp = std::max(1u, std::min(o.spp, probe_spp ? probe_spp : 1u));
    std::vector<float*> outs(n, nullptr);
    int rc = AR_OK;
    for (uint32_t i = 0; i < n && rc == AR_OK; ++i)
        rc = check_hip(c0, hipMalloc((void**)&outs[i], (size_t)po[i].width * po[i].height * 3 * sizeof(float)),
                       "autotune outputs");
    const uint32_t widths[3] = {2, 4, 8};
    const uint32_t tiles[4] = {64, 128, 256, 512};
    float best = INFINITY;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == AR_OK) rc = check_hip(c0, hipEventCreate(&e0), "event");
    if (rc == AR_OK) rc = check_hip(c0, hipEventCreate(&e1), "event");
    for (int wi = 0; wi < 3 && rc == AR_OK; ++wi) {
        for (uint32_t i = 0; i < n && rc == AR_OK; ++i) {
            ar_bvh_opts bo{widths[wi], 3, 12, 0};
            rc = ar_bvh_build(scenes[i], &bo);
        }
        if (rc != AR_OK) break;
        ar_ctx* b = nullptr;
        rc = ar_batch_create(scenes, n, &b);
        if (rc != AR_OK) break;
        for (int ti = 0; ti < 4 && rc == AR_OK; ++ti) {
            ar_set_tile(b, tiles[ti]);
            rc = ar_render_batch(b, n, cams, po.data(), outs.data(), nullptr);   // warm-up (allocations)
            if (rc == AR_OK) rc = check_hip(b, hipDeviceSynchronize(), "autotune");
            if (rc == AR_OK) (void)hipEventRecord(e0, nullptr);
            if (rc == AR_OK) rc = ar_render_batch(b, n, cams, po.data(), outs.data(), nullptr);
            if (rc == AR_OK) (void)hipEventRecord(e1, nullptr);
            if (rc == AR_OK) rc = check_hip(b, hipEventSynchronize(e1), "autotune");
            float ms = 0.0f;
            if (rc == AR_OK) (void)hipEventElapsedTime(&ms, e0, e1);
            out->ms[wi][ti] = ms;
            if (rc == AR_OK && ms < best) { best = ms; out->width = widths[wi]; out->tile = tiles[ti]; }
        }
        if (rc != AR_OK) set_error(c0, rc, ar_last_error(b));
        ar_destroy(b);
    }
    // leave the scenes built at the chosen width
    for (uint32_t i = 0; i < n && rc == AR_OK && out->width != 8; ++i) {
        ar_bvh_opts bo{out->width, 3, 12, 0};
        rc = ar_bvh_build(scenes[i], &bo);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (float* p : outs)
        if (p) (void)hipFree(p);
    return rc;
}

} // extern "C"
