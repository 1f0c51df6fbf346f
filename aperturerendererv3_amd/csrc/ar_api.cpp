// C ABI of libar_hip: context, scene upload, BVH build and the traversal
// entry points.  Declarations and reference citations are in include/ar.h.
#include "ar_internal.hpp"
#include "objects.hpp"

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <new>

namespace ar {

int set_error(ar_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int order_stream(ar_ctx* ctx, hipStream_t s) {
    if (ctx->have_last_stream && ctx->last_stream != s) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hipStreamCaptureStatus cl = hipStreamCaptureStatusNone;
        (void)hipStreamIsCapturing(s, &cs);
        (void)hipStreamIsCapturing(ctx->last_stream, &cl);
        // inside a capture the caller's graph orders the work
        if (cs == hipStreamCaptureStatusNone && cl == hipStreamCaptureStatusNone) {
            if (!ctx->order_event) {
                const int rc = check_hip(ctx, hipEventCreateWithFlags(&ctx->order_event, hipEventDisableTiming),
                                         "hipEventCreate(order)");
                if (rc != AR_OK) return rc;
            }
            int rc = check_hip(ctx, hipEventRecord(ctx->order_event, ctx->last_stream), "hipEventRecord(order)");
            if (rc == AR_OK) rc = check_hip(ctx, hipStreamWaitEvent(s, ctx->order_event, 0), "hipStreamWaitEvent(order)");
            if (rc != AR_OK) return rc;
        }
    }
    ctx->last_stream = s;
    ctx->have_last_stream = true;
    return AR_OK;
}

int check_hip(ar_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipSuccess) return AR_OK;
    return set_error(ctx, e == hipErrorOutOfMemory ? AR_E_OOM : AR_E_DEVICE,
                     std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

void free_bvh(ar_ctx* ctx) {
    invalidate_graphs(ctx);
    if (ctx->bvh.nodes) (void)hipFree(ctx->bvh.nodes);
    if (ctx->bvh.tris) (void)hipFree(ctx->bvh.tris);
    ctx->bvh = DeviceBVH{};
    ctx->bvh_ready = false;
}

struct Converter {
    const BuildResult& br;
    LeafWriter& lw;
    float pad;
    std::vector<GpuNode2> nodes;
    std::vector<float> tris;   // 12 floats per slot
    uint64_t leaves = 0;

    uint32_t leaf_ref(const BuildNode& n) {
        const uint32_t first = (uint32_t)(tris.size() / 12);
        for (uint32_t k = 0; k < n.count; ++k) lw.write(tris, br.order[n.first + k]);
        ++leaves;
        return make_leaf_ref(first, n.count);
    }

    void set_box(GpuNode2& g, int which, const BuildNode& c) {
        const float lx = c.lo[0] - pad, hx = c.hi[0] + pad;
        const float ly = c.lo[1] - pad, hy = c.hi[1] + pad;
        const float lz = c.lo[2] - pad, hz = c.hi[2] + pad;
        float* xy = which == 0 ? g.a : g.b;
        xy[0] = lx; xy[1] = hx; xy[2] = ly; xy[3] = hy;
        g.c[2 * which + 0] = lz;
        g.c[2 * which + 1] = hz;
    }

    uint32_t child(int bi) {
        const BuildNode& n = br.nodes[bi];
        return n.left < 0 ? leaf_ref(n) : interior(bi);
    }

    uint32_t interior(int bi) {
        const uint32_t gi = (uint32_t)nodes.size();
        nodes.emplace_back();
        const BuildNode& n = br.nodes[bi];
        const uint32_t r0 = child(n.left);
        const uint32_t r1 = child(n.right);
        GpuNode2& g = nodes[gi];
        set_box(g, 0, br.nodes[n.left]);
        set_box(g, 1, br.nodes[n.right]);
        g.d[0] = r0; g.d[1] = r1; g.d[2] = 0; g.d[3] = 0;
        return gi;
    }

    void run() {
        const BuildNode& root = br.nodes[0];
        if (root.left < 0) {
            // a single leaf: both children reference it (same answer, tested twice)
            nodes.emplace_back();
            const uint32_t r = leaf_ref(root);
            GpuNode2& g = nodes[0];
            set_box(g, 0, root);
            set_box(g, 1, root);
            g.d[0] = r; g.d[1] = r; g.d[2] = 0; g.d[3] = 0;
        } else {
            interior(0);
        }
    }
};

int ensure_spill(ar_ctx* ctx, uint32_t stack_entries) {
    if (ctx->spill && ctx->spill_stack >= stack_entries) return AR_OK;
    invalidate_graphs(ctx);
    if (ctx->spill) (void)hipFree(ctx->spill);
    ctx->spill = nullptr;
    const int bpc = trace_blocks_per_cu();
    const uint32_t blocks = (uint32_t)ctx->num_cus * (uint32_t)bpc;
    ctx->trace_threads = blocks * kTraceBlock;
    const uint32_t deep = stack_entries > (uint32_t)kLdsStack ? stack_entries - (uint32_t)kLdsStack : 1u;
    ctx->spill_entries = (size_t)deep * ctx->trace_threads;
    int rc = check_hip(ctx, hipMalloc(&ctx->spill, ctx->spill_entries * sizeof(int)), "hipMalloc(stack spill)");
    if (rc != AR_OK) { ctx->spill = nullptr; ctx->spill_stack = 0; return rc; }
    ctx->spill_stack = stack_entries;
    if (!ctx->fetch) {
        rc = check_hip(ctx, hipMalloc((void**)&ctx->fetch, 256), "hipMalloc(refill counter)");
        if (rc != AR_OK) { ctx->fetch = nullptr; return rc; }
    }
    return AR_OK;
}

bool rays_ok(const ar_rays* r) {
    return r && r->ox && r->oy && r->oz && r->tmin && r->dx && r->dy && r->dz && r->tmax;
}

bool hits_ok(const ar_hits* h) { return h && h->prim && h->t && h->u && h->v && h->flags; }

} // namespace

int trace_prologue(ar_ctx* ctx, TraceArgs& ta, size_t n, uint32_t& blocks) {
    if (!ctx->bvh_ready) return set_error(ctx, AR_E_NOSCENE, "no BVH: call ar_scene_set and ar_bvh_build first");
    int rc = ensure_spill(ctx, ctx->bvh.stack_entries);
    if (rc != AR_OK) return rc;
    // resident blocks of the closest-hit and any-hit kernels (their register
    // budgets differ); the spill buffer is laid out for trace_threads threads
    const uint32_t cap_blocks = ctx->trace_threads / kTraceBlock;
    ta.full_closest = std::min(cap_blocks, (uint32_t)ctx->num_cus * (uint32_t)trace_blocks_per_cu((int)ctx->bvh.width, false));
    ta.full_any = std::min(cap_blocks, (uint32_t)ctx->num_cus * (uint32_t)trace_blocks_per_cu((int)ctx->bvh.width, true));
    const size_t need = (n + kTraceBlock - 1) / kTraceBlock;
    blocks = (uint32_t)std::min<size_t>(std::max(ta.full_closest, ta.full_any), std::max<size_t>(need, 1));
    ta.width = ctx->bvh.width;
    ta.stack_entries = ctx->bvh.stack_entries;
    ta.nodes = ctx->bvh.nodes;
    ta.tris = reinterpret_cast<const float4*>(ctx->bvh.tris);
    ta.root = ctx->bvh.root;
    ta.origin_limit = ctx->bvh.origin_limit;
    ta.spill = ctx->spill;
    ta.threads = ctx->trace_threads;
    ta.fetch = ctx->fetch;
    ta.objs = ctx->scene.nobj ? ctx->scene.objs : nullptr;
    // objects follow the triangles (a batch's merged triangles: scene.ntri)
    ta.obj_base = ctx->batch.empty() ? (uint32_t)(ctx->idx.size() / 3) : ctx->scene.ntri;
    ta.vol = ctx->sdf.dev;
    return AR_OK;
}

namespace {

// NULL selects the HIP default (null) stream, which orders with every other
// blocking stream of the process (torch's default stream is that stream).
hipStream_t pick_stream(ar_ctx*, void* stream) {
    return reinterpret_cast<hipStream_t>(stream);
}

// Every float of a material is finite.  The reference takes any value and
// scrubs non-finite path radiance to 0 per sample (core/vector.cuh:189-221);
// a non-finite emission would also poison the light CDF, so it is refused
// here instead.
bool material_finite(const ar_material& m) {
    const float f[] = {m.albedo[0], m.albedo[1], m.albedo[2], m.emission[0], m.emission[1], m.emission[2],
                       m.specular,  m.metallic,  m.roughness, m.ior,         m.sigma_a,     m.abbe,
                       m.tilt};
    for (float v : f)
        if (!std::isfinite(v)) return false;
    return true;
}

} // namespace
} // namespace ar

using namespace ar;

extern "C" {

int ar_abi_version(void) { return AR_ABI_VERSION; }

int ar_create(ar_ctx** out, const ar_device_opts* opts) {
    if (!out) return AR_E_INVALID;
    *out = nullptr;
    ar_ctx* ctx = new (std::nothrow) ar_ctx();
    if (!ctx) return AR_E_OOM;
    ctx->device = opts ? opts->device : 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        delete ctx;
        return AR_E_DEVICE;
    }
    if (ctx->device < 0 || ctx->device >= ndev) {
        delete ctx;
        return AR_E_INVALID;
    }
    if (hipSetDevice(ctx->device) != hipSuccess ||
        hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return AR_E_DEVICE;
    }
    // the no-allocation trace's scratch address on this device; a failure
    // here is reported by that entry point alone
    (void)noalloc_scratch_setup(ctx->device);
    *out = ctx;
    return AR_OK;
}

int ar_destroy(ar_ctx* ctx) {
    if (!ctx) return AR_E_INVALID;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    // captured renders of this context can no longer run
    for (GraphRec* g : ctx->graphs) g->ctx = nullptr;
    ctx->graphs.clear();
    if (ctx->order_event) (void)hipEventDestroy(ctx->order_event);
    free_bvh(ctx);
    free_scene(ctx);
    free_render_buffers(ctx);
    guide_free(ctx);
    mlp_free(ctx);
    smlp_free(ctx);
    free_job(ctx->job);
    ctx->job = nullptr;
    if (ctx->spill) (void)hipFree(ctx->spill);
    if (ctx->fetch) (void)hipFree(ctx->fetch);
    if (ctx->reduce) (void)hipFree(ctx->reduce);
    sdf_free(ctx);
    if (ctx->denoise_feat) (void)hipFree(ctx->denoise_feat);
    if (ctx->host_out) (void)hipFree(ctx->host_out);
    if (ctx->present_buf) (void)hipFree(ctx->present_buf);
    if (ctx->cnt_box) (void)hipFree(ctx->cnt_box);
    if (ctx->cnt_totals) (void)hipFree(ctx->cnt_totals);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return AR_OK;
}

const char* ar_last_error(const ar_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ar_scene_set(ar_ctx* ctx, const ar_mesh* meshes, size_t nmesh) {
    if (!ctx) return AR_E_INVALID;
    if (nmesh && !meshes) return set_error(ctx, AR_E_INVALID, "meshes is NULL");
    std::vector<float> verts;
    std::vector<uint32_t> idx, tri_mat;
    std::vector<ar_material> mats;
    uint64_t total = 0;
    for (size_t m = 0; m < nmesh; ++m) total += meshes[m].ntri;
    if (total > kMaxTriSlots) return set_error(ctx, AR_E_CAPACITY, "more than 2^27-1 triangles");
    for (size_t m = 0; m < nmesh; ++m) {
        const ar_mesh& me = meshes[m];
        if (me.ntri && (!me.vertices || !me.indices))
            return set_error(ctx, AR_E_INVALID, "mesh " + std::to_string(m) + ": NULL vertices or indices");
        const uint32_t vbase = (uint32_t)(verts.size() / 3);
        const uint32_t mbase = (uint32_t)mats.size();
        verts.insert(verts.end(), me.vertices, me.vertices + 3 * (size_t)me.nv);
        for (uint32_t k = 0; k < 3 * me.nv; ++k)
            if (!std::isfinite(me.vertices[k]))
                return set_error(ctx, AR_E_INVALID, "mesh " + std::to_string(m) + ": non-finite vertex");
        for (uint32_t t = 0; t < me.ntri; ++t) {
            for (int k = 0; k < 3; ++k) {
                const uint32_t v = me.indices[3 * (size_t)t + k];
                if (v >= me.nv)
                    return set_error(ctx, AR_E_INVALID, "mesh " + std::to_string(m) + ": index out of range");
                idx.push_back(vbase + v);
            }
            const uint32_t mi = me.tri_material ? me.tri_material[t] : 0;
            if (me.nmaterials && mi >= me.nmaterials)
                return set_error(ctx, AR_E_INVALID, "mesh " + std::to_string(m) + ": material out of range");
            tri_mat.push_back(me.nmaterials ? mbase + mi : 0);
        }
        if (me.nmaterials) {
            if (!me.materials) return set_error(ctx, AR_E_INVALID, "materials is NULL");
            for (uint32_t k = 0; k < me.nmaterials; ++k) {
                if (!material_finite(me.materials[k]))
                    return set_error(ctx, AR_E_INVALID,
                                     "mesh " + std::to_string(m) + ": material " + std::to_string(k) + " has a non-finite value");
                if (me.materials[k].model > AR_MODEL_GGX)
                    return set_error(ctx, me.materials[k].model == AR_MODEL_GLASS ? AR_E_UNSUPPORTED : AR_E_INVALID,
                                     "mesh " + std::to_string(m) +
                                         ": triangle materials are Lambert or GGX (glass is for analytic objects)");
            }
            mats.insert(mats.end(), me.materials, me.materials + me.nmaterials);
        }
    }
    if (mats.empty()) {
        ar_material def{};
        def.albedo[0] = def.albedo[1] = def.albedo[2] = 0.75f;
        def.roughness = 1.0f;
        mats.push_back(def);
    }
    ctx->verts.swap(verts);
    ctx->idx.swap(idx);
    ctx->tri_mat.swap(tri_mat);
    ctx->mats.swap(mats);
    ctx->scene_set = true;
    (void)hipSetDevice(ctx->device);
    free_bvh(ctx);
    free_scene(ctx);
    // probes describe the previous scene's light field
    if (ctx->guide.cap) {
        (void)hipDeviceSynchronize();
        const int rc = guide_clear(ctx, nullptr);
        if (rc != AR_OK) return rc;
    }
    return AR_OK;
}

int ar_scene_set_objects(ar_ctx* ctx, const ar_object* objects, size_t nobj) {
    if (!ctx) return AR_E_INVALID;
    if (nobj && !objects) return set_error(ctx, AR_E_INVALID, "objects is NULL");
    if (nobj > (size_t)(kObjSlot - 1u) || nobj > (size_t)INT32_MAX)
        return set_error(ctx, AR_E_CAPACITY, "too many objects");
    for (size_t k = 0; k < nobj; ++k) {
        const ar_object& o = objects[k];
        const std::string who = "object " + std::to_string(k);
        if (o.type > AR_OBJ_SDF_VOLUME) return set_error(ctx, AR_E_INVALID, who + ": unknown type");
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(o.center[a]) || !std::isfinite(o.extent[a]) || !std::isfinite(o.rotation[a]))
                return set_error(ctx, AR_E_INVALID, who + ": non-finite center, extent or rotation");
        const float ex = std::fabs(o.extent[0]), ey = std::fabs(o.extent[1]), ez = std::fabs(o.extent[2]);
        if (o.type != AR_OBJ_SPHERE ? !(ex > 0.0f && ey > 0.0f && ez > 0.0f) : !(std::max(ex, std::max(ey, ez)) > 0.0f))
            return set_error(ctx, AR_E_INVALID, who + ": empty extent");
        if (o.type == AR_OBJ_SDF_TORUS && !(std::isfinite(o.sdf[0]) && std::isfinite(o.sdf[1])))
            return set_error(ctx, AR_E_INVALID, who + ": non-finite torus radii");
        if (o.type == AR_OBJ_SDF_VOLUME && (o.sdf_volume < 0 || (size_t)o.sdf_volume >= ctx->sdf.off.size()))
            return set_error(ctx, AR_E_INVALID, who + ": no such SDF volume");
        if ((o.type == AR_OBJ_SDF_TORUS || o.type == AR_OBJ_SDF_VOLUME) &&
            (o.material.emission[0] > 0.0f || o.material.emission[1] > 0.0f || o.material.emission[2] > 0.0f))
            return set_error(ctx, AR_E_UNSUPPORTED, who + ": SDF objects cannot be lights");
        if (!material_finite(o.material)) return set_error(ctx, AR_E_INVALID, who + ": non-finite material value");
        if (o.material.model > AR_MODEL_GLASS) return set_error(ctx, AR_E_INVALID, who + ": unknown material model");
        if (o.material.model == AR_MODEL_GLASS && !(o.material.ior > 0.0f))
            return set_error(ctx, AR_E_INVALID, who + ": glass needs ior > 0");
    }
    ctx->objs.assign(objects, objects + nobj);
    (void)hipSetDevice(ctx->device);
    free_bvh(ctx);
    free_scene(ctx);
    if (ctx->guide.cap) {
        (void)hipDeviceSynchronize();
        const int rc = guide_clear(ctx, nullptr);
        if (rc != AR_OK) return rc;
    }
    return AR_OK;
}

int ar_bvh_build(ar_ctx* ctx, const ar_bvh_opts* opts) {
    if (!ctx) return AR_E_INVALID;
    if (!ctx->scene_set) return set_error(ctx, AR_E_NOSCENE, "no scene: call ar_scene_set first");
    ar_bvh_opts o{4, 2, 16, (uint32_t)kMaxDepth};
    if (opts) o = *opts;
    if (o.width != 2 && o.width != 4 && o.width != 8) return set_error(ctx, AR_E_UNSUPPORTED, "BVH width must be 2, 4 or 8");
    if (o.leaf_size < 1 || o.leaf_size > kMaxLeafTris) return set_error(ctx, AR_E_INVALID, "leaf_size must be 1..15");
    if (o.bin_count < 4 || o.bin_count > 64) return set_error(ctx, AR_E_INVALID, "bin_count must be 4..64");
    if (o.max_depth == 0 || o.max_depth > (uint32_t)kMaxDepth) o.max_depth = kMaxDepth;
    (void)hipSetDevice(ctx->device);
    free_bvh(ctx);
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t ntri = (uint32_t)(ctx->idx.size() / 3);
    const uint32_t nobj = (uint32_t)ctx->objs.size();
    const uint32_t n = ntri + nobj;
    ctx->info = ar_bvh_info{};
    ctx->info.width = o.width;
    if (n == 0) {
        ctx->bvh.root = kRefEmpty;
        ctx->bvh_ready = true;
        return upload_scene(ctx);
    }
    std::vector<float> lo(3 * (size_t)n), hi(3 * (size_t)n), cen(3 * (size_t)n);
    float mag = 0.0f;
    for (uint32_t k = 0; k < nobj; ++k) {
        const ar_object& ob = ctx->objs[k];
        const float e[3] = {std::fabs(ob.extent[0]), std::fabs(ob.extent[1]), std::fabs(ob.extent[2])};
        const obj::Basis bs = obj::basis_from_euler(ob.rotation);
        const size_t i = 3 * ((size_t)ntri + k);
        obj::proxy(ob.type, ob.center, e, bs, &lo[i], &hi[i]);
        for (int a = 0; a < 3; ++a) {
            cen[i + a] = ob.center[a];
            mag = std::max(mag, std::max(std::fabs(lo[i + a]), std::fabs(hi[i + a])));
        }
    }
    for (uint32_t t = 0; t < ntri; ++t) {
        for (int a = 0; a < 3; ++a) {
            const float x = ctx->verts[3 * (size_t)ctx->idx[3 * (size_t)t + 0] + a];
            const float y = ctx->verts[3 * (size_t)ctx->idx[3 * (size_t)t + 1] + a];
            const float z = ctx->verts[3 * (size_t)ctx->idx[3 * (size_t)t + 2] + a];
            const float l = std::min(x, std::min(y, z)), h = std::max(x, std::max(y, z));
            lo[3 * (size_t)t + a] = l;
            hi[3 * (size_t)t + a] = h;
            cen[3 * (size_t)t + a] = 0.5f * (l + h);
            mag = std::max(mag, std::max(std::fabs(l), std::fabs(h)));
        }
    }
    BuildResult br = build_bvh2(lo.data(), hi.data(), cen.data(), n, o.leaf_size, o.bin_count, (int)o.max_depth);
    if (br.depth > kMaxDepth)
        return set_error(ctx, AR_E_CAPACITY, "BVH deeper than the traversal stack allows");
    const float pad = mag * kPadRel + 1e-30f;
    std::vector<GpuNode8> nodes8;
    std::vector<GpuNode4> nodes4;
    std::vector<float> tris;
    size_t node_bytes = 0;
    uint64_t leaves = 0;
    int depth = br.depth;
    const void* node_src = nullptr;
    LeafWriter lw{ctx->verts, ctx->idx, ntri, {}};
    Converter cv{br, lw, pad, {}, {}, 0};
    if (o.width == 8) {
        convert_bvh8(br, lw, pad, nodes8, tris, leaves, depth);
        node_src = nodes8.data();
        node_bytes = nodes8.size() * sizeof(GpuNode8);
        ctx->bvh.stack_entries = 7u * (uint32_t)(depth + 1) + 2u;
        ctx->info.node_count = nodes8.size();
    } else if (o.width == 4) {
        convert_bvh4(br, lw, pad, nodes4, tris, leaves, depth);
        node_src = nodes4.data();
        node_bytes = nodes4.size() * sizeof(GpuNode4);
        ctx->bvh.stack_entries = 3u * (uint32_t)(depth + 1) + 2u;
        ctx->info.node_count = nodes4.size();
    } else {
        cv.nodes.reserve(br.nodes.size() / 2 + 1);
        cv.tris.reserve(12 * (size_t)n);
        cv.run();
        tris.swap(cv.tris);
        leaves = cv.leaves;
        node_src = cv.nodes.data();
        node_bytes = cv.nodes.size() * sizeof(GpuNode2);
        ctx->bvh.stack_entries = (uint32_t)br.depth + 2u;
        ctx->info.node_count = cv.nodes.size();
    }
    // the root node (node 0) carries the tree's origin limit in a spare word,
    // so that a batch (whose context limit is the smallest of its scenes')
    // gives every ray its own scene's limit
    const float olim = kOriginLimit * mag;
    uint32_t olim_bits;
    std::memcpy(&olim_bits, &olim, 4);
    if (o.width == 8 && !nodes8.empty()) nodes8[0].pad[0] = olim_bits;
    else if (o.width == 4 && !nodes4.empty()) nodes4[0].pad[0] = olim_bits;
    else if (o.width == 2 && !cv.nodes.empty()) cv.nodes[0].d[2] = olim_bits;
    int rc = check_hip(ctx, hipMalloc((void**)&ctx->bvh.nodes, node_bytes), "hipMalloc(nodes)");
    if (rc != AR_OK) { ctx->bvh.nodes = nullptr; return rc; }
    rc = check_hip(ctx, hipMalloc(&ctx->bvh.tris, tris.size() * sizeof(float)), "hipMalloc(tris)");
    if (rc != AR_OK) { ctx->bvh.tris = nullptr; free_bvh(ctx); return rc; }
    rc = check_hip(ctx, hipMemcpy(ctx->bvh.nodes, node_src, node_bytes, hipMemcpyHostToDevice), "upload nodes");
    if (rc == AR_OK)
        rc = check_hip(ctx, hipMemcpy(ctx->bvh.tris, tris.data(), tris.size() * sizeof(float),
                                      hipMemcpyHostToDevice), "upload triangles");
    if (rc != AR_OK) { free_bvh(ctx); return rc; }
    ctx->bvh.width = o.width;
    ctx->bvh.obj_slot.swap(lw.obj_slot);
    ctx->bvh.node_count = (uint32_t)ctx->info.node_count;
    ctx->bvh.tri_slots = (uint32_t)(tris.size() / 12);
    ctx->bvh.root = 0;
    ctx->bvh.origin_limit = olim;
    ctx->bvh_ready = true;
    const auto t1 = std::chrono::steady_clock::now();
    ctx->info.leaf_count = leaves;
    ctx->info.depth = (uint32_t)depth;
    ctx->info.device_bytes = node_bytes + tris.size() * sizeof(float);
    ctx->info.build_seconds = std::chrono::duration<double>(t1 - t0).count();
    ctx->info.sah_cost = br.sah_cost;
    return upload_scene(ctx);
}

int ar_bvh_info_get(const ar_ctx* ctx, ar_bvh_info* info) {
    if (!ctx || !info) return AR_E_INVALID;
    if (!ctx->bvh_ready) return AR_E_NOSCENE;
    *info = ctx->info;
    return AR_OK;
}

int ar_trace_closest(ar_ctx* ctx, const ar_rays* rays, size_t n, const ar_hits* hits, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (n == 0) return AR_OK;
    if (!rays_ok(rays) || !hits_ok(hits)) return set_error(ctx, AR_E_INVALID, "NULL ray or hit array");
    TraceArgs ta;
    uint32_t blocks;
    int rc = trace_prologue(ctx, ta, n, blocks);
    if (rc != AR_OK) return rc;
    rc = order_stream(ctx, pick_stream(ctx, stream));
    if (rc != AR_OK) return rc;
    return check_hip(ctx, launch_trace_closest(ta, *rays, n, *hits, nullptr, nullptr, blocks, pick_stream(ctx, stream)),
                     "trace_closest launch");
}

int ar_trace_closest_counted(ar_ctx* ctx, const ar_rays* rays, size_t n, const ar_hits* hits,
                             uint32_t* box_tests, uint32_t* tri_tests, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (n == 0) return AR_OK;
    if (!rays_ok(rays) || !hits_ok(hits) || !box_tests || !tri_tests)
        return set_error(ctx, AR_E_INVALID, "NULL ray, hit or counter array");
    TraceArgs ta;
    uint32_t blocks;
    int rc = trace_prologue(ctx, ta, n, blocks);
    if (rc != AR_OK) return rc;
    rc = order_stream(ctx, pick_stream(ctx, stream));
    if (rc != AR_OK) return rc;
    return check_hip(ctx, launch_trace_closest(ta, *rays, n, *hits, box_tests, tri_tests, blocks,
                                               pick_stream(ctx, stream)), "trace_closest_counted launch");
}

int ar_trace_any(ar_ctx* ctx, const ar_rays* rays, size_t n, uint32_t* occluded, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (n == 0) return AR_OK;
    if (!rays_ok(rays) || !occluded) return set_error(ctx, AR_E_INVALID, "NULL ray or occlusion array");
    TraceArgs ta;
    uint32_t blocks;
    int rc = trace_prologue(ctx, ta, n, blocks);
    if (rc != AR_OK) return rc;
    rc = order_stream(ctx, pick_stream(ctx, stream));
    if (rc != AR_OK) return rc;
    return check_hip(ctx, launch_trace_any(ta, *rays, n, occluded, blocks, pick_stream(ctx, stream)),
                     "trace_any launch");
}

int ar_make_shadow_rays(ar_ctx* ctx, const ar_rays* rays, const ar_hits* hits, size_t n, const float target[3],
                        float tmin, float tmax_scale, const ar_rays* out, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (n == 0) return AR_OK;
    if (!rays_ok(rays) || !hits_ok(hits) || !rays_ok(out) || !target)
        return set_error(ctx, AR_E_INVALID, "NULL argument");
    return check_hip(ctx, launch_shadow_rays(*rays, *hits, n, target[0], target[1], target[2], tmin, tmax_scale, *out,
                                             pick_stream(ctx, stream)), "shadow ray launch");
}

int ar_render(ar_ctx* ctx, const ar_camera* cam, const ar_render_opts* opts, float* out_rgb, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (!cam || !opts || !out_rgb) return set_error(ctx, AR_E_INVALID, "NULL camera, options or output");
    if (opts->width == 0 || opts->height == 0 || opts->spp == 0) return set_error(ctx, AR_E_INVALID, "empty image or spp = 0");
    if (opts->max_depth == 0 || opts->max_depth > 1000) return set_error(ctx, AR_E_INVALID, "max_depth must be 1..1000");
    if (opts->sky > AR_SKY_CONSTANT) return set_error(ctx, AR_E_INVALID, "unknown sky mode");
    if (!ctx->bvh_ready) return set_error(ctx, AR_E_NOSCENE, "no BVH: call ar_scene_set and ar_bvh_build first");
    const int rc = order_stream(ctx, pick_stream(ctx, stream));
    if (rc != AR_OK) return rc;
    return render(ctx, cam, opts, out_rgb, pick_stream(ctx, stream));
}

int ar_render_progressive(ar_ctx* ctx, const ar_camera* cam, const ar_render_opts* opts, uint32_t frame,
                          float* accum, float* odd, float* even, float* gbuffer, float* stats, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (!cam || !opts || !accum) return set_error(ctx, AR_E_INVALID, "NULL camera, options or accumulation buffer");
    if ((odd == nullptr) != (even == nullptr)) return set_error(ctx, AR_E_INVALID, "odd and even go together");
    if (opts->width == 0 || opts->height == 0 || opts->spp == 0) return set_error(ctx, AR_E_INVALID, "empty image or spp = 0");
    if (!ctx->bvh_ready) return set_error(ctx, AR_E_NOSCENE, "no BVH: call ar_scene_set and ar_bvh_build first");
    const int rc = order_stream(ctx, pick_stream(ctx, stream));
    if (rc != AR_OK) return rc;
    return render_progressive(ctx, cam, opts, frame, accum, odd, even, gbuffer, stats, pick_stream(ctx, stream));
}

int ar_render_counters(ar_ctx* ctx, int enable) {
    if (!ctx) return AR_E_INVALID;
    (void)hipSetDevice(ctx->device);
    if (!enable) {
        ctx->count_on = false;
        return AR_OK;
    }
    if (!ctx->cnt_totals) {
        const int rc = check_hip(ctx, hipMalloc((void**)&ctx->cnt_totals, 5 * sizeof(unsigned long long)),
                                 "hipMalloc(test totals)");
        if (rc != AR_OK) return rc;
    }
    const unsigned long long zero[5] = {0, 0, 0, 0, 0};
    const int rc = check_hip(ctx, hipMemcpy(ctx->cnt_totals, zero, sizeof zero, hipMemcpyHostToDevice), "zero totals");
    if (rc != AR_OK) return rc;
    invalidate_graphs(ctx);   // graphs captured before counting do not count
    ctx->count_on = true;
    return AR_OK;
}

int ar_render_counters_read(ar_ctx* ctx, uint64_t out[5]) {
    if (!ctx || !out) return AR_E_INVALID;
    if (!ctx->cnt_totals) return set_error(ctx, AR_E_INVALID, "counting was never enabled (ar_render_counters)");
    (void)hipSetDevice(ctx->device);
    unsigned long long h[5];
    int rc = check_hip(ctx, hipDeviceSynchronize(), "counters");
    if (rc == AR_OK) rc = check_hip(ctx, hipMemcpy(h, ctx->cnt_totals, sizeof h, hipMemcpyDeviceToHost), "read totals");
    if (rc != AR_OK) return rc;
    for (int k = 0; k < 5; ++k) out[k] = (uint64_t)h[k];
    const unsigned long long zero[5] = {0, 0, 0, 0, 0};
    return check_hip(ctx, hipMemcpy(ctx->cnt_totals, zero, sizeof zero, hipMemcpyHostToDevice), "zero totals");
}

} // extern "C"
