// Internal types of libar_hip: the context, the device BVH layout and the
// host builder's interface.  Not part of the ABI (include/ar.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ar.h"

namespace ar {

// ---------------------------------------------------------------------------
// Device BVH2 layout ("both child boxes in the parent", one 64-byte node).
//
//   float4 a = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   float4 b = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   float4 c = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   uint4  d = (ref0, ref1, 0, 0)
//
// A child reference with bit 31 clear is an interior node index.  With bit
// 31 set it is a leaf: bits 0..26 = first triangle slot, bits 27..30 =
// triangle count - 1.  kRefEmpty (a leaf of count 16 at the last slot, never
// produced by the builder) marks an empty child and the stack sentinel.
//
// Leaf triangles are stored in leaf order, three float4 per slot:
//   (v0.xyz, prim id bits), (v1.xyz, 0), (v2.xyz, 0).
// Child boxes are padded outward by kPadRel * (largest |coordinate| of the
// scene) so that the box test never culls a triangle the watertight test
// would report (rounding of either test stays far below the pad for ray
// origins within kOriginLimit scene magnitudes; beyond that the kernel adds a
// per-ray pad in t).
// ---------------------------------------------------------------------------
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kRefEmpty = 0xFFFFFFFFu;
constexpr uint32_t kMaxLeafTris = 15;            // count-1 must fit 4 bits and differ from kRefEmpty
constexpr uint32_t kMaxTriSlots = (1u << 27) - 1;
constexpr float kPadRel = 1.0f / 65536.0f;       // 2^-16
constexpr float kOriginLimit = 8.0f;
constexpr int kMaxDepth = 60;                    // tree depth bound (stack <= depth + 1)
constexpr int kStackMax = 64;                    // per-ray stack entries incl. sentinel
constexpr int kLdsStack = 16;                    // entries held in LDS, the rest spill to HBM
constexpr int kTraceBlock = 256;

inline uint32_t make_leaf_ref(uint32_t first, uint32_t count) {
    return kLeafBit | ((count - 1u) << 27) | first;
}

// A leaf slot whose first word carries kObjSlot holds analytic object
// (word & ~kObjSlot) instead of a triangle: (0, 0, 0, word), (0), (0).
// Triangle prim ids stay below 2^27, so the bit is free.
constexpr uint32_t kObjSlot = 0x80000000u;

struct GpuNode2 {
    float a[4];
    float b[4];
    float c[4];
    uint32_t d[4];       // ref0, ref1, 0 (the root: the tree's origin limit, float), 0
};
static_assert(sizeof(GpuNode2) == 64, "GpuNode2 must be 64 bytes");

// ---------------------------------------------------------------------------
// Device BVH8 layout (quantised 8-wide node, 96 bytes = 6 x 16 B of data, padded to 128 B).
//
//   float4 h  = (p.x, p.y, p.z, bits: e.x | e.y << 8 | e.z << 16 | nchild << 24 | axis << 28)
//   uint4 r0  = ref[0..3],  uint4 r1 = ref[4..7]   (same encoding as BVH2 refs)
//   uint4 q0  = lo.x[0..7] bytes, lo.y[0..7] bytes
//   uint4 q1  = lo.z[0..7] bytes, hi.x[0..7] bytes
//   uint4 q2  = hi.y[0..7] bytes, hi.z[0..7] bytes
//
// Child c's box is p + q * 2^(e - 127) per axis (e is a biased exponent, so
// 2^(e-127) is built exactly from bits).  The builder rounds q outward and
// checks the decoded box in float against the padded child box, so the
// decoded box always contains it.
// ---------------------------------------------------------------------------
struct GpuNode8 {
    float p[3];
    uint32_t meta;
    uint32_t ref[8];
    uint8_t q[48];
    uint32_t pad[8];     // to 128 B: one node is one L2 line; pad[0] of the root: the tree's origin limit (float)
};
static_assert(sizeof(GpuNode8) == 128, "GpuNode8 must be 128 bytes");

// Quantised 4-wide node, 64 bytes = 4 x 16 B (half an L2 line):
//
//   float4 h  = (p.x, p.y, p.z, bits: e.x | e.y << 8 | e.z << 16 | nchild << 24 | axis << 28)
//   uint4 r   = ref[0..3]
//   uint4 q0  = lo.x[0..3] bytes, lo.y[0..3], lo.z[0..3], hi.x[0..3]
//   uint4 q1  = hi.y[0..3] bytes, hi.z[0..3], 0, 0
//
// with the 8-wide layout's grid, rounding and child order.
struct GpuNode4 {
    float p[3];
    uint32_t meta;
    uint32_t ref[4];
    uint8_t q[24];
    uint32_t pad[2];     // pad[0] of the root: the tree's origin limit (float)
};
static_assert(sizeof(GpuNode4) == 64, "GpuNode4 must be 64 bytes");

// Host-side binary tree produced by the builder (pre-order, left = parent+1).
struct BuildNode {
    float lo[3], hi[3];
    int32_t left, right;   // -1 for leaves
    uint32_t first, count; // leaf range in the primitive order
};

struct BuildResult {
    std::vector<BuildNode> nodes;
    std::vector<uint32_t> order;  // primitive ids in leaf order
    int depth = 0;
    double sah_cost = 0.0;
};

// Binned-SAH builder (bvh_build.cpp).  lo/hi/cen: n x 3 floats.
BuildResult build_bvh2(const float* lo, const float* hi, const float* cen, uint32_t n,
                       uint32_t leaf_size, uint32_t bins, int max_depth);

// Writes leaf slots (12 floats each) for primitives in leaf order: triangles
// (prim < ntri) as their vertices, objects as a kObjSlot marker; records the
// slot of every object (its solo-trace leaf).
struct LeafWriter {
    const std::vector<float>& verts;
    const std::vector<uint32_t>& idx;
    uint32_t ntri;
    std::vector<uint32_t> obj_slot;
    void write(std::vector<float>& slots, uint32_t prim);
};

// Collapse into quantised 8- / 4-wide nodes (bvh_wide.cpp); depth = wide-tree depth.
bool convert_bvh8(const BuildResult& br, LeafWriter& lw, float pad, std::vector<GpuNode8>& nodes,
                  std::vector<float>& tris, uint64_t& leaves, int& depth);
bool convert_bvh4(const BuildResult& br, LeafWriter& lw, float pad, std::vector<GpuNode4>& nodes,
                  std::vector<float>& tris, uint64_t& leaves, int& depth);

struct DeviceBVH {
    uint32_t width = 2;          // 2: GpuNode2 nodes, 8: GpuNode8 nodes
    GpuNode2* nodes = nullptr;   // device (GpuNode8* when width == 8)
    float* tris = nullptr;       // device, 12 floats per slot
    uint32_t node_count = 0;
    uint32_t tri_slots = 0;
    uint32_t root = kRefEmpty;   // root interior node index (0) or kRefEmpty
    float origin_limit = 0.0f;   // kOriginLimit * scene magnitude
    uint32_t stack_entries = kStackMax;  // per-ray traversal stack bound (incl. sentinel)
    std::vector<uint32_t> obj_slot;      // leaf slot of each analytic object
};

// Shading data of the scene, prim order (uploaded with the BVH).  The
// spectral fields hold the sigmoid coefficients of the albedo and of the
// emission normalised by its largest channel (escale = that channel if > 1).
struct GpuMaterial {
    float albedo[3];
    float specular;
    float emission[3];
    float metallic;
    float roughness;
    uint32_t model;
    float escale;
    float ascale;        // network basis: the albedo's largest channel when above 1
    float calb[4];       // albedo: sigmoid coefficients (3), or the network's RGB basis (4)
    float cem[4];        // emission normalised by escale: likewise
    float ior, sigma_a, abbe, tilt;   // glass
};
static_assert(sizeof(GpuMaterial) == 96, "GpuMaterial must be 96 bytes");

namespace spec {
size_t lut_floats();
const std::vector<float>& uplift_lut(const float* xyz, const float* d65);
void uplift(const float* lut, const float rgb[3], float c[3]);
const float* cie_xyz_table();   // 421 x 3, 360..780 nm
const float* d65_table();       // 421
}

struct SceneDev {
    float4* prims = nullptr;      // 3 per triangle: (v0, material bits), (v1, light index bits), (v2, 0)
    GpuMaterial* mats = nullptr;
    int32_t* light_tri = nullptr;
    float* light_area = nullptr;
    float* light_prob = nullptr;
    float* light_cdf = nullptr;
    uint32_t nlights = 0;
    uint32_t ntri = 0;
    std::vector<GpuMaterial> host_mats;   // host copy (spectral coefficients filled on first spectral render)
    bool spectral_ready = false;
    uint32_t spectral_basis = 0;    // AR_BASIS_* the material records were prepared for
    float4* objs = nullptr;               // analytic objects, obj::kRecord float4 each (objects.hpp)
    uint32_t nobj = 0;
    uint32_t nsdf = 0;                    // SDF objects among them
    std::vector<float4> host_objs;
    float box_lo[3] = {0, 0, 0};          // scene bounds (guiding-network inputs)
    float box_inv[3] = {1, 1, 1};         // 1 / extent
};

// Signed-distance volumes: every volume's voxels back to back (host copy,
// the source of truth), uploaded to `dev` when the scene is uploaded.
struct SdfPool {
    std::vector<float> host;
    std::vector<uint32_t> off;        // first voxel of each volume
    std::vector<uint32_t> dims;       // 3 per volume
    float* dev = nullptr;
    size_t dev_floats = 0;
    bool dirty = false;               // host changed since the last upload
};

// Wavefront state of the integrator, sized for `capacity` paths.
struct RenderBuffers {
    size_t capacity = 0;
    void* block = nullptr;        // one allocation, carved below
    float* rays[2][8];            // queue rays (current / next), SoA
    uint32_t* qpath[2];           // queue slot -> path id
    uint32_t* hit_prim; float* hit_t; float* hit_u; float* hit_v; uint32_t* hit_flags;
    float* srays[8];              // shadow rays
    float* smeps;                 // shadow ray SDF march tolerance (scenes with SDF objects)
    float* fh = nullptr;          // per-path first-hit records of the denoiser G-buffer (allocated on demand)
    size_t fh_cap = 0;            // paths fh holds (its array stride)
    uint32_t* spath;
    float* scontrib[4];
    uint32_t* sflags;             // bit0: the sampled light point faced the surface (both cosines > 0)
    uint32_t* occ;                // shadow occlusion bits
    float* mask[4];
    float* rad[4];
    float* last_pos[3];
    float* last_pdf;
    uint32_t* meta;               // non-delta depth (bits 0..15), previous vertex took NEE (bit 16)
    uint32_t* counters;           // [0] current queue, [1] next queue, [2] shadow queue
    float* lambda0;               // per path: first wavelength (spectral modes)
    float* accum;                 // width*height*3 running sums
    size_t accum_pixels = 0;
    uint32_t* pix;                // per path: pixel, sample index, scene
    uint32_t* smp;
    uint32_t* scn;
    uint32_t* qroot[2];           // queue ray -> BVH root of its scene
    uint32_t* sroot;              // shadow ray -> BVH root
    uint32_t* med;                // medium of the path: 0 none, else object index + 1
    float* mpos[3];               // where the path entered its medium
    float* eps_table;             // INTERSECT_EPS(d) for d < 256
    float* cie;                   // device copies of the CIE 1931 / D65 tables
    float* d65;
};

// Voxel-guiding state: probe pool, key -> probe hash table, this frame's
// accumulators, the collected samples of a pass and the per-path slots.
struct GuideDev {
    uint32_t cap = 0;               // probe capacity
    uint32_t tcap = 0;              // hash table slots (power of two >= 2 cap)
    uint64_t* tkey = nullptr;       // table keys (all ones = empty)
    uint32_t* tval = nullptr;       // probe index per table slot (all ones = none)
    uint64_t* pkey = nullptr;       // key of each probe
    float4* tex = nullptr;          // 86 texels per probe
    uint64_t* cnt = nullptr;        // merged sample count per probe (capped at 1e6)
    float* acc = nullptr;           // this frame's sums, [probe][channel][leaf]
    uint32_t* accn = nullptr;       // this frame's sample count per probe
    uint32_t* touched = nullptr;    // probes with samples this frame
    uint32_t* newslots = nullptr;   // table slots inserted this frame
    uint32_t* counters = nullptr;   // see guiding.hip
    size_t bytes = 0;
    // samples of one pass
    size_t scap = 0;
    uint32_t* s_slot = nullptr;
    uint32_t* s_leaf = nullptr;
    float* s_rad = nullptr;         // [3][scap]
    // per-path collection slots
    size_t pcap = 0;
    float* gm = nullptr;            // slot throughput [5*4][pcap]
    float* gr = nullptr;            // slot radiance [5*4][pcap]
    uint64_t* gkey = nullptr;       // slot probe key [5][pcap]
    uint32_t* gleaf = nullptr;      // slot leaf texels, 6 bits each
    uint32_t* gflag = nullptr;      // bits 0..4: slot received radiance
    float* gx = nullptr;            // learned guiding: slot network inputs [5*kRec][pcap] (features, scene)
    float* gw = nullptr;            // slot sampled direction [5*3][pcap]
    float* gp = nullptr;            // slot direction pdf [5][pcap]
};

// Learned guiding network state (mlp.hip)
struct MlpDev {
    float* p = nullptr;         // FP32 parameters (one allocation: p, m, v, g, scal, wb, wt)
    float* m = nullptr;         // Adam moments
    float* v = nullptr;
    float* g = nullptr;         // gradient
    float* scal = nullptr;      // [0] weight scale, [1] loss, [2] Adam steps taken
    float* part = nullptr;      // training partials (one allocation): per-block gradients [part_blocks][params],
    float* part2 = nullptr;     //   chunk sums, per-block and per-chunk losses, weight-sum partials
    float* ploss = nullptr;
    float* ploss2 = nullptr;
    float* wpart = nullptr;
    uint32_t part_blocks = 0;   // blocks of a training launch at most
    uint16_t* wb = nullptr;     // BF16 weights [in][out]
    uint16_t* wt = nullptr;     // BF16 weights [out][in]
    float lr = 5e-3f;
    uint32_t batch = 1u << 18;  // path vertices per training step
    uint32_t iters = 1;         // training steps per frame
    uint64_t seed = 1;
    // integrator buffers
    size_t cap = 0;
    float* feat = nullptr;      // [cap][kRec] network inputs of the queued vertices (features, scene)
    float* dec = nullptr;       // [cap][48] decoded mixtures
    size_t rcap = 0;
    float* rx = nullptr;        // training records: inputs [rcap][kRec]
    float* rd = nullptr;        // directions [rcap][3]
    float* rw = nullptr;        // weights [rcap]
    uint32_t* rcount = nullptr; // [0] records trained on this frame (min(T, batch)), [1] records T of the frame
    uint32_t* rblk = nullptr;   // per block of 256 selected paths: record count, then its offset
    size_t rblocks = 0;         // blocks rblk holds
};

// Spectral basis network (smlp.hip): parameters, AdamW state, training
// scratch and the uplift tables built from the parameters.
struct SmlpDev {
    float* p = nullptr;          // parameters (one allocation: p, m, v, ema, g)
    float* m = nullptr;
    float* v = nullptr;
    float* ema = nullptr;
    float* g = nullptr;          // gradient of the last batch (after reduction)
    bool have = false;           // parameters set
    uint64_t t = 0;              // AdamW steps taken
    float* part = nullptr;       // per-tile gradient partials [tiles][params]
    float* part2 = nullptr;      // per-chunk sums [chunks][params]
    float* misc = nullptr;       // per-block squared norms, per-tile losses
    size_t part_tiles = 0;
    float* rgb_lut = nullptr;    // [res^3][4], x fastest
    uint32_t rgb_res = 0;
    float* lam_lut = nullptr;    // [res][4]
    uint32_t lam_res = 0;
    uint32_t basis = 0;          // AR_BASIS_* used by spectral renders
};
// per-material network bases (calb, cem, ascale) of n materials on the device
int smlp_material_bases(ar_ctx* ctx, GpuMaterial* mats, uint32_t n);

// One scene of a batch context (ar_batch_create): its place in the merged arrays.
struct BatchEntry {
    uint32_t root;         // BVH root node
    uint32_t light_off;    // first light
    uint32_t light_cnt;
    uint32_t tri_off;      // first triangle (prim id offset)
    uint32_t obj_off;      // first analytic object of the scene in the merged object records
    uint32_t obj_cnt;
    float box_lo[3];       // the scene's box (learned-guiding inputs are normalised to it)
    float box_inv[3];
};

struct RenderJob;           // integrator.hip: a planned render and its device plan

// a captured render (ar_graph_capture)
struct GraphRec {
    ar_ctx* ctx = nullptr;
    RenderJob* job = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t generation = 0;
};

} // namespace ar

struct ar_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int num_cus = 0;

    // scene, host copy
    std::vector<float> verts;         // 3 per vertex
    std::vector<uint32_t> idx;        // 3 per triangle (global vertex ids)
    std::vector<uint32_t> tri_mat;    // per triangle
    std::vector<ar_material> mats;
    std::vector<ar_object> objs;      // analytic objects (prims ntri ..)
    bool scene_set = false;

    ar::DeviceBVH bvh;
    bool bvh_ready = false;
    ar_bvh_info info{};

    ar::SceneDev scene;
    ar::RenderBuffers rb;

    // traversal scratch: per-thread stack spill, sized for the persistent grid
    int* spill = nullptr;
    size_t spill_entries = 0;          // total ints allocated
    uint32_t spill_stack = 0;          // per-thread stack entries the spill buffer covers
    uint32_t* fetch = nullptr;         // ray-refill counter of the traversal kernels
    uint32_t trace_threads = 0;       // persistent grid size in threads

    ar::GuideDev guide;
    uint64_t guide_cap_req = 1u << 20;
    ar::MlpDev mlp;

    std::vector<ar::BatchEntry> batch;   // non-empty for a batch context
    ar::RenderJob* job = nullptr;         // plan of the last render
    uint32_t tile = 256;                  // threads per block of the wavefront kernels
    uint64_t buffer_generation = 0;       // bumped whenever device memory a captured render uses is freed
    std::vector<ar::GraphRec*> graphs;    // live captured renders of this context (detached by ar_destroy)
    hipStream_t last_stream = nullptr;    // stream of the last traversal / render call
    bool have_last_stream = false;
    hipEvent_t order_event = nullptr;     // orders a call on a new stream after the last one
    double* reduce = nullptr;             // partial sums of ar_luminance_mse
    ar::SdfPool sdf;                      // signed-distance volumes
    float* denoise_feat = nullptr;        // per-pixel filter features of ar_denoise (29 planes)
    float* host_out = nullptr;            // ar_render_host's device image (3 floats HDR + 1 word LDR per pixel)
    size_t host_out_pixels = 0;
    // ar_render_counters: per-ray box / triangle test counts of the integrator's
    // traversals and their totals (closest box, closest tri, any box, any tri,
    // shaded vertices) while counting is on
    bool count_on = false;
    uint32_t* cnt_box = nullptr;
    uint32_t* cnt_tri = nullptr;
    size_t cnt_cap = 0;
    unsigned long long* cnt_totals = nullptr;
    float* present_buf = nullptr;         // ar_present's denoised image (3 floats per pixel)
    size_t present_pixels = 0;
    ar::SmlpDev smlp;                     // spectral basis network
    size_t denoise_floats = 0;
};

namespace ar {
int set_error(ar_ctx* ctx, int code, const std::string& msg);
// Captured renders read the context's device buffers by address: every free
// of such a buffer outdates them (ar_graph_launch then refuses to run them).
inline void invalidate_graphs(ar_ctx* ctx) { ctx->buffer_generation += 1; }
// Work of one context shares scratch (refill counter, stack spill, path
// state): a call queued on a stream other than the previous call's is made to
// wait for the previous call's stream first.
int order_stream(ar_ctx* ctx, hipStream_t s);
int check_hip(ar_ctx* ctx, hipError_t e, const char* what);

// trace.hip launchers
struct TraceArgs {
    uint32_t width;
    uint32_t stack_entries;
    const GpuNode2* nodes;
    const float4* tris;
    uint32_t root;
    float origin_limit;
    int* spill;
    uint32_t threads;    // spill-buffer stride (threads it is laid out for)
    uint32_t full_closest = 1, full_any = 1;   // resident blocks of each kernel
    uint32_t* fetch;     // device ray-refill counter
    const uint32_t* roots = nullptr;   // per-ray root node (batched scenes), else `root`
    const float4* objs = nullptr;      // analytic object records; NULL: triangle-only scene
    uint32_t obj_base = 0;             // prim id of object 0 (= triangle count)
    const float* vol = nullptr;        // SDF voxel pool (objects.hpp)
};
// n_dev (optional): device pointer to the ray count, read by the kernel
// (then n is only the capacity used to size the grid).
hipError_t launch_trace_closest(const TraceArgs& ta, const ar_rays& r, size_t n, const ar_hits& h,
                                uint32_t* cnt_box, uint32_t* cnt_tri, uint32_t blocks, hipStream_t s,
                                const uint32_t* n_dev = nullptr);
hipError_t launch_trace_any(const TraceArgs& ta, const ar_rays& r, size_t n, uint32_t* occ,
                            uint32_t blocks, hipStream_t s, const uint32_t* n_dev = nullptr,
                            uint32_t* cnt_box = nullptr, uint32_t* cnt_tri = nullptr);
hipError_t launch_shadow_rays(const ar_rays& r, const ar_hits& h, size_t n, float tx, float ty, float tz,
                              float tmin, float tmax_scale, const ar_rays& out, hipStream_t s);
int trace_blocks_per_cu(int width = 0, bool any = false);
// device address of ar_trace_closest_noalloc's scratch array on `device`
// (the current device); resolved once per device
hipError_t noalloc_scratch_setup(int device);

// integrator.hip
int upload_scene(ar_ctx* ctx);
// sdf.hip: uploads the voxel pool when it changed
int sdf_upload(ar_ctx* ctx);
void sdf_free(ar_ctx* ctx);
void free_scene(ar_ctx* ctx);
void free_render_buffers(ar_ctx* ctx);
int render(ar_ctx* ctx, const ar_camera* cam, const ar_render_opts* opts, float* out, hipStream_t s);
int render_progressive(ar_ctx* ctx, const ar_camera* cam, const ar_render_opts* opts, uint32_t frame, float* accum,
                       float* odd, float* even, float* gbuffer, float* stats, hipStream_t s);
int render_batch(ar_ctx* ctx, uint32_t n, const ar_camera* cams, const ar_render_opts* opts, float* const* outs,
                 hipStream_t s);
int graph_capture(ar_ctx* ctx, uint32_t n, const ar_camera* cams, const ar_render_opts* opts, float* const* outs,
                  GraphRec* g);
void free_job(RenderJob* job);
// batch.cpp
int batch_merge(ar_ctx* out, ar_ctx* const* scenes, uint32_t n);
int trace_prologue(ar_ctx* ctx, TraceArgs& ta, size_t n, uint32_t& blocks);

// guiding.hip
int guide_ensure(ar_ctx* ctx, size_t paths, bool pool);
void guide_free(ar_ctx* ctx);
int guide_clear(ar_ctx* ctx, hipStream_t s);
// merges the samples collected by the paths of one pass (spectral mode and
// first wavelengths give the XYZ weights)
int guide_update(ar_ctx* ctx, uint32_t spectral, const float* lambda0, const float* cie, size_t paths,
                 hipStream_t s);

// mlp.hip
int mlp_ensure(ar_ctx* ctx);
int mlp_init(ar_ctx* ctx, uint64_t seed);
int mlp_set_params(ar_ctx* ctx, const float* host);
void mlp_free(ar_ctx* ctx);
// smlp.hip
void smlp_free(ar_ctx* ctx);
hipError_t mlp_forward(ar_ctx* ctx, const float* x5, int xs, size_t n, const uint32_t* n_dev, float* out, bool decoded,
                       hipStream_t s);
// norm: the loss is normalised by 1/n (kNormCount), by 1/sum(w) computed
// here (kNormSum) or by the normaliser the previous call computed from the
// same weights (kNormKeep); adam: an Adam step is applied with the gradient
// (and the device-side step count advanced)
enum { kNormCount = 0, kNormSum = 1, kNormKeep = 2 };
hipError_t mlp_gradient(ar_ctx* ctx, const float* x5, int xs, const float* dir, const float* w, size_t n,
                        const uint32_t* n_dev, int norm, hipStream_t s, bool adam);
int mlp_ensure_buffers(ar_ctx* ctx, size_t paths);
// training records from the frame's guiding slots, then the frame's Adam steps
int mlp_frame_update(ar_ctx* ctx, uint32_t spectral, const float* lambda0, const float* cie, size_t paths,
                     const uint32_t* ranges, uint32_t nseg, hipStream_t s);
} // namespace ar
