/*
 * ar.h — C ABI of the Aperture MI355X path-tracing core (libar_hip.so).
 *
 * Plain C types only: pointers, sizes, PODs.  No HIP, torch or C++ types
 * cross this boundary.  Every entry point returns an int status
 * (AR_OK = 0, negative AR_E_* on failure); ar_last_error() gives the text.
 * No exception crosses the ABI.
 *
 * Each entry point cites the reference interface it replaces (paths are
 * relative to the reference checkout of ApertureRendererV3).
 *
 * Threading: one ar_ctx per host thread (thread-compatible, not
 * thread-safe).  Device work is queued on the stream passed in (a
 * hipStream_t; NULL = the HIP default stream) and is not waited for unless
 * the entry point says so.  The traversal and render calls of one context
 * share device scratch (ray-refill counter, stack spill, path state): a call
 * on a stream other than the previous call's is queued after the previous
 * call's stream (an event wait), so two streams never race on one context;
 * use one context per stream for concurrency.
 */
#ifndef AR_H
#define AR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AR_ABI_VERSION 2

#define AR_OK 0
#define AR_E_INVALID (-1)     /* bad argument */
#define AR_E_DEVICE (-2)      /* HIP runtime error */
#define AR_E_NOSCENE (-3)     /* no scene / acceleration structure yet */
#define AR_E_OOM (-4)         /* device or host allocation failed */
#define AR_E_UNSUPPORTED (-5) /* option not supported */
#define AR_E_CAPACITY (-6)    /* a fixed capacity would be exceeded */

#define AR_MISS 0xFFFFFFFFu   /* ar_hits.prim of a ray that hit nothing */

/* Hit flags (ar_hits.flags): barycentric weight of v0 / v1 / v2 exactly zero
 * (the ray passed through an edge or a vertex), and det < 0 (the triangle's
 * vertices wind clockwise as seen along the ray). */
#define AR_HIT_W0_ZERO 1u
#define AR_HIT_W1_ZERO 2u
#define AR_HIT_W2_ZERO 4u
#define AR_HIT_BACKFACE 8u
#define AR_HIT_INSIDE 16u      /* analytic object hit from inside (its u = v = 0) */

typedef struct ar_ctx ar_ctx;

typedef struct {
    int device;        /* HIP device ordinal */
    uint32_t flags;    /* reserved, 0 */
} ar_device_opts;

/* Material of a triangle or analytic object (the reference's GBuffer /
 * Material, core/render/object.cuh:36-150). */
typedef struct {
    float albedo[3];
    float emission[3];   /* radiance; > 0 makes the primitive a light (MAT_LIGHT) */
    float specular;      /* UE-style specular (F0 = 0.08*specular for dielectrics) */
    float metallic;
    float roughness;     /* GGX roughness, clamped to [0.01, 1] */
    uint32_t model;      /* AR_MODEL_* */
    float ior;           /* glass: refractive index n_d at 587.56 nm (GBuffer IOR_, default 1.5) */
    float sigma_a;       /* glass: absorption scale of the Beer-Lambert term (sigmaA_, default 100) */
    float abbe;          /* glass: Abbe number V_d, used in spectral modes (abbeNumber_, default 20) */
    float tilt;          /* glass: spectral tilt of sigma_a, -1..1 (tiltNumber_, default 0) */
} ar_material;

#define AR_MODEL_LAMBERT 0u
#define AR_MODEL_GGX 1u      /* GGX + Lambert mix (core/render/material.cuh:47-165) */
#define AR_MODEL_GLASS 2u    /* smooth dielectric (material.cuh:116-161, 206-257); analytic objects only */

/* Analytic primitives (the reference's SceneObject of type OBJ_CUBE /
 * OBJ_SPHERE / OBJ_SDF, core/render/object.cuh:207-305; intersectors
 * core/object/aabb.cuh:9-106, core/object/sphere.cuh:10-85 and
 * core/object/sdf/sdf_base.cuh:54-180).  A signed-distance object lives in
 * its box: unit space u = local / (2 extent) in [-0.5, 0.5]^3, distances
 * scaled by 2 min(extent).  It is sphere-traced (at most 512 steps) with the
 * ray's march tolerance (ar_rays.meps), its normal is the distance gradient,
 * and it cannot be a light. */
#define AR_OBJ_BOX 0u          /* oriented box of half-extents `extent` */
#define AR_OBJ_SPHERE 1u       /* sphere of radius max(extent) */
#define AR_OBJ_SDF_TORUS 2u    /* SDF_TORUS: radii sdf[0] (ring), sdf[1] (tube) in unit space (CreateTorus: 0.375, 0.125) */
#define AR_OBJ_SDF_VOLUME 3u   /* SDF_VOLUME: baked distance grid sdf_volume (ar_sdf_volume_bake / _add) */

typedef struct {
    uint32_t type;        /* AR_OBJ_* */
    float center[3];
    float extent[3];      /* half extents (absolute values are used) */
    float rotation[3];    /* Euler angles in degrees; objectToWorld = Rz * Ry * Rx */
    ar_material material;
    float sdf[4];         /* AR_OBJ_SDF_TORUS: ring and tube radius (sdf[2], sdf[3] unused) */
    int32_t sdf_volume;   /* AR_OBJ_SDF_VOLUME: volume index */
} ar_object;

/* A triangle mesh in host memory.  vertices: nv x 3 floats; indices:
 * ntri x 3 u32; tri_material: ntri u32 indices into materials (NULL: all
 * triangles use materials[0]; materials may be NULL when nmaterials = 0). */
typedef struct {
    const float* vertices;
    uint32_t nv;
    const uint32_t* indices;
    uint32_t ntri;
    const uint32_t* tri_material;
    const ar_material* materials;
    uint32_t nmaterials;
} ar_mesh;

/* Rays and hits in structure-of-arrays form, device pointers, one element
 * per ray.  A ray is valid for t in [tmin, tmax). */
typedef struct {
    const float *ox, *oy, *oz, *tmin, *dx, *dy, *dz, *tmax;
    /* optional: the SDF march tolerance of each ray (the reference's
     * SDF_EPS(depth) = INTERSECT_EPS(depth) / 2); NULL: max(tmin, 1e-3) / 2,
     * which is SDF_EPS(depth) for rays with tmin = INTERSECT_EPS(depth) */
    const float* meps;
} ar_rays;

/* Closest-hit output.  prim = AR_MISS on a miss (then t = +inf, u = v = 0,
 * flags = 0).  Otherwise prim is the scene primitive index: triangles first
 * (meshes in the order given to ar_scene_set, triangles consecutively), then
 * the analytic objects of ar_scene_set_objects.  For a triangle the hit point
 * is (1-u-v)*v0 + u*v1 + v*v2.  For an object t is the distance along
 * normalize(d) (the reference intersectors normalise the direction), so
 * scenes with objects are traced with unit-length directions.  The hit is
 * the one with the smallest t; among equal t, the smallest prim. */
typedef struct {
    uint32_t* prim;
    float* t;
    float* u;
    float* v;
    uint32_t* flags;
} ar_hits;

typedef struct {
    uint32_t width;      /* branching factor: 2 (64 B nodes), 4 (quantised 64 B nodes, default) or 8
                            (quantised 128 B nodes) */
    uint32_t leaf_size;  /* max triangles per leaf, 1..15 (default 2) */
    uint32_t bin_count;  /* SAH bins, 4..64 */
    uint32_t max_depth;  /* tree depth limit, <= 64 */
} ar_bvh_opts;

typedef struct {
    uint64_t node_count;
    uint64_t leaf_count;
    uint32_t depth;
    uint32_t width;
    uint64_t device_bytes;   /* nodes + leaf triangles */
    double build_seconds;    /* host build + upload */
    double sah_cost;
} ar_bvh_info;

/* --- context --------------------------------------------------------------
 * Replaces SceneRenderer's constructor / destructor (core/scene.hpp:94-96,
 * core/scene.cu:609-654, 745-750). */
int ar_create(ar_ctx** out, const ar_device_opts* opts);
int ar_destroy(ar_ctx* ctx);
const char* ar_last_error(const ar_ctx* ctx);
int ar_abi_version(void);

/* --- scene ----------------------------------------------------------------
 * Replaces SceneRenderer(const std::vector<SceneObject>&) + MarkAsDirty
 * (core/scene.hpp:94-101) and the upload in UpdateSceneToDeviceIfDirty
 * (core/scene.cu:656-732).  Copies the meshes; the caller keeps ownership of
 * its arrays.  Invalidates the BVH. */
int ar_scene_set(ar_ctx* ctx, const ar_mesh* meshes, size_t nmesh);

/* Replaces the scene's analytic objects (the box, sphere and SDF
 * SceneObjects of SceneRenderer(const std::vector<SceneObject>&),
 * core/scene.hpp:94-101).  Object k is primitive ntri + k.  A box or sphere
 * with emission is a light, sampled by the facing-facet box sampler or the
 * spherical-cap sampler (core/nee/nee.cuh:153-310); an emissive SDF object
 * is rejected (AR_E_UNSUPPORTED).  Copies the array; invalidates the BVH. */
int ar_scene_set_objects(ar_ctx* ctx, const ar_object* objects, size_t nobj);

/* --- signed-distance volumes ---------------------------------------------
 * Replace SDFCacheManager (core/object/sdf/sdf_cache_manager.cuh:39-114,
 * SceneRenderer::AddVolumeToCacheManager, core/scene.hpp:118-126).  A volume
 * is a grid of dims[0] x dims[1] x dims[2] unit-space distances (x fastest)
 * over the unit cube [-0.5, 0.5]^3; AR_OBJ_SDF_VOLUME objects refer to it by
 * index.  Volumes must exist before ar_scene_set_objects names them. */
#define AR_SDF_SHAPE_TORUS 0u      /* SdfTorus, params (R, r) */
#define AR_SDF_SHAPE_DIAMOND 1u    /* SdfDiamondUnit (sdf_diamond.cuh:21-66), params (d, bias, tableZ, scale);
                                      zeros take 0.94, 0.5, 0.3, 1 (CreateDiamond passes scale 0.5) */

/* FillSDFVolume (sdf_cache_manager.cuh:16-37) on the GPU: voxel (x, y, z) =
 * shape(p01 - 0.5), p01 = (x, y, z) / (dims - 1).  dims: 1..4096 each.
 * Waits for the result. */
int ar_sdf_volume_bake(ar_ctx* ctx, uint32_t shape, const float params[4], const uint32_t dims[3],
                       uint32_t* index);
/* A volume from host data (dims[0] * dims[1] * dims[2] floats, x fastest). */
int ar_sdf_volume_add(ar_ctx* ctx, const float* voxels, const uint32_t dims[3], uint32_t* index);
/* Copies volume `index` to host memory (its dims[0] * dims[1] * dims[2] floats);
 * dims (optional) receives its size. */
int ar_sdf_volume_read(ar_ctx* ctx, uint32_t index, float* voxels, uint32_t dims[3]);
/* Drops every volume (SDFCacheManager::Clear); AR_E_INVALID while the
 * scene's objects refer to one. */
int ar_sdf_volume_clear(ar_ctx* ctx);

/* --- acceleration structure ----------------------------------------------
 * Replaces BuildBVH (core/bvh/bvh.cuh:142-422) as called from
 * UpdateSceneToDeviceIfDirty (core/scene.cu:669-675).  opts may be NULL
 * (defaults).  Synchronous. */
int ar_bvh_build(ar_ctx* ctx, const ar_bvh_opts* opts);
int ar_bvh_info_get(const ar_ctx* ctx, ar_bvh_info* info);

/* --- traversal ------------------------------------------------------------
 * Replace TraverseBVH<false> (closest hit) and TraverseBVH<true> /
 * AnyHitSegmentBVH (any hit) of core/render/intersection.cuh:148-310.
 * All pointers are device pointers; no allocation happens in these calls
 * once the context has traced once. */
int ar_trace_closest(ar_ctx* ctx, const ar_rays* rays, size_t n, const ar_hits* hits, void* stream);

/* occluded: ceil(n/32) u32 words; bit (r & 31) of word r/32 is set iff some
 * triangle is hit with tmin <= t < tmax. */
int ar_trace_any(ar_ctx* ctx, const ar_rays* rays, size_t n, uint32_t* occluded, void* stream);

/* Closest hit that also counts, per ray, the box tests and triangle tests
 * the traversal made (the ops model of the traversal roofline).  Slower
 * than ar_trace_closest; for measurement only.  box_tests/tri_tests: n u32. */
int ar_trace_closest_counted(ar_ctx* ctx, const ar_rays* rays, size_t n, const ar_hits* hits,
                             uint32_t* box_tests, uint32_t* tri_tests, void* stream);

/* ar_trace_closest without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same rays, hits and results as ar_trace_closest, bit for bit.
 * The traversal's scratch (ray-refill counter and stack spill) is a fixed
 * 32 MiB array per device, held until the process ends and shared by every
 * call on that device from any context: calls that overlap on one device
 * race and must be ordered by the caller (one stream, or events).  The grid
 * is limited to what that array holds (the spill needs (stack entries - 16)
 * x 1 KiB per 256-thread block); AR_E_CAPACITY if one block per CU does not
 * fit (an 8-wide BVH deeper than 19 levels) or n > 2^31. */
int ar_trace_closest_noalloc(ar_ctx* ctx, const ar_rays* rays, size_t n, const ar_hits* hits, void* stream);

/* Shadow segments toward a point light for every hit of a closest-hit pass
 * (the traversal benchmark's any-hit workload): for ray r with a hit, the
 * segment starts at o + t*d, points at `target` and ends at
 * dist*(1 - tmax_scale); misses get an empty segment (tmax = 0).  Output
 * rays are device SoA arrays of n entries. */
int ar_make_shadow_rays(ar_ctx* ctx, const ar_rays* rays, const ar_hits* hits, size_t n,
                        const float target[3], float tmin, float tmax_scale,
                        const ar_rays* out, void* stream);

/* --- path integrator ------------------------------------------------------
 * Replaces SceneRenderer::Render (core/scene.hpp:104-110, core/scene.cu:
 * 210-238) with its RenderCamera kernel (core/scene.cu:135-189) and the
 * CalculateRadiance integrator (core/render/render.cuh:460-998), for the
 * triangle scene given to ar_scene_set (the BVH must be built).
 *
 * Camera: the reference convention (core/scene.cu:145-164): right =
 * normalize(dir x +Y), up = normalize(dir x right), pixel (i, j) jittered
 * and mapped to ndc ((u*2-1) * tan(fov/2) * aspect, ((1-v)*2-1) * tan(fov/2)).
 * Row j = 0 therefore looks along +up = normalize(dir x right), which is
 * world -Y for a level camera; the reference's BMP writer (negative height:
 * top-down) stores it as the top row, and so does ar_write_bmp.
 *
 * Random numbers: Philox4x32-10, key = seed, counter = (pixel, sample,
 * bounce, group), so a render is a pure function of its options and
 * sample_offset makes progressive rendering reproducible. */
typedef struct {
    float pos[3];
    float dir[3];
    float vfov_deg;       /* 60 in the reference */
} ar_camera;

#define AR_SKY_NONE 0u
#define AR_SKY_GRADIENT 1u    /* SkyLighting, core/render/render.cuh:58-73 */
#define AR_SKY_CONSTANT 2u    /* env[] radiance in every direction */

typedef struct {
    uint32_t width, height;
    uint32_t spp;
    uint32_t sample_offset;       /* index of the first sample */
    uint64_t seed;
    uint32_t max_depth;           /* bounce-loop cap (MAX_TRACING_DEPTH = 1000 in the reference) */
    uint32_t rr_min_depth;        /* Russian roulette from this non-delta depth (3) */
    uint32_t nee;                 /* next-event estimation on/off */
    uint32_t sky;                 /* AR_SKY_* */
    float env[3];
    uint32_t max_paths_per_pass;  /* 0: 16,777,216 (about 4 GB of path state) */
    uint32_t spectral;            /* AR_SPECTRAL_* */
    uint32_t output;              /* AR_OUTPUT_* (spectral modes) */
    uint32_t guiding;             /* AR_GUIDING_* */
    uint32_t frame_spp;           /* guided renders: samples per pixel per frame, i.e. between guide
                                     updates (0: 1, the reference's frame) */
    uint32_t frame_growth;        /* guided renders: each frame has this many times the samples of the
                                     one before (0 or 1: every frame has frame_spp; 2: frame_spp,
                                     2 frame_spp, 4 frame_spp, ...), the last frame what remains, and no
                                     frame more paths than 2^26 allow (no counterpart in the reference,
                                     whose frames have one sample) */
} ar_render_opts;

/* Path guiding.  VOXEL: the reference's voxel guiding (core/guiding,
 * core/render/render.cuh:345-458, 553-638, 909-964; core/scene.cu:368-570):
 * the render runs frames of one sample per pixel; after each frame the paths'
 * collected radiance (up to 5 non-delta vertices each) is merged into
 * directional probes, and later frames sample the probe of each vertex with
 * probability 0.5 * min(1, frame/64)^0.5 in a one-sample mixture with the
 * BSDF (frame = sample index).  Probes persist in the context across calls
 * until ar_guiding_reset. */
#define AR_GUIDING_OFF 0u
#define AR_GUIDING_VOXEL 1u
/* MLP: learned guiding (no counterpart in the reference): an MLP maps the
 * vertex (position in the scene box, shading normal) to an 8-lobe vMF
 * mixture over directions, sampled in a one-sample mixture with the BSDF
 * (guide or BSDF chosen first, probability 0.75 * min(1, frame/64)^0.5 of
 * the guide) in the voxel probes' hemisphere fold.  After each frame it is
 * trained with BF16 MFMA forward/backward and Adam on records of the
 * frame's path vertices: the direction sampled at a non-delta vertex and the
 * radiance that arrived along it (the next vertex's emission without MIS
 * weight, its NEE and every later bounce) over the direction's pdf; see
 * ar_mlp_configure and ar_mlp_records_get. */
#define AR_GUIDING_MLP 2u

/* Spectral modes.  RGB: three colour channels throughout (the reference with
 * USE_SPECTRUM_RENDERING off).  SINGLE: one wavelength per path in
 * [360, 780] nm drawn from the reference's CIE-fitted mixture
 * (core/render/render.cuh:466-474, core/spectrum/sample_ciexyz.cuh:288-326).
 * HERO4: four hero wavelengths per path, lambda_k = 380 + (u*400 + 100 k)
 * mod 400, each path's estimate divided by the sum of their densities.
 * Reflectances and emitters are uplifted from RGB through the 3 x 64^3
 * sigmoid-coefficient table (ar_spectrum_lut_build); emitters and a
 * constant sky are scaled by D65 normalised to integral D65 * ybar = 1. */
#define AR_SPECTRAL_RGB 0u
#define AR_SPECTRAL_SINGLE 1u
#define AR_SPECTRAL_HERO4 4u

/* Output of a spectral render: CIE XYZ (the reference's spectral output), or
 * linear sRGB through XYZ2SRGBLinearD65 (core/tonemap.cuh:116-126). */
#define AR_OUTPUT_XYZ 0u
#define AR_OUTPUT_SRGB 1u

/* out_rgb: device array of width*height*3 floats (pixel j*width + i),
 * receives the mean over spp of clamp(checkNanInf(L), 0, 1e3) per channel.
 * Queued on `stream`; the call does not wait. */
int ar_render(ar_ctx* ctx, const ar_camera* cam, const ar_render_opts* opts, float* out_rgb, void* stream);

/* Measurement of the integrator's ops model (no counterpart in the
 * reference): while enabled, every render of the context (ar_render,
 * ar_render_batch, ar_render_progressive, graphs captured afterwards) counts
 * the box and triangle tests of its closest-hit and shadow traversals (the
 * counting traversal variant: slower) and the path vertices it shades.
 * ar_render_counters_read waits for the context's work and returns then
 * clears {closest box tests, closest triangle tests, shaded vertices, shadow
 * box tests, shadow triangle tests}. */
int ar_render_counters(ar_ctx* ctx, int enable);
int ar_render_counters_read(ar_ctx* ctx, uint64_t out[5]);

/* --- batched rendering ----------------------------------------------------
 * No counterpart in the reference (one SceneRenderer renders one scene).
 * ar_batch_create merges scene contexts (each with its scene set and BVH
 * built, all at the same BVH width) into a batch context whose render traces
 * the paths of every scene through one set of wavefront queues: each ray
 * carries its scene's BVH root, each path its scene's camera, image,
 * lights and starting medium.  Scenes may hold triangles and analytic
 * objects (boxes, spheres, SDF tori and volumes; each scene's voxel pool is
 * merged).  The scene contexts stay valid and independent.  Scenes are packed
 * into passes in order, whole samples at a time, so each scene's image is
 * bit-identical to rendering it alone with the same options. */
typedef struct ar_graph ar_graph;

int ar_batch_create(ar_ctx* const* scenes, uint32_t n, ar_ctx** out);
int ar_batch_size(const ar_ctx* ctx);
/* cams, opts, outs: n entries (n = the batch's scene count, or 1 for a plain
 * context); per scene: width, height, spp, sample_offset; seed, depth, NEE,
 * sky, spectral, output, guiding, frame_spp and max_paths_per_pass must
 * agree.  Guided batches (at most 256 scenes, one sample_offset) run frames
 * of frame_spp samples of every scene with one guide update per frame over
 * all of them: voxel probes are keyed by scene as well (a batch context's
 * key packs the voxel as three 16-bit offsets (+2^15), the scene index in
 * bits 48..55 and the face in bits 57..61), and the learned network takes
 * the scene index as 8 input bits (rows 5..12 of W0). */
int ar_render_batch(ar_ctx* ctx, uint32_t n, const ar_camera* cams, const ar_render_opts* opts, float* const* outs,
                    void* stream);

/* Threads per block of the wavefront kernels (camera, shade, resolve,
 * accumulate): 64, 128, 256 (default) or 512. */
int ar_set_tile(ar_ctx* ctx, uint32_t threads);

/* Captures a whole render (every pass and bounce) into a HIP graph; each
 * ar_graph_launch replays it on `stream` (same seeds: same image).  The graph
 * uses the context's render buffers: launches must not overlap other renders
 * of the context, and a render that reallocates them (a larger image)
 * invalidates the graph (ar_graph_launch then returns AR_E_INVALID), as
 * does any call that frees memory the graph reads (a new scene or BVH,
 * guiding or table reallocation) and ar_destroy of its context.  Guided
 * renders capture their frames with the guide updates and training steps:
 * each replay renders further frames with the guide's current state (the
 * network's Adam step count is kept on the device). */
int ar_graph_capture(ar_ctx* ctx, uint32_t n, const ar_camera* cams, const ar_render_opts* opts, float* const* outs,
                     ar_graph** out);
int ar_graph_launch(ar_graph* graph, void* stream);
int ar_graph_destroy(ar_graph* graph);

/* Autotuner: for BVH width 2 and 8 (every scene rebuilt, leaf size 3) and
 * tiles of 64..512 threads, times one batched render of the scenes at
 * min(spp, probe_spp) samples and picks the fastest; the scenes are left
 * built at the chosen width.  The caller then creates the batch and sets the
 * tile. */
typedef struct {
    uint32_t width;       /* chosen BVH width */
    uint32_t tile;        /* chosen threads per block */
    float ms[3][4];       /* [width 2, 4, 8][tile 64, 128, 256, 512] */
} ar_tune_result;

int ar_autotune_batch(ar_ctx* const* scenes, uint32_t n, const ar_camera* cams, const ar_render_opts* opts,
                      uint32_t probe_spp, ar_tune_result* out);

/* Replaces the offline SceneRenderer::Render(size, origin, dir, spp)
 * (core/scene.cu:215-237), which allocates its own target and returns a host
 * vector: renders into a device image the context owns and copies it to host
 * memory (waits).  hdr (width*height*3 floats, host, or NULL) receives the
 * image as ar_render defines it; ldr (width*height words, host, or NULL) the
 * ar_tonemap of it with `exposure` and `tone` (spectral XYZ output is
 * converted to sRGB first). */
int ar_render_host(ar_ctx* ctx, const ar_camera* cam, const ar_render_opts* opts, float* hdr, uint32_t* ldr,
                   float exposure, uint32_t tone);

/* --- progressive accumulation and output --------------------------------
 * ar_render_progressive replaces the interactive SceneRenderer::Render
 * (core/scene.cu:240-290 with RenderCamera's running mean and odd/even
 * halves, core/scene.cu:135-189): it renders opts->spp samples per pixel
 * with sample indices frame .. frame + spp - 1 (opts->sample_offset is not
 * used) and folds their mean into accum, the running mean of the frame
 * samples before (accum is overwritten when frame = 0, so it needs no
 * clearing).  odd / even (both or neither) hold the running means of the
 * samples with odd / even index.  All three are width*height*3 floats on the
 * device, in the render's output space (opts->output).  Spectral renders
 * accumulate CIE XYZ (AR_OUTPUT_XYZ) and convert in ar_tonemap (xyz_input),
 * as the reference does: AR_OUTPUT_SRGB clamps each frame's mean at zero
 * before the fold.  Guided frames (opts->guiding) update the guide after
 * every frame_spp samples, as SceneRenderer::Render does after each frame
 * when bUsePathGuiding_ is set (core/scene.cu:240-290). */
int ar_render_progressive(ar_ctx* ctx, const ar_camera* cam, const ar_render_opts* opts, uint32_t frame,
                          float* accum, float* odd, float* even, float* gbuffer, float* stats, void* stream);

/* --- screen-space denoiser -------------------------------------------------
 * The denoiser's per-pixel records, kept by ar_render_progressive when
 * gbuffer / stats are given (NULL: not kept; both restart at frame 0):
 *   gbuffer, AR_GBUFFER_FLOATS per pixel (RecordScreenGBuffer,
 *     core/denoise/denoise.cuh:82-163): running means over the samples that
 *     hit of the first hit's albedo[3], normal[3], distance, metallic and
 *     roughness, the shading-model frequencies over all samples (diffuse,
 *     light, glass, GGX, none), then the hit and sample counts (uint32 bits);
 *   stats, AR_STATS_FLOATS per pixel (RecordScreenStatisticBuffer,
 *     denoise_statistic.cuh:36-69): sample count (uint32 bits), Welford mean,
 *     M2 / n and M3 / n of the Box-Cox (lambda 1/2) transformed samples.
 * Means are kept in float (the reference keeps the G-buffer in half). */
#define AR_GBUFFER_FLOATS 16
#define AR_STATS_FLOATS 12

/* Bilateral16x16<7> (denoise.cuh:352-444): each output pixel averages the
 * 15 x 15 window of hdr in log(1 + c), weighted by the G-buffer similarity
 * (ComputeSVGFWeight), a spatial Gaussian of sigma 2 and a Welch t-test of
 * the two pixels' skewness-corrected Box-Cox means at the 0.5 % level
 * (ComputeTStatisticFinal); borders clamp.  hdr / out: width*height*3
 * floats (device; out must not alias hdr). */
int ar_denoise(ar_ctx* ctx, const float* hdr, const float* gbuffer, const float* stats, float* out, uint32_t width,
               uint32_t height, void* stream);

#define AR_TONE_LINEAR 0u
#define AR_TONE_GAMMA 1u      /* pow(x, 1/2.2) (core/tonemap.cuh:6-9) */
#define AR_TONE_ACES 2u       /* ACES fit (core/tonemap.cuh:66-114) */

/* Tonemap kernel (core/scene.cu:22-133): hdr (width*height*3 floats, device;
 * CIE XYZ when xyz_input, converted by XYZ2SRGBLinearD65, else linear sRGB)
 * times exposure, the tone curve, then 0xff000000 | r << 16 | g << 8 | b
 * per pixel into ldr (device). */
int ar_tonemap(ar_ctx* ctx, const float* hdr, uint32_t* ldr, uint32_t width, uint32_t height, float exposure,
               uint32_t tone, uint32_t xyz_input, void* stream);

/* The output stage of the interactive SceneRenderer::Render (core/scene.cu:
 * 291-365): ldr (width*height words, device) receives one of
 *   AR_VIEW_IMAGE      hdr through exposure and the tone curve (as ar_tonemap)
 *   AR_VIEW_DENOISED   ar_denoise of hdr (into a device image the context
 *                      owns), then the tone curve (sceneSetting_.bDenoise_)
 *   AR_VIEW_ALBEDO .. AR_VIEW_STAT_M3   the debug views of the Tonemap kernel
 *                      (bDebugAlbedo_ .. bDebugStatisticsM3_, core/scene.cu:
 *                      35-107) over the G-buffer / statistics records kept by
 *                      ar_render_progressive: albedo, normal * 0.5 + 0.5, the
 *                      viridis heat map of 1 / (depth + 1), of metallic and of
 *                      roughness (core/render/color.cuh:99-140), log(n + 1) / 100
 *                      in red, and |mean|, |M2/n|, |M3/n| / 4.
 * Each view reads only what it needs (NULL elsewhere is fine).  Queued on
 * `stream`; the call does not wait. */
#define AR_VIEW_IMAGE 0u
#define AR_VIEW_DENOISED 1u
#define AR_VIEW_ALBEDO 2u
#define AR_VIEW_NORMAL 3u
#define AR_VIEW_DEPTH 4u
#define AR_VIEW_METALLIC 5u
#define AR_VIEW_ROUGHNESS 6u
#define AR_VIEW_STAT_COUNT 7u
#define AR_VIEW_STAT_MEAN 8u
#define AR_VIEW_STAT_M2 9u
#define AR_VIEW_STAT_M3 10u
int ar_present(ar_ctx* ctx, const float* hdr, const float* gbuffer, const float* stats, uint32_t* ldr, uint32_t width,
               uint32_t height, float exposure, uint32_t tone, uint32_t xyz_input, uint32_t view, void* stream);

/* DeviceMSELuminance (core/helper/measure.cuh:49-74): mean over npix pixels
 * of (lum(a) - lum(b))^2 with Rec.709 luminance; the reference's displayed
 * variance is 1e6 times this for the odd / even halves.  Waits for the
 * result. */
int ar_luminance_mse(ar_ctx* ctx, const float* a, const float* b, size_t npix, double* out, void* stream);

/* SaveBMP (core/camera.cu:10-62): writes width*height 0xAARRGGBB pixels
 * (host memory, row 0 first) as a 32-bit top-down BMP.  Host only. */
int ar_write_bmp(const char* path, const uint32_t* bgra, uint32_t width, uint32_t height);

/* --- voxel guiding state ------------------------------------------------
 * Probes are keyed by (voxel = floor(p / 0.05), icosahedral face of the
 * reflected direction) and hold an 8x8 equal-area octahedral texture of
 * float4 (XYZ in spectral modes, RGB otherwise; w unused) with 4x4, 2x2, 1x1
 * mips: 85 texels.  A key packs the voxel as three 19-bit offsets (+2^18) and
 * the face in bits 57..61. */
typedef struct {
    uint64_t probes;          /* probes allocated */
    uint64_t capacity;        /* probe capacity */
    uint64_t samples;         /* samples merged by the last frame */
    uint64_t dropped;         /* samples dropped since the last reset because the capacity was reached */
    uint64_t device_bytes;
} ar_guiding_info;

/* Probe capacity (default 1,048,576; about 2.2 KB of device memory per
 * probe).  Takes effect on the next guided render; discards the probes. */
int ar_guiding_configure(ar_ctx* ctx, uint64_t max_probes);

/* Drops every probe; replaces SceneRenderer::MarkNeedResetGuiding /
 * ResetPathGuidingSample (core/scene.hpp:114-116, core/scene.cu:368-407). */
int ar_guiding_reset(ar_ctx* ctx);

int ar_guiding_info_get(ar_ctx* ctx, ar_guiding_info* out);

/* Copies the probes to host memory in allocation order: keys (n), texels
 * (n x 85 x 4 floats), sample counts (n).  cap: room in the arrays (probes);
 * *n receives the probe count (AR_E_CAPACITY if it exceeds cap).  Any array
 * may be NULL.  Synchronises with the context's work. */
int ar_guiding_snapshot(ar_ctx* ctx, uint64_t* keys, float* texels, uint64_t* counts, uint64_t cap, uint64_t* n);

/* Replaces the probes with n probes in the snapshot layout (resume). */
int ar_guiding_restore(ar_ctx* ctx, const uint64_t* keys, const float* texels, const uint64_t* counts, uint64_t n);

/* --- learned guiding network ----------------------------------------------
 * No counterpart in the reference (its guiding is the voxel probes).  The
 * network: 5 inputs (position normalised to the scene box, the shading
 * normal as theta/pi and phi/(2 pi) + 1/2), three hidden layers of 64 ReLU
 * units, 40 outputs = 8 lobes x (logit, log kappa clamped to [-3, 7], mean
 * direction).  Parameters (ar_mlp_params_*): FP32, W0 16x64, b0 64, W1 64x64,
 * b1 64, W2 64x64, b2 64, W3 64x48, b3 48 (W[in][out]; rows 5..12 of W0
 * take the scene index bits of a guided batch and stay zero otherwise, rows
 * 13..15 of W0 and columns 40..47 of W3 are unused and zero): 12,720 floats.  Loss of a sample
 * (x, w, weight): -weight * log q(w | x). */
typedef struct {
    float learning_rate;        /* Adam step (default 5e-3) */
    uint32_t batch;             /* records per training step (default 2^18); a frame's records
                                   come from one path per record slot, the slots shared equally by
                                   the frame's scenes */
    uint32_t iters_per_frame;   /* training steps after each frame (default 1) */
    uint64_t seed;              /* He-normal initialisation seed */
} ar_mlp_opts;

/* Sets the options and re-initialises the parameters. */
int ar_mlp_configure(ar_ctx* ctx, const ar_mlp_opts* opts);
int ar_mlp_params_get(ar_ctx* ctx, float* host_out, size_t count);
int ar_mlp_params_set(ar_ctx* ctx, const float* host_in, size_t count);
/* Raw network outputs (n x 40, device) of inputs (n x 5, device). */
int ar_mlp_forward(ar_ctx* ctx, const float* x5, size_t n, float* out40, void* stream);
/* Gradient (12,720 floats, device) and loss (1 float, device, may be NULL) of
 * the mean over the n samples of -weight * log q(dir | x); dir: n x 3. */
int ar_mlp_gradient(ar_ctx* ctx, const float* x5, const float* dir, const float* weight, size_t n, float* grad,
                    float* loss, void* stream);
/* `steps` Adam steps on the weighted loss normalised by the sum of weights;
 * loss (device, may be NULL) receives the last step's loss. */
int ar_mlp_train(ar_ctx* ctx, const float* x5, const float* dir, const float* weight, size_t n, uint32_t steps,
                 float* loss, void* stream);
/* The training records of the last learned-guided frame, copied to host
 * after the device is idle (no counterpart in the reference; for inspecting
 * what the guide is trained on): *count = the records trained on (at most
 * cap are copied), *total = the frame's records before the batch was drawn;
 * x: count x 6 (the network inputs and the scene index), dir: count x 3,
 * weight: count (luminance of the radiance that arrived along dir, over the
 * pdf dir was sampled with).  Any output may be NULL. */
int ar_mlp_records_get(ar_ctx* ctx, float* x, float* dir, float* weight, size_t cap, uint32_t* count,
                       uint32_t* total);

/* --- spectral uplift table --------------------------------------------------
 * Replaces PrecomputeRGBLUTKernel / SetupSpectrumLUT (core/spectrum/
 * spectrum_lut.cuh:12-78, core/scene.cu:591-607): instead of evaluating the
 * reference's RGB->basis network on a 512^3 grid (2 GiB), builds the
 * 3 x 64^3 x 3 sigmoid-coefficient table once per process on the host.  Needs
 * no device.  out: 3*64*64*64*3 floats, layout [largest channel][z][y][x][c].
 * Returns AR_E_INVALID if count is smaller. */
int ar_spectrum_lut_build(float* out, size_t count);

/* Sigmoid coefficients for n RGB colours (rows of 3, channels clamped to
 * [0, 1]) from the table (builds it on first use).  No device needed. */
int ar_spectrum_uplift(const float* rgb, size_t n, float* coeffs);

/* --- spectral basis network -------------------------------------------------
 * The reference's RGB -> spectrum uplift network (core/spectrum/
 * spectrum_basis.cuh:2273-2470): RGBEncode (sin / cos features of r, g, b
 * over their maximum s + 0.001 and of s, Linear 32 -> 32 + ReLU, 7 residual
 * blocks h += ReLU(W h + b), Linear 32 -> 4, times s) and LambdaEncode
 * (sin / cos features of (lambda - 360) / 420, 7 residual blocks, Linear
 * 32 -> 4); reflectance clamp(dot, 0, 1).  Parameters: AR_SMLP_PARAMS
 * floats in the header's order (RGB_IN_W, RGB_IN_B, RGB_BLKk_W, RGB_BLKk_B
 * for k = 0..6, RGB_OUT_W, LAM_BLKk_W, LAM_BLKk_B, LAM_OUT_W; matrices
 * row-major [out][in]).  Evaluated and trained with FP32 MFMA. */
#define AR_SMLP_PARAMS 16096

/* Spectral uplift of spectral renders: AR_BASIS_SIGMOID (default) uses the
 * Jakob-Hanika sigmoid coefficients; AR_BASIS_NETWORK the reference's path
 * (core/render/material.cuh:69-358, render.cuh:290-738): reflectance and
 * emission are Interact(QueryLambda(lambda), RGB) through the tables of
 * ar_spectral_mlp_build_tables, which must exist.  A new table build
 * refreshes the material bases at the next render. */
#define AR_BASIS_SIGMOID 0u
#define AR_BASIS_NETWORK 1u
int ar_spectral_set_basis(ar_ctx* ctx, uint32_t basis);

/* Sets the parameters (host array) and restarts the optimiser state; the
 * EMA copy starts at the parameters. */
int ar_spectral_mlp_set(ar_ctx* ctx, const float* params);
/* Reads the parameters (ema = 0) or their EMA (ema = 1) into a host array;
 * waits for the context's work. */
int ar_spectral_mlp_get(ar_ctx* ctx, float* params, int ema);
/* SpectralInteract(RGBEncode(rgb), LambdaEncode(lambda)) of n rows
 * x = (r, g, b, lambda nm) (device, n x 4) into out (device, n). */
int ar_spectral_mlp_eval(ar_ctx* ctx, const float* x, float* out, size_t n, void* stream);
/* PrecomputeSpectrumLUTsToDevice (spectrum_lut.cuh:12-78): RGB table of
 * rgb_res^3 entries at voxel centres (the reference uses 512) and wavelength
 * table of lambda_res entries (840), kept in the context. */
int ar_spectral_mlp_build_tables(ar_ctx* ctx, uint32_t rgb_res, uint32_t lambda_res, void* stream);
/* Copies the tables to host arrays (rgb_res^3 x 4, lambda_res x 4 floats;
 * either may be NULL); waits for the device. */
int ar_spectral_mlp_read_tables(ar_ctx* ctx, float* rgb_table, float* lambda_table);
/* Interact (spectrum_lut.cuh:223-235) of n rows x = (r, g, b, lambda nm)
 * through the tables: trilinear RGB and linear wavelength queries; an RGB
 * above 1 is divided by its largest channel and the result scaled back. */
int ar_spectral_mlp_query(ar_ctx* ctx, const float* x, float* out, size_t n, void* stream);

/* Training (pytorch/spectrum/train_spectrum.py:636-752).  Step k of a call
 * uses dataset rows ((first_step + k) batch + j) mod n, j < batch, and the
 * reference's training-mode noise when noise != 0 (standard deviations
 * 0.025 / 256 on r, g, b and 0.025 / 420 on the normalised wavelength of the
 * first half of the batch, 0.001 on every block's Linear input; Box-Muller
 * pairs drawn from a counter hash of seed, step, row in batch and slot
 * pair).  Loss: mean over the batch and three channels of
 * ((pred - y)(0.25 + max(cie_c(lambda), 0)))^2
 * with the analytic CIE fits.  Then gradient clipping to clip_norm (0: none),
 * AdamW (decoupled weight decay on weight matrices only) at
 * lr * warmup-cosine(first_step + k of total_steps; 0 = first_step + steps)
 * and the EMA update. */
typedef struct {
    uint32_t steps;          /* optimiser steps of this call */
    uint32_t batch;          /* multiple of 32, at most 262144 (reference 4096) */
    uint32_t first_step;     /* global step of the first one (schedule, rows, noise) */
    uint32_t total_steps;    /* schedule length; 0 = first_step + steps */
    uint32_t seed;
    uint32_t noise;          /* training-mode noise on / off */
    float lr;                /* 5e-4 in the reference's config.ini */
    float weight_decay;      /* 1e-3 */
    float beta1, beta2;      /* 0.9, 0.95 */
    float eps;               /* 1e-8 */
    float clip_norm;         /* 1.0 */
    float warmup_ratio;      /* 0.001 */
    float min_lr_ratio;      /* the reference passes lr * 0.01 = 5e-6 */
    float ema_decay;         /* 0.999 */
} ar_smlp_train_opts;

/* Runs opts->steps optimiser steps on the dataset x (device, n x 4: r, g, b,
 * lambda nm), y (device, n); loss_history (device, steps floats, or NULL)
 * receives each step's loss. */
int ar_spectral_mlp_train(ar_ctx* ctx, const float* x, const float* y, size_t n, const ar_smlp_train_opts* opts,
                          float* loss_history, void* stream);
/* The loss gradient of one batch (rows as above for global step `step`)
 * into grad (device, AR_SMLP_PARAMS floats) and the loss into *loss (device,
 * or NULL); parameters unchanged. */
int ar_spectral_mlp_gradient(ar_ctx* ctx, const float* x, const float* y, size_t n, uint32_t batch, uint32_t step,
                             uint32_t seed, int noise, float* grad, float* loss, void* stream);
/* A training set for the uplift this library ships (no measured spectra
 * needed): n rows drawn like the reference's synthetic branch
 * (pytorch/spectrum/spectrum_dataset.py:331-361) with the Jakob-Hanika
 * sigmoid reflectance of each RGB at its wavelength as the target, written
 * to x (device, n x 4) and y (device, n); waits for the copies. */
int ar_spectral_uplift_dataset(ar_ctx* ctx, size_t n, uint32_t seed, float* x, float* y, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* AR_H */
