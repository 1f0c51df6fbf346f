/*
 * CPU oracle for the Aperture path-tracing core.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and there only as the checker / CPU baseline.
 *
 * It restates, in plain C99, the algorithms of the reference renderer that
 * sit on the hot path (citations are relative to the reference checkout):
 *
 *   orc_bvh_build        core/bvh/bvh.cuh:142-422   binned-SAH BVH2 builder
 *   box_hit (static)     core/bvh/bvh.cuh:424-470   FastAABBHit(WithEnter)
 *   orc_trace_*          core/render/intersection.cuh:148-310 TraverseBVH
 *   orc_tri_test         (new: the reference has no triangles; the
 *                        watertight test of Woop, Benthin & Wald,
 *                        JCGT 2(1) 2013, is restated from the paper)
 *
 * Parity pins (what checks each function, see tests/):
 *   orc_tri_test         closed-form FP64 ray/plane hits, known-answer rays,
 *                        watertightness on closed meshes (tests/test_oracle_trace.py)
 *   orc_bvh_build        structural invariants (each primitive in exactly one
 *                        leaf, bounds containment, pre-order left=parent+1,
 *                        leaf size, SAH accept rule)
 *   orc_trace_closest    equality with orc_trace_brute on every ray
 *   orc_trace_any        equality with (closest hit t < tmax)
 *   orc_object_*         FP64 quadratic / slab solutions, sampled surfaces,
 *                        proxy extreme points (tests/test_oracle_objects.py)
 *   SDF objects          the torus's closed-form ray intersection (quartic
 *                        roots in FP64) and normal, the march against the
 *                        analytic torus through a baked volume, the bake's
 *                        grid values against the closed form
 *                        (tests/test_oracle_sdf.py)
 *   object lights, glass white / diffuse furnaces, n^2 radiance inside glass,
 *                        Beer-Lambert series, sampler pairing and solid-angle
 *                        integrals (tests/test_oracle_render_objects.py)
 *
 *   integrator           Lambert furnace, polygon-irradiance direct light,
 *                        NEE on / off, FP32 = FP64 paths
 *                        (tests/test_oracle_render.py)
 *   GGX BSDF, samplers   D's projected-area integral and tangent form, Smith
 *                        G's Lambda form, a numpy microfacet model,
 *                        reciprocity, Schlick limits; bounded-VNDF, cosine and
 *                        mixture samplers: pdf pairing, chi-square against
 *                        the pdf, surviving mass, unbiased albedo estimate
 *   light pick, RR       area * lum(Le) select probabilities and pick
 *                        chi-square, NEE on / off with unequal lights; RR
 *                        survivors at exactly 1/q of the RR-free path value,
 *                        survival fraction q, deep mean unchanged
 *                        (tests/test_oracle_bsdf.py)
 *
 *   output stage         linear / gamma curves and byte packing against
 *                        numpy, the XYZ -> sRGB matrix on D65 white, ACES
 *                        properties, the MSE and running mean against numpy
 *                        (tests/test_oracle_output.py)
 *   denoiser             numpy moments of the Box-Cox samples, scipy's Welch
 *                        t-test, the reference's t-quantile table
 *                        (tests/test_oracle_denoise.py)
 *   spectral network     the reference's golden vectors under TF32,
 *                        torch.autograd, torch's clipping and AdamW
 *                        (tests/test_oracle_spectral_mlp.py)
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp, no fast-math).
 */
#ifndef AR_ORACLE_H
#define AR_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MISS 0xFFFFFFFFu

/* 32-bit reference node layout (core/bvh/bvh.cuh:36-43): leaf iff count > 0,
 * pre-order, left child = parent + 1. */
typedef struct {
    float bmin[3];
    float bmax[3];
    int32_t left, right, first, count;
} orc_node;

/* Binned-SAH build over primitive proxies (bvh.cuh:142-422).  Arrays are
 * n x 3 floats.  nodes_out must hold 2n entries, indices_out n entries.
 * Returns the node count (0 for an empty input). */
int orc_bvh_build(const float* pmin, const float* pmax, const float* pcen,
                  int n, int leaf_size, int bin_count,
                  orc_node* nodes_out, int32_t* indices_out);

/* Proxies of a triangle mesh: box of the three vertices and its centre. */
void orc_tri_proxies(const float* verts, const uint32_t* idx, int ntri,
                     float* pmin, float* pmax, float* pcen);

/* Watertight ray/triangle test.  Returns 1 when the ray's line crosses the
 * triangle (any t; callers apply their own [tmin, tmax) rule) and writes
 * t, the barycentric weights of v1 and v2, and flags: bit0/1/2 = weight of
 * v0/v1/v2 exactly zero (edge or vertex hit), bit3 = det < 0. */
int orc_tri_test(const float org[3], const float dir[3],
                 const float v0[3], const float v1[3], const float v2[3],
                 float* t, float* u, float* v, uint32_t* flags);

/* Ray batch in structure-of-arrays form (same as the library's ar_rays). */
typedef struct {
    const float *ox, *oy, *oz, *tmin, *dx, *dy, *dz, *tmax;
    const float* meps;   /* optional SDF march tolerance per ray (NULL: max(tmin, 1e-3) / 2) */
} orc_rays;

typedef struct {
    uint32_t* prim; float* t; float* u; float* v; uint32_t* flags;
} orc_hits;

/* mode 0 = canonical: padded (conservative) box test and the rule
 *          "smallest t, then smallest primitive id" (the brute-force answer);
 * mode 1 = reference-faithful: FastAABBHit as written and the first hit
 *          found in traversal order wins (strict <).
 * boxes/tris: optional per-ray counters of box and triangle tests. */
void orc_trace_closest(const orc_node* nodes, int node_count, const int32_t* indices,
                       const float* verts, const uint32_t* idx,
                       orc_rays rays, size_t n, orc_hits hits, int mode,
                       uint32_t* boxes, uint32_t* tris, int threads);

/* occluded bit r set iff some triangle has tmin <= t < tmax. */
void orc_trace_any(const orc_node* nodes, int node_count, const int32_t* indices,
                   const float* verts, const uint32_t* idx,
                   orc_rays rays, size_t n, uint32_t* occluded_bits, int mode,
                   uint32_t* boxes, uint32_t* tris, int threads);

/* O(N*M) closest hit over all triangles, canonical rule. */
void orc_trace_brute(const float* verts, const uint32_t* idx, int ntri,
                     orc_rays rays, size_t n, orc_hits hits, int threads);

void orc_trace_any_brute(const float* verts, const uint32_t* idx, int ntri,
                         orc_rays rays, size_t n, uint32_t* occluded_bits, int threads);

/* ---------------- analytic objects (trace.c) ---------------- */

#define ORC_OBJ_BOX 0u        /* OBJ_CUBE: oriented box of half-extents `extent` */
#define ORC_OBJ_SPHERE 1u     /* OBJ_SPHERE: radius = largest extent */
#define ORC_OBJ_SDF_TORUS 2u  /* OBJ_SDF + SDF_TORUS: radii sdf[0], sdf[1] in unit space */
#define ORC_OBJ_SDF_VOLUME 3u /* OBJ_SDF + SDF_VOLUME: distance grid vol (unit-space values) */
#define ORC_HIT_INSIDE 16u    /* hit flag: the ray started inside the object */

typedef struct {
    uint32_t type;
    float center[3];
    float extent[3];     /* absolute values */
    float w2o[9];        /* world-to-object rotation, row-major */
    float sdf[4];        /* torus radii */
    const float* vol;    /* volume voxels, x fastest (dims[0] * dims[1] * dims[2]) */
    uint32_t dims[3];
    float vmax;          /* voxelSizeMax_: |nearest| above it skips the trilinear lookup */
} orc_object;

/* SceneObject(center, extent, material, rotation, type) (core/render/object.cuh:286-305) */
void orc_object_init(orc_object* ob, uint32_t type, const float center[3], const float extent[3],
                     const float rot_deg[3]);
void orc_object_proxy(const orc_object* ob, float pmin[3], float pmax[3], float pcen[3]);
float orc_object_area(const orc_object* ob);
/* first crossing with t > 0 along normalize(d); 1 on a hit */
int orc_object_test(const orc_object* ob, const float o[3], const float d[3], float* t, int* inside);
void orc_object_surface(const orc_object* ob, const float o[3], const float d[3], float t, int inside,
                        float pos[3], float n[3]);
int orc_object_overlap(const orc_object* ob, const float p[3]);
/* SDF objects: params (torus radii) and the volume of an SDF_VOLUME object;
 * vmax = |(1/dx, 1/dy, 1/dz)| * 2.5 (SDFVolumeInfo::ComputeVoxelSize) */
void orc_object_set_sdf(orc_object* ob, const float params[4], const float* vol, const uint32_t dims[3]);
/* as orc_object_test / _surface with the SDF march tolerance meps (the
 * reference's SDF_EPS(depth) = INTERSECT_EPS(depth) / 2; unused by boxes and
 * spheres); the plain forms use meps = 5e-4 */
int orc_object_test_eps(const orc_object* ob, const float o[3], const float d[3], float meps, float* t, int* inside);
void orc_object_surface_eps(const orc_object* ob, const float o[3], const float d[3], float t, int inside,
                            float meps, float pos[3], float n[3]);
/* unit-space distance of an SDF object's shape at u (|u_k| <= 0.5 inside its box) */
float orc_sdf_value(const orc_object* ob, const float u[3]);
/* FillSDFVolume (sdf_cache_manager.cuh:16-37): out[x + dims0 (y + dims1 z)] =
 * shape(p01 - 0.5), p01 = (x, y, z) / (dims - 1); shape 0 = torus (params R, r),
 * 1 = diamond (SdfDiamondUnit, sdf_diamond.cuh:21-66; params d, bias, tableZ,
 * scale, zeros taking the defaults 0.94, 0.5, 0.3, 1) */
void orc_sdf_bake(uint32_t shape, const float params[4], const uint32_t dims[3], float* out);

/* Scene geometry: primitives 0..ntri-1 are triangles, ntri..ntri+nobj-1 objects. */
typedef struct {
    const float* verts;
    const uint32_t* idx;
    uint32_t ntri;
    const orc_object* objs;
    uint32_t nobj;
} orc_geom;

/* proxies of every primitive of g (arrays of (ntri + nobj) x 3 floats) */
void orc_geom_proxies(const orc_geom* g, float* pmin, float* pmax, float* pcen);
void orc_trace_closest_g(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
                         orc_rays rays, size_t n, orc_hits hits, int mode,
                         uint32_t* boxes, uint32_t* tris, int threads);
void orc_trace_any_g(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
                     orc_rays rays, size_t n, uint32_t* occluded_bits, int mode,
                     uint32_t* boxes, uint32_t* tris, int threads);
void orc_trace_brute_g(const orc_geom* g, orc_rays rays, size_t n, orc_hits hits, int threads);
void orc_trace_any_brute_g(const orc_geom* g, orc_rays rays, size_t n, uint32_t* occluded_bits, int threads);
int orc_closest1_g(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
                   const float o[3], const float d[3], float tmin, float tmax,
                   uint32_t* prim, float* t, float* u, float* v, uint32_t* flags);
int orc_any1_g(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
               const float o[3], const float d[3], float tmin, float tmax);
/* the same with an explicit SDF march tolerance */
int orc_closest1_ge(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
                    const float o[3], const float d[3], float tmin, float tmax, float meps,
                    uint32_t* prim, float* t, float* u, float* v, uint32_t* flags);
int orc_any1_ge(const orc_node* nodes, int node_count, const int32_t* indices, const orc_geom* g,
                const float o[3], const float d[3], float tmin, float tmax, float meps);

/* Maximum depth (root = 0) of a built tree. */
int orc_bvh_depth(const orc_node* nodes, int node_count);

/* Single-ray closest / any hit (canonical rule), for the oracle integrator. */
int orc_closest1(const orc_node* nodes, int node_count, const int32_t* indices, const float* verts,
                 const uint32_t* idx, const float o[3], const float d[3], float tmin, float tmax,
                 uint32_t* prim, float* t, float* u, float* v);
int orc_any1(const orc_node* nodes, int node_count, const int32_t* indices, const float* verts,
             const uint32_t* idx, const float o[3], const float d[3], float tmin, float tmax);

/* ---------------- random numbers (render.c) ---------------- */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 4 x u32 from a 128-bit
 * counter and a 64-bit key. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* 4*ngroups uniforms in [0,1) for counter (pix, smp, bounce, group g),
 * key = seed: u = (x >> 8) * 2^-24. */
void orc_philox_uniforms(uint64_t seed, uint32_t pix, uint32_t smp, uint32_t bounce, int ngroups, float* out);

#define ORC_CAMERA_BOUNCE 0xFFFFFFFFu

/* ---------------- integrator (render.c) ---------------- */

/* Same layout as ar_material (include/ar.h). */
typedef struct {
    float albedo[3];
    float emission[3];
    float specular, metallic, roughness;
    uint32_t model;       /* 0 Lambert, 1 GGX + Lambert, 2 glass (objects only) */
    float ior, sigma_a, abbe, tilt;   /* glass: n_d, absorption scale, Abbe number, spectral tilt */
} orc_material;

typedef struct {
    const float* verts;           /* nv x 3 */
    const uint32_t* idx;          /* ntri x 3 */
    uint32_t ntri;
    const uint32_t* tri_mat;      /* ntri, or NULL: material 0 */
    const orc_material* mats;
    /* filled by orc_render */
    const orc_node* nodes;
    int node_count;
    const int32_t* indices;
    const int32_t* light_index;   /* ntri + nobj: index into the light table or -1 */
    const orc_object* objs;       /* analytic objects, prims ntri .. ntri + nobj - 1 */
    uint32_t nobj;
    const orc_material* obj_mats; /* material of each object */
    uint32_t nmats;               /* mesh materials; object k's spectral coefficients are row nmats + k */
    const float* mat_coef;        /* spectral modes, 12 per material: albedo coefficients / basis [4],
                                     emission [4] (normalised by its largest channel above 1), escale, ascale */
    float env_coef[8];            /* spectral constant sky: coefficients / basis [4], scale */
    const float* lam_tab;         /* network basis: wavelength table, lam_res x 4 (NULL: sigmoid uplift) */
    int lam_res;
} orc_render_scene;

typedef struct {
    int count;
    const int32_t* tri;    /* light primitive ids (triangles, then objects at ntri + k) */
    const float* area;
    const float* prob;     /* select probability = w / sum(w), w = max(area * lum(Le), 1e-20) */
    const float* cdf;      /* running FP32 sum of prob */
} orc_render_lights;

/* ---------------- voxel path guiding (guide.c) ---------------- */

#define ORC_NO_KEY 0xFFFFFFFFFFFFFFFFull

/* Clarberg equal-area octahedral map between [0,1]^2 and the unit sphere. */
void orc_uv_to_dir(float u, float v, float d[3]);
void orc_dir_to_uv(const float d[3], float uv[2]);
/* icosahedral face (argmax of the 20 face normals) of a direction */
int orc_face20(const float n[3]);
const float* orc_face20_normals(void);
/* probe key of (position, direction): 0.05 voxel and face, packed 19/19/19/5
 * bits; ORC_NO_KEY outside +-2^18 voxels */
uint64_t orc_voxel_key(const float p[3], const float dir[3]);
/* leaf texel (y*8 + x) of a direction */
int orc_leaf_pixel(const float dir[3]);

/* A probe texture is 86 float4 texels: 64 leaves (8x8), 16, 4, 1 mips, pad. */
void orc_probe_mips(float* tex);
float orc_probe_pdf(const float* tex, const float dir[3], const float q[3]);
void orc_probe_sample(const float* tex, const float q[3], const float u[5], float dir[3], float* pdf);
float orc_probe_hemi_pdf(const float* tex, const float dir[3], const float n[3], const float q[3]);
void orc_probe_hemi_sample(const float* tex, const float n[3], const float q[3], const float u[5], float dir[3],
                           float* pdf);

typedef struct orc_guide orc_guide;
orc_guide* orc_guide_create(int cap);
void orc_guide_destroy(orc_guide* g);
int orc_guide_count(const orc_guide* g);
int orc_guide_find(const orc_guide* g, uint64_t key);
const float* orc_guide_texels(const orc_guide* g, int i);
/* keys (count), texels (count x 85 x 4) and sample counts (count) in probe order */
void orc_guide_export(const orc_guide* g, uint64_t* keys, float* tex, uint64_t* counts);
/* one frame: ns samples (key, direction, radiance) accumulated in order */
void orc_guide_update(orc_guide* g, int ns, const uint64_t* keys, const float* dirs, const float* rad);

/* guiding mixture probability of frame f: 0.5 * (min(1, f/64))^0.5 */
float orc_guide_alpha(uint32_t frame);

/* ---------------- learned guiding distribution (mlp.c) ---------------- */

/* parameters: W0 16x64, b0 64, W1 64x64, b1 64, W2 64x64, b2 64, W3 64x48, b3 48 (W[in][out]) */
int orc_mlp_param_count(void);
void orc_mlp_forward(const double* params, const float* x5, double* out40);
double orc_vmf_mixture_pdf(const double* out40, const double* w);
void orc_vmf_mixture_sample(const double* out40, const double* u3, double* w);
/* mean over n samples of -weight * log q(dir | x); grad (param count) or NULL */
double orc_mlp_loss_grad(const double* params, int n, const float* x5, const float* dir, const float* weight,
                         double* grad);
/* the same with BF16 rounding where the GPU stores BF16 (inputs, weights,
 * post-ReLU activations, loss gradient, masked backward activations) */
void orc_mlp_forward_bf16(const double* params, const float* x5, double* out40);
double orc_mlp_loss_grad_bf16(const double* params, int n, const float* x5, const float* dir, const float* weight,
                              double* grad);

typedef struct {
    uint32_t width, height, spp, sample_offset;
    uint64_t seed;
    int max_depth;
    int rr_min_depth;
    int nee;
    int sky;               /* 0 none, 1 reference sky gradient, 2 constant env */
    float env[3];
    float cam_pos[3], cam_dir[3];
    float vfov_deg;
    int spectral;          /* 0 RGB, 1 one wavelength (reference mixture), 4 hero wavelengths */
    int output;            /* spectral modes: 0 XYZ, 1 linear sRGB */
    const float* cie;      /* 421 x 3 CIE 1931 2-degree table, 360..780 nm at 1 nm */
    const float* d65;      /* 421 D65 values */
    const float* lut;      /* 3 x 64^3 x 3 sigmoid-coefficient table (spectral modes) */
    int guiding;           /* 1: voxel guiding, one sample per pixel per frame, probes updated per frame */
    orc_guide* guide;      /* probe store used and updated when guiding (kept across calls) */
    const orc_object* objects;     /* analytic objects (may be NULL) */
    uint32_t nobj;
    const orc_material* obj_mats;  /* nobj materials */
    float* gbuffer;                /* optional: 16 floats per pixel, folded per sample (orc_gbuffer_record) */
    float* stats;                  /* optional: 12 floats per pixel, folded per sample (orc_stats_record) */
    const float* smlp;             /* spectral modes: basis-network parameters (NULL: sigmoid uplift);
                                      reflectance clamp(dot(QueryRGBBasis, QueryLambda), 0, 1) through
                                      tables of smlp_rgb_res^3 and smlp_lambda_res entries */
    int smlp_rgb_res, smlp_lambda_res;
    /* optional crop (crop_w > 0, unguided renders): the output holds pixels
     * [crop_x0, crop_x0 + crop_w) x [crop_y0, crop_y0 + crop_h) of the
     * width x height image, each traced exactly as in the full render (same
     * camera ray and random numbers) */
    uint32_t crop_x0, crop_y0, crop_w, crop_h;
} orc_render_opts;

/* Sigmoid coefficients of an RGB colour (channels clamped to [0, 1]) from the
 * 3 x 64^3 x 3 table: largest channel l, its value z on the smoothstep^2
 * scale, the other two channels divided by z; trilinear in the coefficients. */
void orc_uplift(const float* lut, const float rgb[3], float c[3]);

/* FP32 wavelength sampling pieces (sample_ciexyz.cuh): normal quantile,
 * mixture density and sample, and the sigmoid spectrum at a wavelength. */
float orc_ndtri_f(float u);
float orc_lambda_pdf_f(float l);
float orc_sample_lambda_f(float u0, float u1, float u2);
float orc_sigmoid_f(const float c[3], float l);

/* Box / sphere light sampler (a point y, its normal and area density without
 * the select probability; 0 when nothing faces x) and the solid-angle
 * density NEE assigns to a light point, select probability included
 * (NEESampleBoxAreaFacing / NEESampleSphereCapArea and
 * NEEPDFOmegaFromHit_Box / _Sphere); FP64 instances. */
int orc_obj_light_sample_d(const orc_object* ob, const double x[3], double ub, double u0, double u1, double y[3],
                           double ny[3], double* pa);
double orc_obj_light_pdf_d(const orc_object* ob, double select, const double x[3], const double y[3],
                           const double ny[3]);

/* Light table of orc_render (core/scene.cu:677-701): every emissive
 * primitive (triangles, then objects) with select probability area *
 * lum(Le) / sum, its running FP32 CDF; light_index: per primitive, -1 or its
 * light.  Returns the light count. */
int orc_light_table(const float* verts, const uint32_t* idx, uint32_t ntri, const uint32_t* tri_mat,
                    const orc_material* mats, const orc_object* objects, uint32_t nobj, const orc_material* obj_mats,
                    int32_t* light_index, int32_t* ltri, float* larea, float* lprob, float* lcdf);

/* FP64 instances of the BSDF pieces (brdf.cuh, sample.cuh, material.cuh; a
 * material in RGB mode), the light pick of NEESampleLight and the RR
 * survival probability, for the tests.  rd points toward the surface. */
void orc_ggx_d_d(double alpha, size_t count, const double* noh, double* out);
void orc_g_correlated_d(double alpha, size_t count, const double* nol, const double* nov, double* out);
void orc_sample_vndf_d(const double v[3], const double n[3], double alpha, size_t count, const double* r,
                       double* l, double* pdf);
void orc_vndf_pdf_d(const double v[3], const double n[3], double alpha, size_t count, const double* l, double* out);
void orc_sample_cosine_d(const double n[3], size_t count, const double* e, double* l, double* pdf);
double orc_pspec_d(const orc_material* m);
void orc_eval_bsdf_d(const orc_material* m, const double rd[3], const double n[3], size_t count, const double* l,
                     double* out);
void orc_bsdf_pdf_d(const orc_material* m, const double rd[3], const double n[3], size_t count, const double* l,
                    double* out);
void orc_sample_bsdf_d(const orc_material* m, const double rd[3], const double n[3], size_t count, const double* u,
                       double* l, double* pdf, int32_t* ok);
void orc_pick_light_d(const float* cdf, int nlights, size_t count, const double* u, int32_t* out);
double orc_rr_survival_d(const double mask[3]);

/* ---------------- output stage (output.c) ---------------- */

/* ACES display curve (core/tonemap.cuh:66-114), linear sRGB in */
void orc_aces(const float in[3], float out[3]);
/* Tonemap (core/scene.cu:22-133): hdr npix x 3 (XYZ when xyz_input, else
 * linear sRGB) -> 0xff000000 | r << 16 | g << 8 | b; tone 0 linear, 1 gamma
 * 2.2, 2 ACES */
void orc_tonemap(const float* hdr, uint32_t* ldr, uint32_t npix, float exposure, int tone, int xyz_input);
/* mean over pixels of (lum(a) - lum(b))^2, luminance in float, difference in double */
double orc_luminance_mse(const float* a, const float* b, size_t npix);
/* running mean: acc (mean of n_before samples) takes in the mean of n_new more */
void orc_running_mean(float* acc, const float* mean_new, size_t nfloat, uint32_t n_before, uint32_t n_new);
/* HeatMapViridis (core/render/color.cuh:99-140) */
void orc_viridis(float t, float out[3]);
/* debug views 2..10 of ar_present (AR_VIEW_ALBEDO .. AR_VIEW_STAT_M3) over
 * G-buffer (16 floats) / statistics (12 floats) records, packed as
 * orc_tonemap packs */
void orc_view(const float* gbuf, const float* stats, uint32_t* ldr, uint32_t npix, int view);

/* ---------------- denoiser (denoise.c) ---------------- */

/* Records per pixel (uint32 counts stored as their bit patterns):
 *   G-buffer (16 floats): albedo[3] normal[3] depth metallic roughness
 *     model[5] (diffuse, light, glass, GGX, none) valid-frame count, frame count
 *   statistics (12 floats): n mean[3] M2/n[3] M3/n[3], 2 unused */
void orc_gbuffer_record(float* g, int hit, const float albedo[3], const float normal[3], float depth, float metallic,
                        float roughness, uint32_t model);
void orc_stats_record(float* st, const float x[3]);
float orc_t_quantile(int nu);
float orc_svgf_weight(const float* ga, const float* gb);
float orc_tstat_gate(const float* sa, const float* sb);
void orc_bilateral(const float* hdr, const float* gbuf, const float* stats, float* out, uint32_t width,
                   uint32_t height);

/* Renders the mean of spp samples per pixel into out (width*height*3,
 * row-major, pixel index = j*width + i).  precision 32: float out, FP32
 * arithmetic with the same triangle test and hit rule as the GPU; 64:
 * double out, FP64 arithmetic throughout.  Returns 0 on success. */
int orc_render(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t ntri, const uint32_t* tri_mat,
               const orc_material* mats, uint32_t nmats, const orc_render_opts* opts, int precision,
               void* out, int threads);

/* ---------------- spectral basis network (spectral_mlp.c) ---------------- */

/* 16096 parameters in the reference header's layer order; 1 when entry p is
 * a weight-matrix entry (weight decay applies), 0 for a bias */
int orc_smlp_param_count(void);
int orc_smlp_decays(int p);
/* FP32 inference in the reference's operation order */
void orc_smlp_rgb_encode(const float* w, const float rgb[3], float out[4]);
void orc_smlp_lambda_encode(const float* w, float lam, float out[4]);
float orc_smlp_interact(const float a[4], const float b[4]);
void orc_smlp_eval(const float* w, size_t n, const float* x, float* out);
/* PyTorch evaluation with TF32 matmuls, as the reference made its golden vectors */
void orc_smlp_eval_torch_tf32(const float* w, size_t n, const float* x, float* out);
/* table queries at a resolution (entries evaluated directly) */
void orc_smlp_query_rgb(const float* w, int res, const float rgb[3], float out[4]);
void orc_smlp_lambda_entry(const float* w, int res, int i, float out[4]);
void orc_smlp_query_lambda(const float* w, int res, float lam, float out[4]);
float orc_smlp_lut_interact(const float* w, int rgb_res, int lam_res, const float rgb[3], float lam);
/* training pieces (FP64) */
double orc_smlp_normal(uint32_t seed, uint32_t step, uint32_t sample, uint32_t slot);
double orc_smlp_importance2(double lam);
double orc_smlp_loss_grad(const float* w, uint32_t B, const float* x, const float* y, int noise, uint32_t seed,
                          uint32_t step, double* grad);
double orc_smlp_clip(double* grad, double max_norm);
void orc_smlp_adamw(double* p, double* m, double* v, const double* g, int t, double lr, double b1, double b2,
                    double eps, double wd, double* ema, double ema_decay);
double orc_smlp_lr_factor(int step, int total, double warmup_ratio, double min_ratio);

#ifdef __cplusplus
}
#endif
#endif
