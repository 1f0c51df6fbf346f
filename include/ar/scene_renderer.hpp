// scene_renderer.hpp — header-only C++ front end of the C ABI (include/ar.h)
// with the reference's host API, so code written for the reference's
// SceneRenderer / Camera (core/scene.hpp:18-135, core/camera.hpp:14-99) and
// its scene description (core/render/object.cuh:10-316) builds against this
// core with its includes changed.
//
// Kept: SceneSetting, EObjectType, ESDFType, EShadingModel, ETextureType,
// GBuffer, Texture, Material (+ its Create* helpers), SDFInfo,
// AdditionalObjectInfo, SceneObject (+ GetArea, GetEmissivePower),
// sdf::CreateTorus / CreateDiamond, SceneRenderer (constructor from objects,
// SetExposure, MarkAsDirty, MarkNeedResetGuiding, both Render overloads,
// UpdatePathGuidingSample, ResetPathGuidingSample, AddVolumeToCacheManager,
// ClearCacheManager) and Camera (constructor, position / direction
// accessors, UpdateCamera, RenderToFile, Render, variance_).
//
// Differences a caller sees:
// - Vectors are the POD ar_float3 / ar_float4 / ar_int2 / ar_int3 (the
//   layout of HIP's float3 / float4 / int2 / int3); no HIP or CUDA header is
//   included.
// - Failures throw ar::Error (the ABI's status code and ar_last_error text);
//   the reference prints and continues.
// - SceneObject keeps its Euler angles (rotation_) instead of the two
//   matrices; the library builds objectToWorld = Rz * Ry * Rx from them as
//   the reference's constructor does.
// - Only MAT_LIGHT objects emit (the reference's light list and integrator
//   do the same); their albedo is ignored.
// - The compile-time switches of the reference (core/config.cuh: spectral
//   rendering, sky, tracing depth, seed) are run-time fields of
//   SceneRenderer::config_.
// - Device buffers of the interactive Render (hdr, ldr, odd, even, G-buffer,
//   statistics) are the caller's, as in the reference; the G-buffer and
//   statistics records are AR_GBUFFER_FLOATS / AR_STATS_FLOATS floats per
//   pixel.  The path-guiding sample buffer the reference takes is owned by
//   the context here (the argument is accepted and unused), and each guided
//   frame merges its samples into the probes inside Render, so
//   UpdatePathGuidingSample has nothing left to do.
// - sceneSetting_.bUseBVH_ = false still traverses the BVH: the linear mode
//   and the BVH give the same closest hit (smallest t, then smallest
//   primitive), so only the time differs.
// - The interactive Render is queued on stream_ (a hipStream_t; NULL: the
//   default stream) and does not wait; the offline Render, RenderToFile and
//   Camera::Render's variance wait for their results.
#ifndef AR_SCENE_RENDERER_HPP
#define AR_SCENE_RENDERER_HPP

#include <cmath>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../ar.h"

struct ar_float3 {
    float x, y, z;
};
struct ar_float4 {
    float x, y, z, w;
};
struct ar_int2 {
    int x, y;
};
struct ar_int3 {
    int x, y, z;
};
static_assert(sizeof(ar_float3) == 12 && sizeof(ar_float4) == 16 && sizeof(ar_int2) == 8 && sizeof(ar_int3) == 12,
              "vector layouts");

namespace ar {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const noexcept { return code_; }

private:
    int code_;
};

inline void check(int rc, const ar_ctx* ctx, const char* call) {
    if (rc != AR_OK)
        throw Error(rc, std::string(call) + " failed (" + std::to_string(rc) + "): " + (ctx ? ar_last_error(ctx) : ""));
}

// tone curves of RenderToFile / Render (the reference passes ACES / Gamma as functions, or toneType 0..2)
constexpr uint32_t Linear = AR_TONE_LINEAR;
constexpr uint32_t Gamma = AR_TONE_GAMMA;
constexpr uint32_t ACES = AR_TONE_ACES;

using ToneMapFunction = std::function<ar_float3(const ar_float3)>;

// ---------------------------------------------------------------- scene description (core/render/object.cuh)

struct SceneSetting {
    bool bUseBVH_ = true;
    bool bUseNEE_ = true;
    bool bUsePathGuiding_ = true;
    bool bDenoise_ = false;

    bool bDebugAlbedo_ = false;
    bool bDebugNormal_ = false;
    bool bDebugDepth_ = false;
    bool bDebugMetallic_ = false;
    bool bDebugRoughness_ = false;

    bool bDebugStatisticsNum_ = false;
    bool bDebugStatisticsMean_ = false;
    bool bDebugStatisticsM2_ = false;
    bool bDebugStatisticsM3_ = false;
};

enum class EObjectType { OBJ_CUBE = 0, OBJ_SPHERE = 1, OBJ_SDF = 2 };
enum class ESDFType { SDF_TORUS = 0, SDF_VOLUME = 1, SDF_UNDEFINED = 2, SDF_MAX = 3 };
enum class EShadingModel { MAT_DIFFUSE = 0, MAT_LIGHT = 1, MAT_GLASS = 2, MAT_GGX = 3, MAT_MAX = 4 };
enum class ETextureType { TEX_PURE_COLOR = 0 };

struct GBuffer {
    ar_float3 albedo_ = {1.0f, 1.0f, 1.0f};
    ar_float3 emissive_ = {0.0f, 0.0f, 0.0f};
    float specular_ = 0.0f;
    float metallic_ = 0.0f;
    float roughness_ = 0.0f;
    float IOR_ = 1.5f;
    float sigmaA_ = 100.0f;
    ar_float3 normal_ = {0.0f, 0.0f, 1.0f};
    float abbeNumber_ = 20.0f;
    float tiltNumber_ = 0.0f;
};

struct Texture {
    ETextureType textureType_ = ETextureType::TEX_PURE_COLOR;
    GBuffer baseTexture_ = {};
};

inline float clamp01(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }

struct Material {
    EShadingModel shadingModel_ = EShadingModel::MAT_DIFFUSE;
    Texture texture_ = {};

    static Material CreateDiffusePureColor(const ar_float3 color, const ar_float3 emissive = {0.0f, 0.0f, 0.0f}) {
        Material m;
        m.texture_.baseTexture_.albedo_ = color;
        m.texture_.baseTexture_.emissive_ = emissive;
        return m;
    }
    static Material CreateLight(const ar_float3 emissive) {
        Material m;
        m.shadingModel_ = EShadingModel::MAT_LIGHT;
        m.texture_.baseTexture_.albedo_ = {0.0f, 0.0f, 0.0f};
        m.texture_.baseTexture_.emissive_ = emissive;
        return m;
    }
    static Material CreateGlassPureColor(const ar_float3 color, const float IOR, const float sigmaA,
                                         const float abbeNumber = 20.0f, const float tiltNumber = 0.0f) {
        Material m;
        m.shadingModel_ = EShadingModel::MAT_GLASS;
        GBuffer& g = m.texture_.baseTexture_;
        g.albedo_ = color;
        g.IOR_ = IOR;
        g.sigmaA_ = sigmaA;
        g.abbeNumber_ = abbeNumber;
        g.tiltNumber_ = tiltNumber;
        return m;
    }
    // the glass BSDF is smooth: roughness is stored and not used, as in the reference
    static Material CreateGGXGlassPureColor(const ar_float3 color, const float IOR, const float sigmaA,
                                            const float roughness) {
        Material m = CreateGlassPureColor(color, IOR, sigmaA);
        m.texture_.baseTexture_.roughness_ = roughness;
        return m;
    }
    static Material CreateGGXPureColor(const ar_float3 baseColor, const float roughness, const float metallic,
                                       const float specular, const ar_float3 emissive = {0.0f, 0.0f, 0.0f}) {
        Material m;
        m.shadingModel_ = EShadingModel::MAT_GGX;
        GBuffer& g = m.texture_.baseTexture_;
        g.albedo_ = baseColor;
        g.roughness_ = clamp01(roughness, 0.01f, 1.0f);
        g.metallic_ = clamp01(metallic, 0.0f, 1.0f);
        g.specular_ = clamp01(specular, 0.0f, 1.0f);
        g.emissive_ = emissive;
        return m;
    }

    ar_material ToAr() const {
        const GBuffer& g = texture_.baseTexture_;
        const bool light = shadingModel_ == EShadingModel::MAT_LIGHT;
        ar_material a{};
        a.albedo[0] = light ? 0.0f : g.albedo_.x;
        a.albedo[1] = light ? 0.0f : g.albedo_.y;
        a.albedo[2] = light ? 0.0f : g.albedo_.z;
        a.emission[0] = light ? g.emissive_.x : 0.0f;
        a.emission[1] = light ? g.emissive_.y : 0.0f;
        a.emission[2] = light ? g.emissive_.z : 0.0f;
        a.specular = g.specular_;
        a.metallic = g.metallic_;
        a.roughness = g.roughness_;
        a.model = shadingModel_ == EShadingModel::MAT_GGX     ? AR_MODEL_GGX
                  : shadingModel_ == EShadingModel::MAT_GLASS ? AR_MODEL_GLASS
                                                              : AR_MODEL_LAMBERT;
        a.ior = g.IOR_;
        a.sigma_a = g.sigmaA_;
        a.abbe = g.abbeNumber_;
        a.tilt = g.tiltNumber_;
        return a;
    }
};

struct SDFInfo {
    ar_float4 sdfInfo0_ = {0.0f, 0.0f, 0.0f, 0.0f};
    int cachedVolumeIndex_ = -1;
};

struct AdditionalObjectInfo {
    ESDFType sdfType_ = ESDFType::SDF_TORUS;
    SDFInfo sdfInfo_ = {};
};

struct SceneObject {
    ar_float3 center_ = {0.0f, 0.0f, 0.0f};
    ar_float3 extent_ = {0.0f, 0.0f, 0.0f};
    Material material_ = {};
    ar_float3 rotation_ = {0.0f, 0.0f, 0.0f};   // Euler angles in degrees: objectToWorld = Rz * Ry * Rx
    EObjectType type_ = EObjectType::OBJ_CUBE;
    AdditionalObjectInfo additionalObjectInfo_ = {};

    SceneObject() = default;
    SceneObject(const ar_float3 center, const ar_float3 extent, const Material& material, const ar_float3 rotation,
                const EObjectType type = EObjectType::OBJ_CUBE, const AdditionalObjectInfo& info = {})
        : center_(center), extent_{std::fabs(extent.x), std::fabs(extent.y), std::fabs(extent.z)}, material_(material),
          rotation_(rotation), type_(type), additionalObjectInfo_(info) {}

    // surface area: a box's 8 (xy + xz + yz) of its half extents, a sphere's 4 pi x^2
    float GetArea() const {
        const float ex = std::fabs(extent_.x), ey = std::fabs(extent_.y), ez = std::fabs(extent_.z);
        if (type_ == EObjectType::OBJ_SPHERE) return 4.0f * 3.14159265358979f * ex * ex;
        return 8.0f * (ex * ey + ex * ez + ey * ez);
    }
    // Rec. 709 luminance of the emission
    float GetEmissivePower() const {
        const ar_float3 e = material_.texture_.baseTexture_.emissive_;
        return 0.2126f * e.x + 0.7152f * e.y + 0.0722f * e.z;
    }

    ar_object ToAr() const {
        ar_object o{};
        switch (type_) {
        case EObjectType::OBJ_CUBE: o.type = AR_OBJ_BOX; break;
        case EObjectType::OBJ_SPHERE: o.type = AR_OBJ_SPHERE; break;
        case EObjectType::OBJ_SDF:
            if (additionalObjectInfo_.sdfType_ == ESDFType::SDF_TORUS) {
                o.type = AR_OBJ_SDF_TORUS;
                o.sdf[0] = additionalObjectInfo_.sdfInfo_.sdfInfo0_.x;
                o.sdf[1] = additionalObjectInfo_.sdfInfo_.sdfInfo0_.y;
            } else if (additionalObjectInfo_.sdfType_ == ESDFType::SDF_VOLUME) {
                o.type = AR_OBJ_SDF_VOLUME;
                o.sdf_volume = additionalObjectInfo_.sdfInfo_.cachedVolumeIndex_;
            } else {
                throw Error(AR_E_INVALID, "SceneObject: an OBJ_SDF object must be SDF_TORUS or SDF_VOLUME");
            }
            break;
        }
        o.sdf_volume = o.type == AR_OBJ_SDF_VOLUME ? o.sdf_volume : -1;
        o.center[0] = center_.x; o.center[1] = center_.y; o.center[2] = center_.z;
        o.extent[0] = extent_.x; o.extent[1] = extent_.y; o.extent[2] = extent_.z;
        o.rotation[0] = rotation_.x; o.rotation[1] = rotation_.y; o.rotation[2] = rotation_.z;
        o.material = material_.ToAr();
        return o;
    }
};

namespace sdf {

// CreateTorus (core/object/sdf/sdf_torus.cuh:9-12): radii R, r in unit space, times scale
inline AdditionalObjectInfo CreateTorus(const float scale = 1.0f, const float R = 0.375f, const float r = 0.125f) {
    AdditionalObjectInfo a;
    a.sdfType_ = ESDFType::SDF_TORUS;
    a.sdfInfo_.sdfInfo0_ = {R * scale, r * scale, 0.0f, 0.0f};
    return a;
}

// CreateDiamond (core/object/sdf/sdf_diamond.cuh:9-12): the parameters of the
// baked diamond volume; objects use it as SDF_VOLUME with the volume's index
inline AdditionalObjectInfo CreateDiamond(const float scale = 1.0f, const float d = 0.94f, const float bias = 0.5f,
                                          const float tableZ = 0.3f) {
    AdditionalObjectInfo a;
    a.sdfType_ = ESDFType::SDF_UNDEFINED;
    a.sdfInfo_.sdfInfo0_ = {d, bias, tableZ, scale * 0.5f};
    return a;
}

// Shapes AddVolumeToCacheManager bakes inside the library (any other functor
// is evaluated on the host: float operator()(ar_float3 u, const SDFInfo&) with
// u in [-0.5, 0.5]^3)
struct SDFTorusFunctor {};
struct SDFDiamondFunctor {};

} // namespace sdf

// ---------------------------------------------------------------- run-time configuration

// The reference fixes these at compile time (core/config.cuh); defaults follow it.
struct RenderConfig {
    uint64_t seed = 0;
    uint32_t max_depth = 1000;              // MAX_TRACING_DEPTH
    uint32_t rr_min_depth = 3;
    uint32_t sky = AR_SKY_GRADIENT;         // SkyLighting on misses
    float env[3] = {0.0f, 0.0f, 0.0f};      // AR_SKY_CONSTANT radiance
    float vfov_deg = 60.0f;
    uint32_t spectral = AR_SPECTRAL_RGB;    // USE_SPECTRUM_RENDERING: AR_SPECTRAL_SINGLE or AR_SPECTRAL_HERO4
    uint32_t output = AR_OUTPUT_XYZ;        // spectral output space (XYZ, converted by the tonemap)
    uint32_t guiding = AR_GUIDING_VOXEL;    // the guide bUsePathGuiding_ turns on
    ar_bvh_opts bvh = {4, 2, 16, 64};
};

// ---------------------------------------------------------------- SceneRenderer (core/scene.hpp:18-135)

class SceneRenderer {
public:
    SceneSetting sceneSetting_;
    RenderConfig config_;
    void* stream_ = nullptr;   // hipStream_t of the interactive Render (NULL: the default stream)

    explicit SceneRenderer(const std::vector<SceneObject>& sceneObjects, int device = 0)
        : sceneObjects_(sceneObjects) {
        ar_device_opts d{device, 0};
        check(ar_create(&ctx_, &d), nullptr, "ar_create");
    }
    // no counterpart: objects plus triangle meshes (copied now; the caller keeps its arrays)
    SceneRenderer(const std::vector<SceneObject>& sceneObjects, const std::vector<ar_mesh>& meshes, int device = 0)
        : SceneRenderer(sceneObjects, device) {
        SetMeshes(meshes);
    }
    ~SceneRenderer() {
        if (ctx_) (void)ar_destroy(ctx_);
    }
    SceneRenderer(const SceneRenderer&) = delete;
    SceneRenderer& operator=(const SceneRenderer&) = delete;

    void SetExposure(float exp) { exposure_ = exp; }
    void MarkAsDirty() { dirty_ = true; }
    void MarkNeedResetGuiding() { resetGuiding_ = true; }

    std::vector<SceneObject>& Objects() { return sceneObjects_; }   // edit, then MarkAsDirty
    const std::vector<SceneObject>& Objects() const { return sceneObjects_; }
    void SetMeshes(const std::vector<ar_mesh>& meshes) {
        check(ar_scene_set(ctx_, meshes.data(), meshes.size()), ctx_, "ar_scene_set");
        meshesSet_ = true;
        dirty_ = true;
    }
    ar_ctx* Context() const { return ctx_; }

    // Offline render (core/scene.cu:215-237): sampleNum samples per pixel,
    // returned as a host image, pixel j * size.x + i.
    std::vector<ar_float3> Render(const ar_int2 size, const ar_float3 cameraOrigin, const ar_float3 cameraDirection,
                                  int sampleNum = 1024) {
        if (size.x <= 0 || size.y <= 0) throw Error(AR_E_INVALID, "SceneRenderer: empty image");
        std::vector<ar_float3> out((size_t)size.x * size.y);
        RenderHost(size, cameraOrigin, cameraDirection, sampleNum, &out[0].x, nullptr, 1.0f, AR_TONE_LINEAR);
        return out;
    }

    // Interactive frame (core/scene.cu:240-365): one sample per pixel with
    // sample index frameIndex folded into the running means hdrTarget (and
    // oddTarget / evenTarget), the guide updated (or reset when guiding is
    // off), then ldrTarget written from hdrTarget through the exposure and
    // toneType (0 linear, 1 gamma, 2 ACES), denoised first with bDenoise_, or
    // one of the debug views.  All buffers are device memory.
    void Render(ar_float3* hdrTarget, unsigned int* ldrTarget, const ar_int2 size, const ar_float3 cameraOrigin,
                const ar_float3 cameraDirection, int frameIndex, int toneType, ar_float3* oddTarget = nullptr,
                ar_float3* evenTarget = nullptr, void* radianceSampleBuffer = nullptr, float* screenGBuffer = nullptr,
                float* screenStatisticsBuffer = nullptr) {
        (void)radianceSampleBuffer;
        UpdateSceneToDeviceIfDirty();
        if (resetGuiding_ || !sceneSetting_.bUsePathGuiding_) {
            check(ar_guiding_reset(ctx_), ctx_, "ar_guiding_reset");
            resetGuiding_ = false;
        }
        const ar_camera cam = Cam(cameraOrigin, cameraDirection);
        ar_render_opts o = Opts(size, 1);
        o.guiding = sceneSetting_.bUsePathGuiding_ ? config_.guiding : AR_GUIDING_OFF;
        check(ar_render_progressive(ctx_, &cam, &o, (uint32_t)frameIndex, &hdrTarget->x,
                                    oddTarget ? &oddTarget->x : nullptr, evenTarget ? &evenTarget->x : nullptr,
                                    screenGBuffer, screenStatisticsBuffer, stream_),
              ctx_, "ar_render_progressive");
        if (ldrTarget)
            check(ar_present(ctx_, &hdrTarget->x, screenGBuffer, screenStatisticsBuffer, ldrTarget, (uint32_t)size.x,
                             (uint32_t)size.y, exposure_, (uint32_t)toneType, XyzOutput() ? 1u : 0u, View(), stream_),
                  ctx_, "ar_present");
    }

    void UpdatePathGuidingSample() {}
    void ResetPathGuidingSample(void* stream = nullptr) {
        (void)stream;
        check(ar_guiding_reset(ctx_), ctx_, "ar_guiding_reset");
    }

    // SDFCacheManager::AddVolume (core/object/sdf/sdf_cache_manager.cuh:63-107):
    // a dims grid of shape(p01 - 0.5), p01 = (x, y, z) / (dims - 1); returns
    // the volume index (SDFInfo::cachedVolumeIndex_ of the objects that use it).
    // pMin / pMax are unused, as in the reference.
    template <class SDFFunctor>
    uint32_t AddVolumeToCacheManager(ar_int3 dims, ar_float3 pMin, ar_float3 pMax, const SDFFunctor& sdfFn,
                                     const SDFInfo& sdfInfo) {
        (void)pMin;
        (void)pMax;
        if (dims.x <= 0 || dims.y <= 0 || dims.z <= 0) throw Error(AR_E_INVALID, "Invalid volume dims");
        const uint32_t d[3] = {(uint32_t)dims.x, (uint32_t)dims.y, (uint32_t)dims.z};
        const uint32_t index = AddVolume(d, sdfFn, sdfInfo);
        dirty_ = true;
        return index;
    }
    void ClearCacheManager() {
        check(ar_sdf_volume_clear(ctx_), ctx_, "ar_sdf_volume_clear");
    }

    // Offline render into host memory: hdr (3 floats per pixel) and / or ldr
    // (0xAARRGGBB per pixel through exposure and tone).  Used by Render and
    // Camera::RenderToFile.
    void RenderHost(const ar_int2 size, const ar_float3 origin, const ar_float3 dir, int sampleNum, float* hdr,
                    uint32_t* ldr, float exposure, uint32_t tone) {
        UpdateSceneToDeviceIfDirty();
        if (resetGuiding_) {
            check(ar_guiding_reset(ctx_), ctx_, "ar_guiding_reset");
            resetGuiding_ = false;
        }
        const ar_camera cam = Cam(origin, dir);
        ar_render_opts o = Opts(size, (uint32_t)sampleNum);
        o.guiding = sceneSetting_.bUsePathGuiding_ ? config_.guiding : AR_GUIDING_OFF;
        check(ar_render_host(ctx_, &cam, &o, hdr, ldr, exposure, tone), ctx_, "ar_render_host");
    }

    bool XyzOutput() const { return config_.spectral != AR_SPECTRAL_RGB && config_.output == AR_OUTPUT_XYZ; }

private:
    ar_ctx* ctx_ = nullptr;
    std::vector<SceneObject> sceneObjects_;
    bool dirty_ = true;
    bool meshesSet_ = false;
    bool resetGuiding_ = true;
    float exposure_ = 1.0f;

    // UpdateSceneToDeviceIfDirty (core/scene.cu:656-732): objects, then the BVH
    void UpdateSceneToDeviceIfDirty() {
        if (!dirty_) return;
        if (!meshesSet_) {
            check(ar_scene_set(ctx_, nullptr, 0), ctx_, "ar_scene_set");
            meshesSet_ = true;
        }
        std::vector<ar_object> objs;
        objs.reserve(sceneObjects_.size());
        for (const SceneObject& so : sceneObjects_) objs.push_back(so.ToAr());
        check(ar_scene_set_objects(ctx_, objs.data(), objs.size()), ctx_, "ar_scene_set_objects");
        check(ar_bvh_build(ctx_, &config_.bvh), ctx_, "ar_bvh_build");
        dirty_ = false;
        resetGuiding_ = true;
    }

    ar_camera Cam(const ar_float3 origin, const ar_float3 dir) const {
        return ar_camera{{origin.x, origin.y, origin.z}, {dir.x, dir.y, dir.z}, config_.vfov_deg};
    }

    ar_render_opts Opts(const ar_int2 size, uint32_t spp) const {
        if (size.x <= 0 || size.y <= 0) throw Error(AR_E_INVALID, "SceneRenderer: empty image");
        ar_render_opts o{};
        o.width = (uint32_t)size.x;
        o.height = (uint32_t)size.y;
        o.spp = spp;
        o.seed = config_.seed;
        o.max_depth = config_.max_depth;
        o.rr_min_depth = config_.rr_min_depth;
        o.nee = sceneSetting_.bUseNEE_ ? 1u : 0u;
        o.sky = config_.sky;
        for (int k = 0; k < 3; ++k) o.env[k] = config_.env[k];
        o.spectral = config_.spectral;
        o.output = config_.output;
        o.frame_spp = 1;
        return o;
    }

    // the reference's output switch: debug views first, then the denoiser
    uint32_t View() const {
        const SceneSetting& s = sceneSetting_;
        if (s.bDebugAlbedo_) return AR_VIEW_ALBEDO;
        if (s.bDebugNormal_) return AR_VIEW_NORMAL;
        if (s.bDebugDepth_) return AR_VIEW_DEPTH;
        if (s.bDebugMetallic_) return AR_VIEW_METALLIC;
        if (s.bDebugRoughness_) return AR_VIEW_ROUGHNESS;
        if (s.bDebugStatisticsNum_) return AR_VIEW_STAT_COUNT;
        if (s.bDebugStatisticsMean_) return AR_VIEW_STAT_MEAN;
        if (s.bDebugStatisticsM2_) return AR_VIEW_STAT_M2;
        if (s.bDebugStatisticsM3_) return AR_VIEW_STAT_M3;
        return s.bDenoise_ ? AR_VIEW_DENOISED : AR_VIEW_IMAGE;
    }

    uint32_t Bake(uint32_t shape, const uint32_t d[3], const SDFInfo& info) {
        const float prm[4] = {info.sdfInfo0_.x, info.sdfInfo0_.y, info.sdfInfo0_.z, info.sdfInfo0_.w};
        uint32_t index = 0;
        check(ar_sdf_volume_bake(ctx_, shape, prm, d, &index), ctx_, "ar_sdf_volume_bake");
        return index;
    }
    uint32_t AddVolume(const uint32_t d[3], const sdf::SDFTorusFunctor&, const SDFInfo& info) {
        return Bake(AR_SDF_SHAPE_TORUS, d, info);
    }
    uint32_t AddVolume(const uint32_t d[3], const sdf::SDFDiamondFunctor&, const SDFInfo& info) {
        return Bake(AR_SDF_SHAPE_DIAMOND, d, info);
    }
    template <class F>
    uint32_t AddVolume(const uint32_t d[3], const F& fn, const SDFInfo& info) {
        std::vector<float> v((size_t)d[0] * d[1] * d[2]);
        auto at = [](uint32_t i, uint32_t n) { return n > 1 ? (float)i / (float)(n - 1) : 0.0f; };
        for (uint32_t z = 0; z < d[2]; ++z)
            for (uint32_t y = 0; y < d[1]; ++y)
                for (uint32_t x = 0; x < d[0]; ++x)
                    v[((size_t)z * d[1] + y) * d[0] + x] =
                        fn(ar_float3{at(x, d[0]) - 0.5f, at(y, d[1]) - 0.5f, at(z, d[2]) - 0.5f}, info);
        uint32_t index = 0;
        check(ar_sdf_volume_add(ctx_, v.data(), d, &index), ctx_, "ar_sdf_volume_add");
        return index;
    }
};

// ---------------------------------------------------------------- Camera (core/camera.hpp:14-99)

class Camera {
protected:
    ar_float3 position_;
    ar_float3 direction_;

public:
    SceneRenderer* sceneRenderer_;
    int resolutionX_;
    int resolutionY_;
    std::string name_;
    mutable double variance_ = 0.0;

    Camera(SceneRenderer& sceneRenderer, const std::string& name = "test camera", const int resolutionX = 1080,
           int resolutionY = 720, const ar_float3 position = {0.0f, 0.75f, 0.75f},
           const ar_float3 direction = {0.0f, 0.0f, -0.75f})
        : position_(position), direction_(direction), sceneRenderer_(&sceneRenderer), resolutionX_(resolutionX),
          resolutionY_(resolutionY), name_(name) {
        SetDirection(direction);
    }

    void SetPosition(const ar_float3 position) { position_ = position; }
    ar_float3 GetPosition() const { return position_; }
    void SetDirection(const ar_float3 d) {
        const float n = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
        direction_ = n > 1e-20f ? ar_float3{d.x / n, d.y / n, d.z / n} : ar_float3{0.0f, 0.0f, 0.0f};
    }
    ar_float3 GetDirection() const { return direction_; }
    void SetPositionAndDirection(const ar_float3 position, const ar_float3 direction) {
        SetPosition(position);
        SetDirection(direction);
    }
    void SetVolume(SceneRenderer& sceneRenderer) { sceneRenderer_ = &sceneRenderer; }

    // Mouse look and WASD motion (core/camera.hpp:65-96): yaw / pitch of the
    // view direction turned by 0.0025 rad per mouse unit, pitch held within
    // +-89 degrees, then a move of 0.5 per unit along the new forward and
    // along up x forward.
    void UpdateCamera(const double deltaMouseX, const double deltaMouseY, const double deltaWS, const double deltaAD) {
        const ar_float3 d = GetDirection();
        float yaw = std::atan2(d.x, d.z);
        float pitch = std::asin(std::fmin(1.0f, std::fmax(-1.0f, d.y)));
        yaw -= (float)deltaMouseX * 0.0025f;
        pitch -= (float)deltaMouseY * 0.0025f;
        const float limit = 89.0f * 3.1415926f / 180.0f;
        pitch = std::fmin(limit, std::fmax(-limit, pitch));
        const ar_float3 fwd = {std::sin(yaw) * std::cos(pitch), std::sin(pitch), std::cos(yaw) * std::cos(pitch)};
        const ar_float3 right = {fwd.z, 0.0f, -fwd.x};   // (0, 1, 0) x fwd
        const float ws = 0.5f * (float)deltaWS, ad = 0.5f * (float)deltaAD;
        const ar_float3 p = GetPosition();
        SetPositionAndDirection({p.x + fwd.x * ws + right.x * ad, p.y + fwd.y * ws + right.y * ad,
                                 p.z + fwd.z * ws + right.z * ad},
                                fwd);
    }

    // RenderToFile (core/camera.cu:64-89): sampleNum samples per pixel, the
    // tone curve, written to path + ".bmp" (32-bit, row 0 first).
    void RenderToFile(const std::string& path, const int sampleNum = 1, const uint32_t tone = AR_TONE_ACES,
                      const float exp = 1.0f) const {
        std::vector<uint32_t> ldr((size_t)resolutionX_ * resolutionY_);
        sceneRenderer_->RenderHost({resolutionX_, resolutionY_}, position_, direction_, sampleNum, nullptr, ldr.data(),
                                   exp, tone);
        Write(path, ldr);
    }
    // the same with a host tone curve on linear sRGB (spectral XYZ converted first)
    void RenderToFile(const std::string& path, const int sampleNum, const ToneMapFunction& tonemapFunction,
                      const float exp = 1.0f) const {
        const std::vector<ar_float3> hdr = sceneRenderer_->Render({resolutionX_, resolutionY_}, position_, direction_,
                                                                  sampleNum);
        const bool xyz = sceneRenderer_->XyzOutput();
        std::vector<uint32_t> ldr(hdr.size());
        auto sat = [](float x) { return std::fmin(std::fmax(x, 0.0f), 1.0f); };
        for (size_t i = 0; i < hdr.size(); ++i) {
            ar_float3 v = {hdr[i].x * exp, hdr[i].y * exp, hdr[i].z * exp};
            if (xyz) {   // XYZ2SRGBLinearD65 (core/tonemap.cuh:116-126)
                const float r = 3.2406f * v.x - 1.5372f * v.y - 0.4986f * v.z;
                const float g = -0.9689f * v.x + 1.8758f * v.y + 0.0415f * v.z;
                const float b = 0.0557f * v.x - 0.204f * v.y + 1.057f * v.z;
                v = {std::fmax(0.0f, r), std::fmax(0.0f, g), std::fmax(0.0f, b)};
            }
            const ar_float3 t = tonemapFunction(v);
            ldr[i] = 0xff000000u | ((uint32_t)(sat(t.x) * 255.0f) << 16) | ((uint32_t)(sat(t.y) * 255.0f) << 8) |
                     (uint32_t)(sat(t.z) * 255.0f);
        }
        Write(path, ldr);
    }

    // Camera::Render (core/camera.cu:91-98): one interactive frame; with odd /
    // even halves, variance_ = 1e6 x their luminance MSE (waits).
    void Render(ar_float3* target, ar_float3* oddTarget, ar_float3* evenTarget, void* sampleTarget,
                float* screenGBuffer, float* screenStatisticsBuffer, unsigned int* target2, const ar_int2 size,
                const int frame, const int tone = 2) const {
        sceneRenderer_->Render(target, target2, size, position_, direction_, frame, tone, oddTarget, evenTarget,
                               sampleTarget, screenGBuffer, screenStatisticsBuffer);
        if (oddTarget && evenTarget && size.x * size.y > 0) {
            double mse = 0.0;
            ar_ctx* c = sceneRenderer_->Context();
            check(ar_luminance_mse(c, &oddTarget->x, &evenTarget->x, (size_t)size.x * size.y, &mse,
                                   sceneRenderer_->stream_),
                  c, "ar_luminance_mse");
            variance_ = mse * 1000000.0;
        }
    }

private:
    void Write(const std::string& path, const std::vector<uint32_t>& ldr) const {
        const std::string file = path + ".bmp";
        const int rc = ar_write_bmp(file.c_str(), ldr.data(), (uint32_t)resolutionX_, (uint32_t)resolutionY_);
        if (rc != AR_OK) throw Error(rc, "ar_write_bmp(" + file + ") failed");
    }
};

} // namespace ar

#ifndef AR_SCENE_RENDERER_NO_GLOBAL_NAMES
// the reference's names, unqualified, for code written against it
using ar::ACES;
using ar::AdditionalObjectInfo;
using ar::Camera;
using ar::EObjectType;
using ar::ESDFType;
using ar::EShadingModel;
using ar::ETextureType;
using ar::GBuffer;
using ar::Gamma;
using ar::Material;
using ar::SceneObject;
using ar::SceneRenderer;
using ar::SceneSetting;
using ar::SDFInfo;
using ar::Texture;
using ar::ToneMapFunction;
namespace sdf = ar::sdf;
#endif

#endif // AR_SCENE_RENDERER_HPP
