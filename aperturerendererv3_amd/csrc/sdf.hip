// Signed-distance volumes: the voxel pool of AR_OBJ_SDF_VOLUME objects and
// the GPU bake of the built-in shapes.
//
//   ar_sdf_volume_bake   FillSDFVolume (core/object/sdf/sdf_cache_manager.cuh:16-37)
//                        with SdfTorus (sdf_torus.cuh:14-18) or SdfDiamondUnit
//                        (sdf_diamond.cuh:21-66)
//   ar_sdf_volume_add / _read / _clear   SDFCacheManager::AddVolume / Clear
//                        (sdf_cache_manager.cuh:39-114, sdf_cache_manager.cu:36-48)
//
// The host copy of the pool is the source of truth; the scene upload copies
// it to the device when it changed.  Volumes are appended (offsets never
// move), so records built earlier stay valid.
#include "ar_internal.hpp"
#include "objects.hpp"

#include <cmath>
#include <cstring>

namespace ar {
namespace {

constexpr uint32_t kMaxDim = 4096;

// PlaneSDF (sdf_diamond.cuh:15-19)
__device__ __forceinline__ float plane_sdf(float px, float py, float pz, float nx, float ny, float nz, float d) {
    const float len = fmaxf(sqrtf(nx * nx + ny * ny + nz * nz), 1e-8f);
    return (px * nx + py * ny + pz * nz - d) / len;
}

// SdfDiamondUnit: the facet planes of a brilliant cut in the azimuth sector of p
__device__ float sdf_diamond(float ux, float uy, float uz, float4 prm) {
    const float dd = prm.x != 0.0f ? prm.x : 0.94f;
    const float bias = prm.y != 0.0f ? prm.y : 0.5f;
    const float table = prm.z != 0.0f ? prm.z : 0.3f;
    const float scale = prm.w != 0.0f ? prm.w : 1.0f;
    const float si = 1.0f / scale;
    const float px = ux * si, py = uy * si, pz = uz * si;
    const float af2 = 4.0f / 3.14159265358979f;
    const float s = atan2f(py, px);
    const float sf = floorf(s * af2 + bias) / af2;
    const float sf2 = floorf(s * af2) / af2;
    const float csf = cosf(sf), ssf = sinf(sf);
    const float c1 = cosf(sf + 0.21f), s1 = sinf(sf + 0.21f);
    const float c2 = cosf(sf - 0.21f), s2 = sinf(sf - 0.21f);
    const float c8 = cosf(sf2 + 0.393f), s8 = sinf(sf2 + 0.393f);
    float m = -1e20f;
    m = fmaxf(m, plane_sdf(px, py, pz, csf, ssf, 1.444f, dd));        // crown, bezel facets
    m = fmaxf(m, plane_sdf(px, py, pz, csf, ssf, -1.072f, dd));       // pavilion facets
    m = fmaxf(m, plane_sdf(px, py, pz, 0.0f, 0.0f, 1.0f, table));     // table
    m = fmaxf(m, plane_sdf(px, py, pz, 0.0f, 0.0f, -1.0f, 0.865f));   // culet
    m = fmaxf(m, plane_sdf(px, py, pz, cosf(s), sinf(s), 0.0f, 0.911f));   // girdle
    m = fmaxf(m, plane_sdf(px, py, pz, c1, s1, -1.02f, 0.9193f));     // lower-girdle facets
    m = fmaxf(m, plane_sdf(px, py, pz, c2, s2, -1.02f, 0.9193f));
    m = fmaxf(m, plane_sdf(px, py, pz, c2, s2, 1.03f, 0.912f));       // upper-girdle facets
    m = fmaxf(m, plane_sdf(px, py, pz, c1, s1, 1.03f, 0.912f));
    m = fmaxf(m, plane_sdf(px, py, pz, c8, s8, 2.21f, 1.131f));       // star facets
    return m * scale;
}

// One thread per voxel; x along the wave so that stores coalesce.
__global__ __launch_bounds__(256) void k_sdf_bake(float* __restrict__ out, uint32_t nx, uint32_t ny, uint32_t nz,
                                                   uint32_t shape, float4 prm) {
    const uint32_t x = blockIdx.x * 64u + threadIdx.x;
    const uint32_t y = blockIdx.y * 4u + threadIdx.y;
    const uint32_t z = blockIdx.z;
    if (x >= nx || y >= ny || z >= nz) return;
    const float ux = (nx > 1 ? (float)x / (float)(nx - 1) : 0.0f) - 0.5f;
    const float uy = (ny > 1 ? (float)y / (float)(ny - 1) : 0.0f) - 0.5f;
    const float uz = (nz > 1 ? (float)z / (float)(nz - 1) : 0.0f) - 0.5f;
    out[((size_t)z * ny + y) * nx + x] =
        shape == AR_SDF_SHAPE_TORUS ? obj::sdf_torus(prm, ux, uy, uz) : sdf_diamond(ux, uy, uz, prm);
}

int check_dims(ar_ctx* ctx, const uint32_t* dims, size_t& voxels) {
    if (!dims) return set_error(ctx, AR_E_INVALID, "dims is NULL");
    for (int a = 0; a < 3; ++a)
        if (dims[a] < 1 || dims[a] > kMaxDim) return set_error(ctx, AR_E_INVALID, "volume dims must be 1..4096");
    voxels = (size_t)dims[0] * dims[1] * dims[2];
    if (ctx->sdf.host.size() + voxels > (size_t)UINT32_MAX)
        return set_error(ctx, AR_E_CAPACITY, "SDF voxel pool limited to 2^32 - 1 voxels");
    return AR_OK;
}

uint32_t append(ar_ctx* ctx, const float* voxels, size_t n, const uint32_t* dims) {
    SdfPool& p = ctx->sdf;
    p.off.push_back((uint32_t)p.host.size());
    p.host.insert(p.host.end(), voxels, voxels + n);
    p.dims.insert(p.dims.end(), dims, dims + 3);
    p.dirty = true;
    return (uint32_t)(p.off.size() - 1);
}

} // namespace

int sdf_upload(ar_ctx* ctx) {
    SdfPool& p = ctx->sdf;
    if (!p.dirty) return AR_OK;
    invalidate_graphs(ctx);
    if (p.dev) (void)hipFree(p.dev);
    p.dev = nullptr;
    p.dev_floats = 0;
    p.dirty = false;
    if (p.host.empty()) return AR_OK;
    int rc = check_hip(ctx, hipMalloc((void**)&p.dev, p.host.size() * sizeof(float)), "hipMalloc(sdf pool)");
    if (rc != AR_OK) { p.dev = nullptr; p.dirty = true; return rc; }
    rc = check_hip(ctx, hipMemcpy(p.dev, p.host.data(), p.host.size() * sizeof(float), hipMemcpyHostToDevice),
                   "upload sdf pool");
    if (rc != AR_OK) { p.dirty = true; return rc; }
    p.dev_floats = p.host.size();
    return AR_OK;
}

void sdf_free(ar_ctx* ctx) {
    invalidate_graphs(ctx);
    SdfPool& p = ctx->sdf;
    if (p.dev) (void)hipFree(p.dev);
    p.dev = nullptr;
    p.dev_floats = 0;
    p.dirty = !p.host.empty();
}

} // namespace ar

using namespace ar;

extern "C" int ar_sdf_volume_bake(ar_ctx* ctx, uint32_t shape, const float params[4], const uint32_t dims[3],
                                  uint32_t* index) {
    if (!ctx || !index) return AR_E_INVALID;
    if (shape != AR_SDF_SHAPE_TORUS && shape != AR_SDF_SHAPE_DIAMOND)
        return set_error(ctx, AR_E_INVALID, "unknown SDF shape");
    size_t n = 0;
    int rc = check_dims(ctx, dims, n);
    if (rc != AR_OK) return rc;
    const float4 prm = params ? make_float4(params[0], params[1], params[2], params[3]) : make_float4(0, 0, 0, 0);
    (void)hipSetDevice(ctx->device);
    float* d = nullptr;
    rc = check_hip(ctx, hipMalloc((void**)&d, n * sizeof(float)), "hipMalloc(sdf bake)");
    if (rc != AR_OK) return rc;
    hipLaunchKernelGGL(k_sdf_bake, dim3((dims[0] + 63u) / 64u, (dims[1] + 3u) / 4u, dims[2]), dim3(64, 4, 1), 0,
                       nullptr, d, dims[0], dims[1], dims[2], shape, prm);
    rc = check_hip(ctx, hipGetLastError(), "sdf bake launch");
    std::vector<float> h;
    if (rc == AR_OK) {
        h.resize(n);
        rc = check_hip(ctx, hipMemcpy(h.data(), d, n * sizeof(float), hipMemcpyDeviceToHost), "sdf bake readback");
    }
    (void)hipFree(d);
    if (rc != AR_OK) return rc;
    *index = append(ctx, h.data(), n, dims);
    return AR_OK;
}

extern "C" int ar_sdf_volume_add(ar_ctx* ctx, const float* voxels, const uint32_t dims[3], uint32_t* index) {
    if (!ctx || !index) return AR_E_INVALID;
    if (!voxels) return set_error(ctx, AR_E_INVALID, "voxels is NULL");
    size_t n = 0;
    const int rc = check_dims(ctx, dims, n);
    if (rc != AR_OK) return rc;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(voxels[i])) return set_error(ctx, AR_E_INVALID, "non-finite voxel");
    *index = append(ctx, voxels, n, dims);
    return AR_OK;
}

extern "C" int ar_sdf_volume_read(ar_ctx* ctx, uint32_t index, float* voxels, uint32_t dims[3]) {
    if (!ctx) return AR_E_INVALID;
    const SdfPool& p = ctx->sdf;
    if (index >= p.off.size()) return set_error(ctx, AR_E_INVALID, "no such volume");
    const uint32_t* dm = &p.dims[3 * (size_t)index];
    if (dims) { dims[0] = dm[0]; dims[1] = dm[1]; dims[2] = dm[2]; }
    if (voxels) std::memcpy(voxels, p.host.data() + p.off[index], (size_t)dm[0] * dm[1] * dm[2] * sizeof(float));
    return AR_OK;
}

extern "C" int ar_sdf_volume_clear(ar_ctx* ctx) {
    if (!ctx) return AR_E_INVALID;
    for (const ar_object& o : ctx->objs)
        if (o.type == AR_OBJ_SDF_VOLUME)
            return set_error(ctx, AR_E_INVALID, "the scene's objects refer to a volume: set the objects first");
    (void)hipSetDevice(ctx->device);
    SdfPool& p = ctx->sdf;
    p.host.clear(); p.host.shrink_to_fit();
    p.off.clear();
    p.dims.clear();
    sdf_free(ctx);
    p.dirty = false;
    return AR_OK;
}
