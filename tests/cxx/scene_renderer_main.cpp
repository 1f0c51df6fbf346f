// A program written against the SceneRenderer / Camera wrapper
// (include/ar/scene_renderer.hpp), as the reference's main.cu is written
// against its SceneRenderer.  Built twice by the tests: against libar_hip.so
// (AR_TEST_HIP: the image buffers of the interactive Render are device
// memory) and against the oracle's libar_cpu.so (host memory).  It renders the
// object Cornell box and writes what it rendered to <out_dir> for the tests to
// compare with the oracle:
//
//   offline.f32        SceneRenderer::Render, 48 x 48, 16 spp (seed 3, depth 5, no sky)
//   cornell.bmp        Camera::RenderToFile of the same view (ACES, exposure 1)
//   frames_*.f32/u32   Camera::Render frames 0..7: running mean, odd and even
//                      halves, G-buffer and statistics records, display image
//   denoised_*         frame 8 with bDenoise_: running mean, records, display
//   albedo_*           frame 9 with bDebugAlbedo_: G-buffer, display
//   guided.f32         the offline render with voxel guiding, 64 spp
//   results.txt        variance_, SDF volume checks, error path, UpdateCamera
#include "ar/scene_renderer.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#ifdef AR_TEST_HIP
#include <hip/hip_runtime_api.h>
static void* buf_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) throw ar::Error(AR_E_OOM, "hipMalloc");
    return p;
}
static void buf_free(void* p) { (void)hipFree(p); }
static void to_host(void* dst, const void* src, size_t bytes) {
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        throw ar::Error(AR_E_DEVICE, "hipMemcpy");
}
#else
static void* buf_alloc(size_t bytes) {
    void* p = std::calloc(1, bytes);
    if (!p) throw ar::Error(AR_E_OOM, "calloc");
    return p;
}
static void buf_free(void* p) { std::free(p); }
static void to_host(void* dst, const void* src, size_t bytes) { std::memcpy(dst, src, bytes); }
#endif

static std::string g_dir;

static void save_host(const std::string& name, const void* data, size_t bytes) {
    FILE* f = std::fopen((g_dir + "/" + name).c_str(), "wb");
    if (!f || std::fwrite(data, 1, bytes, f) != bytes) throw ar::Error(AR_E_INVALID, "cannot write " + name);
    std::fclose(f);
}

// a buffer of the interactive Render (device memory with AR_TEST_HIP)
static void save(const std::string& name, const void* buffer, size_t bytes) {
    std::vector<unsigned char> h(bytes);
    to_host(h.data(), buffer, bytes);
    save_host(name, h.data(), bytes);
}

// the object Cornell box (CreateDiffuseCornellBox with a 0.25 x 0.25 light and two boxes)
static std::vector<SceneObject> cornell() {
    const float t = 0.01f;
    const ar_float3 grey{0.75f, 0.75f, 0.75f}, zero{0.0f, 0.0f, 0.0f};
    std::vector<SceneObject> s;
    s.emplace_back(ar_float3{0.0f, -0.5f - t, 0.0f}, ar_float3{0.5f, t, 0.5f}, Material::CreateDiffusePureColor(grey),
                   zero, EObjectType::OBJ_CUBE);
    s.emplace_back(ar_float3{0.0f, 0.5f + t, 0.0f}, ar_float3{0.5f, t, 0.5f}, Material::CreateDiffusePureColor(grey),
                   zero, EObjectType::OBJ_CUBE);
    s.emplace_back(ar_float3{0.0f, 0.5f - t, 0.0f}, ar_float3{0.125f, t, 0.125f},
                   Material::CreateLight(ar_float3{17.0f, 12.0f, 4.0f}), zero, EObjectType::OBJ_CUBE);
    s.emplace_back(ar_float3{0.0f, 0.0f, -0.5f - t}, ar_float3{0.5f, 0.5f, t}, Material::CreateDiffusePureColor(grey),
                   zero, EObjectType::OBJ_CUBE);
    s.emplace_back(ar_float3{-0.5f - t, 0.0f, 0.0f}, ar_float3{t, 0.5f, 0.5f},
                   Material::CreateDiffusePureColor(ar_float3{0.63f, 0.065f, 0.05f}), zero, EObjectType::OBJ_CUBE);
    s.emplace_back(ar_float3{0.5f + t, 0.0f, 0.0f}, ar_float3{t, 0.5f, 0.5f},
                   Material::CreateDiffusePureColor(ar_float3{0.14f, 0.45f, 0.09f}), zero, EObjectType::OBJ_CUBE);
    s.emplace_back(ar_float3{-0.16f, -0.2f, -0.14f}, ar_float3{0.15f, 0.3f, 0.15f},
                   Material::CreateDiffusePureColor(grey), ar_float3{0.0f, 17.0f, 0.0f}, EObjectType::OBJ_CUBE);
    s.emplace_back(ar_float3{0.17f, -0.35f, 0.12f}, ar_float3{0.15f, 0.15f, 0.15f},
                   Material::CreateDiffusePureColor(grey), ar_float3{0.0f, -17.0f, 0.0f}, EObjectType::OBJ_CUBE);
    return s;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s out_dir\n", argv[0]);
        return 2;
    }
    g_dir = argv[1];
    try {
        const int W = 48, H = 48;
        const ar_int2 size{W, H};
        const ar_float3 pos{0.0f, 0.0f, 1.1f}, dir{0.0f, 0.0f, -1.0f};
        FILE* res = std::fopen((g_dir + "/results.txt").c_str(), "w");
        if (!res) throw ar::Error(AR_E_INVALID, "cannot write results.txt");

        SceneRenderer r(cornell());
        r.sceneSetting_.bUsePathGuiding_ = false;
        r.config_.seed = 3;
        r.config_.max_depth = 5;
        r.config_.sky = AR_SKY_NONE;

        // offline render and RenderToFile
        const std::vector<ar_float3> img = r.Render(size, pos, dir, 16);
        save_host("offline.f32", img.data(), img.size() * sizeof(ar_float3));
        Camera cam(r, "cornell", W, H, pos, dir);
        cam.RenderToFile(g_dir + "/cornell", 16);
        std::printf("offline render and RenderToFile done\n");
        std::fflush(stdout);

        // interactive frames
        const size_t np = (size_t)W * H;
        ar_float3* hdr = (ar_float3*)buf_alloc(np * sizeof(ar_float3));
        ar_float3* odd = (ar_float3*)buf_alloc(np * sizeof(ar_float3));
        ar_float3* even = (ar_float3*)buf_alloc(np * sizeof(ar_float3));
        float* gb = (float*)buf_alloc(np * AR_GBUFFER_FLOATS * sizeof(float));
        float* st = (float*)buf_alloc(np * AR_STATS_FLOATS * sizeof(float));
        unsigned* ldr = (unsigned*)buf_alloc(np * sizeof(unsigned));
        for (int f = 0; f < 8; ++f) cam.Render(hdr, odd, even, nullptr, gb, st, ldr, size, f, 2);
        std::fprintf(res, "variance %.17g\n", cam.variance_);
        save("frames_hdr.f32", hdr, np * sizeof(ar_float3));
        save("frames_odd.f32", odd, np * sizeof(ar_float3));
        save("frames_even.f32", even, np * sizeof(ar_float3));
        save("frames_gbuf.f32", gb, np * AR_GBUFFER_FLOATS * sizeof(float));
        save("frames_stats.f32", st, np * AR_STATS_FLOATS * sizeof(float));
        save("frames_ldr.u32", ldr, np * sizeof(unsigned));
        r.sceneSetting_.bDenoise_ = true;
        cam.Render(hdr, odd, even, nullptr, gb, st, ldr, size, 8, 2);
        save("denoised_hdr.f32", hdr, np * sizeof(ar_float3));
        save("denoised_gbuf.f32", gb, np * AR_GBUFFER_FLOATS * sizeof(float));
        save("denoised_stats.f32", st, np * AR_STATS_FLOATS * sizeof(float));
        save("denoised_ldr.u32", ldr, np * sizeof(unsigned));
        r.sceneSetting_.bDenoise_ = false;
        r.sceneSetting_.bDebugAlbedo_ = true;
        cam.Render(hdr, odd, even, nullptr, gb, st, ldr, size, 9, 2);
        save("albedo_gbuf.f32", gb, np * AR_GBUFFER_FLOATS * sizeof(float));
        save("albedo_ldr.u32", ldr, np * sizeof(unsigned));
        r.sceneSetting_.bDebugAlbedo_ = false;
        for (void* p : {(void*)hdr, (void*)odd, (void*)even, (void*)gb, (void*)st, (void*)ldr}) buf_free(p);
        std::printf("interactive frames done\n");
        std::fflush(stdout);

        // voxel-guided offline render
        r.sceneSetting_.bUsePathGuiding_ = true;
        r.MarkNeedResetGuiding();
        const std::vector<ar_float3> guided = r.Render(size, pos, dir, 64);
        save_host("guided.f32", guided.data(), guided.size() * sizeof(ar_float3));
        ar_guiding_info gi{};
        ar::check(ar_guiding_info_get(r.Context(), &gi), r.Context(), "ar_guiding_info_get");
        std::fprintf(res, "probes %llu\n", (unsigned long long)gi.probes);
        std::printf("guided render done\n");
        std::fflush(stdout);

        // SDF volumes: a host functor and the library's bake of the same torus
        {
            SceneRenderer s(cornell());
            const AdditionalObjectInfo torus = sdf::CreateTorus();
            auto host_torus = [](const ar_float3 u, const SDFInfo& info) {
                const float q = std::sqrt(u.x * u.x + u.z * u.z) - info.sdfInfo0_.x;
                return std::sqrt(q * q + u.y * u.y) - info.sdfInfo0_.y;
            };
            const uint32_t a = s.AddVolumeToCacheManager(ar_int3{17, 9, 13}, ar_float3{0, 0, 0}, ar_float3{1, 1, 1},
                                                         host_torus, torus.sdfInfo_);
            const uint32_t b = s.AddVolumeToCacheManager(ar_int3{17, 9, 13}, ar_float3{0, 0, 0}, ar_float3{1, 1, 1},
                                                         sdf::SDFTorusFunctor{}, torus.sdfInfo_);
            std::vector<float> va(17 * 9 * 13), vb(va.size());
            ar::check(ar_sdf_volume_read(s.Context(), a, va.data(), nullptr), s.Context(), "ar_sdf_volume_read");
            ar::check(ar_sdf_volume_read(s.Context(), b, vb.data(), nullptr), s.Context(), "ar_sdf_volume_read");
            double dmax = 0.0;
            for (size_t i = 0; i < va.size(); ++i) dmax = std::fmax(dmax, std::fabs((double)va[i] - vb[i]));
            std::fprintf(res, "sdf_volumes %u %u\nsdf_max_diff %.9g\n", a, b, dmax);
            AdditionalObjectInfo vol;
            vol.sdfType_ = ESDFType::SDF_VOLUME;
            vol.sdfInfo_.cachedVolumeIndex_ = (int)b;
            s.Objects().emplace_back(ar_float3{0.0f, -0.2f, 0.0f}, ar_float3{0.2f, 0.2f, 0.2f},
                                     Material::CreateGlassPureColor(ar_float3{1, 1, 1}, 1.5f, 0.0f),
                                     ar_float3{30.0f, 0.0f, 0.0f}, EObjectType::OBJ_SDF, vol);
            s.Objects().emplace_back(ar_float3{0.25f, 0.2f, 0.0f}, ar_float3{0.1f, 0.1f, 0.1f},
                                     Material::CreateGGXPureColor(ar_float3{0.8f, 0.6f, 0.2f}, 0.3f, 1.0f, 0.5f),
                                     ar_float3{0.0f, 0.0f, 0.0f}, EObjectType::OBJ_SDF, torus);
            s.MarkAsDirty();
            s.sceneSetting_.bUsePathGuiding_ = false;
            s.config_.max_depth = 6;
            const std::vector<ar_float3> si = s.Render(ar_int2{16, 16}, pos, dir, 2);
            double sum = 0.0;
            for (const ar_float3& p : si) sum += p.x + p.y + p.z;
            std::fprintf(res, "sdf_render_mean %.9g\n", sum / (3.0 * si.size()));
        }

        // error path: an SDF object naming a volume that does not exist
        {
            SceneRenderer e(cornell());
            AdditionalObjectInfo bad;
            bad.sdfType_ = ESDFType::SDF_VOLUME;
            bad.sdfInfo_.cachedVolumeIndex_ = 7;
            e.Objects().emplace_back(ar_float3{0, 0, 0}, ar_float3{0.1f, 0.1f, 0.1f},
                                     Material::CreateDiffusePureColor(ar_float3{1, 1, 1}), ar_float3{0, 0, 0},
                                     EObjectType::OBJ_SDF, bad);
            int code = 0;
            try {
                (void)e.Render(ar_int2{4, 4}, pos, dir, 1);
            } catch (const ar::Error& err) {
                code = err.code();
                std::fprintf(res, "error_text %s\n", err.what());
            }
            std::fprintf(res, "error_code %d\n", code);
        }

        // UpdateCamera: a quarter turn of yaw, then one unit forward
        Camera c2(r, "c2", 8, 8, ar_float3{0.0f, 0.0f, 1.0f}, ar_float3{0.0f, 0.0f, -2.0f});
        c2.UpdateCamera(3.14159265358979 / 2.0 / 0.0025, 0.0, 1.0, 0.0);
        const ar_float3 p2 = c2.GetPosition(), d2 = c2.GetDirection();
        std::fprintf(res, "camera %.9g %.9g %.9g %.9g %.9g %.9g\n", p2.x, p2.y, p2.z, d2.x, d2.y, d2.z);
        std::fprintf(res, "area %.9g power %.9g\n", cornell()[0].GetArea(), cornell()[2].GetEmissivePower());
        std::fclose(res);
        std::printf("done\n");
        return 0;
    } catch (const ar::Error& e) {
        std::fprintf(stderr, "ar::Error %d: %s\n", e.code(), e.what());
        return 1;
    }
}
