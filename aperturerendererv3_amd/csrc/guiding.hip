// Voxel path guiding: the per-frame probe update on the GPU and the guiding
// C ABI.
//
// The reference (core/guiding/guiding.cuh:70-517, core/scene.cu:409-570)
// compacts the frame's samples, sorts and uniques their (voxel, face) keys,
// grows a dense voxel x 20 index volume, allocates probes for new keys,
// scatters float atomics into a zeroed copy of every probe and then merges
// every probe with one 32-thread block whose lane 0 rebuilds the mips
// serially, with host synchronisations in between.  Here the same update is
// four kernels on the caller's stream with no host round trip:
//   collect  - one thread per path: its valid slots become samples; keys are
//              inserted into an open-addressing hash table (atomicCAS), new
//              table slots are listed;
//   alloc    - one wave per new key: takes a probe from the pool, zeroes it;
//   accum    - one thread per sample: float atomics of radiance * 4 pi into
//              the probe's leaf accumulator, the first sample of a probe
//              lists it as touched;
//   merge    - one wave64 per touched probe, one lane per 8x8 leaf texel:
//              FP64 running mean, count cap 1e6, the 1e-5 floor, and the
//              4x4 / 2x2 / 1x1 mips by cross-lane shuffles; then clears the
//              probe's accumulators.
// Only touched probes are visited, and nothing is sorted.
#include "ar_internal.hpp"
#include "guide.hpp"
#include "mlp.hpp"
#include "path_common.hpp"
#include "spectral.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace ar {
namespace {

using guide::kNoKey;
using guide::kNoProbe;
using guide::kTexels;

constexpr int kBlock = 256;
constexpr uint32_t kGridStrideBlocks = 2048;

// counters: [0] probes allocated (may run past cap; read min(., cap)),
// [1] new table slots this frame, [2] samples this frame, [3] touched
// probes this frame, [4] samples dropped since the reset
enum { kCProbes = 0, kCNew = 1, kCSamples = 2, kCTouched = 3, kCDropped = 4, kCounters = 8 };

__global__ void k_fill_u64(uint64_t* p, size_t n, uint64_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f32(float* p, size_t n, float v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_guide_begin(uint32_t* c) {
    c[kCNew] = 0;
    c[kCSamples] = 0;
    c[kCTouched] = 0;
}

// find-or-insert; returns the table slot, or kNoProbe if the probe run is full
__device__ __forceinline__ uint32_t table_insert(uint64_t* tkey, uint32_t tmask, uint64_t key, uint32_t* newslots,
                                                 uint32_t* counters) {
    uint32_t h = (uint32_t)guide::mix64(key) & tmask;
    for (int s = 0; s < guide::kMaxProbeSteps; ++s, h = (h + 1) & tmask) {
        const uint64_t k = __atomic_load_n(&tkey[h], __ATOMIC_RELAXED);
        if (k == key) return h;
        if (k == kNoKey) {
            const unsigned long long old =
                atomicCAS((unsigned long long*)&tkey[h], (unsigned long long)kNoKey, (unsigned long long)key);
            if (old == kNoKey) {
                newslots[atomicAdd(&counters[kCNew], 1u)] = h;
                return h;
            }
            if (old == key) return h;
        }
    }
    return kNoProbe;
}

struct SlotArgs {
    const float* gr;        // [5*4][stride]
    const uint64_t* gkey;   // [5][stride]
    const uint32_t* gleaf;
    const uint32_t* gflag;
    uint32_t stride;
};

// Samples of each path's valid slots (render.cuh:909-964): spectral modes
// weight the slot radiance by the CMFs over the wavelength density, keep it
// if min(max channel, 1e5) >= 1e-5 and store max(XYZ, 0); RGB keeps it if
// the max channel >= 1e-5.
__global__ __launch_bounds__(kBlock) void k_guide_collect(SlotArgs S, uint32_t npaths, uint32_t spectral,
                                                          const float* __restrict__ lambda0,
                                                          const float* __restrict__ cie, uint64_t* tkey,
                                                          uint32_t tmask, uint32_t* newslots, uint32_t* counters,
                                                          uint32_t* s_slot, uint32_t* s_leaf, float* s_rad,
                                                          size_t scap) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    const bool live = p < npaths;
    const uint32_t flags = live ? S.gflag[p] : 0u;
    const uint32_t leaves = live ? S.gleaf[p] : 0u;
    Lam L = make_lam(spectral, live && spectral != AR_SPECTRAL_RGB ? lambda0[p] : 0.0f);
    for (int i = 0; i < guide::kSlots; ++i) {
        bool ok = live && ((flags >> i) & 1u);
        float x = 0.0f, y = 0.0f, z = 0.0f;
        uint32_t slot = kNoProbe;
        if (ok) {
            const float* r = S.gr + (size_t)(4 * i) * S.stride + p;
            const float r0 = r[0], r1 = r[S.stride], r2 = r[2 * (size_t)S.stride], r3 = r[3 * (size_t)S.stride];
            float strength;
            if (spectral == AR_SPECTRAL_RGB) {
                x = r0; y = r1; z = r2;
                strength = fmaxf(x, fmaxf(y, z));
            } else {
                if (spectral == AR_SPECTRAL_SINGLE) {
                    float cx, cy, cz;
                    spec::cie_xyz(cie, L.l[0], cx, cy, cz);
                    const float pdf = spec::lambda_pdf(L.l[0]);
                    x = cx * r0 / pdf; y = cy * r0 / pdf; z = cz * r0 / pdf;
                } else {
                    const float rr[4] = {r0, r1, r2, r3};
                    for (int k = 0; k < 4; ++k) {
                        float cx, cy, cz;
                        spec::cie_xyz(cie, L.l[k], cx, cy, cz);
                        x = x + cx * rr[k]; y = y + cy * rr[k]; z = z + cz * rr[k];
                    }
                    const float psum = 4.0f / 400.0f;
                    x = x / psum; y = y / psum; z = z / psum;
                }
                strength = fminf(fmaxf(x, fmaxf(y, z)), 1.0e5f);
            }
            const uint64_t key = S.gkey[(size_t)i * S.stride + p];
            ok = strength >= 1.0e-5f && key != kNoKey;
            if (ok) {
                slot = table_insert(tkey, tmask, key, newslots, counters);
                if (slot == kNoProbe) { atomicAdd(&counters[kCDropped], 1u); ok = false; }
            }
        }
        const uint32_t w = block_append<kBlock / 64>(ok, &counters[kCSamples]);
        if (ok && w < scap) {
            s_slot[w] = slot;
            s_leaf[w] = (leaves >> (6 * i)) & 63u;
            s_rad[w] = fmaxf(x, 0.0f);
            s_rad[scap + w] = fmaxf(y, 0.0f);
            s_rad[2 * scap + w] = fmaxf(z, 0.0f);
        }
    }
}

struct Pool {
    uint32_t cap;
    uint32_t* tval;
    const uint64_t* tkey;
    uint64_t* pkey;
    float4* tex;
    uint64_t* cnt;
};

// one wave per new key: a probe from the pool, zeroed
__global__ __launch_bounds__(kBlock) void k_guide_alloc(Pool P, const uint32_t* __restrict__ newslots,
                                                        uint32_t* counters) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nwaves = gridDim.x * (kBlock / 64);
    const uint32_t n = counters[kCNew];
    for (uint32_t w = blockIdx.x * (kBlock / 64) + threadIdx.x / 64; w < n; w += nwaves) {
        const uint32_t h = newslots[w];
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(&counters[kCProbes], 1u);
        idx = __shfl(idx, 0);
        if (idx >= P.cap) continue;
        for (int t = (int)lane; t < kTexels; t += 64) P.tex[(size_t)idx * kTexels + t] = make_float4(0, 0, 0, 0);
        if (lane == 0) {
            P.cnt[idx] = 0;
            P.pkey[idx] = P.tkey[h];
            P.tval[h] = idx;
        }
    }
}

struct Accum {
    const uint32_t* tval;
    float* acc;          // [probe][3][64]
    uint32_t* accn;
    uint32_t* touched;
};

// AccumulateSamplesToOctaLeafKernel (guiding.cuh:371-410)
__global__ __launch_bounds__(kBlock) void k_guide_accum(Accum A, const uint32_t* __restrict__ s_slot,
                                                        const uint32_t* __restrict__ s_leaf,
                                                        const float* __restrict__ s_rad, size_t scap,
                                                        uint32_t* counters) {
    const uint32_t n = (uint32_t)std::min<size_t>(counters[kCSamples], scap);
    const float J = 4.0f * guide::kPiF;
    for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) {
        const uint32_t idx = A.tval[s_slot[s]];
        if (idx == kNoProbe) { atomicAdd(&counters[kCDropped], 1u); continue; }
        const uint32_t leaf = s_leaf[s];
        float* a = A.acc + (size_t)idx * 192 + leaf;
        atomicAdd(a, s_rad[s] * J);
        atomicAdd(a + 64, s_rad[scap + s] * J);
        atomicAdd(a + 128, s_rad[2 * scap + s] * J);
        if (atomicAdd(&A.accn[idx], 1u) == 0u) A.touched[atomicAdd(&counters[kCTouched], 1u)] = idx;
    }
}

__device__ __forceinline__ float4 shfl4(float4 v, int src) {
    return make_float4(__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src), __shfl(v.w, src));
}
__device__ __forceinline__ float4 avg4(float4 a, float4 b, float4 c, float4 d) {
    return make_float4((a.x + b.x + c.x + d.x) * 0.25f, (a.y + b.y + c.y + d.y) * 0.25f,
                       (a.z + b.z + c.z + d.z) * 0.25f, (a.w + b.w + c.w + d.w) * 0.25f);
}

// MergeProbeKernel + GenerateMipmaps (guiding.cuh:429-517,
// octahedron.cuh:455-510): one wave per touched probe, lane = leaf texel.
__global__ __launch_bounds__(kBlock) void k_guide_merge(Accum A, float4* tex, uint64_t* cnt,
                                                        const uint32_t* counters) {
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t nwaves = gridDim.x * (kBlock / 64);
    const uint32_t n = counters[kCTouched];
    for (uint32_t w = blockIdx.x * (kBlock / 64) + threadIdx.x / 64; w < n; w += nwaves) {
        const uint32_t idx = A.touched[w];
        const uint64_t nn = A.accn[idx], no = cnt[idx];
        const double inv = 1.0 / (double)(no + nn);
        const double alpha = (double)no * inv;
        float4* t = tex + (size_t)idx * kTexels;
        float* a = A.acc + (size_t)idx * 192;
        float4 v = t[lane];
        v.x = (float)((double)v.x * alpha + (double)a[lane] * inv);
        v.y = (float)((double)v.y * alpha + (double)a[64 + lane] * inv);
        v.z = (float)((double)v.z * alpha + (double)a[128 + lane] * inv);
        v.x = fmaxf(v.x, 1.0e-5f); v.y = fmaxf(v.y, 1.0e-5f); v.z = fmaxf(v.z, 1.0e-5f); v.w = fmaxf(v.w, 1.0e-5f);
        t[lane] = v;
        // level 1 (4x4): texel (x, y) averages leaves (2x, 2y), (2x+1, 2y), (2x, 2y+1), (2x+1, 2y+1)
        const int x1 = lane & 3, y1 = (lane >> 2) & 3;
        const int b1 = 16 * y1 + 2 * x1;
        const float4 l1 = avg4(shfl4(v, b1), shfl4(v, b1 + 1), shfl4(v, b1 + 8), shfl4(v, b1 + 9));
        if (lane < 16) t[64 + lane] = l1;
        // level 2 (2x2) from level 1 held in lanes 0..15 (index 4 y + x)
        const int x2 = lane & 1, y2 = (lane >> 1) & 1;
        const int b2 = 8 * y2 + 2 * x2;
        const float4 l2 = avg4(shfl4(l1, b2), shfl4(l1, b2 + 1), shfl4(l1, b2 + 4), shfl4(l1, b2 + 5));
        if (lane < 4) t[80 + lane] = l2;
        const float4 l3 = avg4(shfl4(l2, 0), shfl4(l2, 1), shfl4(l2, 2), shfl4(l2, 3));
        if (lane == 0) {
            t[84] = l3;
            t[85] = make_float4(0, 0, 0, 0);
            cnt[idx] = no + nn < 1000000ull ? no + nn : 1000000ull;
        }
        a[lane] = 0.0f;
        a[64 + lane] = 0.0f;
        a[128 + lane] = 0.0f;
        if (lane == 0) A.accn[idx] = 0u;
    }
}

// resume: probe i gets key keys[i]
__global__ void k_guide_restore(uint64_t* tkey, uint32_t* tval, uint32_t tmask, const uint64_t* __restrict__ pkey,
                                uint32_t n, uint32_t* newslots, uint32_t* counters) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t h = table_insert(tkey, tmask, pkey[i], newslots, counters);
        if (h != kNoProbe) tval[h] = i;
    }
}

template <typename T>
int dev_alloc(ar_ctx* ctx, T** p, size_t n, size_t& bytes) {
    if (n == 0) n = 1;
    const hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(ctx, AR_E_OOM, "guiding: device allocation failed");
    }
    bytes += n * sizeof(T);
    return AR_OK;
}

template <typename T>
void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

uint32_t grid_for(size_t n) { return (uint32_t)std::max<size_t>(1, std::min<size_t>((n + kBlock - 1) / kBlock, 65535u * 16u)); }

void free_slots(GuideDev& g) {
    dev_free(g.s_slot); dev_free(g.s_leaf); dev_free(g.s_rad);
    dev_free(g.gm); dev_free(g.gr); dev_free(g.gkey); dev_free(g.gleaf); dev_free(g.gflag);
    dev_free(g.gx); dev_free(g.gw); dev_free(g.gp);
    g.scap = g.pcap = 0;
}

void free_pool(GuideDev& g) {
    dev_free(g.tkey); dev_free(g.tval); dev_free(g.pkey); dev_free(g.tex); dev_free(g.cnt);
    dev_free(g.acc); dev_free(g.accn); dev_free(g.touched); dev_free(g.newslots); dev_free(g.counters);
    g.cap = g.tcap = 0;
    g.bytes = 0;
}

} // namespace

void guide_free(ar_ctx* ctx) {
    invalidate_graphs(ctx);
    free_slots(ctx->guide);
    free_pool(ctx->guide);
}

int guide_clear(ar_ctx* ctx, hipStream_t s) {
    GuideDev& g = ctx->guide;
    if (!g.cap) return AR_OK;
    hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(kBlock), 0, s, g.tkey, (size_t)g.tcap, kNoKey);
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(kBlock), 0, s, g.tval, (size_t)g.tcap, kNoProbe);
    hipLaunchKernelGGL(k_fill_f32, dim3(2048), dim3(kBlock), 0, s, g.acc, (size_t)g.cap * 192, 0.0f);
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(kBlock), 0, s, g.accn, (size_t)g.cap, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, g.counters, (size_t)kCounters, 0u);
    return check_hip(ctx, hipGetLastError(), "guiding reset");
}

int guide_ensure(ar_ctx* ctx, size_t paths, bool pool) {
    GuideDev& g = ctx->guide;
    (void)hipSetDevice(ctx->device);
    if (pool && g.cap != ctx->guide_cap_req) {
        invalidate_graphs(ctx);
        free_pool(g);
        const uint32_t cap = (uint32_t)ctx->guide_cap_req;
        uint32_t tcap = 1;
        while (tcap < 2u * cap) tcap <<= 1;
        size_t b = 0;
        int rc = AR_OK;
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.tkey, tcap, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.tval, tcap, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.newslots, tcap, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.pkey, cap, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.tex, (size_t)cap * kTexels, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.cnt, cap, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.acc, (size_t)cap * 192, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.accn, cap, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.touched, cap, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.counters, kCounters, b);
        if (rc != AR_OK) {
            free_pool(g);
            return rc;
        }
        g.cap = cap;
        g.tcap = tcap;
        g.bytes = b;
        rc = guide_clear(ctx, nullptr);
        if (rc != AR_OK) return rc;
        if (hipDeviceSynchronize() != hipSuccess) return set_error(ctx, AR_E_DEVICE, "guiding: clear failed");
    }
    if (g.pcap < paths) {
        invalidate_graphs(ctx);
        free_slots(g);
        size_t b = 0;
        int rc = AR_OK;
        const size_t scap = (size_t)guide::kSlots * paths;
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.s_slot, scap, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.s_leaf, scap, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.s_rad, 3 * scap, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.gm, 20 * paths, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.gr, 20 * paths, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.gkey, 5 * paths, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.gleaf, paths, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.gflag, paths, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.gx, 5 * (size_t)mlp::kRec * paths, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.gw, 15 * paths, b);
        if (rc == AR_OK) rc = dev_alloc(ctx, &g.gp, 5 * paths, b);
        if (rc != AR_OK) {
            free_slots(g);
            return rc;
        }
        g.scap = scap;
        g.pcap = paths;
    }
    return AR_OK;
}

int guide_update(ar_ctx* ctx, uint32_t spectral, const float* lambda0, const float* cie, size_t paths,
                 hipStream_t s) {
    GuideDev& g = ctx->guide;
    SlotArgs sa{g.gr, g.gkey, g.gleaf, g.gflag, (uint32_t)g.pcap};
    hipLaunchKernelGGL(k_guide_begin, dim3(1), dim3(1), 0, s, g.counters);
    hipLaunchKernelGGL(k_guide_collect, dim3(grid_for(paths)), dim3(kBlock), 0, s, sa, (uint32_t)paths, spectral,
                       lambda0, cie, g.tkey, g.tcap - 1, g.newslots, g.counters, g.s_slot, g.s_leaf, g.s_rad,
                       g.scap);
    Pool pool{g.cap, g.tval, g.tkey, g.pkey, g.tex, g.cnt};
    hipLaunchKernelGGL(k_guide_alloc, dim3(kGridStrideBlocks), dim3(kBlock), 0, s, pool, g.newslots, g.counters);
    Accum A{g.tval, g.acc, g.accn, g.touched};
    hipLaunchKernelGGL(k_guide_accum, dim3(kGridStrideBlocks), dim3(kBlock), 0, s, A, g.s_slot, g.s_leaf, g.s_rad,
                       g.scap, g.counters);
    hipLaunchKernelGGL(k_guide_merge, dim3(kGridStrideBlocks), dim3(kBlock), 0, s, A, g.tex, g.cnt, g.counters);
    return check_hip(ctx, hipGetLastError(), "guiding update");
}

} // namespace ar

using namespace ar;

extern "C" {

int ar_guiding_configure(ar_ctx* ctx, uint64_t max_probes) {
    if (!ctx) return AR_E_INVALID;
    if (max_probes == 0 || max_probes > (1ull << 30))
        return set_error(ctx, AR_E_INVALID, "max_probes must be in [1, 2^30]");
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    free_pool(ctx->guide);
    ctx->guide_cap_req = max_probes;
    return AR_OK;
}

int ar_guiding_reset(ar_ctx* ctx) {
    if (!ctx) return AR_E_INVALID;
    (void)hipSetDevice(ctx->device);
    if (!ctx->guide.cap) return AR_OK;
    (void)hipDeviceSynchronize();
    const int rc = guide_clear(ctx, nullptr);
    if (rc != AR_OK) return rc;
    return check_hip(ctx, hipDeviceSynchronize(), "guiding reset");
}

int ar_guiding_info_get(ar_ctx* ctx, ar_guiding_info* out) {
    if (!ctx || !out) return AR_E_INVALID;
    std::memset(out, 0, sizeof *out);
    out->capacity = ctx->guide.cap ? ctx->guide.cap : ctx->guide_cap_req;
    if (!ctx->guide.cap) return AR_OK;
    (void)hipSetDevice(ctx->device);
    uint32_t c[kCounters];
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(c, ctx->guide.counters, sizeof c, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return check_hip(ctx, e, "guiding info");
    out->probes = std::min<uint64_t>(c[kCProbes], ctx->guide.cap);
    out->samples = c[kCSamples];
    out->dropped = c[kCDropped];
    out->device_bytes = ctx->guide.bytes;
    return AR_OK;
}

int ar_guiding_snapshot(ar_ctx* ctx, uint64_t* keys, float* texels, uint64_t* counts, uint64_t cap, uint64_t* n) {
    if (!ctx || !n) return AR_E_INVALID;
    ar_guiding_info info;
    int rc = ar_guiding_info_get(ctx, &info);
    if (rc != AR_OK) return rc;
    *n = info.probes;
    if (info.probes > cap) return set_error(ctx, AR_E_CAPACITY, "snapshot arrays too small");
    if (info.probes == 0) return AR_OK;
    GuideDev& g = ctx->guide;
    hipError_t e = hipSuccess;
    if (keys) e = hipMemcpy(keys, g.pkey, info.probes * sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess && counts) e = hipMemcpy(counts, g.cnt, info.probes * sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess && texels) {
        std::vector<float4> t((size_t)info.probes * kTexels);
        e = hipMemcpy(t.data(), g.tex, t.size() * sizeof(float4), hipMemcpyDeviceToHost);
        for (size_t i = 0; e == hipSuccess && i < info.probes; ++i)
            std::memcpy(texels + i * 85 * 4, &t[i * kTexels], 85 * sizeof(float4));
    }
    return check_hip(ctx, e, "guiding snapshot");
}

int ar_guiding_restore(ar_ctx* ctx, const uint64_t* keys, const float* texels, const uint64_t* counts, uint64_t n) {
    if (!ctx || (n && (!keys || !texels || !counts))) return AR_E_INVALID;
    int rc = guide_ensure(ctx, 0, true);
    if (rc != AR_OK) return rc;
    GuideDev& g = ctx->guide;
    if (n > g.cap) return set_error(ctx, AR_E_CAPACITY, "more probes than the configured capacity");
    (void)hipDeviceSynchronize();
    rc = guide_clear(ctx, nullptr);
    if (rc != AR_OK) return rc;
    if (n == 0) return check_hip(ctx, hipDeviceSynchronize(), "guiding restore");
    std::vector<float4> t((size_t)n * kTexels);
    for (size_t i = 0; i < n; ++i) {
        std::memcpy(&t[i * kTexels], texels + i * 85 * 4, 85 * sizeof(float4));
        t[i * kTexels + 85] = make_float4(0, 0, 0, 0);
    }
    hipError_t e = hipMemcpy(g.tex, t.data(), t.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g.pkey, keys, n * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g.cnt, counts, n * sizeof(uint64_t), hipMemcpyHostToDevice);
    const uint32_t nn = (uint32_t)n;
    if (e == hipSuccess) e = hipMemcpy(&g.counters[kCProbes], &nn, sizeof nn, hipMemcpyHostToDevice);
    if (e != hipSuccess) return check_hip(ctx, e, "guiding restore upload");
    hipLaunchKernelGGL(k_guide_restore, dim3(grid_for(n)), dim3(kBlock), 0, nullptr, g.tkey, g.tval, g.tcap - 1,
                       g.pkey, nn, g.newslots, g.counters);
    hipLaunchKernelGGL(k_guide_begin, dim3(1), dim3(1), 0, nullptr, g.counters);
    return check_hip(ctx, hipDeviceSynchronize(), "guiding restore");
}

} // extern "C"
