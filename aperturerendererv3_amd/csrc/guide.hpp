// Voxel path guiding: directional probes keyed by (0.05 voxel, icosahedral
// face), each an equal-area octahedral texture of 8x8 float4 leaves plus
// 4x4, 2x2 and 1x1 mips (86 texels with one pad).
//
//  * Clarberg 2008 equal-area octahedral map (core/guiding/octahedron.cuh:
//    13-134), Jacobian 4 pi;
//  * 20-face binning by argmax over the icosahedron face normals and the
//    voxel floor(p * 20) (core/guiding/guiding_bin.cuh:40-113);
//  * SampleOctahedron / OctahedronPDF / hemisphere fold (octahedron.cuh:
//    248-445): three-level descent weighted by dot(texel.xyz, query), then
//    uniform within the leaf;
//  * probe lookup through an open-addressing hash table of 64-bit keys (the
//    reference keeps a dense voxel x face index volume; the hash gives the same
//    probe per key with memory proportional to the probes).
// All arithmetic is FP32 in the order the reference uses.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace ar {
namespace guide {

constexpr uint64_t kNoKey = ~0ull;
constexpr uint32_t kNoProbe = 0xFFFFFFFFu;
constexpr int kTexels = 86;          // 64 + 16 + 4 + 1 + pad
constexpr int kSlots = 5;            // PATH_GUIDING_COLLECT_DEPTH
constexpr int kMaxProbeSteps = 64;   // longest linear-probing run examined
constexpr float kPiF = 3.141592654f;

#define AR_HD __host__ __device__ __forceinline__

AR_HD float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

AR_HD void uv_to_dir(float u01, float v01, float& dx, float& dy, float& dz) {
    const float u = 2.0f * u01 - 1.0f, v = 2.0f * v01 - 1.0f;
    const float a = v + u, b = v - u;
    float r, phi4, zs;
    if (v >= 0.0f) {
        if (u >= 0.0f) {
            if (a <= 1.0f) { r = a; zs = 1.0f; phi4 = r == 0.0f ? 0.0f : v / r; }
            else { r = 2.0f - a; zs = -1.0f; phi4 = (1.0f - u) / r; }
        } else {
            if (b <= 1.0f) { r = b; zs = 1.0f; phi4 = 1.0f - u / r; }
            else { r = 2.0f - b; zs = -1.0f; phi4 = 1.0f + (1.0f - v) / r; }
        }
    } else {
        if (u < 0.0f) {
            if (a >= -1.0f) { r = -a; zs = 1.0f; phi4 = 2.0f - v / r; }
            else { r = 2.0f + a; zs = -1.0f; phi4 = 2.0f + (1.0f + u) / r; }
        } else {
            if (b >= -1.0f) { r = -b; zs = 1.0f; phi4 = 3.0f + u / r; }
            else { r = 2.0f + b; zs = -1.0f; phi4 = 3.0f + (1.0f + v) / r; }
        }
    }
    if (r == 0.0f) phi4 = 0.0f;
    const float r2 = r * r;
    const float phi = phi4 * (0.5f * kPiF);
    const float st = r * sqrtf(fmaxf(0.0f, 2.0f - r2));
    dx = st * cosf(phi);
    dy = st * sinf(phi);
    dz = zs * (1.0f - r2);
}

AR_HD void dir_to_uv(float dx, float dy, float dz, float& uo, float& vo) {
    const float p = atan2f(dy, dx) * (2.0f / kPiF);
    float u, v;
    if (dz < 0.0f) {
        const float r = sqrtf(fmaxf(0.0f, 1.0f + dz));
        if (p >= 0.0f) {
            if (p <= 1.0f) { u = 1.0f - r * p; v = 2.0f - r - u; }
            else { u = r * (2.0f - p) - 1.0f; v = 2.0f - r + u; }
        } else {
            if (p >= -1.0f) { u = r * p + 1.0f; v = r - 2.0f + u; }
            else { u = r * (2.0f + p) - 1.0f; v = r - 2.0f - u; }
        }
    } else {
        const float r = sqrtf(fmaxf(0.0f, 1.0f - dz));
        if (p >= 0.0f) {
            if (p < 1.0f) { v = r * p; u = r - v; }
            else { v = r * (2.0f - p); u = v - r; }
        } else {
            if (p > -1.0f) { v = r * p; u = r + v; }
            else { v = -r * (2.0f + p); u = -(r + v); }
        }
    }
    uo = 0.5f * (u + 1.0f);
    vo = 0.5f * (v + 1.0f);
}

// icosahedron face normals in the reference's order (guiding_bin.cuh:40-66)
AR_HD void face_normal(int f, float& x, float& y, float& z) {
    const float a = 0.577350269f, b = 0.934172359f, c = 0.356822090f;
    switch (f) {
    case 0: x = -a; y = a; z = a; break;
    case 1: x = 0.0f; y = b; z = c; break;
    case 2: x = 0.0f; y = b; z = -c; break;
    case 3: x = -a; y = a; z = -a; break;
    case 4: x = -b; y = c; z = 0.0f; break;
    case 5: x = a; y = a; z = a; break;
    case 6: x = -c; y = 0.0f; z = b; break;
    case 7: x = -b; y = -c; z = 0.0f; break;
    case 8: x = -c; y = 0.0f; z = -b; break;
    case 9: x = a; y = a; z = -a; break;
    case 10: x = a; y = -a; z = a; break;
    case 11: x = 0.0f; y = -b; z = c; break;
    case 12: x = 0.0f; y = -b; z = -c; break;
    case 13: x = a; y = -a; z = -a; break;
    case 14: x = b; y = -c; z = 0.0f; break;
    case 15: x = c; y = 0.0f; z = b; break;
    case 16: x = -a; y = -a; z = a; break;
    case 17: x = -a; y = -a; z = -a; break;
    case 18: x = c; y = 0.0f; z = -b; break;
    default: x = b; y = c; z = 0.0f; break;
    }
}

AR_HD void normalize3(float& x, float& y, float& z) {
    const float l2 = x * x + y * y + z * z;
    if (l2 <= 1e-20f) { x = y = z = 0.0f; return; }
    const float inv = 1.0f / sqrtf(l2);
    x *= inv; y *= inv; z *= inv;
}

AR_HD int face20(float x, float y, float z) {
    normalize3(x, y, z);
    int best = 0;
    float bd = -1e30f;
    for (int f = 0; f < 20; ++f) {
        float nx, ny, nz;
        face_normal(f, nx, ny, nz);
        const float v = x * nx + y * ny + z * nz;
        if (v > bd) { bd = v; best = f; }
    }
    return best;
}

// (voxel, face) packed 19/19/19/5 bits; kNoKey beyond +-2^18 voxels
AR_HD uint64_t voxel_key(float px, float py, float pz, float dx, float dy, float dz) {
    const float s[3] = {floorf(px * 20.0f), floorf(py * 20.0f), floorf(pz * 20.0f)};
    uint64_t k = 0;
    for (int a = 0; a < 3; ++a) {
        if (!(s[a] >= -262144.0f && s[a] < 262144.0f)) return kNoKey;
        k |= ((uint64_t)((int64_t)s[a] + 262144) & 0x7FFFFull) << (19 * a);
    }
    return k | ((uint64_t)face20(dx, dy, dz) << 57);
}

// The key of a batched scene: (voxel, scene, face) packed 16/16/16/8/_/5
// bits (scene in bits 48..55, face in 57..61), so that scenes never share a
// probe; kNoKey beyond +-2^15 voxels or past scene 255.
AR_HD uint64_t voxel_key_scene(float px, float py, float pz, float dx, float dy, float dz, uint32_t scene) {
    if (scene > 255u) return kNoKey;
    const float s[3] = {floorf(px * 20.0f), floorf(py * 20.0f), floorf(pz * 20.0f)};
    uint64_t k = 0;
    for (int a = 0; a < 3; ++a) {
        if (!(s[a] >= -32768.0f && s[a] < 32768.0f)) return kNoKey;
        k |= ((uint64_t)((int64_t)s[a] + 32768) & 0xFFFFull) << (16 * a);
    }
    return k | ((uint64_t)scene << 48) | ((uint64_t)face20(dx, dy, dz) << 57);
}

AR_HD int leaf_of(float dx, float dy, float dz) {
    normalize3(dx, dy, dz);
    float u, v;
    dir_to_uv(dx, dy, dz, u, v);
    int ix = (int)floorf(u * 8.0f), iy = (int)floorf(v * 8.0f);
    ix = ix < 0 ? 0 : (ix > 7 ? 7 : ix);
    iy = iy < 0 ? 0 : (iy > 7 ? 7 : iy);
    return iy * 8 + ix;
}

AR_HD uint64_t mix64(uint64_t x) {
    x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27; x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

AR_HD int texel_index(int x, int y, int level) {
    const int off = level == 0 ? 0 : (level == 1 ? 64 : (level == 2 ? 80 : 84));
    return off + y * (8 >> level) + x;
}

#undef AR_HD

// ------------------------------------------------------------ device side

struct Probes {
    const uint64_t* __restrict__ tkey;   // hash table keys (kNoKey = empty)
    const uint32_t* __restrict__ tval;   // probe index of each table slot
    uint32_t tmask;                      // table size - 1
    const float4* __restrict__ tex;      // kTexels per probe
    uint32_t count;                      // probes (0: guiding has no data yet)
};

__device__ __forceinline__ uint32_t find(const Probes& P, uint64_t key) {
    if (key == kNoKey) return kNoProbe;
    uint32_t h = (uint32_t)mix64(key) & P.tmask;
    for (int s = 0; s < kMaxProbeSteps; ++s, h = (h + 1) & P.tmask) {
        const uint64_t k = P.tkey[h];
        if (k == key) return P.tval[h];
        if (k == kNoKey) return kNoProbe;
    }
    return kNoProbe;
}

__device__ __forceinline__ float wdot(float4 t, const float q[3]) { return t.x * q[0] + t.y * q[1] + t.z * q[2]; }

__device__ __forceinline__ float probe_pdf(const float4* __restrict__ t, float dx, float dy, float dz, const float q[3]) {
    float u, v;
    dir_to_uv(dx, dy, dz, u, v);
    int ix = (int)floorf(u * 8.0f), iy = (int)floorf(v * 8.0f);
    ix = ix < 0 ? 0 : (ix > 7 ? 7 : ix);
    iy = iy < 0 ? 0 : (iy > 7 ? 7 : iy);
    const float wl = wdot(t[texel_index(ix, iy, 0)], q);
    const float wa = wdot(t[texel_index(0, 0, 3)], q);
    const float inv_g = (float)(1.0 / (double)(4.0f * kPiF));
    if (!(wa > 0.0f)) return inv_g;
    return fmaxf((wl / wa) * inv_g, 1.0e-12f);
}

// ReflectHemisphereNormal (octahedron.cuh:270-273)
__device__ __forceinline__ void mirror_of(float wx, float wy, float wz, float nx, float ny, float nz, float& mx,
                                          float& my, float& mz) {
    normalize3(wx, wy, wz);
    normalize3(nx, ny, nz);
    const float d = nx * wx + ny * wy + nz * wz;
    mx = wx - 2.0f * d * nx;
    my = wy - 2.0f * d * ny;
    mz = wz - 2.0f * d * nz;
}

__device__ __forceinline__ float hemi_pdf(const float4* __restrict__ t, float dx, float dy, float dz, float nx, float ny,
                                          float nz, const float q[3]) {
    if (dx * nx + dy * ny + dz * nz <= 0.0f) return 0.0f;
    float mx, my, mz;
    mirror_of(dx, dy, dz, nx, ny, nz, mx, my, mz);
    return probe_pdf(t, dx, dy, dz, q) + probe_pdf(t, mx, my, mz, q);
}

// u[0..2]: level picks (top first), u[3..4]: position within the leaf
__device__ __forceinline__ void hemi_sample(const float4* __restrict__ t, float nx, float ny, float nz, const float q[3],
                                            const float u[5], float& dx, float& dy, float& dz, float& pdf) {
    int cx = 0, cy = 0;
    float pt = 1.0f;
#pragma unroll
    for (int level = 3, k = 0; level > 0; --level, ++k) {
        const float ws = 4.0f * wdot(t[texel_index(cx, cy, level)], q);
        const int bx = cx << 1, by = cy << 1;
        const float r = u[k];
        if (ws <= 0.0f) {
            const int pick = min(3, (int)(r * 4.0f));
            cx = bx + (pick & 1);
            cy = by + (pick >> 1);
            pt *= 0.25f;
        } else {
            // children in the order (0,0) (0,1) (1,0) (1,1) of (x, y)
            const float w0 = wdot(t[texel_index(bx, by, level - 1)], q) / ws;
            const float w1 = wdot(t[texel_index(bx, by + 1, level - 1)], q) / ws;
            const float w2 = wdot(t[texel_index(bx + 1, by, level - 1)], q) / ws;
            const float w3 = wdot(t[texel_index(bx + 1, by + 1, level - 1)], q) / ws;
            float cdf = w0;
            if (r <= cdf) { cy = by; cx = bx; pt *= w0; }
            else {
                cdf += w1;
                if (r <= cdf) { cx = bx; cy = by + 1; pt *= w1; }
                else {
                    cdf += w2;
                    if (r <= cdf) { cx = bx + 1; cy = by; pt *= w2; }
                    else { cx = bx + 1; cy = by + 1; pt *= w3; }
                }
            }
        }
    }
    const float su = ((float)cx + (1.0f - u[3])) / 8.0f, sv = ((float)cy + (1.0f - u[4])) / 8.0f;
    uv_to_dir(clamp01(su), clamp01(sv), dx, dy, dz);
    pdf = fmaxf(pt * 64.0f / (4.0f * kPiF), 1.0e-12f);
    float mx, my, mz;
    mirror_of(dx, dy, dz, nx, ny, nz, mx, my, mz);
    pdf += probe_pdf(t, mx, my, mz, q);
    if (dx * nx + dy * ny + dz * nz < 0.0f) { dx = mx; dy = my; dz = mz; }
}

} // namespace guide
} // namespace ar
