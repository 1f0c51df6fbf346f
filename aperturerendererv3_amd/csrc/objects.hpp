// Analytic primitives: oriented box, sphere and signed-distance objects (the
// reference's OBJ_CUBE / OBJ_SPHERE / OBJ_SDF SceneObjects).  Host set-up and
// device intersection, shared by the traversal and the integrator.
//
//   constructor / proxy / area   core/render/object.cuh:239-305
//   box intersector              core/object/aabb.cuh:9-89
//   sphere intersector           core/object/sphere.cuh:10-68
//   SDF intersector              core/object/sdf/sdf_base.cuh:54-167 (RaymarchLocal,
//                                IntersectSDFObject, sdfNormalUnit)
//   torus / volume distance      sdf_torus.cuh:14-18, sdf_volume.cuh:75-160
//   overlap                      aabb.cuh:91-105, sphere.cuh:70-84, sdf_base.cuh:169-180
//
// Device record, 6 x float4 (96 B) per object:
//   c  = (center, type bits)       e  = (extent, radius = max extent)
//   r0 = (worldToObject row 0, material id bits)
//   r1 = (row 1, light index bits, -1 = not a light)
//   r2 = (row 2, leaf reference of the object's own slot: a traversal started
//         there tests this object alone, the reference's solo trace inside a
//         medium, render.cuh:89-92)
//   sd = torus: (R, r, 0, 0) in unit space; volume: (first voxel bits,
//         nx | ny << 16 bits, nz bits, voxelSizeMax_) into the context's voxel pool
// An SDF object is sphere-traced between the entry and exit of its box with
// the per-ray tolerance meps (the reference's SDF_EPS(depth) =
// INTERSECT_EPS(depth) / 2); its normal is the outward distance gradient.
// The arithmetic is written to round like the CPU oracle: no contraction
// (the library is built with -ffp-contract=off), IEEE division and square
// root, dot products summed x + y then + z.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace ar {
namespace obj {

constexpr int kRecord = 6;   // float4 per object
constexpr int kSdfSteps = 512;   // SDF_STEPS

struct Basis {
    float w2o[9];   // world-to-object rows
};

// objectToWorld = Rz (Ry Rx) from Euler angles in degrees (float3x3::rotX/Y/Z
// and ::mul, core/vector.cuh:338-425); worldToObject is its transpose.
inline Basis basis_from_euler(const float rot_deg[3]) {
    const float k = 3.141592654f / 180.0f;
    const float ax = rot_deg[0] * k, ay = rot_deg[1] * k, az = rot_deg[2] * k;
    const float cx = std::cos(ax), sx = std::sin(ax), cy = std::cos(ay), sy = std::sin(ay);
    const float cz = std::cos(az), sz = std::sin(az);
    const float Rx[9] = {1, 0, 0, 0, cx, -sx, 0, sx, cx};
    const float Ry[9] = {cy, 0, sy, 0, 1, 0, -sy, 0, cy};
    const float Rz[9] = {cz, -sz, 0, sz, cz, 0, 0, 0, 1};
    auto mul = [](const float* A, const float* B, float* Cm) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) Cm[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
    };
    float ryrx[9], o2w[9];
    mul(Ry, Rx, ryrx);
    mul(Rz, ryrx, o2w);
    Basis b;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) b.w2o[3 * r + c] = o2w[3 * c + r];
    return b;
}

inline float max3h(const float v[3]) { return std::fmax(std::fmax(v[0], v[1]), v[2]); }

__host__ __device__ inline bool is_sdf(uint32_t type) { return type == AR_OBJ_SDF_TORUS || type == AR_OBJ_SDF_VOLUME; }

// UpdateProxy: box and SDF = centre +- sum_k |axis_k| extent_k, sphere = centre +- r
inline void proxy(uint32_t type, const float c[3], const float e[3], const Basis& b, float lo[3], float hi[3]) {
    float half[3];
    if (type != AR_OBJ_SPHERE) {
        for (int a = 0; a < 3; ++a)
            half[a] = std::fabs(b.w2o[a]) * e[0] + std::fabs(b.w2o[3 + a]) * e[1] + std::fabs(b.w2o[6 + a]) * e[2];
    } else {
        const float r = max3h(e);
        half[0] = half[1] = half[2] = r;
    }
    for (int a = 0; a < 3; ++a) { lo[a] = c[a] - half[a]; hi[a] = c[a] + half[a]; }
}

// GetArea: the box's (and an SDF object's box's) 8 (ex ey + ex ez + ey ez), the sphere's 4 pi ex^2
inline float area(uint32_t type, const float e[3]) {
    if (type != AR_OBJ_SPHERE) return 8.0f * (e[0] * e[1] + e[0] * e[2] + e[1] * e[2]);
    return 4.0f * 3.141592654f * e[0] * e[0];
}

// ---------------------------------------------------------------- SDF (host and device)

__host__ __device__ inline int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// SdfTorus: q = (|u.xz| - R, u.y), |q| - r
__host__ __device__ inline float sdf_torus(float4 sd, float ux, float uy, float uz) {
    const float qx = sqrtf(ux * ux + uz * uz) - sd.x;
    return sqrtf(qx * qx + uy * uy) - sd.y;
}

// SampleSDFAdaptive01 at u + 0.5: the nearest voxel when it is farther than
// voxelSizeMax_, else trilinear over the 8 neighbours (indices clamped)
__host__ __device__ inline float sdf_volume(float4 sd, const float* __restrict__ pool, float ux, float uy, float uz) {
    uint32_t off, nxy, nzb;
    __builtin_memcpy(&off, &sd.x, 4);
    __builtin_memcpy(&nxy, &sd.y, 4);
    __builtin_memcpy(&nzb, &sd.z, 4);
    const int nx = (int)(nxy & 0xFFFFu), ny = (int)(nxy >> 16), nz = (int)nzb;
    const float* v = pool + off;
    const float gx = (ux + 0.5f) * (float)(nx - 1), gy = (uy + 0.5f) * (float)(ny - 1), gz = (uz + 0.5f) * (float)(nz - 1);
    auto at = [&](int x, int y, int z) { return v[((size_t)z * (size_t)ny + (size_t)y) * (size_t)nx + (size_t)x]; };
    const int xi = clampi((int)floorf(gx + 0.5f), nx - 1);
    const int yi = clampi((int)floorf(gy + 0.5f), ny - 1);
    const int zi = clampi((int)floorf(gz + 0.5f), nz - 1);
    const float dn = at(xi, yi, zi);
    if (fabsf(dn) > sd.w) return dn;
    int x0 = (int)floorf(gx), y0 = (int)floorf(gy), z0 = (int)floorf(gz);
    const int x1 = clampi(x0 + 1, nx - 1), y1 = clampi(y0 + 1, ny - 1), z1 = clampi(z0 + 1, nz - 1);
    x0 = clampi(x0, nx - 1); y0 = clampi(y0, ny - 1); z0 = clampi(z0, nz - 1);
    const float fx = fminf(fmaxf(gx - (float)x0, 0.0f), 1.0f);
    const float fy = fminf(fmaxf(gy - (float)y0, 0.0f), 1.0f);
    const float fz = fminf(fmaxf(gz - (float)z0, 0.0f), 1.0f);
    const float c00 = at(x0, y0, z0) * (1.0f - fx) + at(x1, y0, z0) * fx;
    const float c10 = at(x0, y1, z0) * (1.0f - fx) + at(x1, y1, z0) * fx;
    const float c01 = at(x0, y0, z1) * (1.0f - fx) + at(x1, y0, z1) * fx;
    const float c11 = at(x0, y1, z1) * (1.0f - fx) + at(x1, y1, z1) * fx;
    const float c0 = c00 * (1.0f - fy) + c10 * fy;
    const float c1 = c01 * (1.0f - fy) + c11 * fy;
    return c0 * (1.0f - fz) + c1 * fz;
}

__host__ __device__ inline float sdf_unit(uint32_t type, float4 sd, const float* __restrict__ pool, float ux, float uy,
                                          float uz) {
    return type == AR_OBJ_SDF_TORUS ? sdf_torus(sd, ux, uy, uz) : sdf_volume(sd, pool, ux, uy, uz);
}

// OverlapAABBObject / OverlapSphereObject: strict containment of a point;
// OverlapSDFObject: distance <= 0 (host side, on the device record layout;
// pool = the host copy of the voxel pool)
inline bool overlap_host(const float4* r, const float p[3], const float* pool) {
    const float ox = p[0] - r[0].x, oy = p[1] - r[0].y, oz = p[2] - r[0].z;
    const float lx = r[2].x * ox + r[2].y * oy + r[2].z * oz;
    const float ly = r[3].x * ox + r[3].y * oy + r[3].z * oz;
    const float lz = r[4].x * ox + r[4].y * oy + r[4].z * oz;
    uint32_t type;
    std::memcpy(&type, &r[0].w, 4);
    const float4 e = r[1];
    if (is_sdf(type)) {
        const float ux = lx / (2.0f * std::fmax(e.x, 1e-8f)), uy = ly / (2.0f * std::fmax(e.y, 1e-8f)),
                    uz = lz / (2.0f * std::fmax(e.z, 1e-8f));
        return sdf_unit(type, r[5], pool, ux, uy, uz) <= 0.0f;
    }
    if (type == AR_OBJ_BOX)
        return -e.x < lx && -e.y < ly && -e.z < lz && e.x > lx && e.y > ly && e.z > lz;
    return lx * lx + ly * ly + lz * lz < e.w * e.w;
}

// ---------------------------------------------------------------- device

struct Local {
    float ox, oy, oz;   // origin in object space
    float dx, dy, dz;   // normalised direction in object space
};

__device__ __forceinline__ Local to_local(const float4* __restrict__ ob, float px, float py, float pz, float qx,
                                          float qy, float qz) {
    const float4 c = ob[0], r0 = ob[2], r1 = ob[3], r2 = ob[4];
    const float ox = px - c.x, oy = py - c.y, oz = pz - c.z;
    Local l;
    l.ox = r0.x * ox + r0.y * oy + r0.z * oz;
    l.oy = r1.x * ox + r1.y * oy + r1.z * oz;
    l.oz = r2.x * ox + r2.y * oy + r2.z * oz;
    float dx = r0.x * qx + r0.y * qy + r0.z * qz;
    float dy = r1.x * qx + r1.y * qy + r1.z * qz;
    float dz = r2.x * qx + r2.y * qy + r2.z * qz;
    const float l2 = dx * dx + dy * dy + dz * dz;
    if (l2 <= 1e-20f) {
        dx = dy = dz = 0.0f;
    } else {
        const float inv = 1.0f / sqrtf(l2);
        dx = dx * inv; dy = dy * inv; dz = dz * inv;
    }
    l.dx = dx; l.dy = dy; l.dz = dz;
    return l;
}

// invSafe (core/vector.cuh:151-154)
__device__ __forceinline__ float inv_safe_ref(float x) {
    return x == 0.0f ? (x >= 0.0f ? FLT_MAX : -FLT_MAX) : 1.0f / x;
}

// RaymarchLocal: sphere tracing from max(0, te) to tx in object space; a
// ray that starts inside reports the first point with 0 < d < meps, one that
// starts outside the first with -meps < d < 0 (d = unit distance * sMin).
__device__ inline bool sdf_march(uint32_t type, float4 e, float4 sd, const float* __restrict__ pool, const Local& l,
                                       float te, float tx, float meps, float& t_hit, bool& inside) {
    const float min_step = meps * 0.5f;
    const float step_scale = type == AR_OBJ_SDF_VOLUME ? 0.9f : 1.0f;
    const float smin = 2.0f * fminf(e.x, fminf(e.y, e.z));
    const float sx = 2.0f * fmaxf(e.x, 1e-8f), sy = 2.0f * fmaxf(e.y, 1e-8f), sz = 2.0f * fmaxf(e.z, 1e-8f);
    const bool start_in = sdf_unit(type, sd, pool, l.ox / sx, l.oy / sy, l.oz / sz) <= 0.0f;
    inside = start_in;
    float t = fmaxf(0.0f, te);
    for (int i = 0; i < kSdfSteps; ++i) {
        const float tl = t;
        const float x = l.ox + l.dx * t, y = l.oy + l.dy * t, z = l.oz + l.dz * t;
        const float dl = sdf_unit(type, sd, pool, x / sx, y / sy, z / sz) * smin;
        const float ad = fabsf(dl);
        if (start_in ? (dl > 0.0f && dl < meps) : (dl < 0.0f && -dl < meps)) {
            t_hit = tl;
            return true;
        }
        t = t + fmaxf(ad * step_scale, min_step);
        if (t > tx) break;
    }
    return false;
}

// First crossing with t > 0 along the normalised direction (meps: SDF march
// tolerance; pool: voxel pool of SDF volumes).
__device__ __forceinline__ bool test(const float4* __restrict__ ob, float px, float py, float pz, float qx, float qy,
                                     float qz, float meps, const float* __restrict__ pool, float& t, bool& inside) {
    const Local l = to_local(ob, px, py, pz, qx, qy, qz);
    const float4 c = ob[0], e = ob[1];
    inside = false;
    const uint32_t type = __float_as_uint(c.w);
    if (type != AR_OBJ_SPHERE) {
        const float ix = inv_safe_ref(l.dx), iy = inv_safe_ref(l.dy), iz = inv_safe_ref(l.dz);
        const float x1 = (-e.x - l.ox) * ix, x2 = (e.x - l.ox) * ix;
        const float y1 = (-e.y - l.oy) * iy, y2 = (e.y - l.oy) * iy;
        const float z1 = (-e.z - l.oz) * iz, z2 = (e.z - l.oz) * iz;
        const float te = fmaxf(fmaxf(fminf(x1, x2), fminf(y1, y2)), fminf(z1, z2));
        const float tx = fminf(fminf(fmaxf(x1, x2), fmaxf(y1, y2)), fmaxf(z1, z2));
        if (te > tx) return false;
        if (tx <= 0.0f) return false;
        if (type != AR_OBJ_BOX) return sdf_march(type, e, ob[5], pool, l, te, tx, meps, t, inside);
        inside = te < 0.0f;
        t = inside ? tx : te;
        return true;
    }
    const float r = e.w;
    const float b = l.ox * l.dx + l.oy * l.dy + l.oz * l.dz;
    const float cc = (l.ox * l.ox + l.oy * l.oy + l.oz * l.oz) - r * r;
    const float disc = b * b - cc;
    if (disc < 0.0f) return false;
    const float s = sqrtf(disc);
    float tt = -b - s;
    if (tt <= 0.0f) { tt = -b + s; inside = true; }
    if (tt <= 0.0f) return false;
    t = tt;
    return true;
}

// sdfNormalUnit + UnitNormalToLocal: central differences of step meps / sMin
// in unit space, normalised, mapped to object space by 1 / (2 extent) and
// normalised again
__device__ inline void sdf_normal(uint32_t type, float4 e, float4 sd, const float* __restrict__ pool, float lx,
                                        float ly, float lz, float meps, float& nx, float& ny, float& nz) {
    const float sx = 2.0f * fmaxf(e.x, 1e-8f), sy = 2.0f * fmaxf(e.y, 1e-8f), sz = 2.0f * fmaxf(e.z, 1e-8f);
    const float ux = lx / sx, uy = ly / sy, uz = lz / sz;
    const float h = meps / fmaxf(2.0f * fminf(e.x, fminf(e.y, e.z)), 1e-8f);
    float gx = sdf_unit(type, sd, pool, ux + h, uy, uz) - sdf_unit(type, sd, pool, ux - h, uy, uz);
    float gy = sdf_unit(type, sd, pool, ux, uy + h, uz) - sdf_unit(type, sd, pool, ux, uy - h, uz);
    float gz = sdf_unit(type, sd, pool, ux, uy, uz + h) - sdf_unit(type, sd, pool, ux, uy, uz - h);
    for (int pass = 0; pass < 2; ++pass) {
        const float l2 = gx * gx + gy * gy + gz * gz;
        if (l2 <= 1e-20f) {
            gx = gy = gz = 0.0f;
        } else {
            const float inv = 1.0f / sqrtf(l2);
            gx = gx * inv; gy = gy * inv; gz = gz * inv;
        }
        if (pass == 0) { gx = gx / sx; gy = gy / sy; gz = gz / sz; }
    }
    nx = gx; ny = gy; nz = gz;
}

// Hit point and normal (box and sphere: outward, flipped for a hit from
// inside; SDF: the outward distance gradient), world space.
__device__ __forceinline__ void surface(const float4* __restrict__ ob, float px, float py, float pz, float qx,
                                        float qy, float qz, float t, bool inside, float meps,
                                        const float* __restrict__ pool, float pos[3], float n[3]) {
    const Local l = to_local(ob, px, py, pz, qx, qy, qz);
    const float4 c = ob[0], e = ob[1], r0 = ob[2], r1 = ob[3], r2 = ob[4];
    const float lx = l.ox + l.dx * t, ly = l.oy + l.dy * t, lz = l.oz + l.dz * t;
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
    const uint32_t type = __float_as_uint(c.w);
    if (is_sdf(type)) {
        sdf_normal(type, e, ob[5], pool, lx, ly, lz, meps, nx, ny, nz);
    } else if (type == AR_OBJ_BOX) {
        const float rx = fabsf(lx) / e.x, ry = fabsf(ly) / e.y, rz = fabsf(lz) / e.z;
        if (rx >= ry && rx >= rz) nx = lx > 0.0f ? 1.0f : -1.0f;
        else if (ry >= rz) ny = ly > 0.0f ? 1.0f : -1.0f;
        else nz = lz > 0.0f ? 1.0f : -1.0f;
        if (inside) { nx = -nx; ny = -ny; nz = -nz; }
    } else {
        const float l2 = lx * lx + ly * ly + lz * lz;
        if (l2 > 1e-20f) {
            const float inv = 1.0f / sqrtf(l2);
            nx = lx * inv; ny = ly * inv; nz = lz * inv;
        }
        const float sg = inside ? -1.0f : 1.0f;
        nx = nx * sg; ny = ny * sg; nz = nz * sg;
    }
    // objectToWorld rows = columns of worldToObject
    pos[0] = (r0.x * lx + r1.x * ly + r2.x * lz) + c.x;
    pos[1] = (r0.y * lx + r1.y * ly + r2.y * lz) + c.y;
    pos[2] = (r0.z * lx + r1.z * ly + r2.z * lz) + c.z;
    float wx = r0.x * nx + r1.x * ny + r2.x * nz;
    float wy = r0.y * nx + r1.y * ny + r2.y * nz;
    float wz = r0.z * nx + r1.z * ny + r2.z * nz;
    const float w2 = wx * wx + wy * wy + wz * wz;
    if (w2 <= 1e-20f) {
        wx = wy = wz = 0.0f;
    } else {
        const float inv = 1.0f / sqrtf(w2);
        wx = wx * inv; wy = wy * inv; wz = wz * inv;
    }
    n[0] = wx; n[1] = wy; n[2] = wz;
}

} // namespace obj
} // namespace ar
