// Analytic primitives: oriented box and sphere (the reference's OBJ_CUBE /
// OBJ_SPHERE SceneObjects).  Host set-up and device intersection, shared by
// the traversal and the integrator.
//
//   constructor / proxy / area   core/render/object.cuh:239-305
//   box intersector              core/object/aabb.cuh:9-89
//   sphere intersector           core/object/sphere.cuh:10-68
//   overlap                      aabb.cuh:91-105, sphere.cuh:70-84
//
// Device record, 5 x float4 (80 B) per object:
//   c  = (center, type bits)       e  = (extent, radius = max extent)
//   r0 = (worldToObject row 0, material id bits)
//   r1 = (row 1, light index bits, -1 = not a light)
//   r2 = (row 2, leaf reference of the object's own slot: a traversal started
//         there tests this object alone, the reference's solo trace inside a
//         medium, render.cuh:89-92)
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

constexpr int kRecord = 5;   // float4 per object

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

// UpdateProxy: box = centre +- sum_k |axis_k| extent_k, sphere = centre +- r
inline void proxy(uint32_t type, const float c[3], const float e[3], const Basis& b, float lo[3], float hi[3]) {
    float half[3];
    if (type == AR_OBJ_BOX) {
        for (int a = 0; a < 3; ++a)
            half[a] = std::fabs(b.w2o[a]) * e[0] + std::fabs(b.w2o[3 + a]) * e[1] + std::fabs(b.w2o[6 + a]) * e[2];
    } else {
        const float r = max3h(e);
        half[0] = half[1] = half[2] = r;
    }
    for (int a = 0; a < 3; ++a) { lo[a] = c[a] - half[a]; hi[a] = c[a] + half[a]; }
}

// GetArea: the box's 8 (ex ey + ex ez + ey ez), the sphere's 4 pi ex^2
inline float area(uint32_t type, const float e[3]) {
    if (type == AR_OBJ_BOX) return 8.0f * (e[0] * e[1] + e[0] * e[2] + e[1] * e[2]);
    return 4.0f * 3.141592654f * e[0] * e[0];
}

// OverlapAABBObject / OverlapSphereObject: strict containment of a point
// (host side, on the device record layout)
inline bool overlap_host(const float4* r, const float p[3]) {
    const float ox = p[0] - r[0].x, oy = p[1] - r[0].y, oz = p[2] - r[0].z;
    const float lx = r[2].x * ox + r[2].y * oy + r[2].z * oz;
    const float ly = r[3].x * ox + r[3].y * oy + r[3].z * oz;
    const float lz = r[4].x * ox + r[4].y * oy + r[4].z * oz;
    uint32_t type;
    std::memcpy(&type, &r[0].w, 4);
    const float4 e = r[1];
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

// First crossing with t > 0 along the normalised direction.
__device__ __forceinline__ bool test(const float4* __restrict__ ob, float px, float py, float pz, float qx, float qy,
                                     float qz, float& t, bool& inside) {
    const Local l = to_local(ob, px, py, pz, qx, qy, qz);
    const float4 c = ob[0], e = ob[1];
    inside = false;
    if (__float_as_uint(c.w) == AR_OBJ_BOX) {
        const float ix = inv_safe_ref(l.dx), iy = inv_safe_ref(l.dy), iz = inv_safe_ref(l.dz);
        const float x1 = (-e.x - l.ox) * ix, x2 = (e.x - l.ox) * ix;
        const float y1 = (-e.y - l.oy) * iy, y2 = (e.y - l.oy) * iy;
        const float z1 = (-e.z - l.oz) * iz, z2 = (e.z - l.oz) * iz;
        const float te = fmaxf(fmaxf(fminf(x1, x2), fminf(y1, y2)), fminf(z1, z2));
        const float tx = fminf(fminf(fmaxf(x1, x2), fmaxf(y1, y2)), fmaxf(z1, z2));
        if (te > tx) return false;
        if (tx <= 0.0f) return false;
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

// Hit point and normal (outward, flipped for a hit from inside), world space.
__device__ __forceinline__ void surface(const float4* __restrict__ ob, float px, float py, float pz, float qx,
                                        float qy, float qz, float t, bool inside, float pos[3], float n[3]) {
    const Local l = to_local(ob, px, py, pz, qx, qy, qz);
    const float4 c = ob[0], e = ob[1], r0 = ob[2], r1 = ob[3], r2 = ob[4];
    const float lx = l.ox + l.dx * t, ly = l.oy + l.dy * t, lz = l.oz + l.dz * t;
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
    if (__float_as_uint(c.w) == AR_OBJ_BOX) {
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
