// Wavefront path integrator for gfx950.
//
// Replaces the reference's megakernel RenderCamera -> CalculateRadiance
// (core/scene.cu:135-189, core/render/render.cuh:460-998) with one kernel per
// stage and compacted queues between them:
//   camera -> [ trace closest -> shade -> trace any (shadow) -> resolve ] x depth
//   -> accumulate.
// Queues are structure-of-arrays; shade appends surviving paths and shadow
// segments with one atomic per wave (ballot + popcount prefix), so every
// traversal launch sees a dense ray array.  Ray counts live in device memory
// and the traversal kernels read them, so the host loop issues a fixed
// sequence of launches with no synchronisation (it can be captured in a
// HIP graph).
//
// Per-path arithmetic follows the reference step by step (see the citations
// on each stage) in the same floating-point order as the CPU oracle, so at
// equal seed the GPU and the oracle trace the same paths except where a
// transcendental (sinf/cosf) differs in its last bit.
#include "ar_internal.hpp"
#include "smlp.hpp"
#include "spectral.hpp"
#include "path_common.hpp"
#include "guide.hpp"
#include "mlp.hpp"
#include "objects.hpp"
#include "denoise.hpp"

#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace ar {
namespace {

constexpr uint32_t kCameraBounce = 0xFFFFFFFFu;
constexpr int kBlock = 256;        // default threads per block of the wavefront kernels
constexpr int kMaxBlock = 512;     // largest block the wavefront kernels accept
constexpr uint32_t kMetaNee = 1u << 16;
constexpr uint32_t kAccumFloats = 9;   // per pixel: sums of all, odd-index and even-index samples
constexpr size_t kFirstHitFloats = 10; // per path: model bits, albedo[3], normal[3], distance, metallic, roughness
constexpr uint32_t kMetaHeroOnly = 1u << 17;   // secondary hero wavelengths dropped at a dispersive interface

// ---------------------------------------------------------------- helpers
struct F3 { float x, y, z; };
__device__ __forceinline__ F3 f3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 mul(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 scale(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(F3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ F3 normalize(F3 v) {
    const float l2 = dot(v, v);
    if (l2 <= 1e-20f) return {0.0f, 0.0f, 0.0f};
    const float inv = 1.0f / sqrtf(l2);
    return scale(v, inv);
}
__device__ __forceinline__ float maxr(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float minr(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float clampr(float x, float a, float b) { return minr(maxr(x, a), b); }
__device__ __forceinline__ float lum(F3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }
__device__ __forceinline__ float max3(F3 c) { return maxr(c.x, maxr(c.y, c.z)); }
__device__ __forceinline__ float fixnan(float x) { return (isnan(x) || isinf(x)) ? 0.0f : x; }
__device__ __forceinline__ F3 xyz(float4 v) { return {v.x, v.y, v.z}; }

constexpr float kPi = 3.14159265358979323846f;

// Philox4x32-10 (Salmon et al., SC'11)
__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                       uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ void uniforms(uint64_t seed, uint32_t pix, uint32_t smp, uint32_t bounce, uint32_t group,
                                         float* u) {
    uint32_t x[4];
    philox(pix, smp, bounce, group, (uint32_t)seed, (uint32_t)(seed >> 32), x);
#pragma unroll
    for (int k = 0; k < 4; ++k) u[k] = (float)(x[k] >> 8) * 0x1p-24f;
}

// ---------------------------------------------------------------- sampling
// BuildTangentBasisRandom (core/vector.cuh:284-294)
__device__ __forceinline__ void tangent_basis_random(F3 n, float e, F3& t, F3& b) {
    const F3 a = (fabsf(n.z) < 0.999f) ? f3(0, 0, 1) : f3(1, 0, 0);
    const F3 t0 = normalize(cross(a, n));
    const F3 b0 = cross(n, t0);
    const float ang = e * kPi * 2.0f;
    const float s = sinf(ang), c = cosf(ang);
    t = add(scale(t0, c), scale(b0, s));
    b = add(scale(t0, -s), scale(b0, c));
}

// ConcentricSampleDisk (core/render/sample.cuh:63-86)
__device__ __forceinline__ void concentric_disk(float u0, float u1, float& dx, float& dy) {
    const float sx = 2.0f * u0 - 1.0f, sy = 2.0f * u1 - 1.0f;
    if (sx == 0.0f && sy == 0.0f) { dx = 0.0f; dy = 0.0f; return; }
    float r, phi;
    if (fabsf(sx) > fabsf(sy)) {
        r = sx;
        phi = (kPi / 4.0f) * (sy / maxr(fabsf(sx), 1e-20f));
    } else {
        r = sy;
        phi = (kPi / 2.0f) - (kPi / 4.0f) * (sx / maxr(fabsf(sy), 1e-20f));
    }
    dx = r * cosf(phi);
    dy = r * sinf(phi);
}

// SampleCosineWeightedHemisphere (core/render/sample.cuh:88-108)
__device__ __forceinline__ F3 sample_cosine(float e0, float e1, float e2, F3 n, float& pdf) {
    F3 t, b;
    tangent_basis_random(n, e2, t, b);
    float x, y;
    concentric_disk(e0, e1, x, y);
    const float z = sqrtf(maxr(0.0f, 1.0f - x * x - y * y));
    const F3 dir = add(add(scale(t, x), scale(b, y)), scale(n, z));
    pdf = z / kPi;
    return normalize(dir);
}

// GGX_D (core/render/brdf.cuh:82-87)
__device__ __forceinline__ float ggx_d(float alpha, float noh) {
    const float a2 = alpha * alpha;
    const float d = noh * noh * (a2 - 1.0f) + 1.0f;
    return a2 / maxr(kPi * d * d, 1e-12f);
}

// SampleGGXBoundedVNDF (core/render/sample.cuh:179-224)
__device__ __forceinline__ F3 sample_vndf(F3 v, F3 n, float alpha, float r0, float r1, float r2, float& pdf) {
    F3 X, Y;
    tangent_basis_random(n, r2, X, Y);
    const float xov = dot(X, v), yov = dot(Y, v), nov = dot(n, v);
    const float nov2 = minr(1.0f, nov * nov);
    const float a2 = alpha * alpha;
    const float inv = 1.0f / sqrtf(a2 + (1.0f - a2) * nov2);
    const F3 vstd = f3(xov * alpha * inv, yov * alpha * inv, nov * inv);
    const float phi = 2.0f * kPi * r0;
    const float s = 1.0f + sqrtf(1.0f - nov2);
    const float s2 = s * s;
    const float k = (1.0f - a2) * s2 / (s2 + a2 * nov2);
    const float b = nov > 0.0f ? k * vstd.z : vstd.z;
    const float z = (1.0f - r1) * (1.0f + b) - b;
    const float st = sqrtf(clampr(1.0f - z * z, 0.0f, 1.0f));
    const F3 mstd = f3(vstd.x + st * cosf(phi), vstd.y + st * sinf(phi), vstd.z + z);
    const F3 h = normalize(f3(mstd.x * alpha, mstd.y * alpha, mstd.z));
    const F3 lv = f3(xov, yov, nov);
    const float d2 = 2.0f * dot(lv, h);
    const F3 ll = f3(d2 * h.x - xov, d2 * h.y - yov, d2 * h.z - nov);
    const F3 dir = add(add(scale(X, ll.x), scale(Y, ll.y)), scale(n, ll.z));
    const float ndf = ggx_d(alpha, h.z);
    const float len2 = (1.0f - nov2) * a2;
    const float t = sqrtf(len2 + nov2);
    pdf = ndf / (2.0f * (k * nov + t));
    return dir;
}

// GGXBoundedVNDF_PDF (core/render/sample.cuh:226-248)
__device__ __forceinline__ float vndf_pdf(F3 v, F3 n, F3 l, float alpha) {
    const float nov = dot(n, v), nol = dot(n, l);
    if (nov <= 0.0f || nol <= 0.0f) return 0.0f;
    const F3 h = normalize(add(v, l));
    const float nov2 = minr(1.0f, nov * nov);
    const float a2 = alpha * alpha;
    const float s = 1.0f + sqrtf(1.0f - nov2);
    const float s2 = s * s;
    const float k = (1.0f - a2) * s2 / (s2 + a2 * nov2);
    const float noh = maxr(0.0f, dot(n, h));
    const float ndf = ggx_d(alpha, noh);
    const float len2 = (1.0f - nov2) * a2;
    const float t = sqrtf(len2 + nov2);
    return maxr(ndf / (2.0f * (k * nov + t)), 1e-12f);
}

// ---------------------------------------------------------------- BSDF
struct Mat {
    F3 albedo, emission;
    float specular, metallic, roughness, escale, ascale;
    uint32_t model;
    float ca[4], ce[4];
    float ior, sigma_a, abbe, tilt;
};

__device__ __forceinline__ Mat load_mat(const GpuMaterial* __restrict__ mats, uint32_t id) {
    const float4* m4 = reinterpret_cast<const float4*>(mats + id);
    const float4 a = m4[0], b = m4[1], c = m4[2], d = m4[3], e = m4[4], g = m4[5];
    Mat m;
    m.albedo = f3(a.x, a.y, a.z);
    m.specular = a.w;
    m.emission = f3(b.x, b.y, b.z);
    m.metallic = b.w;
    m.roughness = c.x;
    m.model = __float_as_uint(c.y);
    m.escale = c.z;
    m.ascale = c.w;
    m.ca[0] = d.x; m.ca[1] = d.y; m.ca[2] = d.z; m.ca[3] = d.w;
    m.ce[0] = e.x; m.ce[1] = e.y; m.ce[2] = e.z; m.ce[3] = e.w;
    m.ior = g.x; m.sigma_a = g.y; m.abbe = g.z; m.tilt = g.w;
    return m;
}

// luminance and largest channel of a throughput (RGB: Rec.709 luminance;
// one wavelength: the value itself; hero wavelengths: mean and max)
__device__ __forceinline__ float lum4(const S4& c, uint32_t mode) {
    if (mode == AR_SPECTRAL_RGB) return c.v[0] * 0.2126f + c.v[1] * 0.7152f + c.v[2] * 0.0722f;
    if (mode == AR_SPECTRAL_SINGLE) return c.v[0];
    return (c.v[0] + c.v[1] + c.v[2] + c.v[3]) * 0.25f;
}
__device__ __forceinline__ float max4(const S4& c, uint32_t mode) {
    if (mode == AR_SPECTRAL_RGB) return maxr(c.v[0], maxr(c.v[1], c.v[2]));
    if (mode == AR_SPECTRAL_SINGLE) return c.v[0];
    return maxr(maxr(c.v[0], c.v[1]), maxr(c.v[2], c.v[3]));
}

__device__ __forceinline__ float f0_ue(float spec) { return 0.08f * clampr(spec, 0.0f, 1.0f); }

__device__ __forceinline__ float g_correlated(float alpha, float nol, float nov) {
    const float a2 = alpha * alpha;
    const float gv = nol * sqrtf(a2 + (1.0f - a2) * nov * nov);
    const float gl = nov * sqrtf(a2 + (1.0f - a2) * nol * nol);
    const float den = gv + gl;
    return den > 0.0f ? (2.0f * nol * nov) / den : 0.0f;
}

__device__ __forceinline__ float alpha_of(const Mat& m) {
    const float r = clampr(m.roughness, 0.01f, 1.0f);
    return r * r;
}

// lobe selection probability of the GGX mix: clamp(lum(F0), 0.1, 0.9) in RGB,
// clamp(F0(lambda), 0.1, 0.9) of the first wavelength otherwise
__device__ __forceinline__ float pspec_of(const Mat& m, const S4& alb, uint32_t mode) {
    const float f = f0_ue(m.specular);
    if (mode == AR_SPECTRAL_RGB) {
        const F3 f0 = f3(f + (m.albedo.x - f) * m.metallic, f + (m.albedo.y - f) * m.metallic,
                         f + (m.albedo.z - f) * m.metallic);
        return clampr(lum(f0), 0.1f, 0.9f);
    }
    return clampr(f + (alb.v[0] - f) * m.metallic, 0.1f, 0.9f);
}

// EvaluateBSDFWithCos (core/render/material.cuh:47-165), rd toward the surface
__device__ __forceinline__ S4 eval_bsdf(const Mat& m, const S4& alb, F3 rd, F3 l, F3 n) {
    const float nol = maxr(0.0f, dot(n, l));
    if (m.model == AR_MODEL_LAMBERT) return s4scale(alb, nol / kPi);
    const F3 v = normalize(scale(rd, -1.0f));
    const float nov = maxr(0.0f, dot(n, v));
    if (nov <= 0.0f || nol <= 0.0f) return s4(0, 0, 0, 0);
    const float alpha = alpha_of(m);
    const float f = f0_ue(m.specular);
    const float kd = 1.0f - m.metallic;
    const F3 h = normalize(add(v, l));
    const float noh = maxr(0.0f, dot(n, h));
    const float loh = maxr(0.0f, dot(l, h));
    const float c = clampr(1.0f - loh, 0.0f, 1.0f);
    const float c5 = c * c * c * c * c;
    const float D = ggx_d(alpha, noh);
    const float G = g_correlated(alpha, nol, nov);
    const float sden = maxr(4.0f * nov, 1e-7f);
    const float dcos = nol / kPi;
    S4 r;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float f0 = f + (alb.v[k] - f) * m.metallic;
        const float F = f0 + (1.0f - f0) * c5;
        r.v[k] = F * D * G / sden + kd * (1.0f - F) * alb.v[k] * dcos;
    }
    return r;
}

// BSDF_PDF (core/render/material.cuh:316-371)
__device__ __forceinline__ float bsdf_pdf(const Mat& m, float pspec, F3 rd, F3 l, F3 n) {
    const float nol = maxr(0.0f, dot(n, l));
    if (m.model == AR_MODEL_LAMBERT) return nol > 0.0f ? nol * (1.0f / kPi) : 0.0f;
    const F3 v = scale(rd, -1.0f);
    const float nov = maxr(0.0f, dot(n, v));
    if (nov <= 0.0f || nol <= 0.0f) return 0.0f;
    const float alpha = alpha_of(m);
    const float ps = vndf_pdf(v, n, l, alpha);
    const float pd = nol * (1.0f / kPi);
    return maxr(0.0f, pspec * ps + (1.0f - pspec) * pd);
}

// SampleBSDF (core/render/material.cuh:167-314); false ends the path
__device__ __forceinline__ bool sample_bsdf(const Mat& m, float pspec, F3 rd, F3 n, const float* u, F3& next,
                                            float& pdf) {
    if (m.model == AR_MODEL_LAMBERT) {
        next = sample_cosine(u[1], u[2], u[3], n, pdf);
        return true;
    }
    const F3 v = scale(rd, -1.0f);
    const float nov = maxr(0.0f, dot(n, v));
    if (nov <= 0.0f) return false;
    const float alpha = alpha_of(m);
    F3 l;
    float ps, pd;
    if (u[0] < pspec) {
        l = sample_vndf(v, n, alpha, u[1], u[2], u[3], ps);
        const float nol = maxr(0.0f, dot(n, l));
        if (nol <= 0.0f) return false;
        pd = nol / kPi;
    } else {
        l = sample_cosine(u[1], u[2], u[3], n, pd);
        ps = vndf_pdf(v, n, l, alpha);
    }
    next = l;
    pdf = pspec * ps + (1.0f - pspec) * pd;
    return true;
}

__device__ __forceinline__ float power_heuristic(float pa, float pb) {
    const float a = pa * pa, b = pb * pb;
    return a / maxr(1e-24f, a + b);
}

// ---------------------------------------------------------------- box / sphere lights
// Facet f of a box light (nee.cuh:153-183): outward normal +-normalised
// worldToObject row f/2, centre c +- that axis times the extent, tangent and
// bitangent axes and half sides.
struct Facet {
    F3 n, c, t, b;
    float s1, s2;
};

__device__ __forceinline__ Facet box_facet(const float4* __restrict__ ob, int f) {
    const float4 c4 = ob[0], e = ob[1], r0 = ob[2], r1 = ob[3], r2 = ob[4];
    const F3 ux = normalize(f3(r0.x, r0.y, r0.z)), uy = normalize(f3(r1.x, r1.y, r1.z)),
             uz = normalize(f3(r2.x, r2.y, r2.z));
    const int a = f >> 1;
    const F3 ax = a == 0 ? ux : (a == 1 ? uy : uz);
    const float ea = a == 0 ? e.x : (a == 1 ? e.y : e.z);
    const F3 c = xyz(c4);
    Facet r;
    r.n = (f & 1) ? scale(ax, -1.0f) : ax;
    r.c = (f & 1) ? sub(c, scale(ax, ea)) : add(c, scale(ax, ea));
    r.t = a == 0 ? uy : ux;
    r.b = a == 2 ? uy : uz;
    r.s1 = a == 0 ? e.y : e.x;
    r.s2 = a == 2 ? e.y : e.z;
    return r;
}

// NEEFacingWeight_Box (nee.cuh:55-61)
__device__ __forceinline__ float facing_weight(const Facet& f, F3 x) {
    const float area = 4.0f * f.s1 * f.s2;
    return area * maxr(0.0f, dot(f.n, normalize(sub(x, f.c))));
}

// AreaToSolidAnglePDF (nee.cuh:35-46)
__device__ __forceinline__ float area_to_solid(F3 x, F3 y, F3 ny, float pa_total) {
    const F3 s2l = sub(y, x);
    const float dist2 = maxr(dot(s2l, s2l), 1e-20f);
    const F3 dir = scale(s2l, 1.0f / sqrtf(dist2));
    const float cosy = maxr(0.0f, dot(ny, scale(dir, -1.0f)));
    if (cosy <= 0.0f) return 0.0f;
    return pa_total * (dist2 / cosy);
}

// NEEPDFOmegaFromHit_Box / _Sphere (nee.cuh:63-151, 253-271): solid-angle
// density with which NEE from x samples the light point y (normal ny),
// select probability included.
__device__ __forceinline__ float obj_light_pdf(const float4* __restrict__ ob, float select, F3 x, F3 y, F3 ny) {
    if (__float_as_uint(ob[0].w) == AR_OBJ_BOX) {
        int hit = -1;
        for (int f = 0; f < 6; ++f)
            if (dot(box_facet(ob, f).n, ny) > 0.999f) { hit = f; break; }
        if (hit < 0) {
            float best = 1e30f;
            for (int f = 0; f < 6; ++f) {
                const Facet fc = box_facet(ob, f);
                const float dd = fabsf(dot(fc.n, sub(y, fc.c)));
                if (dd < best) { best = dd; hit = f; }
            }
        }
        float sum = 0.0f, whit = 0.0f, ahit = 0.0f;
        for (int f = 0; f < 6; ++f) {
            const Facet fc = box_facet(ob, f);
            const float w = facing_weight(fc, x);
            sum += w;
            if (f == hit) { whit = w; ahit = 4.0f * fc.s1 * fc.s2; }
        }
        if (sum <= 0.0f || whit <= 0.0f) return 0.0f;
        const float pa = select * (1.0f / ahit) * (whit / sum);
        return area_to_solid(x, y, ny, pa);
    }
    const F3 c = xyz(ob[0]);
    const float r = ob[1].w;
    const float dist = length(sub(x, c));
    if (dist <= r) return 0.0f;
    const float cosc = clampr(r / dist, 0.0f, 1.0f);
    const float cap = 2.0f * kPi * r * r * (1.0f - cosc);
    if (cap <= 1e-20f) return 0.0f;
    return area_to_solid(x, y, ny, select / cap);
}

// NEESampleBoxAreaFacing / NEESampleSphereCapArea (nee.cuh:153-251,
// 273-310): light point y, its normal, its area density (facet or cap,
// without the select probability).  False when nothing faces x.
__device__ __forceinline__ bool obj_light_sample(const float4* __restrict__ ob, F3 x, float ub, float u0, float u1,
                                                 F3& y, F3& ny, float& pa) {
    if (__float_as_uint(ob[0].w) == AR_OBJ_BOX) {
        float sum = 0.0f;
        for (int f = 0; f < 6; ++f) sum += facing_weight(box_facet(ob, f), x);
        // the valid facets in order; the first whose running sum reaches ub * sum
        int pick = -1, first = -1;
        float wpick = 0.0f, acc = 0.0f;
        const float r = ub * sum;
        for (int f = 0; f < 6; ++f) {
            const float w = facing_weight(box_facet(ob, f), x);
            if (w > 0.0f) {
                if (first < 0) { first = f; wpick = w; }
                acc += w;
                if (pick < 0 && r <= acc) { pick = f; wpick = w; }
            }
        }
        if (first < 0) return false;
        if (pick < 0) {
            pick = first;
            wpick = facing_weight(box_facet(ob, first), x);
        }
        const Facet fc = box_facet(ob, pick);
        const float a1 = 2.0f * u0 - 1.0f, a2 = 2.0f * u1 - 1.0f;
        y = add(add(fc.c, scale(fc.t, a1 * fc.s1)), scale(fc.b, a2 * fc.s2));
        ny = fc.n;
        pa = (wpick / sum) * (1.0f / (4.0f * fc.s1 * fc.s2));
        return true;
    }
    const F3 c = xyz(ob[0]);
    const float r = ob[1].w;
    const F3 pc = sub(x, c);
    const float dist = length(pc);
    if (dist <= r) return false;
    const F3 axis = scale(pc, 1.0f / dist);
    const float cosc = clampr(r / dist, 0.0f, 1.0f);
    const float cap = 2.0f * kPi * r * r * (1.0f - cosc);
    const float cost = cosc * (1.0f - u0) + u0;
    const float sint = sqrtf(maxr(0.0f, 1.0f - cost * cost));
    const float phi = 2.0f * kPi * u1;
    const F3 a = (fabsf(axis.z) < 0.999f) ? f3(0, 0, 1) : f3(1, 0, 0);
    const F3 t = normalize(cross(a, axis));
    const F3 b = cross(axis, t);
    const F3 nn = add(add(scale(t, cosf(phi) * sint), scale(b, sinf(phi) * sint)), scale(axis, cost));
    y = add(c, scale(nn, r));
    ny = nn;
    pa = 1.0f / cap;
    return true;
}

// ---------------------------------------------------------------- glass
// IORFromAbbeNM and SigmaAFromTiltNM (core/spectrum/spectrum.cuh:16-54)
__device__ __forceinline__ float ior_at(float nd, float vd, float lambda) {
    if (!(lambda > 0.0f)) return nd;
    const float invf = 1.0f / 486.1327f, invc = 1.0f / 656.2725f, invd = 1.0f / 587.5618f;
    const float k = vd > 0.0f ? (nd - 1.0f) / (vd * (invf - invc)) : 0.0f;
    return nd + k * (1.0f / lambda - invd);
}
__device__ __forceinline__ float sigma_at(float sa, float lambda, float tilt) {
    const float invf = 1.0f / 486.1327f, invc = 1.0f / 656.2725f, invd = 1.0f / 587.5618f;
    const float u = 1.0f / maxr(lambda, 1e-6f);
    const float du = (u - invd) / (invf - invc);
    return maxr(0.0f, sa * (1.0f + tilt * du));
}
// FresnelDielectricExact (brdf.cuh:25-31), eta = n_in / n_out
__device__ __forceinline__ float fresnel_dielectric(float cosi, float cost, float eta) {
    const float inve = 1.0f / eta;
    const float rpar = (inve * cosi - cost) / (inve * cosi + cost);
    const float rperp = (cosi - inve * cost) / (cosi + inve * cost);
    return 0.5f * (rpar * rpar + rperp * rperp);
}
__device__ __forceinline__ F3 reflect3(F3 v, F3 n) { return sub(v, scale(n, 2.0f * dot(v, n))); }
// refract (core/vector.cuh:264-272)
__device__ __forceinline__ F3 refract3(F3 in, F3 normal, float eta) {
    const F3 n = dot(normal, in) > 0.0f ? scale(normal, -1.0f) : normal;
    const float cosi = -dot(n, in);
    const float k = 1.0f - (eta * eta) * maxr(0.0f, 1.0f - cosi * cosi);
    const float cost = sqrtf(maxr(0.0f, k));
    return add(scale(in, eta), scale(n, eta * cosi - cost));
}

// SkyLighting (core/render/render.cuh:58-73)
__device__ __forceinline__ F3 sky_gradient(F3 d) {
    const F3 white = f3(1, 1, 1), blue = f3(0.05f, 0.45f, 1.0f);
    const F3 down = scale(f3(0.2f, 0.6f, 1.0f), 0.125f);
    const float t1 = powf(maxr(d.y, 0.0f), 0.75f);
    const F3 a = add(blue, scale(sub(white, blue), t1));
    const float t2 = minr(maxr(-d.y * 0.5f + 0.5f, 0.0f), 1.0f);
    const F3 e = add(a, scale(sub(down, a), t2));
    return scale(e, 0.5f);
}

// One scene of a render (a batch entry): image, samples, camera basis
// (core/scene.cu:145-164), BVH root and light range in the merged arrays.
struct SceneView {
    uint32_t width, height, npix, spp;
    uint32_t sample_offset;      // index of the scene's first sample
    uint32_t out_base;           // first pixel of the scene in the accumulation buffer
    uint32_t root;               // BVH root of the scene
    uint32_t light_off, light_cnt;
    uint32_t medium0;            // object containing the camera + 1 (0: none)
    uint32_t root0;              // root of the camera rays (the medium's own leaf inside one)
    uint32_t prog;               // progressive: fold the mean into out, the running mean of sample_offset samples
    float* odd;                  // progressive running means of the odd / even sample indices (or null)
    float* even;
    float* gbuf;                 // progressive: denoiser G-buffer records (AR_GBUFFER_FLOATS per pixel) or null
    float* stats;                // progressive: denoiser sample statistics (AR_STATS_FLOATS per pixel) or null
    float cam_pos[3];
    float fwd[3], right[3], up[3];
    float thx, thy;
    float* out;                  // output image, width*height*3 floats
    float box_lo[3], box_inv[3]; // the scene's box (learned-guiding inputs are normalised to it)
};

// Paths [base, base + count * npix) of a pass: samples sample0 .. sample0 +
// count - 1 of one scene, sample-major; pixels [pixbase, pixbase + npix) of
// the pass's accumulation grid.
struct Segment {
    uint32_t scene, sample0, count, base, pixbase, pad[3];
};

struct RenderParams {
    uint64_t seed;
    uint32_t capacity;           // paths in this pass
    uint32_t npixels;            // pixels of this pass's segments
    uint32_t max_depth;
    uint32_t rr_min_depth;
    uint32_t nee;
    uint32_t sky;
    uint32_t spectral;           // AR_SPECTRAL_*
    uint32_t output;             // AR_OUTPUT_*
    uint32_t guiding;            // AR_GUIDING_*
    uint32_t batched;            // a batch context: voxel-guiding keys carry the scene index
    float alpha;                 // guide/BSDF mixture probability of this frame
    float env[3];
    float env_c[3], env_scale;   // sigmoid coefficients of the constant sky (spectral modes)
    const SceneView* views;
    const Segment* segs;
    uint32_t nseg;
};

// segment of path p (segments sorted by base)
__device__ __forceinline__ uint32_t segment_of(const Segment* segs, uint32_t nseg, uint32_t p) {
    uint32_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (segs[mid].base <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}

struct QueueRays {
    float *ox, *oy, *oz, *tmin, *dx, *dy, *dz, *tmax;
    uint32_t* root;              // BVH root of the ray's scene
    float* meps;                 // SDF march tolerance (shadow queue of scenes with SDF objects; else null)
};

struct SceneArgs {
    const float4* __restrict__ prims;
    const float4* __restrict__ objs;   // analytic objects (obj::kRecord float4 each), prims ntri ..
    uint32_t ntri;
    const GpuMaterial* __restrict__ mats;
    const int32_t* __restrict__ light_tri;
    const float* __restrict__ light_area;
    const float* __restrict__ light_prob;
    const float* __restrict__ light_cdf;
    const float* __restrict__ cie;
    const float* __restrict__ d65;
    const float* __restrict__ vol;     // SDF voxel pool
    uint32_t sdf;                      // the scene has SDF objects
    uint32_t basis;                    // spectral uplift: 0 sigmoid coefficients, 1 basis network
    const float* __restrict__ lam_lut; // basis network tables: wavelength [lam_res][4], RGB [rgb_res^3][4]
    uint32_t lam_res;
    const float* __restrict__ rgb_lut;
    uint32_t rgb_res;
};

// Reflectance of an uplift record at wavelength l: the sigmoid of the
// coefficients, or (network basis) SpectralInteract of the material's RGB
// basis with the wavelength table's basis at l (spectrum_lut.cuh:183-235).
__device__ __forceinline__ float uplift_at(const SceneArgs& sc, const float c[4], float l) {
    if (!sc.basis) return spec::sigmoid_poly(c[0], c[1], c[2], l);
    return sb::dot_clamp(make_float4(c[0], c[1], c[2], c[3]), sb::query_lambda(sc.lam_lut, (int)sc.lam_res, l));
}

// Reflectance of an RGB albedo in this path's channels.
__device__ __forceinline__ S4 albedo_of(const Mat& m, const Lam& L, const SceneArgs& sc) {
    if (L.mode == AR_SPECTRAL_RGB) return s4(m.albedo.x, m.albedo.y, m.albedo.z, 0.0f);
    S4 r;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        r.v[k] = (L.mode == AR_SPECTRAL_HERO4 || k == 0) ? uplift_at(sc, m.ca, L.l[k]) * m.ascale : 0.0f;
    return r;
}

// Emitted radiance (Interact * D65Norm of render.cuh:657-700 in spectral modes)
__device__ __forceinline__ S4 emission_of(const Mat& m, const Lam& L, const SceneArgs& sc) {
    if (L.mode == AR_SPECTRAL_RGB) return s4(m.emission.x, m.emission.y, m.emission.z, 0.0f);
    S4 r;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        r.v[k] = (L.mode == AR_SPECTRAL_HERO4 || k == 0)
                     ? uplift_at(sc, m.ce, L.l[k]) * m.escale * spec::d65_norm(sc.d65, L.l[k])
                     : 0.0f;
    return r;
}


struct PathArgs {
    float* mask0; float* mask1; float* mask2; float* mask3;
    float* rad0; float* rad1; float* rad2; float* rad3;
    float* lp0; float* lp1; float* lp2;
    float* last_pdf;
    uint32_t* meta;
    float* lambda0;
    uint32_t* pix;               // pixel of the path
    uint32_t* smp;               // sample index of the path
    uint32_t* scn;               // scene of the path
    uint32_t* med;               // medium: 0 none, else object index + 1
    float* mp0; float* mp1; float* mp2;   // where the path entered its medium
    float* fh;                   // first vertex of each path for the denoiser (kFirstHit arrays of fh_stride) or null
    uint32_t fh_stride;
};

// Guiding inputs of the integrator.  Voxel guiding: the probes (read-only
// during a frame) and the per-path collection slots (CollectRadiance,
// render.cuh:345-410): throughput since the slot's vertex (gm), collected
// radiance (gr), probe key and leaf texel of the slot's sampled direction.
// Learned guiding: a slot is opened at a non-delta vertex (inputs, sampled
// direction, its pdf) and completed when the path's next vertex is shaded,
// with the path's throughput there (gm) and its radiance before anything
// that vertex adds, less the vertex's own emission without MIS weight (gr);
// at the end of the frame (radiance - gr) / gm is the radiance that arrived
// along the slot's direction: emission, NEE and every later bounce.
struct GuideArgs {
    guide::Probes pr;
    const uint32_t* probe_count;   // device counter; 0 = no probes yet
    float* gm;                     // [5*4][stride]
    float* gr;                     // [5*4][stride]
    uint64_t* gkey;                // [5][stride]
    uint32_t* gleaf;               // 6 bits per slot
    uint32_t* gflag;               // bit i: slot i received radiance (voxel) / is complete (learned);
                                   // bit 8 + i (learned): slot i awaits its next vertex
    uint32_t stride;
    // learned guiding
    const float* dec;              // mixture per queue slot [q][kPacked]: per lobe (pi, kappa, octahedral mu)
    const float* feat;             // network inputs per queue slot [q][kRec] (k_mlp_features)
    float* gx;                     // slot network inputs [5*kRec][stride] (features, scene index)
    float* gw;                     // slot direction [5*3][stride]
    float* gp;                     // slot direction pdf [5][stride]
};

__device__ __forceinline__ bool s4zero(const S4& c, uint32_t mode) {
    if (mode == AR_SPECTRAL_RGB) return c.v[0] == 0.0f && c.v[1] == 0.0f && c.v[2] == 0.0f;
    if (mode == AR_SPECTRAL_SINGLE) return c.v[0] == 0.0f;
    return c.v[0] == 0.0f && c.v[1] == 0.0f && c.v[2] == 0.0f && c.v[3] == 0.0f;
}

__device__ __forceinline__ S4 gload(const float* a, uint32_t stride, int i, uint32_t p) {
    const float* b = a + (size_t)(4 * i) * stride + p;
    return s4(b[0], b[stride], b[2 * (size_t)stride], b[3 * (size_t)stride]);
}
__device__ __forceinline__ void gstore(float* a, uint32_t stride, int i, uint32_t p, const S4& v) {
    float* b = a + (size_t)(4 * i) * stride + p;
    b[0] = v.v[0]; b[stride] = v.v[1]; b[2 * (size_t)stride] = v.v[2]; b[3 * (size_t)stride] = v.v[3];
}

// CollectRadiance with an endpoint emission: every open slot started at an
// earlier non-delta vertex receives throughput * Le (render.cuh:364-376).
__device__ __forceinline__ void g_endpoint(const GuideArgs& G, uint32_t p, const S4& le, uint32_t non_delta,
                                           uint32_t mode) {
    if (s4zero(le, mode)) return;
    const int hi = non_delta < 5u ? (int)non_delta : 5;
    uint32_t flags = 0;
    for (int i = 0; i < hi; ++i) {
        const S4 m = gload(G.gm, G.stride, i, p);
        if (!s4zero(m, mode)) {
            gstore(G.gr, G.stride, i, p, s4add(gload(G.gr, G.stride, i, p), s4mul(m, le)));
            flags |= 1u << i;
        }
    }
    if (flags) G.gflag[p] |= flags;
}

// CollectRadiance at a surviving vertex (render.cuh:378-409): opens the slot
// of this non-delta depth, then scales the open slots by contrib / q.
__device__ __forceinline__ void g_advance(const GuideArgs& G, uint32_t p, uint32_t non_delta, uint32_t mode,
                                          uint64_t key, int leaf, const S4& contrib, float qrr, bool start) {
    if (start && non_delta < 5u) {
        const int i = (int)non_delta;
        if (s4zero(gload(G.gm, G.stride, i, p), mode)) {
            gstore(G.gm, G.stride, i, p, s4(1.0f, 1.0f, 1.0f, 1.0f));
            G.gkey[(size_t)i * G.stride + p] = key;
            G.gleaf[p] = (G.gleaf[p] & ~(63u << (6 * i))) | ((uint32_t)leaf << (6 * i));
        }
    }
    const float invq = (float)(1.0 / (double)qrr);
    const S4 c = s4scale(contrib, invq);
    const int hi = non_delta + 1u < 5u ? (int)non_delta + 1 : 5;
    for (int i = 0; i < hi; ++i) {
        const S4 m = gload(G.gm, G.stride, i, p);
        if (!s4zero(m, mode)) gstore(G.gm, G.stride, i, p, s4mul(m, c));
    }
}

// Learned guiding: opens slot i at a non-delta vertex; it is completed when
// the path's next vertex is shaded (see GuideArgs).
__device__ __forceinline__ uint32_t g_open_learned(const GuideArgs& G, uint32_t p, uint32_t non_delta, uint32_t flags,
                                                   const float* xr, F3 dir, float pdf) {
    if (non_delta >= 5u) return flags;
    const int i = (int)non_delta;
    if ((flags >> i) & 0x101u) return flags;
    for (int k = 0; k < mlp::kRec; ++k) G.gx[(size_t)(mlp::kRec * i + k) * G.stride + p] = xr[k];
    G.gw[(size_t)(3 * i + 0) * G.stride + p] = dir.x;
    G.gw[(size_t)(3 * i + 1) * G.stride + p] = dir.y;
    G.gw[(size_t)(3 * i + 2) * G.stride + p] = dir.z;
    G.gp[(size_t)i * G.stride + p] = pdf;
    return flags | (1u << (8 + i));
}

// maskForProbeQuery (render.cuh:893-900): max(w, 1e-6) / (sum + 1e-6) with w
// the throughput (RGB) or the throughput times the CMFs of the wavelength(s).
__device__ __forceinline__ void probe_query(const S4& mask, const Lam& L, const float* __restrict__ cie, float q[3]) {
    float w0, w1, w2;
    if (L.mode == AR_SPECTRAL_RGB) {
        w0 = mask.v[0]; w1 = mask.v[1]; w2 = mask.v[2];
    } else if (L.mode == AR_SPECTRAL_SINGLE) {
        float cx, cy, cz;
        spec::cie_xyz(cie, L.l[0], cx, cy, cz);
        w0 = mask.v[0] * cx; w1 = mask.v[0] * cy; w2 = mask.v[0] * cz;
    } else {
        w0 = w1 = w2 = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float cx, cy, cz;
            spec::cie_xyz(cie, L.l[k], cx, cy, cz);
            w0 = w0 + mask.v[k] * cx; w1 = w1 + mask.v[k] * cy; w2 = w2 + mask.v[k] * cz;
        }
    }
    q[0] = fmaxf(w0, 1e-6f); q[1] = fmaxf(w1, 1e-6f); q[2] = fmaxf(w2, 1e-6f);
    const float sum = (float)((double)(q[0] + q[1] + q[2]) + 1e-6);
    q[0] = q[0] / sum; q[1] = q[1] / sum; q[2] = q[2] / sum;
}

__device__ __forceinline__ S4 load4(const float* a, const float* b, const float* c, const float* d, uint32_t p) {
    return s4(a[p], b[p], c[p], d[p]);
}

// ---------------------------------------------------------------- kernels

// RenderCamera's ray generation (core/scene.cu:135-164) for every path slot
// of the pass (slot p = s_local * npix + pixel), plus the path's first
// wavelength in spectral modes (render.cuh:466-474).
__global__ __launch_bounds__(kMaxBlock) void k_camera(RenderParams P, PathArgs S, GuideArgs G, QueueRays q,
                                                   uint32_t* __restrict__ qpath, uint32_t* __restrict__ counters,
                                                   const float* __restrict__ eps_table) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) { counters[0] = P.capacity; counters[1] = 0; counters[2] = 0; }
    if (p >= P.capacity) return;
    const Segment sg = P.segs[segment_of(P.segs, P.nseg, p)];
    const SceneView& V = P.views[sg.scene];
    const uint32_t local = p - sg.base;
    const uint32_t pix = local % V.npix, smp = sg.sample0 + local / V.npix;
    const uint32_t i = pix % V.width, j = pix / V.width;
    float u[4];
    uniforms(P.seed, pix, smp, kCameraBounce, 0, u);
    const float uu = ((float)i + u[0]) / (float)V.width;
    const float vv = ((float)j + u[1]) / (float)V.height;
    const float ndx = uu * 2.0f - 1.0f, ndy = (1.0f - vv) * 2.0f - 1.0f;
    const F3 fwd = f3(V.fwd[0], V.fwd[1], V.fwd[2]);
    const F3 right = f3(V.right[0], V.right[1], V.right[2]);
    const F3 up = f3(V.up[0], V.up[1], V.up[2]);
    F3 d = normalize(add(add(fwd, scale(right, ndx * V.thx)), scale(up, ndy * V.thy)));
    d = normalize(d);   // CalculateRadiance normalises the incoming direction again
    float l0 = 0.0f;
    if (P.spectral != AR_SPECTRAL_RGB) {
        float w[4];
        uniforms(P.seed, pix, smp, kCameraBounce, 1, w);
        l0 = P.spectral == AR_SPECTRAL_SINGLE ? spec::sample_lambda(w[0], w[1], w[2]) : 380.0f + 400.0f * w[0];
    }
    S.pix[p] = pix;
    S.smp[p] = smp;
    S.scn[p] = sg.scene;
    q.root[p] = V.root0;
    S.med[p] = V.medium0;
    S.mp0[p] = V.cam_pos[0]; S.mp1[p] = V.cam_pos[1]; S.mp2[p] = V.cam_pos[2];
    q.ox[p] = V.cam_pos[0]; q.oy[p] = V.cam_pos[1]; q.oz[p] = V.cam_pos[2];
    q.tmin[p] = eps_table[0];
    q.dx[p] = d.x; q.dy[p] = d.y; q.dz[p] = d.z;
    q.tmax[p] = __int_as_float(0x7f800000);
    qpath[p] = p;
    S.mask0[p] = 1.0f; S.mask1[p] = 1.0f; S.mask2[p] = 1.0f; S.mask3[p] = 1.0f;
    S.rad0[p] = 0.0f; S.rad1[p] = 0.0f; S.rad2[p] = 0.0f; S.rad3[p] = 0.0f;
    S.last_pdf[p] = 0.0f;
    S.lp0[p] = 0.0f; S.lp1[p] = 0.0f; S.lp2[p] = 0.0f;
    S.meta[p] = 0u;
    S.lambda0[p] = l0;
    if (P.guiding == AR_GUIDING_VOXEL) {
        for (int i = 0; i < guide::kSlots; ++i) {
            gstore(G.gm, G.stride, i, p, s4(0, 0, 0, 0));
            gstore(G.gr, G.stride, i, p, s4(0, 0, 0, 0));
            G.gkey[(size_t)i * G.stride + p] = guide::kNoKey;
        }
        G.gleaf[p] = 0u;
    }
    if (P.guiding) G.gflag[p] = 0u;
}

// Network inputs of every queued hit (the learned guide's inference batch):
// the hit point (in the box of the path's scene) and the shading normal
// facing the ray, as k_shade computes them, and the scene index.
__global__ __launch_bounds__(kMaxBlock) void k_mlp_features(const float4* __restrict__ prims,
                                                         const float4* __restrict__ objs, const float* __restrict__ vol,
                                                         uint32_t ntri, QueueRays cur, const uint32_t* __restrict__ qcur,
                                                         const uint32_t* __restrict__ scn,
                                                         const SceneView* __restrict__ views,
                                                         const uint32_t* __restrict__ hprim,
                                                         const float* __restrict__ ht, const float* __restrict__ hu,
                                                         const float* __restrict__ hv,
                                                         const uint32_t* __restrict__ hflags,
                                                         const uint32_t* __restrict__ counters,
                                                         float* __restrict__ feat) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= counters[0]) return;
    float x[mlp::kRec] = {0, 0, 0, 0, 0, 0};
    const uint32_t scene = scn[qcur[q]];
    const SceneView& V = views[scene];
    const float* lo = V.box_lo;
    const float* inv = V.box_inv;
    x[mlp::kUsedIn] = (float)scene;
    const uint32_t prim = hprim[q];
    if (prim != AR_MISS && prim >= ntri) {
        float pp[3], np[3];
        obj::surface(objs + (size_t)obj::kRecord * (prim - ntri), cur.ox[q], cur.oy[q], cur.oz[q], cur.dx[q], cur.dy[q],
                     cur.dz[q], ht[q], (hflags[q] & AR_HIT_INSIDE) != 0u, 0.5f * cur.tmin[q], vol, pp, np);
        mlp::features(pp[0], pp[1], pp[2], np[0], np[1], np[2], lo, inv, x);
    } else if (prim != AR_MISS) {
        const F3 a = xyz(prims[3 * (size_t)prim + 0]), b = xyz(prims[3 * (size_t)prim + 1]),
                 c = xyz(prims[3 * (size_t)prim + 2]);
        const float bu = hu[q], bv = hv[q];
        const float bw = 1.0f - bu - bv;
        const F3 pos = add(add(scale(a, bw), scale(b, bu)), scale(c, bv));
        const F3 ng = normalize(cross(sub(b, a), sub(c, a)));
        const F3 d = f3(cur.dx[q], cur.dy[q], cur.dz[q]);
        const F3 nn = dot(ng, d) < 0.0f ? ng : scale(ng, -1.0f);
        mlp::features(pos.x, pos.y, pos.z, nn.x, nn.y, nn.z, lo, inv, x);
    }
#pragma unroll
    for (int k = 0; k < mlp::kRec; ++k) feat[mlp::kRec * (size_t)q + k] = x[k];
}

// k_shade<AR_GUIDING_MLP>'s dynamic LDS: one row of kMixRow floats per
// thread for its vertex's decoded mixture
constexpr int kMixRow = mlp::kDecoded + 1;
extern __shared__ float shade_mix_lds[];
__device__ __forceinline__ float* shade_mix_row() { return shade_mix_lds + threadIdx.x * kMixRow; }

// One bounce of CalculateRadiance for every queued path: emission with MIS
// (render.cuh:664-710), BSDF sampling and the hit-point offset
// (render.cuh:75-163), NEE light sampling (render.cuh:205-343, 766-827;
// the shadow segment is queued and resolved after the any-hit pass), Russian
// roulette (render.cuh:838-859) and the next ray.
template <uint32_t GM>
__device__ __forceinline__ void shade_one(
    const RenderParams& P, uint32_t depth, const SceneArgs& sc, const PathArgs& S, const GuideArgs& G,
    const QueueRays& cur, const uint32_t* __restrict__ qcur,
    const uint32_t* __restrict__ hprim, const float* __restrict__ ht, const float* __restrict__ hu,
    const float* __restrict__ hv, const uint32_t* __restrict__ hflags,
    const QueueRays& nxt, uint32_t* __restrict__ qnext,
    const QueueRays& sh, uint32_t* __restrict__ spath, float* __restrict__ sc0, float* __restrict__ sc1,
    float* __restrict__ sc2, float* __restrict__ sc3, uint32_t* __restrict__ sflags,
    uint32_t* __restrict__ counters, const float* __restrict__ eps_table, uint32_t q, uint32_t n) {
    bool cont = false, shadow = false;
    uint32_t p = 0, root = 0;
    F3 next_o = f3(0, 0, 0), next_d = f3(0, 0, 0);
    float next_tmin = 0.0f;
    F3 s_o = f3(0, 0, 0), s_d = f3(0, 0, 1);
    S4 s_contrib = s4(0, 0, 0, 0);
    float s_tmax = 0.0f;
    uint32_t s_flag = 0;
    if (q < n) {
        p = qcur[q];
        const uint32_t pix = S.pix[p], smp = S.smp[p];
        const SceneView& V = P.views[S.scn[p]];
        root = V.root;
        float u[12];
        uniforms(P.seed, pix, smp, depth, 0, u);
        uniforms(P.seed, pix, smp, depth, 1, u + 4);
        uniforms(P.seed, pix, smp, depth, 2, u + 8);
        const uint32_t meta = S.meta[p];
        const uint32_t non_delta = meta & 0xFFFFu;
        const bool prev_nee = (meta & kMetaNee) != 0;
        const Lam L = make_lam(P.spectral, S.lambda0[p]);
        S4 mask = load4(S.mask0, S.mask1, S.mask2, S.mask3, p);
        S4 rad = load4(S.rad0, S.rad1, S.rad2, S.rad3, p);
        constexpr bool learned = GM == AR_GUIDING_MLP;
        // learned guiding: the slot opened at the previous non-delta vertex is
        // completed here with the entry throughput and radiance
        const uint32_t lflags_in = learned ? G.gflag[p] : 0u;
        uint32_t lflags = lflags_in;
        const S4 mask_in = mask, rad_in = rad;
        S4 le_unw = s4(0, 0, 0, 0), le_add = s4(0, 0, 0, 0);   // this vertex's emission: unweighted, as added
        const bool guided = learned ? P.alpha > 0.0f && depth + 1 < P.max_depth
                                    : (GM != AR_GUIDING_OFF && *G.probe_count > 0u);
        float gq[3] = {1.0f, 1.0f, 1.0f};
        if (GM != AR_GUIDING_OFF && depth > 0) probe_query(mask, L, sc.cie, gq);
        const F3 o = f3(cur.ox[q], cur.oy[q], cur.oz[q]);
        const F3 d = f3(cur.dx[q], cur.dy[q], cur.dz[q]);
        const float eps = eps_table[non_delta < 255 ? non_delta : 255];
        const uint32_t prim = hprim[q];
        uint32_t new_meta = meta & ~kMetaNee;
        uint32_t medium = S.med[p];
        if (depth == 0 && S.fh && prim == AR_MISS) S.fh[p] = __uint_as_float(4u);   // MAT_MAX: no first hit
        if (prim == AR_MISS && medium != 0u) {
            // inside a medium with no exit found (bMissThinSurfaceInsideMedia_,
            // render.cuh:107-124): leave the medium and continue straight from
            // the same origin as a delta vertex of weight 1
            float qrr = 1.0f;
            bool alive = true;
            if (non_delta >= P.rr_min_depth) {
                qrr = clampr(lum4(mask, P.spectral) * 0.5f + max4(mask, P.spectral) * 0.5f, 0.05f, 0.95f);
                if (u[9] > qrr) alive = false;
            }
            new_meta = meta & ~kMetaNee;
            if (alive && depth + 1 < P.max_depth) {
                if (GM == AR_GUIDING_VOXEL)
                    g_advance(G, p, non_delta, P.spectral, guide::kNoKey, 0, s4(1, 1, 1, 1), qrr, false);
                mask = s4scale(mask, 1.0f / qrr);
                cont = true;
                next_o = o;
                next_d = d;
                next_tmin = eps;
                root = V.root;
                new_meta = meta;
            }
            S.med[p] = 0u;
        } else if (prim == AR_MISS) {
            S4 e = s4(0, 0, 0, 0);
            if (P.sky == AR_SKY_GRADIENT) {
                const F3 eg = sky_gradient(d);
                e = s4(eg.x, eg.y, eg.z, 0.0f);
                rad = s4add(rad, s4mul(mask, e));
            } else if (P.sky == AR_SKY_CONSTANT) {
                if (P.spectral == AR_SPECTRAL_RGB) {
                    e = s4(P.env[0], P.env[1], P.env[2], 0.0f);
                } else {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        float v = 0.0f;
                        if (P.spectral == AR_SPECTRAL_HERO4 || k == 0)
                            v = (sc.basis ? sb::lut_interact(sc.rgb_lut, (int)sc.rgb_res,
                                                             sb::query_lambda(sc.lam_lut, (int)sc.lam_res, L.l[k]),
                                                             P.env[0], P.env[1], P.env[2])
                                          : spec::sigmoid_poly(P.env_c[0], P.env_c[1], P.env_c[2], L.l[k]) * P.env_scale) *
                                spec::d65_norm(sc.d65, L.l[k]);
                        e.v[k] = v;
                    }
                }
                rad = s4add(rad, s4mul(mask, e));
            }
            if (GM == AR_GUIDING_VOXEL) g_endpoint(G, p, e, non_delta, P.spectral);
            if (learned && P.sky != AR_SKY_NONE) { le_unw = e; le_add = s4mul(mask, e); }
        } else {
            const bool is_obj = prim >= sc.ntri;
            const float4* ob = is_obj ? sc.objs + (size_t)obj::kRecord * (prim - sc.ntri) : nullptr;
            F3 pos, ng, nn;
            uint32_t mat_id;
            int light_id;
            float ndd;
            if (is_obj) {
                // the intersector's normal: outward, or inward for a box or sphere hit
                // from inside; an SDF normal is the outward gradient, its
                // difference step the march tolerance of the ray (SDF_EPS of its depth)
                float pp[3], np[3];
                obj::surface(ob, o.x, o.y, o.z, d.x, d.y, d.z, ht[q], (hflags[q] & AR_HIT_INSIDE) != 0u,
                             0.5f * cur.tmin[q], sc.vol, pp, np);
                pos = f3(pp[0], pp[1], pp[2]);
                ng = f3(np[0], np[1], np[2]);
                nn = ng;
                ndd = dot(ng, d);
                mat_id = __float_as_uint(ob[2].w);
                light_id = __float_as_int(ob[3].w);
            } else {
                const float4 a4 = sc.prims[3 * (size_t)prim + 0];
                const float4 b4 = sc.prims[3 * (size_t)prim + 1];
                const float4 c4 = sc.prims[3 * (size_t)prim + 2];
                const F3 a = xyz(a4), b = xyz(b4), c = xyz(c4);
                mat_id = __float_as_uint(a4.w);
                light_id = __float_as_int(b4.w);
                const float bu = hu[q], bv = hv[q];
                const float bw = 1.0f - bu - bv;
                pos = add(add(scale(a, bw), scale(b, bu)), scale(c, bv));
                ng = normalize(cross(sub(b, a), sub(c, a)));
                ndd = dot(ng, d);
                nn = ndd < 0.0f ? ng : scale(ng, -1.0f);
            }
            const Mat m = load_mat(sc.mats, mat_id);
            if (depth == 0 && S.fh) {
                // RecordScreenGBuffer's inputs (render.cuh:640-643): albedo, the
                // intersector's normal (facing the ray for triangles), distance,
                // metallic, roughness and MAT_DIFFUSE / LIGHT / GLASS / GGX
                const uint32_t model = light_id >= 0 ? 1u : (m.model == AR_MODEL_GLASS ? 2u : (m.model == AR_MODEL_GGX ? 3u : 0u));
                float* f = S.fh + p;
                const size_t st = S.fh_stride;
                f[0] = __uint_as_float(model);
                f[st * 1] = m.albedo.x; f[st * 2] = m.albedo.y; f[st * 3] = m.albedo.z;
                f[st * 4] = nn.x; f[st * 5] = nn.y; f[st * 6] = nn.z;
                f[st * 7] = ht[q];
                f[st * 8] = m.metallic;
                f[st * 9] = m.roughness;
            }
            if (light_id >= 0) {
                // triangle lights emit from the front face, box and sphere
                // lights from every hit (MAT_LIGHT objects, render.cuh:664-710)
                if (is_obj || ndd < 0.0f) {
                    float w = 1.0f;
                    if (P.nee && depth > 0 && prev_nee) {
                        const F3 lp = f3(S.lp0[p], S.lp1[p], S.lp2[p]);
                        float pl;
                        if (is_obj) {
                            pl = obj_light_pdf(ob, sc.light_prob[light_id], lp, pos, ng);
                        } else {
                            const F3 tl = sub(pos, lp);
                            const float dist2 = maxr(dot(tl, tl), 1e-20f);
                            const F3 dir = scale(tl, 1.0f / sqrtf(dist2));
                            const float cosy = maxr(0.0f, -dot(ng, dir));
                            const float pa = sc.light_prob[light_id] / sc.light_area[light_id];
                            pl = cosy <= 0.0f ? 0.0f : pa * (dist2 / cosy);
                        }
                        const float last_pdf = S.last_pdf[p];
                        if (pl >= 0.0f && last_pdf > 0.0f) w = power_heuristic(last_pdf, pl);
                    }
                    const S4 le = emission_of(m, L, sc);
                    rad = s4add(rad, s4scale(s4mul(mask, le), w));
                    if (GM == AR_GUIDING_VOXEL) g_endpoint(G, p, s4scale(le, w), non_delta, P.spectral);
                    if (learned) { le_unw = le; le_add = s4scale(s4mul(mask, le), w); }
                }
            } else if (m.model == AR_MODEL_GLASS) {
                // Fresnel reflection or refraction (SampleBSDF MAT_GLASS,
                // material.cuh:206-257) weighted by EvaluateBSDFWithCos
                // (material.cuh:116-161), Beer-Lambert over the path inside
                const bool ins = medium != 0u;
                float ior = m.ior;
                if (P.spectral != AR_SPECTRAL_RGB) ior = ior_at(m.ior, m.abbe, L.l[0]);
                if (P.spectral == AR_SPECTRAL_HERO4 && m.abbe > 0.0f && !(meta & kMetaHeroOnly)) {
                    // dispersion: the hero wavelength continues alone with the
                    // weight of all four
                    mask = s4(mask.v[0] * 4.0f, 0.0f, 0.0f, 0.0f);
                    new_meta |= kMetaHeroOnly;
                }
                S4 tr = s4(1, 1, 1, 1);
                if (ins) {
                    const float dist = length(sub(f3(S.mp0[p], S.mp1[p], S.mp2[p]), pos));
                    const S4 alb = albedo_of(m, L, sc);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const float sa = P.spectral == AR_SPECTRAL_RGB ? m.sigma_a : sigma_at(m.sigma_a, L.l[k], m.tilt);
                        const float av = minr(maxr(alb.v[k], 1e-12f), 1.0f);
                        tr.v[k] = expf(dist * sa * logf(av));
                    }
                }
                const F3 nin = dot(ng, d) > 0.0f ? scale(ng, -1.0f) : ng;
                const float eta = ins ? ior : 1.0f / ior;
                const float cosi = -dot(nin, d);
                const float sin2i = maxr(0.0f, 1.0f - cosi * cosi);
                const float kk = 1.0f - sin2i * (eta * eta);
                F3 next;
                float pdf;
                S4 fcos;
                uint32_t next_medium = medium;
                if (kk <= 0.0f) {
                    next = reflect3(d, nin);
                    pdf = 1.0f;
                    fcos = s4scale(tr, 0.99f);
                } else {
                    const float cost = sqrtf(kk);
                    const float fr = fresnel_dielectric(cosi, cost, eta);
                    const F3 rdir = reflect3(d, nin);
                    const F3 tdir = refract3(d, nin, eta);
                    if (u[0] < fr) {
                        next = rdir;
                        pdf = fr;
                    } else {
                        next = tdir;
                        pdf = 1.0f - fr;
                        next_medium = ins ? 0u : prim - sc.ntri + 1u;
                    }
                    if (dot(next, rdir) >= dot(next, tdir)) {
                        fcos = s4scale(tr, fr);
                    } else {
                        const float wgt = (eta * eta) * (fabsf(cost) / maxr(1e-6f, cosi));
                        fcos = s4scale(tr, (1.0f - fr) * wgt);
                    }
                }
                S.mp0[p] = pos.x; S.mp1[p] = pos.y; S.mp2[p] = pos.z;
                medium = next_medium;
                S.med[p] = medium;
                const float off = eps * maxr(1.0f, length(pos)) * (u[4] + 0.5f);
                const float dn = dot(ng, next);
                const float sgn = dn > 0.0f ? 1.0f : (dn < 0.0f ? -1.0f : 0.0f);
                pos = add(pos, scale(ng, sgn * off));
                S.last_pdf[p] = pdf;
                S.lp0[p] = pos.x; S.lp1[p] = pos.y; S.lp2[p] = pos.z;
                const S4 contrib = s4scale(fcos, 1.0f / maxr(pdf, 1e-12f));
                float qrr = 1.0f;
                bool alive = true;
                if (non_delta >= P.rr_min_depth) {
                    const float base =
                        clampr(lum4(mask, P.spectral) * 0.5f + max4(mask, P.spectral) * 0.5f, 0.05f, 0.95f);
                    const float tl = medium != 0u ? 0.99f : 0.0f;
                    qrr = base * (1.0f - tl) + tl;
                    if (u[9] > qrr) alive = false;
                }
                if (alive && depth + 1 < P.max_depth) {
                    if (GM == AR_GUIDING_VOXEL)
                        g_advance(G, p, non_delta, P.spectral, guide::kNoKey, 0, contrib, qrr, false);
                    mask = s4scale(s4mul(mask, contrib), 1.0f / qrr);
                    cont = true;
                    next_o = pos;
                    next_d = next;
                    next_tmin = eps;
                    if (medium != 0u) {
                        const float4* mo = sc.objs + (size_t)obj::kRecord * (medium - 1u);
                        root = __float_as_uint(mo[4].w);
                    } else {
                        root = V.root;
                    }
                }
            } else {
                const S4 alb = albedo_of(m, L, sc);
                const float pspec = m.model == AR_MODEL_LAMBERT ? 0.0f : pspec_of(m, alb, P.spectral);
                F3 next;
                float pdf;
                // the learned guide: choose the guide (probability alpha) or the
                // BSDF first and sample only the chosen one, so the direction's
                // density is exactly the one-sample mixture (1 - alpha) p_bsdf +
                // alpha p_guide however sharp the guide's lobes are
                // the vertex's mixture (kPacked words from inference),
                // unpacked once into this lane's own LDS row (kMixRow floats,
                // an odd stride: no bank conflicts) that the evaluations below
                // read several times
                const float* pmix = nullptr;
                bool via_guide = false, sampled;
                if (learned && guided) {
                    float* row = shade_mix_row();
                    const float4* src = reinterpret_cast<const float4*>(G.dec + (size_t)q * mlp::kPacked);
#pragma unroll
                    for (int k = 0; k < mlp::kPacked / 12; ++k) {   // four lobes per three 16-byte loads
                        const float4 a = src[3 * k], b = src[3 * k + 1], c = src[3 * k + 2];
                        float* l = row + 24 * k;
                        mlp::unpack_lobe(a.x, a.y, __float_as_uint(a.z), l);
                        mlp::unpack_lobe(a.w, b.x, __float_as_uint(b.y), l + 6);
                        mlp::unpack_lobe(b.z, b.w, __float_as_uint(c.x), l + 12);
                        mlp::unpack_lobe(c.y, c.z, __float_as_uint(c.w), l + 18);
                    }
                    pmix = row;
                    via_guide = u[10] < P.alpha;
                }
                if (via_guide) {
                    float ug[4];
                    uniforms(P.seed, pix, smp, depth, 3, ug);
                    float gx, gy, gz;
                    mlp::hemi_sample(pmix, nn.x, nn.y, nn.z, ug, gx, gy, gz, pdf);
                    next = f3(gx, gy, gz);
                    sampled = dot(next, nn) > 0.0f && pdf > 0.0f;
                } else {
                    sampled = sample_bsdf(m, pspec, d, nn, u, next, pdf);
                }
                if (sampled) {
                    S4 fcos = eval_bsdf(m, alb, d, next, nn);
                    if (pmix) {
                        const float pb = via_guide ? bsdf_pdf(m, pspec, d, next, nn) : pdf;
                        const float pg = via_guide ? pdf : mlp::hemi_pdf(pmix, next.x, next.y, next.z, nn.x, nn.y, nn.z);
                        pdf = (1.0f - P.alpha) * pb + P.alpha * pg;
                    }
                    float xr[mlp::kRec] = {0, 0, 0, 0, 0, 0};
                    if (learned) {
                        if (guided) {
                            // k_mlp_features took them from the same hit, the same way
#pragma unroll
                            for (int k = 0; k < mlp::kRec; ++k) xr[k] = G.feat[(size_t)mlp::kRec * q + k];
                        } else {
                            mlp::features(pos.x, pos.y, pos.z, nn.x, nn.y, nn.z, V.box_lo, V.box_inv, xr);
                            xr[mlp::kUsedIn] = (float)S.scn[p];
                        }
                    }
                    S.mp0[p] = pos.x; S.mp1[p] = pos.y; S.mp2[p] = pos.z;
                    const float off = eps * maxr(1.0f, length(pos)) * (u[4] + 0.5f);
                    const float dn = dot(nn, next);
                    const float sgn = dn > 0.0f ? 1.0f : (dn < 0.0f ? -1.0f : 0.0f);
                    pos = add(pos, scale(nn, sgn * off));
                    // voxel guide/BSDF one-sample mixture (render.cuh:553-638)
                    const float4* ptex = nullptr;
                    uint64_t gkey = guide::kNoKey;
                    if (GM == AR_GUIDING_VOXEL) {
                        const F3 rv = sub(d, scale(nn, 2.0f * dot(d, nn)));
                        gkey = P.batched ? guide::voxel_key_scene(pos.x, pos.y, pos.z, rv.x, rv.y, rv.z, S.scn[p])
                                         : guide::voxel_key(pos.x, pos.y, pos.z, rv.x, rv.y, rv.z);
                    }
                    if (GM == AR_GUIDING_VOXEL && guided) {
                        const uint32_t pi = guide::find(G.pr, gkey);
                        if (pi != guide::kNoProbe) {
                            ptex = G.pr.tex + (size_t)pi * guide::kTexels;
                            const float pb = pdf;
                            const float pp = guide::hemi_pdf(ptex, next.x, next.y, next.z, nn.x, nn.y, nn.z, gq);
                            if (pb > 1e-12f && pp > 1e-12f) {
                                if (u[10] < P.alpha) {
                                    float ug[5];
                                    ug[0] = u[11];
                                    uniforms(P.seed, pix, smp, depth, 3, ug + 1);
                                    float gx, gy, gz, gp;
                                    guide::hemi_sample(ptex, nn.x, nn.y, nn.z, gq, ug, gx, gy, gz, gp);
                                    const F3 gv = f3(gx, gy, gz);
                                    if (dot(gv, nn) > 1e-12f) {
                                        const float pbs = bsdf_pdf(m, pspec, d, gv, nn);
                                        next = gv;
                                        fcos = eval_bsdf(m, alb, d, gv, nn);
                                        pdf = (1.0f - P.alpha) * maxr(pbs, 1e-12f) + P.alpha * gp;
                                    } else {
                                        ptex = nullptr;
                                    }
                                } else {
                                    pdf = (1.0f - P.alpha) * pb + P.alpha * pp;
                                }
                            } else {
                                ptex = nullptr;
                            }
                        }
                    }
                    bool nee_now = false;
                    if (P.nee && V.light_cnt > 0 && depth + 1 < P.max_depth && non_delta < 10 && medium == 0u &&
                        u[5] < sqrtf(clampr(lum4(mask, P.spectral), 0.0f, 1.0f))) {
                        // pick by the scene's CDF (first entry with u < cdf; binary search)
                        uint32_t lo = V.light_off, hi = V.light_off + V.light_cnt - 1;
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (u[6] < sc.light_cdf[mid]) hi = mid; else lo = mid + 1;
                        }
                        const uint32_t li = lo;
                        const uint32_t lp = (uint32_t)sc.light_tri[li];
                        F3 qpt, ln;
                        float pa_total = 1.0f;
                        bool valid = true;
                        Mat lm;
                        if (lp < sc.ntri) {
                            const float4 la4 = sc.prims[3 * (size_t)lp + 0];
                            const F3 la = xyz(la4), lb = xyz(sc.prims[3 * (size_t)lp + 1]),
                                     lc = xyz(sc.prims[3 * (size_t)lp + 2]);
                            const float su = sqrtf(u[7]);
                            qpt = add(add(scale(la, 1.0f - su), scale(lb, su * (1.0f - u[8]))), scale(lc, su * u[8]));
                            ln = normalize(cross(sub(lb, la), sub(lc, la)));
                            lm = load_mat(sc.mats, __float_as_uint(la4.w));
                            pa_total = maxr(1e-20f, sc.light_prob[li] / sc.light_area[li]);
                        } else {
                            // box facet pick from a fifth group of uniforms
                            const float4* lo4 = sc.objs + (size_t)obj::kRecord * (lp - sc.ntri);
                            float ux[4];
                            uniforms(P.seed, pix, smp, depth, 4, ux);
                            float pa;
                            valid = obj_light_sample(lo4, pos, ux[0], u[7], u[8], qpt, ln, pa);
                            if (valid) pa_total = maxr(1e-20f, pa * sc.light_prob[li]);
                            lm = load_mat(sc.mats, __float_as_uint(lo4[2].w));
                        }
                        if (valid) {
                            const F3 lu = sub(qpt, pos);
                            const float dist2 = maxr(dot(lu, lu), 1e-12f);
                            const F3 ldir = scale(lu, 1.0f / sqrtf(dist2));
                            const float dist = sqrtf(dot(lu, lu));
                            const float cosx = maxr(0.0f, dot(nn, ldir));
                            const float cosy = maxr(0.0f, -dot(ln, ldir));
                            const bool faces = cosx > 0.0f && cosy > 0.0f;
                            S4 contrib_w = s4(0, 0, 0, 0);
                            if (faces) {
                                const S4 f = eval_bsdf(m, alb, d, ldir, nn);
                                const float g = cosy / dist2;
                                const S4 contrib = s4scale(s4mul(f, emission_of(lm, L, sc)), g / pa_total);
                                const float pl = pa_total * (dist2 / cosy);
                                float pb = maxr(bsdf_pdf(m, pspec, d, ldir, nn), 0.0f);
                                if (ptex)
                                    pb = (1.0f - P.alpha) * pb +
                                         P.alpha * guide::hemi_pdf(ptex, ldir.x, ldir.y, ldir.z, nn.x, nn.y, nn.z, gq);
                                else if (pmix)
                                    pb = (1.0f - P.alpha) * pb +
                                         P.alpha * mlp::hemi_pdf(pmix, ldir.x, ldir.y, ldir.z, nn.x, nn.y, nn.z);
                                const float w = power_heuristic(pl, pb);
                                contrib_w = s4mul(mask, s4scale(contrib, w));
                            }
                            // AnyHitSegmentBVH (intersection.cuh:285-310)
                            float tmx = 0.0f;
                            F3 sdir = f3(0, 0, 1);
                            if (dist > 0.0f) {
                                sdir = scale(lu, 1.0f / dist);
                                const float seps = eps_table[depth < 255 ? depth : 255] *
                                                   maxr(1.0f, maxr(length(pos), length(qpt)));
                                tmx = maxr(0.0f, dist - seps);
                            }
                            if (tmx > 0.0f) {
                                shadow = true;
                                s_o = pos; s_d = sdir; s_tmax = tmx; s_contrib = contrib_w; s_flag = faces ? 1u : 0u;
                            } else {
                                // nothing to test: visible
                                if (faces) { rad = s4add(rad, contrib_w); nee_now = true; }
                            }
                        }
                    }
                    // the resolve pass sets the NEE bit for queued shadow segments
                    new_meta = (meta & kMetaHeroOnly) | non_delta | (nee_now ? kMetaNee : 0u);
                    S.last_pdf[p] = pdf;
                    S.lp0[p] = pos.x; S.lp1[p] = pos.y; S.lp2[p] = pos.z;
                    const float inv_pdf = 1.0f / maxr(pdf, 1e-12f);
                    const S4 contrib = s4scale(fcos, inv_pdf);
                    float qrr = 1.0f;
                    bool alive = true;
                    if (non_delta >= P.rr_min_depth) {
                        qrr = clampr(lum4(mask, P.spectral) * 0.5f + max4(mask, P.spectral) * 0.5f, 0.05f, 0.95f);
                        if (medium != 0u) qrr = qrr * (1.0f - 0.99f) + 0.99f;
                        if (u[9] > qrr) alive = false;
                    }
                    if (alive && depth + 1 < P.max_depth) {
                        if (GM == AR_GUIDING_VOXEL)
                            g_advance(G, p, non_delta, P.spectral, gkey, guide::leaf_of(next.x, next.y, next.z),
                                      contrib, qrr, medium == 0u);
                        if (learned && medium == 0u) lflags = g_open_learned(G, p, non_delta, lflags, xr, next, pdf);
                        mask = s4scale(s4mul(mask, contrib), 1.0f / qrr);
                        cont = true;
                        next_o = pos;
                        next_d = next;
                        const uint32_t nd = non_delta + 1;
                        next_tmin = eps_table[nd < 255 ? nd : 255];
                        new_meta = (new_meta & (kMetaNee | kMetaHeroOnly)) | nd;
                        if (medium != 0u) {
                            const float4* mo = sc.objs + (size_t)obj::kRecord * (medium - 1u);
                            root = __float_as_uint(mo[4].w);
                        }
                    }
                }
            }
        }
        if (learned) {
            const uint32_t pend = (lflags_in >> 8) & 31u;
            if (pend) {
                const int j = __ffs(pend) - 1;
                const S4 me = s4mul(mask_in, le_unw);
                const S4 base = s4(rad_in.v[0] + le_add.v[0] - me.v[0], rad_in.v[1] + le_add.v[1] - me.v[1],
                                   rad_in.v[2] + le_add.v[2] - me.v[2], rad_in.v[3] + le_add.v[3] - me.v[3]);
                gstore(G.gr, G.stride, j, p, base);
                gstore(G.gm, G.stride, j, p, mask_in);
                lflags = (lflags & ~(1u << (8 + j))) | (1u << j);
            }
            if (lflags != lflags_in) G.gflag[p] = lflags;
        }
        S.mask0[p] = mask.v[0]; S.mask1[p] = mask.v[1]; S.mask2[p] = mask.v[2]; S.mask3[p] = mask.v[3];
        S.rad0[p] = rad.v[0]; S.rad1[p] = rad.v[1]; S.rad2[p] = rad.v[2]; S.rad3[p] = rad.v[3];
        S.meta[p] = new_meta;
    }
    uint32_t slot, ss;
    block_append2<kMaxBlock / 64>(cont, shadow, &counters[1], &counters[2], slot, ss);
    if (cont) {
        nxt.ox[slot] = next_o.x; nxt.oy[slot] = next_o.y; nxt.oz[slot] = next_o.z;
        nxt.tmin[slot] = next_tmin;
        nxt.dx[slot] = next_d.x; nxt.dy[slot] = next_d.y; nxt.dz[slot] = next_d.z;
        nxt.tmax[slot] = __int_as_float(0x7f800000);
        nxt.root[slot] = root;
        qnext[slot] = p;
    }
    if (shadow) {
        sh.ox[ss] = s_o.x; sh.oy[ss] = s_o.y; sh.oz[ss] = s_o.z; sh.tmin[ss] = 0.0f;
        sh.dx[ss] = s_d.x; sh.dy[ss] = s_d.y; sh.dz[ss] = s_d.z; sh.tmax[ss] = s_tmax;
        // SegmentTrace(depth, ...): SDF_EPS of the bounce index
        if (sc.sdf) sh.meps[ss] = 0.5f * eps_table[depth < 255 ? depth : 255];
        spath[ss] = p;
        sh.root[ss] = root;
        sc0[ss] = s_contrib.v[0]; sc1[ss] = s_contrib.v[1]; sc2[ss] = s_contrib.v[2]; sc3[ss] = s_contrib.v[3];
        sflags[ss] = s_flag;
    }
}

// One thread per entry of the current queue (grid sized for the pass);
// specialised per guiding mode so that plain renders carry no guiding state.
// Register budget: the plain variant runs at 4 waves per SIMD (128 VGPRs,
// about 60 bytes of spills per lane: 3 % faster on the room than 3 waves
// without spills); the guided variants spill 90-150 bytes at 4 and run
// faster at 3 (138-155 VGPRs).
template <uint32_t GM>
__global__ __launch_bounds__(kMaxBlock)
__attribute__((amdgpu_waves_per_eu(GM == AR_GUIDING_OFF ? 4 : 3, 8))) void k_shade(
    RenderParams P, uint32_t depth, SceneArgs sc, PathArgs S, GuideArgs G,
    QueueRays cur, const uint32_t* __restrict__ qcur,
    const uint32_t* __restrict__ hprim, const float* __restrict__ ht, const float* __restrict__ hu,
    const float* __restrict__ hv, const uint32_t* __restrict__ hflags,
    QueueRays nxt, uint32_t* __restrict__ qnext,
    QueueRays sh, uint32_t* __restrict__ spath, float* __restrict__ sc0, float* __restrict__ sc1,
    float* __restrict__ sc2, float* __restrict__ sc3, uint32_t* __restrict__ sflags,
    uint32_t* __restrict__ counters, const float* __restrict__ eps_table) {
    shade_one<GM>(P, depth, sc, S, G, cur, qcur, hprim, ht, hu, hv, hflags, nxt, qnext, sh, spath, sc0, sc1, sc2, sc3,
                  sflags, counters, eps_table, blockIdx.x * blockDim.x + threadIdx.x, counters[0]);
}

// Shadow results: an occluded segment still counts as a taken NEE sample
// (render.cuh:297-305); an unoccluded one adds its MIS-weighted
// contribution when the light point faced the surface.
__global__ __launch_bounds__(kMaxBlock) void k_resolve(PathArgs S, const uint32_t* __restrict__ spath,
                                                    const float* __restrict__ sc0, const float* __restrict__ sc1,
                                                    const float* __restrict__ sc2, const float* __restrict__ sc3,
                                                    const uint32_t* __restrict__ sflags,
                                                    const uint32_t* __restrict__ occ,
                                                    const uint32_t* __restrict__ counters) {
    const uint32_t n = counters[2];
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const uint32_t p = spath[s];
        const bool occluded = (occ[s >> 5] >> (s & 31u)) & 1u;
        const bool faces = sflags[s] & 1u;
        if (occluded || faces) S.meta[p] |= kMetaNee;
        if (!occluded && faces) {
            S.rad0[p] = S.rad0[p] + sc0[s];
            S.rad1[p] = S.rad1[p] + sc1[s];
            S.rad2[p] = S.rad2[p] + sc2[s];
            S.rad3[p] = S.rad3[p] + sc3[s];
        }
    }
}

// Next bounce: the appended queue becomes the current one.
// Ops-model counting (ar_render_counters): adds the box and triangle tests of
// the n_dev[0] rays just traced to tot[0], tot[1] and, when vert, the ray
// count (the vertices shaded next) to tot[2].  Integer sums, so the atomic
// order does not matter.
__global__ __launch_bounds__(256) void k_count_sum(const uint32_t* __restrict__ box, const uint32_t* __restrict__ tri,
                                                   const uint32_t* __restrict__ n_dev, unsigned long long* tot,
                                                   int vert) {
    __shared__ unsigned long long sb[256], st[256];
    const uint32_t n = *n_dev;
    unsigned long long a = 0, c = 0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) { a += box[i]; c += tri[i]; }
    sb[threadIdx.x] = a;
    st[threadIdx.x] = c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) { sb[threadIdx.x] += sb[threadIdx.x + off]; st[threadIdx.x] += st[threadIdx.x + off]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&tot[0], sb[0]);
        atomicAdd(&tot[1], st[0]);
        if (vert && blockIdx.x == 0) atomicAdd(&tot[2], (unsigned long long)n);
    }
}

__global__ void k_advance(uint32_t* counters) {
    counters[0] = counters[1];
    counters[1] = 0;
    counters[2] = 0;
}

// A path's pixel value: RGB clamp(checkNanInf(L), 0, 1e3), or in spectral
// modes clamp(XYZ(lambda) * checkNanInf(L) / p(lambda), 0, 1e3)
// (render.cuh:966-997), summed over the hero wavelengths over the sum of
// their densities.
__device__ __forceinline__ F3 path_value(const RenderParams& P, const SceneArgs& sc, const PathArgs& S, uint32_t p) {
    if (P.spectral == AR_SPECTRAL_RGB)
        return f3(clampr(fixnan(S.rad0[p]), 0.0f, 1e3f), clampr(fixnan(S.rad1[p]), 0.0f, 1e3f),
                  clampr(fixnan(S.rad2[p]), 0.0f, 1e3f));
    const Lam L = make_lam(P.spectral, S.lambda0[p]);
    const float r[4] = {S.rad0[p], S.rad1[p], S.rad2[p], S.rad3[p]};
    float X = 0.0f, Y = 0.0f, Z = 0.0f;
    if (P.spectral == AR_SPECTRAL_SINGLE) {
        float x, y, z;
        spec::cie_xyz(sc.cie, L.l[0], x, y, z);
        const float pdf = spec::lambda_pdf(L.l[0]);
        const float v = fixnan(r[0]);
        X = x * v / pdf; Y = y * v / pdf; Z = z * v / pdf;
    } else {
        const float psum = 4.0f / 400.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float x, y, z;
            spec::cie_xyz(sc.cie, L.l[k], x, y, z);
            const float v = fixnan(r[k]);
            X = X + x * v; Y = Y + y * v; Z = Z + z * v;
        }
        X = X / psum; Y = Y / psum; Z = Z / psum;
    }
    return f3(clampr(X, 0.0f, 1e3f), clampr(Y, 0.0f, 1e3f), clampr(Z, 0.0f, 1e3f));
}

// Per-pixel sum over the pass's samples of each scene in sample order, onto
// the scene's running sum; the scene's last pass writes the mean
// (XYZ2SRGBLinearD65, core/tonemap.cuh:116-126, for sRGB output).
__global__ __launch_bounds__(kMaxBlock) void k_accumulate(RenderParams P, SceneArgs sc, PathArgs S,
                                                          float* __restrict__ accum) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P.npixels) return;
    uint32_t lo = 0, hi = P.nseg - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (P.segs[mid].pixbase <= t) lo = mid; else hi = mid - 1;
    }
    const Segment sg = P.segs[lo];
    const SceneView& V = P.views[sg.scene];
    const uint32_t pix = t - sg.pixbase;
    const bool first = sg.sample0 == V.sample_offset;
    const bool last = sg.sample0 + sg.count == V.sample_offset + V.spp;
    // per pixel: sum of all samples, of the odd-index and of the even-index ones
    float* acc = accum + kAccumFloats * ((size_t)V.out_base + pix);
    float ax = first ? 0.0f : acc[0];
    float ay = first ? 0.0f : acc[1];
    float az = first ? 0.0f : acc[2];
    float h[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) h[k] = first ? 0.0f : acc[3 + k];
    // the denoiser's records start empty with the first progressive frame
    float* gb = V.gbuf ? V.gbuf + kGBufferFloats * (size_t)pix : nullptr;
    float* stt = V.stats ? V.stats + kStatsFloats * (size_t)pix : nullptr;
    if (first && V.sample_offset == 0) {
        if (gb)
            for (int k = 0; k < kGBufferFloats; ++k) gb[k] = 0.0f;
        if (stt)
            for (int k = 0; k < kStatsFloats; ++k) stt[k] = 0.0f;
    }
    for (uint32_t s = 0; s < sg.count; ++s) {
        const uint32_t pp = sg.base + s * V.npix + pix;
        const F3 r = path_value(P, sc, S, pp);
        if (gb) {
            const size_t st = S.fh_stride;
            const float* f = S.fh + pp;
            const uint32_t model = __float_as_uint(f[0]);
            const float alb[3] = {f[st * 1], f[st * 2], f[st * 3]};
            const float nrm[3] = {f[st * 4], f[st * 5], f[st * 6]};
            dn::gbuffer_record(gb, model != 4u, alb, nrm, f[st * 7], f[st * 8], f[st * 9], model);
        }
        if (stt) {
            const float x[3] = {r.x, r.y, r.z};
            dn::stats_record(stt, x);
        }
        ax = ax + r.x; ay = ay + r.y; az = az + r.z;
        if (V.odd) {
            float* hh = ((sg.sample0 + s) & 1u) ? h : h + 3;
            hh[0] = hh[0] + r.x; hh[1] = hh[1] + r.y; hh[2] = hh[2] + r.z;
        }
    }
    if (last) {
        auto finish = [&](float sx, float sy, float sz, float n, float* dst, uint32_t before) {
            float x = sx / n, y = sy / n, z = sz / n;
            if (P.spectral != AR_SPECTRAL_RGB && P.output == AR_OUTPUT_SRGB) {
                const float r = 3.2406f * x + -1.5372f * y + -0.4986f * z;
                const float g = -0.9689f * x + 1.8758f * y + 0.0415f * z;
                const float b = 0.0557f * x + -0.204f * y + 1.057f * z;
                x = maxr(0.0f, r); y = maxr(0.0f, g); z = maxr(0.0f, b);
            }
            if (V.prog && before > 0) {
                // running mean: lerp(target, result, n / (before + n))
                const float total = (float)before + n;
                const float wa = (float)before / total, wb = n / total;
                x = dst[3 * pix + 0] * wa + x * wb;
                y = dst[3 * pix + 1] * wa + y * wb;
                z = dst[3 * pix + 2] * wa + z * wb;
            }
            dst[3 * pix + 0] = x;
            dst[3 * pix + 1] = y;
            dst[3 * pix + 2] = z;
        };
        finish(ax, ay, az, (float)V.spp, V.out, V.prog ? V.sample_offset : 0u);
        if (V.odd) {
            const uint32_t f = V.sample_offset, e = f + V.spp;
            const uint32_t odd_before = f / 2, odd_new = e / 2 - f / 2;
            const uint32_t even_before = (f + 1) / 2, even_new = (e + 1) / 2 - (f + 1) / 2;
            if (odd_new) finish(h[0], h[1], h[2], (float)odd_new, V.odd, odd_before);
            if (even_new) finish(h[3], h[4], h[5], (float)even_new, V.even, even_before);
        }
    } else {
        acc[0] = ax; acc[1] = ay; acc[2] = az;
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[3 + k] = h[k];
    }
}

uint32_t grid_for(size_t n, uint32_t block = kBlock) { return (uint32_t)((n + block - 1) / block); }

} // namespace

// ---------------------------------------------------------------- host side

void free_scene(ar_ctx* ctx) {
    invalidate_graphs(ctx);
    SceneDev& s = ctx->scene;
    for (void* ptr : {(void*)s.prims, (void*)s.mats, (void*)s.light_tri, (void*)s.light_area, (void*)s.light_prob,
                      (void*)s.light_cdf, (void*)s.objs})
        if (ptr) (void)hipFree(ptr);
    s = SceneDev{};
}

void free_render_buffers(ar_ctx* ctx) {
    invalidate_graphs(ctx);
    if (ctx->rb.fh) (void)hipFree(ctx->rb.fh);
    ctx->rb.fh = nullptr;
    ctx->rb.fh_cap = 0;
    if (ctx->rb.block) (void)hipFree(ctx->rb.block);
    if (ctx->rb.accum) (void)hipFree(ctx->rb.accum);
    if (ctx->rb.eps_table) (void)hipFree(ctx->rb.eps_table);
    if (ctx->rb.cie) (void)hipFree(ctx->rb.cie);
    if (ctx->rb.d65) (void)hipFree(ctx->rb.d65);
    ctx->rb = RenderBuffers{};
}

static float tri_area_host(const float* a, const float* b, const float* c) {
    const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const float cx = e1[1] * e2[2] - e1[2] * e2[1];
    const float cy = e1[2] * e2[0] - e1[0] * e2[2];
    const float cz = e1[0] * e2[1] - e1[1] * e2[0];
    return 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
}

// Uploads prim-order triangles, materials and the light table (the light
// selection CDF of core/scene.cu:677-701: w = max(area * lum(Le), 1e-20),
// prob = w / sum(w), cdf = running FP32 sum of prob).
int upload_scene(ar_ctx* ctx) {
    free_scene(ctx);
    SceneDev& s = ctx->scene;
    const uint32_t n = (uint32_t)(ctx->idx.size() / 3);
    s.ntri = n;
    std::vector<float> prims(12 * (size_t)std::max(n, 1u), 0.0f);
    std::vector<int32_t> ltri;
    std::vector<float> larea, lprob, lcdf;
    float total = 0.0f;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t t = 0; t < n; ++t) {
        const float* v[3];
        for (int k = 0; k < 3; ++k) v[k] = &ctx->verts[3 * (size_t)ctx->idx[3 * (size_t)t + k]];
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], v[k][a]); hi[a] = std::max(hi[a], v[k][a]); }
        const ar_material& m = ctx->mats[ctx->tri_mat[t]];
        int32_t light = -1;
        if (m.emission[0] > 0.0f || m.emission[1] > 0.0f || m.emission[2] > 0.0f) {
            const float area = tri_area_host(v[0], v[1], v[2]);
            const float l = m.emission[0] * 0.2126f + m.emission[1] * 0.7152f + m.emission[2] * 0.0722f;
            const float w = std::max(area * l, 1.0e-20f);
            light = (int32_t)ltri.size();
            ltri.push_back((int32_t)t);
            larea.push_back(area);
            lprob.push_back(w);
            total += w;
        }
        for (int k = 0; k < 3; ++k) {
            float* dst = &prims[12 * (size_t)t + 4 * k];
            dst[0] = v[k][0]; dst[1] = v[k][1]; dst[2] = v[k][2];
        }
        std::memcpy(&prims[12 * (size_t)t + 3], &ctx->tri_mat[t], 4);
        std::memcpy(&prims[12 * (size_t)t + 7], &light, 4);
    }
    // analytic objects: materials after the mesh materials, lights after the triangles
    const uint32_t nobj = (uint32_t)ctx->objs.size();
    uint32_t nsdf = 0;
    std::vector<float4> objs((size_t)obj::kRecord * std::max(nobj, 1u), make_float4(0, 0, 0, 0));
    for (uint32_t k = 0; k < nobj; ++k) {
        const ar_object& ob = ctx->objs[k];
        const float e[3] = {std::fabs(ob.extent[0]), std::fabs(ob.extent[1]), std::fabs(ob.extent[2])};
        const obj::Basis bs = obj::basis_from_euler(ob.rotation);
        float plo[3], phi[3];
        obj::proxy(ob.type, ob.center, e, bs, plo, phi);
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], plo[a]); hi[a] = std::max(hi[a], phi[a]); }
        const ar_material& m = ob.material;
        int32_t light = -1;
        if (m.emission[0] > 0.0f || m.emission[1] > 0.0f || m.emission[2] > 0.0f) {
            const float ar = obj::area(ob.type, e);
            const float l = m.emission[0] * 0.2126f + m.emission[1] * 0.7152f + m.emission[2] * 0.0722f;
            const float w = std::max(ar * l, 1.0e-20f);
            light = (int32_t)ltri.size();
            ltri.push_back((int32_t)(n + k));
            larea.push_back(ar);
            lprob.push_back(w);
            total += w;
        }
        const uint32_t mat_id = (uint32_t)ctx->mats.size() + k;
        const uint32_t solo = k < ctx->bvh.obj_slot.size() ? make_leaf_ref(ctx->bvh.obj_slot[k], 1) : kRefEmpty;
        float4* r = &objs[(size_t)obj::kRecord * k];
        uint32_t tb = ob.type;
        r[0] = make_float4(ob.center[0], ob.center[1], ob.center[2], 0.0f);
        std::memcpy(&r[0].w, &tb, 4);
        r[1] = make_float4(e[0], e[1], e[2], std::fmax(std::fmax(e[0], e[1]), e[2]));
        r[2] = make_float4(bs.w2o[0], bs.w2o[1], bs.w2o[2], 0.0f);
        std::memcpy(&r[2].w, &mat_id, 4);
        r[3] = make_float4(bs.w2o[3], bs.w2o[4], bs.w2o[5], 0.0f);
        std::memcpy(&r[3].w, &light, 4);
        r[4] = make_float4(bs.w2o[6], bs.w2o[7], bs.w2o[8], 0.0f);
        std::memcpy(&r[4].w, &solo, 4);
        r[5] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (ob.type == AR_OBJ_SDF_TORUS) {
            r[5] = make_float4(ob.sdf[0], ob.sdf[1], 0.0f, 0.0f);
            nsdf += 1;
        } else if (ob.type == AR_OBJ_SDF_VOLUME) {
            // validated by ar_scene_set_objects
            const uint32_t v = (uint32_t)ob.sdf_volume;
            const uint32_t* dm = &ctx->sdf.dims[3 * (size_t)v];
            const uint32_t nxy = dm[0] | (dm[1] << 16);
            // voxelSizeMax_ = |1 / max(dims, 1)| * SDF_VOLUME_USE_NEAREST_SCALE (sdf_volume.cuh:23-45)
            const float ix = 1.0f / std::fmax((float)dm[0], 1.0f), iy = 1.0f / std::fmax((float)dm[1], 1.0f),
                        iz = 1.0f / std::fmax((float)dm[2], 1.0f);
            r[5] = make_float4(0.0f, 0.0f, 0.0f, std::sqrt(ix * ix + iy * iy + iz * iz) * 2.5f);
            std::memcpy(&r[5].x, &ctx->sdf.off[v], 4);
            std::memcpy(&r[5].y, &nxy, 4);
            std::memcpy(&r[5].z, &dm[2], 4);
            nsdf += 1;
        }
    }
    const bool any_prim = n + nobj > 0;
    for (int a = 0; a < 3; ++a) {
        const float ext = any_prim ? hi[a] - lo[a] : 0.0f;
        s.box_lo[a] = any_prim ? lo[a] : 0.0f;
        s.box_inv[a] = ext > 0.0f ? 1.0f / ext : 1.0f;
    }
    float acc = 0.0f;
    for (size_t i = 0; i < lprob.size(); ++i) {
        lprob[i] = lprob[i] / total;
        acc += lprob[i];
        lcdf.push_back(acc);
    }
    std::vector<GpuMaterial> mats(ctx->mats.size() + nobj);
    for (size_t i = 0; i < mats.size(); ++i) {
        const ar_material& m = i < ctx->mats.size() ? ctx->mats[i] : ctx->objs[i - ctx->mats.size()].material;
        GpuMaterial g{};
        for (int k = 0; k < 3; ++k) { g.albedo[k] = m.albedo[k]; g.emission[k] = m.emission[k]; }
        g.specular = m.specular; g.metallic = m.metallic; g.roughness = m.roughness; g.model = m.model;
        g.ior = m.ior; g.sigma_a = m.sigma_a; g.abbe = m.abbe; g.tilt = m.tilt;
        mats[i] = g;
    }
    s.nlights = (uint32_t)ltri.size();
    s.nobj = nobj;
    s.nsdf = nsdf;
    s.host_objs = objs;
    s.host_mats = mats;
    s.spectral_ready = false;
    auto up = [&](void** dst, const void* src, size_t bytes, const char* what) -> int {
        if (bytes == 0) bytes = 16;
        int rc = check_hip(ctx, hipMalloc(dst, bytes), what);
        if (rc != AR_OK) { *dst = nullptr; return rc; }
        if (src) rc = check_hip(ctx, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), what);
        return rc;
    };
    int rc = up((void**)&s.prims, prims.data(), prims.size() * sizeof(float), "upload prims");
    if (rc == AR_OK) rc = up((void**)&s.mats, mats.data(), mats.size() * sizeof(GpuMaterial), "upload materials");
    if (rc == AR_OK) rc = up((void**)&s.light_tri, ltri.empty() ? nullptr : ltri.data(), ltri.size() * 4, "upload lights");
    if (rc == AR_OK) rc = up((void**)&s.light_area, larea.empty() ? nullptr : larea.data(), larea.size() * 4, "upload lights");
    if (rc == AR_OK) rc = up((void**)&s.light_prob, lprob.empty() ? nullptr : lprob.data(), lprob.size() * 4, "upload lights");
    if (rc == AR_OK) rc = up((void**)&s.light_cdf, lcdf.empty() ? nullptr : lcdf.data(), lcdf.size() * 4, "upload lights");
    if (rc == AR_OK) rc = up((void**)&s.objs, objs.data(), objs.size() * sizeof(float4), "upload objects");
    if (rc == AR_OK) rc = sdf_upload(ctx);
    if (rc != AR_OK) free_scene(ctx);
    return rc;
}

static int ensure_render_buffers(ar_ctx* ctx, size_t cap, size_t npix) {
    RenderBuffers& b = ctx->rb;
    if (!b.eps_table) {
        float eps[256];
        for (int d = 0; d < 256; ++d) eps[d] = 1.0e-3f * std::exp(std::min((float)d * 0.5f, 4.0f));  // INTERSECT_EPS
        int rc = check_hip(ctx, hipMalloc((void**)&b.eps_table, sizeof eps), "hipMalloc(eps)");
        if (rc != AR_OK) { b.eps_table = nullptr; return rc; }
        rc = check_hip(ctx, hipMemcpy(b.eps_table, eps, sizeof eps, hipMemcpyHostToDevice), "upload eps");
        if (rc != AR_OK) return rc;
        rc = check_hip(ctx, hipMalloc((void**)&b.cie, 421 * 3 * sizeof(float)), "hipMalloc(cie)");
        if (rc != AR_OK) { b.cie = nullptr; return rc; }
        rc = check_hip(ctx, hipMemcpy(b.cie, spec::cie_xyz_table(), 421 * 3 * sizeof(float), hipMemcpyHostToDevice),
                       "upload cie");
        if (rc != AR_OK) return rc;
        rc = check_hip(ctx, hipMalloc((void**)&b.d65, 421 * sizeof(float)), "hipMalloc(d65)");
        if (rc != AR_OK) { b.d65 = nullptr; return rc; }
        rc = check_hip(ctx, hipMemcpy(b.d65, spec::d65_table(), 421 * sizeof(float), hipMemcpyHostToDevice), "upload d65");
        if (rc != AR_OK) return rc;
    }
    if (b.accum_pixels < npix) {
        ctx->buffer_generation += 1;
        if (b.accum) (void)hipFree(b.accum);
        b.accum = nullptr;
        int rc = check_hip(ctx, hipMalloc((void**)&b.accum, npix * kAccumFloats * sizeof(float)), "hipMalloc(accum)");
        if (rc != AR_OK) { b.accum_pixels = 0; return rc; }
        b.accum_pixels = npix;
    }
    if (b.capacity >= cap) return AR_OK;
    if (b.block) (void)hipFree(b.block);
    b.block = nullptr;
    b.capacity = 0;
    // 4-byte arrays of cap entries: 2x8 queue rays, 2 qpath, 5 hits, 8 shadow rays,
    // spath, 4 scontrib, sflags, 4 mask, 4 rad, 3 last_pos, last_pdf, meta, lambda0,
    // pix, smp, scn, 2 qroot, sroot, med, 3 mpos, smeps = 62
    const size_t arrays = 62;
    ctx->buffer_generation += 1;
    const size_t occ_words = (cap + 31) / 32 + 64;
    const size_t bytes = arrays * cap * 4 + occ_words * 4 + 64;
    int rc = check_hip(ctx, hipMalloc(&b.block, bytes), "hipMalloc(render state)");
    if (rc != AR_OK) { b.block = nullptr; return rc; }
    char* base = (char*)b.block;
    auto take = [&](size_t count) { void* r = base; base += count * 4; return r; };
    for (int k = 0; k < 2; ++k)
        for (int f = 0; f < 8; ++f) b.rays[k][f] = (float*)take(cap);
    b.qpath[0] = (uint32_t*)take(cap);
    b.qpath[1] = (uint32_t*)take(cap);
    b.hit_prim = (uint32_t*)take(cap);
    b.hit_t = (float*)take(cap);
    b.hit_u = (float*)take(cap);
    b.hit_v = (float*)take(cap);
    b.hit_flags = (uint32_t*)take(cap);
    for (int f = 0; f < 8; ++f) b.srays[f] = (float*)take(cap);
    b.spath = (uint32_t*)take(cap);
    for (int k = 0; k < 4; ++k) b.scontrib[k] = (float*)take(cap);
    b.sflags = (uint32_t*)take(cap);
    for (int k = 0; k < 4; ++k) b.mask[k] = (float*)take(cap);
    for (int k = 0; k < 4; ++k) b.rad[k] = (float*)take(cap);
    for (int k = 0; k < 3; ++k) b.last_pos[k] = (float*)take(cap);
    b.last_pdf = (float*)take(cap);
    b.meta = (uint32_t*)take(cap);
    b.lambda0 = (float*)take(cap);
    b.pix = (uint32_t*)take(cap);
    b.smp = (uint32_t*)take(cap);
    b.scn = (uint32_t*)take(cap);
    b.qroot[0] = (uint32_t*)take(cap);
    b.qroot[1] = (uint32_t*)take(cap);
    b.sroot = (uint32_t*)take(cap);
    b.med = (uint32_t*)take(cap);
    for (int k = 0; k < 3; ++k) b.mpos[k] = (float*)take(cap);
    b.smeps = (float*)take(cap);
    b.occ = (uint32_t*)take(occ_words);
    b.counters = (uint32_t*)take(16);
    b.capacity = cap;
    return AR_OK;
}

// Uplift records of every material: sigmoid coefficients (albedo; emission
// normalised by its largest channel when that exceeds 1, as Interact does,
// spectrum_lut.cuh:223-235), or with the basis network the RGB bases that
// QueryRGBBasis returns for them, computed on the device from the tables.
static int prepare_spectral_materials(ar_ctx* ctx) {
    SceneDev& s = ctx->scene;
    const uint32_t basis = ctx->smlp.basis;
    if (s.spectral_ready && s.spectral_basis == basis) return AR_OK;
    const size_t bytes = s.host_mats.size() * sizeof(GpuMaterial);
    if (basis == AR_BASIS_NETWORK) {
        if (!ctx->smlp.rgb_lut || !ctx->smlp.lam_lut)
            return set_error(ctx, AR_E_INVALID, "basis network tables not built (ar_spectral_mlp_build_tables)");
        int rc = check_hip(ctx, hipMemcpy(s.mats, s.host_mats.data(), bytes, hipMemcpyHostToDevice),
                           "upload materials");
        if (rc == AR_OK) rc = smlp_material_bases(ctx, s.mats, (uint32_t)s.host_mats.size());
        if (rc == AR_OK)
            rc = check_hip(ctx, hipMemcpy(s.host_mats.data(), s.mats, bytes, hipMemcpyDeviceToHost),
                           "read material bases");
        if (rc != AR_OK) return rc;
    } else {
        const std::vector<float>& lut = spec::uplift_lut(spec::cie_xyz_table(), spec::d65_table());
        for (GpuMaterial& g : s.host_mats) {
            spec::uplift(lut.data(), g.albedo, g.calb);
            g.calb[3] = 0.0f;
            g.ascale = 1.0f;
            const float mx = std::max(g.emission[0], std::max(g.emission[1], g.emission[2]));
            float e[3] = {g.emission[0], g.emission[1], g.emission[2]};
            g.escale = 1.0f;
            if (mx > 1.0f) {
                for (float& x : e) x /= mx;
                g.escale = mx;
            }
            spec::uplift(lut.data(), e, g.cem);
            g.cem[3] = 0.0f;
        }
        int rc = check_hip(ctx, hipMemcpy(s.mats, s.host_mats.data(), bytes, hipMemcpyHostToDevice),
                           "upload spectral materials");
        if (rc != AR_OK) return rc;
    }
    s.spectral_ready = true;
    s.spectral_basis = basis;
    return AR_OK;
}

// ------------------------------------------------------------ render plans
//
// A render (one scene, or a batch of scenes merged into one context) is
// planned on the host as passes of at most `cap` paths; each pass is a list
// of segments, each a run of whole samples of one scene.  Scenes are packed
// in order, so every scene's samples are accumulated in sample order and a
// batched image equals the scene's own render bit for bit.

struct RenderPlan {
    std::vector<SceneView> views;
    std::vector<Segment> segs;       // all passes, concatenated
    struct Pass { uint32_t seg0, nseg, paths, pixels, sample_base; };
    std::vector<Pass> passes;
    size_t cap = 0;                  // largest pass (paths)
    size_t pixels = 0;               // accumulation pixels (all scenes)
    size_t pass_pixels = 0;          // largest pass (pixels)
};

struct RenderJob {
    RenderPlan plan;
    void* dev = nullptr;             // device copy of views + segments
    size_t dev_bytes = 0;
    void* host = nullptr;            // pinned staging copy (stream-ordered uploads)
    size_t host_bytes = 0;
    hipEvent_t uploaded = nullptr;   // the last upload from `host` has been read
    RenderParams base{};
    bool guiding = false, learned = false;
};

// The first object of [k0, k1) containing the camera (Overlap,
// intersection.cuh:312-327; CalculateRadiance:536-538): camera rays start
// inside it and trace it alone.
static void camera_medium(const ar_ctx* ctx, const float pos[3], uint32_t k0, uint32_t k1, SceneView& v) {
    v.medium0 = 0;
    v.root0 = v.root;
    const SceneDev& s = ctx->scene;
    for (uint32_t k = k0; k < k1 && k < s.nobj; ++k) {
        const float4* r = &s.host_objs[(size_t)obj::kRecord * k];
        if (obj::overlap_host(r, pos, ctx->sdf.host.data())) {
            v.medium0 = k + 1;
            std::memcpy(&v.root0, &r[4].w, 4);
            return;
        }
    }
}

static void camera_basis(const ar_camera* cam, uint32_t width, uint32_t height, SceneView& v) {
    // the reference computes the basis per pixel (core/scene.cu:154-164):
    // FOV = vfov * 3.1415926 / 180
    const float fx = cam->dir[0], fy = cam->dir[1], fz = cam->dir[2];
    float rx = fy * 0.0f - fz * 1.0f, ry = fz * 0.0f - fx * 0.0f, rz = fx * 1.0f - fy * 0.0f;
    float l2 = rx * rx + ry * ry + rz * rz;
    float inv = l2 <= 1e-20f ? 0.0f : 1.0f / std::sqrt(l2);
    rx *= inv; ry *= inv; rz *= inv;
    if (l2 <= 1e-20f) { rx = ry = rz = 0.0f; }
    float ux = fy * rz - fz * ry, uy = fz * rx - fx * rz, uz = fx * ry - fy * rx;
    l2 = ux * ux + uy * uy + uz * uz;
    inv = l2 <= 1e-20f ? 0.0f : 1.0f / std::sqrt(l2);
    ux *= inv; uy *= inv; uz *= inv;
    for (int k = 0; k < 3; ++k) { v.cam_pos[k] = cam->pos[k]; v.fwd[k] = cam->dir[k]; }
    v.right[0] = rx; v.right[1] = ry; v.right[2] = rz;
    v.up[0] = ux; v.up[1] = uy; v.up[2] = uz;
    const float fov = cam->vfov_deg * 3.1415926f / 180.0f;
    const float aspect = (float)width / std::max(1.0f, (float)height);
    v.thy = std::tan(0.5f * fov);
    v.thx = v.thy * aspect;
}

// shared settings must agree across a batch
static bool same_settings(const ar_render_opts& a, const ar_render_opts& b) {
    return a.seed == b.seed && a.max_depth == b.max_depth && a.rr_min_depth == b.rr_min_depth && a.nee == b.nee &&
           a.sky == b.sky && a.env[0] == b.env[0] && a.env[1] == b.env[1] && a.env[2] == b.env[2] &&
           a.spectral == b.spectral && a.output == b.output && a.guiding == b.guiding && a.frame_spp == b.frame_spp &&
           a.frame_growth == b.frame_growth && a.max_paths_per_pass == b.max_paths_per_pass;
}

static int make_job(ar_ctx* ctx, uint32_t n, const ar_camera* cams, const ar_render_opts* opts, float* const* outs,
                    RenderJob& J) {
    const ar_render_opts& o = opts[0];
    if (o.guiding != AR_GUIDING_OFF && o.guiding != AR_GUIDING_VOXEL && o.guiding != AR_GUIDING_MLP)
        return set_error(ctx, AR_E_INVALID, "guiding must be AR_GUIDING_OFF, _VOXEL or _MLP");
    if (o.spectral != AR_SPECTRAL_RGB && o.spectral != AR_SPECTRAL_SINGLE && o.spectral != AR_SPECTRAL_HERO4)
        return set_error(ctx, AR_E_INVALID, "spectral must be AR_SPECTRAL_RGB, _SINGLE or _HERO4");
    if (o.spectral != AR_SPECTRAL_RGB && o.sky == AR_SKY_GRADIENT)
        return set_error(ctx, AR_E_UNSUPPORTED, "the gradient sky is RGB-only; use AR_SKY_CONSTANT in spectral modes");
    J.guiding = o.guiding != AR_GUIDING_OFF;
    J.learned = o.guiding == AR_GUIDING_MLP;
    if (J.guiding && n > mlp::kMaxGuidedScenes)
        return set_error(ctx, AR_E_CAPACITY, "guided batches hold at most 256 scenes");
    for (uint32_t i = 1; i < n && J.guiding; ++i)
        if (opts[i].sample_offset != o.sample_offset)
            return set_error(ctx, AR_E_INVALID, "guided batch scenes must share sample_offset (their frames run together)");
    const size_t max_paths = o.max_paths_per_pass ? o.max_paths_per_pass : ((size_t)1 << 24);
    RenderPlan& L = J.plan;
    L = RenderPlan{};
    size_t max_npix = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const ar_render_opts& oi = opts[i];
        if (i > 0 && !same_settings(o, oi))
            return set_error(ctx, AR_E_INVALID, "batched scenes must share seed, depth, NEE, sky, spectral and guiding settings");
        if (oi.width == 0 || oi.height == 0 || oi.spp == 0) return set_error(ctx, AR_E_INVALID, "empty image or spp = 0");
        if (oi.max_depth == 0 || oi.max_depth > 1000) return set_error(ctx, AR_E_INVALID, "max_depth must be in [1, 1000]");
        const size_t npix = (size_t)oi.width * oi.height;
        if (npix > ((size_t)1 << 26)) return set_error(ctx, AR_E_CAPACITY, "image larger than 2^26 pixels");
        if ((uint64_t)oi.sample_offset + oi.spp > 0xFFFFFFFFull) return set_error(ctx, AR_E_INVALID, "sample index overflow");
        if (!outs[i]) return set_error(ctx, AR_E_INVALID, "null output image");
        SceneView v{};
        v.width = oi.width; v.height = oi.height; v.npix = (uint32_t)npix; v.spp = oi.spp;
        v.sample_offset = oi.sample_offset;
        v.out_base = (uint32_t)L.pixels;
        if (ctx->batch.empty()) {
            v.root = ctx->bvh.root; v.light_off = 0; v.light_cnt = ctx->scene.nlights;
            camera_medium(ctx, cams[i].pos, 0, ctx->scene.nobj, v);
            for (int a = 0; a < 3; ++a) { v.box_lo[a] = ctx->scene.box_lo[a]; v.box_inv[a] = ctx->scene.box_inv[a]; }
        } else {
            const BatchEntry& be = ctx->batch[i];
            v.root = be.root; v.light_off = be.light_off; v.light_cnt = be.light_cnt;
            camera_medium(ctx, cams[i].pos, be.obj_off, be.obj_off + be.obj_cnt, v);
            for (int a = 0; a < 3; ++a) { v.box_lo[a] = be.box_lo[a]; v.box_inv[a] = be.box_inv[a]; }
        }
        camera_basis(&cams[i], oi.width, oi.height, v);
        v.out = outs[i];
        L.views.push_back(v);
        L.pixels += npix;
        max_npix = std::max(max_npix, npix);
        if (L.pixels > 0xFFFFFFFFull) return set_error(ctx, AR_E_CAPACITY, "batch larger than 2^32 pixels");
    }
    // guiding: frames of frame_spp samples per pixel of every scene (times
    // frame_growth from one frame to the next) with a guide update after
    // each; one pass per frame, of at most 2^26 paths
    const uint32_t frame_spp = std::max(1u, o.frame_spp);
    const uint32_t growth = std::max(1u, o.frame_growth);
    constexpr size_t kMaxFramePaths = (size_t)1 << 26;
    if (J.guiding && L.pixels * frame_spp > kMaxFramePaths)
        return set_error(ctx, AR_E_CAPACITY, "guided frame larger than 2^26 paths");
    const size_t cap = J.guiding ? (size_t)L.pixels * frame_spp : std::max(max_paths, max_npix);
    RenderPlan::Pass cur{0, 0, 0, 0, 0};
    auto close = [&]() {
        if (cur.nseg) {
            L.passes.push_back(cur);
            L.cap = std::max<size_t>(L.cap, cur.paths);
            L.pass_pixels = std::max<size_t>(L.pass_pixels, cur.pixels);
        }
        cur = RenderPlan::Pass{(uint32_t)L.segs.size(), 0, 0, 0, 0};
    };
    if (J.guiding) {
        // frame f: samples f * frame_spp .. of every scene that has them
        uint32_t max_spp = 0;
        for (const SceneView& v : L.views) max_spp = std::max(max_spp, v.spp);
        uint32_t fs = frame_spp;
        for (uint32_t done = 0; done < max_spp;) {
            size_t px = 0;   // pixels of the scenes this frame renders
            for (const SceneView& v : L.views)
                if (done < v.spp) px += v.npix;
            const uint32_t f = (uint32_t)std::max<size_t>(frame_spp, std::min<size_t>(fs, kMaxFramePaths / px));
            for (uint32_t i = 0; i < n; ++i) {
                const SceneView& v = L.views[i];
                if (done >= v.spp) continue;
                const uint32_t k = std::min(f, v.spp - done);
                Segment sg{};
                sg.scene = i; sg.sample0 = v.sample_offset + done; sg.count = k;
                sg.base = cur.paths; sg.pixbase = cur.pixels;
                if (cur.nseg == 0) cur.sample_base = sg.sample0;
                L.segs.push_back(sg);
                cur.nseg += 1;
                cur.paths += k * v.npix;
                cur.pixels += v.npix;
            }
            close();
            done += f;
            fs = (uint32_t)std::min<size_t>((size_t)fs * growth, kMaxFramePaths);
        }
    }
    for (uint32_t i = 0; i < n && !J.guiding; ++i) {
        const SceneView& v = L.views[i];
        uint32_t done = 0;
        while (done < v.spp) {
            const size_t fit = (cap - cur.paths) / v.npix;
            if (fit == 0) { close(); continue; }
            const uint32_t k = (uint32_t)std::min<size_t>(fit, v.spp - done);
            Segment sg{};
            sg.scene = i; sg.sample0 = v.sample_offset + done; sg.count = k;
            sg.base = cur.paths; sg.pixbase = cur.pixels;
            if (cur.nseg == 0) cur.sample_base = sg.sample0;
            L.segs.push_back(sg);
            cur.nseg += 1;
            cur.paths += k * v.npix;
            cur.pixels += v.npix;
            done += k;
        }
    }
    close();
    RenderParams& P = J.base;
    P = RenderParams{};
    P.seed = o.seed;
    P.max_depth = o.max_depth;
    P.rr_min_depth = o.rr_min_depth;
    P.nee = o.nee;
    P.sky = o.sky;
    P.spectral = o.spectral;
    P.output = o.output;
    P.guiding = o.guiding;
    P.batched = ctx->batch.empty() ? 0u : 1u;
    float e[3] = {o.env[0], o.env[1], o.env[2]};
    for (int k = 0; k < 3; ++k) P.env[k] = o.env[k];
    const float mx = std::max(e[0], std::max(e[1], e[2]));
    P.env_scale = 1.0f;
    if (mx > 1.0f) { for (float& x : e) x /= mx; P.env_scale = mx; }
    if (o.spectral != AR_SPECTRAL_RGB) {
        const std::vector<float>& lut = spec::uplift_lut(spec::cie_xyz_table(), spec::d65_table());
        spec::uplift(lut.data(), e, P.env_c);
    }
    return AR_OK;
}

// allocations and uploads; nothing here may be captured into a graph.
// stream == nullptr with sync: synchronous upload (graphs); else the plan is
// copied on `stream`, ordered after the previous render on that stream.
static bool wants_first_hits(const RenderPlan& L) {
    for (const SceneView& v : L.views)
        if (v.gbuf) return true;
    return false;
}

static int prepare_job(ar_ctx* ctx, RenderJob& J, hipStream_t stream, bool sync) {
    (void)hipSetDevice(ctx->device);
    const RenderPlan& L = J.plan;
    int rc = ensure_render_buffers(ctx, L.cap, L.pixels);
    if (rc != AR_OK) return rc;
    if (ctx->count_on && ctx->cnt_cap < L.cap) {
        invalidate_graphs(ctx);
        if (ctx->cnt_box) (void)hipFree(ctx->cnt_box);
        ctx->cnt_box = ctx->cnt_tri = nullptr;
        ctx->cnt_cap = 0;
        rc = check_hip(ctx, hipMalloc((void**)&ctx->cnt_box, 2 * L.cap * sizeof(uint32_t)), "hipMalloc(test counts)");
        if (rc != AR_OK) return rc;
        ctx->cnt_tri = ctx->cnt_box + L.cap;
        ctx->cnt_cap = L.cap;
    }
    if (wants_first_hits(L)) {
        // per-path first-hit records for the denoiser's G-buffer
        RenderBuffers& b = ctx->rb;
        if (b.fh_cap < b.capacity) {
            if (b.fh) (void)hipFree(b.fh);
            b.fh = nullptr;
            b.fh_cap = 0;
            rc = check_hip(ctx, hipMalloc((void**)&b.fh, b.capacity * kFirstHitFloats * sizeof(float)),
                           "hipMalloc(first hits)");
            if (rc != AR_OK) return rc;
            b.fh_cap = b.capacity;
        }
    }
    if (J.guiding) {
        rc = guide_ensure(ctx, L.cap, !J.learned);
        if (rc != AR_OK) return rc;
    }
    if (J.learned) {
        rc = mlp_ensure(ctx);
        if (rc == AR_OK) rc = mlp_ensure_buffers(ctx, L.cap);
        if (rc != AR_OK) return rc;
        // the learned shading kernel's LDS rows: above 64 KB for blocks of more than 334 threads
        rc = check_hip(ctx,
                       hipFuncSetAttribute(reinterpret_cast<const void*>(&k_shade<AR_GUIDING_MLP>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(kMaxBlock * kMixRow * sizeof(float))),
                       "hipFuncSetAttribute(k_shade)");
        if (rc != AR_OK) return rc;
    }
    if (J.base.spectral != AR_SPECTRAL_RGB) {
        rc = prepare_spectral_materials(ctx);
        if (rc != AR_OK) return rc;
    }
    TraceArgs ta;
    uint32_t tblocks;
    rc = trace_prologue(ctx, ta, L.cap, tblocks);
    if (rc != AR_OK) return rc;
    const size_t vb = L.views.size() * sizeof(SceneView);
    const size_t sb = L.segs.size() * sizeof(Segment);
    // per segment: its first path and path count (the learned guide's record selection)
    std::vector<uint32_t> ranges(2 * L.segs.size());
    for (size_t k = 0; k < L.segs.size(); ++k) {
        ranges[2 * k] = L.segs[k].base;
        ranges[2 * k + 1] = L.segs[k].count * L.views[L.segs[k].scene].npix;
    }
    const size_t rb = ranges.size() * sizeof(uint32_t);
    const size_t bytes = vb + sb + rb;
    if (J.dev_bytes < bytes) {
        if (J.dev) (void)hipFree(J.dev);
        J.dev = nullptr;
        J.dev_bytes = 0;
        rc = check_hip(ctx, hipMalloc(&J.dev, bytes), "hipMalloc(render plan)");
        if (rc != AR_OK) return rc;
        J.dev_bytes = bytes;
    }
    if (sync) {
        rc = check_hip(ctx, hipMemcpy(J.dev, L.views.data(), vb, hipMemcpyHostToDevice), "upload render plan");
        if (rc == AR_OK)
            rc = check_hip(ctx, hipMemcpy((char*)J.dev + vb, L.segs.data(), sb, hipMemcpyHostToDevice),
                           "upload render plan");
        if (rc == AR_OK)
            rc = check_hip(ctx, hipMemcpy((char*)J.dev + vb + sb, ranges.data(), rb, hipMemcpyHostToDevice),
                           "upload render plan");
        return rc;
    }
    if (!J.uploaded) {
        rc = check_hip(ctx, hipEventCreateWithFlags(&J.uploaded, hipEventDisableTiming), "hipEventCreate");
        if (rc != AR_OK) return rc;
    } else {
        (void)hipEventSynchronize(J.uploaded);
    }
    if (J.host_bytes < bytes) {
        if (J.host) (void)hipHostFree(J.host);
        J.host = nullptr;
        J.host_bytes = 0;
        rc = check_hip(ctx, hipHostMalloc(&J.host, bytes, hipHostMallocDefault), "hipHostMalloc(render plan)");
        if (rc != AR_OK) return rc;
        J.host_bytes = bytes;
    }
    std::memcpy(J.host, L.views.data(), vb);
    std::memcpy((char*)J.host + vb, L.segs.data(), sb);
    std::memcpy((char*)J.host + vb + sb, ranges.data(), rb);
    rc = check_hip(ctx, hipMemcpyAsync(J.dev, J.host, bytes, hipMemcpyHostToDevice, stream), "upload render plan");
    if (rc == AR_OK) rc = check_hip(ctx, hipEventRecord(J.uploaded, stream), "hipEventRecord");
    return rc;
}

// kernel launches only (graph-capturable)
static int enqueue_job(ar_ctx* ctx, RenderJob& J, hipStream_t stream) {
    const RenderPlan& L = J.plan;
    TraceArgs ta;
    uint32_t tblocks;
    int rc = trace_prologue(ctx, ta, L.cap, tblocks);
    if (rc != AR_OK) return rc;
    RenderBuffers& b = ctx->rb;
    const uint32_t blk = ctx->tile;
    const bool fh = wants_first_hits(L);
    PathArgs S{b.mask[0], b.mask[1], b.mask[2], b.mask[3], b.rad[0], b.rad[1], b.rad[2], b.rad[3],
               b.last_pos[0], b.last_pos[1], b.last_pos[2], b.last_pdf, b.meta, b.lambda0, b.pix, b.smp, b.scn,
               b.med, b.mpos[0], b.mpos[1], b.mpos[2], fh ? b.fh : nullptr, (uint32_t)b.fh_cap};
    const bool sdf = ctx->scene.nsdf > 0;
    const bool net = ctx->smlp.basis == AR_BASIS_NETWORK && J.base.spectral != AR_SPECTRAL_RGB;
    SceneArgs sc{ctx->scene.prims, ctx->scene.objs, ctx->scene.ntri, ctx->scene.mats, ctx->scene.light_tri,
                 ctx->scene.light_area, ctx->scene.light_prob, ctx->scene.light_cdf, b.cie, b.d65, ctx->sdf.dev,
                 sdf ? 1u : 0u, net ? 1u : 0u, ctx->smlp.lam_lut, ctx->smlp.lam_res, ctx->smlp.rgb_lut,
                 ctx->smlp.rgb_res};
    QueueRays q[2] = {{b.rays[0][0], b.rays[0][1], b.rays[0][2], b.rays[0][3], b.rays[0][4], b.rays[0][5], b.rays[0][6],
                       b.rays[0][7], b.qroot[0], nullptr},
                      {b.rays[1][0], b.rays[1][1], b.rays[1][2], b.rays[1][3], b.rays[1][4], b.rays[1][5], b.rays[1][6],
                       b.rays[1][7], b.qroot[1], nullptr}};
    QueueRays shq{b.srays[0], b.srays[1], b.srays[2], b.srays[3], b.srays[4], b.srays[5], b.srays[6], b.srays[7],
                  b.sroot, sdf ? b.smeps : nullptr};
    ar_hits hits{b.hit_prim, b.hit_t, b.hit_u, b.hit_v, b.hit_flags};
    GuideDev& gd = ctx->guide;
    GuideArgs G{};
    if (J.guiding) {
        G.pr = guide::Probes{gd.tkey, gd.tval, gd.tcap ? gd.tcap - 1 : 0u, gd.tex, 0u};
        G.probe_count = gd.counters;
        G.gm = gd.gm; G.gr = gd.gr; G.gkey = gd.gkey; G.gleaf = gd.gleaf; G.gflag = gd.gflag;
        G.stride = (uint32_t)gd.pcap;
        G.dec = ctx->mlp.dec;
        G.feat = ctx->mlp.feat;
        G.gx = gd.gx; G.gw = gd.gw; G.gp = gd.gp;
    }
    RenderParams P = J.base;
    P.views = (const SceneView*)J.dev;
    const Segment* segs = (const Segment*)((const char*)J.dev + L.views.size() * sizeof(SceneView));
    const uint32_t* ranges = (const uint32_t*)(segs + L.segs.size());
    const uint32_t max_depth = P.max_depth;
    for (const RenderPlan::Pass& pass : L.passes) {
        P.segs = segs + pass.seg0;
        P.nseg = pass.nseg;
        P.capacity = pass.paths;
        P.npixels = pass.pixels;
        if (J.guiding) {
            // GetPathGuidingProb (render.cuh:195-200): lerp(0, 0.5, min(1, frame/64)^0.5);
            // the learned guide ramps to kGuideProb instead
            const float t = std::pow(std::fmin(1.0f, (float)pass.sample_base / 64.0f), 0.5f);
            const float amax = J.learned ? mlp::kGuideProb : 0.5f;
            P.alpha = std::fmin(std::fmax(0.0f * (1.0f - t) + amax * t, 0.0f), 1.0f);
        }
        // grid-stride kernels: enough blocks to fill every CU several times over
        const uint32_t stride_grid = std::min<uint32_t>(grid_for(P.capacity, blk), (uint32_t)ctx->num_cus * (2048u / blk) * 2u);
        hipLaunchKernelGGL(k_camera, dim3(grid_for(P.capacity, blk)), dim3(blk), 0, stream, P, S, G, q[0], b.qpath[0],
                           b.counters, b.eps_table);
        int c = 0;
        for (uint32_t depth = 0; depth < max_depth; ++depth) {
            ar_rays rr{q[c].ox, q[c].oy, q[c].oz, q[c].tmin, q[c].dx, q[c].dy, q[c].dz, q[c].tmax, nullptr};
            TraceArgs tq = ta;
            tq.roots = q[c].root;
            const bool count = ctx->count_on && ctx->cnt_cap >= P.capacity;
            hipError_t e = launch_trace_closest(tq, rr, P.capacity, hits, count ? ctx->cnt_box : nullptr,
                                                count ? ctx->cnt_tri : nullptr, tblocks, stream, &b.counters[0]);
            if (e != hipSuccess) return check_hip(ctx, e, "integrator trace");
            if (count)
                hipLaunchKernelGGL(k_count_sum, dim3(stride_grid), dim3(256), 0, stream, ctx->cnt_box, ctx->cnt_tri,
                                   &b.counters[0], ctx->cnt_totals, 1);
            if (J.learned && P.alpha > 0.0f && depth + 1 < max_depth) {
                // (the last bounce samples no direction and takes no NEE sample)
                const SceneDev& sd = ctx->scene;
                hipLaunchKernelGGL(k_mlp_features, dim3(grid_for(P.capacity, blk)), dim3(blk), 0, stream, sd.prims,
                                   sd.objs, ctx->sdf.dev, sd.ntri, q[c], b.qpath[c], b.scn, P.views, b.hit_prim,
                                   b.hit_t, b.hit_u, b.hit_v, b.hit_flags, b.counters, ctx->mlp.feat);
                e = mlp_forward(ctx, ctx->mlp.feat, mlp::kRec, P.capacity, &b.counters[0], ctx->mlp.dec, true, stream);
                if (e != hipSuccess) return check_hip(ctx, e, "guiding network inference");
            }
#define AR_SHADE(GM)                                                                                           \
    hipLaunchKernelGGL(k_shade<GM>, dim3(grid_for(P.capacity, blk)), dim3(blk),                                 \
                       GM == AR_GUIDING_MLP ? blk * kMixRow * sizeof(float) : 0, stream, P, depth, sc, S, G, q[c], \
                       b.qpath[c], b.hit_prim, b.hit_t, b.hit_u, b.hit_v, b.hit_flags, q[1 - c], b.qpath[1 - c], shq,  \
                       b.spath,                                                                                     \
                       b.scontrib[0], b.scontrib[1], b.scontrib[2], b.scontrib[3], b.sflags, b.counters, b.eps_table)
            if (P.guiding == AR_GUIDING_VOXEL) AR_SHADE(AR_GUIDING_VOXEL);
            else if (P.guiding == AR_GUIDING_MLP) AR_SHADE(AR_GUIDING_MLP);
            else AR_SHADE(AR_GUIDING_OFF);
#undef AR_SHADE
            if (P.nee) {
                ar_rays sr{shq.ox, shq.oy, shq.oz, shq.tmin, shq.dx, shq.dy, shq.dz, shq.tmax, shq.meps};
                TraceArgs ts = ta;
                ts.roots = shq.root;
                e = launch_trace_any(ts, sr, P.capacity, b.occ, tblocks, stream, &b.counters[2],
                                     count ? ctx->cnt_box : nullptr, count ? ctx->cnt_tri : nullptr);
                if (e != hipSuccess) return check_hip(ctx, e, "integrator shadow trace");
                if (count)
                    hipLaunchKernelGGL(k_count_sum, dim3(stride_grid), dim3(256), 0, stream, ctx->cnt_box, ctx->cnt_tri,
                                       &b.counters[2], ctx->cnt_totals + 3, 0);
                hipLaunchKernelGGL(k_resolve, dim3(stride_grid), dim3(blk), 0, stream, S, b.spath,
                                   b.scontrib[0], b.scontrib[1], b.scontrib[2], b.scontrib[3], b.sflags, b.occ,
                                   b.counters);
            }
            hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, stream, b.counters);
            c = 1 - c;
        }
        if (J.guiding) {
            rc = J.learned ? mlp_frame_update(ctx, P.spectral, b.lambda0, b.cie, P.capacity, ranges + 2 * pass.seg0,
                                              pass.nseg, stream)
                           : guide_update(ctx, P.spectral, b.lambda0, b.cie, P.capacity, stream);
            if (rc != AR_OK) return rc;
        }
        hipLaunchKernelGGL(k_accumulate, dim3(grid_for(P.npixels, blk)), dim3(blk), 0, stream, P, sc, S, b.accum);
    }
    return check_hip(ctx, hipGetLastError(), "integrator launch");
}

int render_batch(ar_ctx* ctx, uint32_t n, const ar_camera* cams, const ar_render_opts* opts, float* const* outs,
                 hipStream_t stream) {
    if (!ctx->job) ctx->job = new RenderJob();
    RenderJob& J = *ctx->job;
    int rc = make_job(ctx, n, cams, opts, outs, J);
    if (rc == AR_OK) rc = prepare_job(ctx, J, stream, false);
    if (rc == AR_OK) rc = enqueue_job(ctx, J, stream);
    return rc;
}

int render(ar_ctx* ctx, const ar_camera* cam, const ar_render_opts* opts, float* out, hipStream_t stream) {
    return render_batch(ctx, 1, cam, opts, &out, stream);
}

int render_progressive(ar_ctx* ctx, const ar_camera* cam, const ar_render_opts* opts, uint32_t frame, float* accum,
                       float* odd, float* even, float* gbuffer, float* stats, hipStream_t stream) {
    if (!ctx->job) ctx->job = new RenderJob();
    RenderJob& J = *ctx->job;
    ar_render_opts o = *opts;
    o.sample_offset = frame;
    float* out = accum;
    int rc = make_job(ctx, 1, cam, &o, &out, J);
    if (rc != AR_OK) return rc;
    SceneView& v = J.plan.views[0];
    v.prog = 1u;
    v.odd = odd;
    v.even = even;
    v.gbuf = gbuffer;
    v.stats = stats;
    rc = prepare_job(ctx, J, stream, false);
    if (rc == AR_OK) rc = enqueue_job(ctx, J, stream);
    return rc;
}

void free_job(RenderJob* J) {
    if (!J) return;
    if (J->uploaded) { (void)hipEventSynchronize(J->uploaded); (void)hipEventDestroy(J->uploaded); }
    if (J->dev) (void)hipFree(J->dev);
    if (J->host) (void)hipHostFree(J->host);
    delete J;
}

// ------------------------------------------------------------ graphs

int graph_capture(ar_ctx* ctx, uint32_t n, const ar_camera* cams, const ar_render_opts* opts, float* const* outs,
                  GraphRec* g) {
    g->job = new RenderJob();
    int rc = make_job(ctx, n, cams, opts, outs, *g->job);
    if (rc == AR_OK) rc = prepare_job(ctx, *g->job, nullptr, true);
    if (rc != AR_OK) return rc;
    rc = check_hip(ctx, hipDeviceSynchronize(), "graph capture");
    if (rc != AR_OK) return rc;
    hipStream_t cs = nullptr;
    rc = check_hip(ctx, hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "graph capture stream");
    if (rc != AR_OK) return rc;
    rc = check_hip(ctx, hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (rc == AR_OK) {
        const int rce = enqueue_job(ctx, *g->job, cs);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &graph);
        rc = rce != AR_OK ? rce : check_hip(ctx, e, "hipStreamEndCapture");
        if (rc == AR_OK) rc = check_hip(ctx, hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        if (graph) (void)hipGraphDestroy(graph);
    }
    (void)hipStreamDestroy(cs);
    g->generation = ctx->buffer_generation;
    return rc;
}

} // namespace ar
