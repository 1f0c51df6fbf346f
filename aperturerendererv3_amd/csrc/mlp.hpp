// Learned guiding distribution: a 5-64-64-64-40 ReLU MLP whose outputs are an
// 8-lobe von Mises-Fisher mixture over directions (per lobe: logit, log kappa,
// mean direction).  Inputs: the vertex position normalised to the scene box
// and the shading normal as (theta / pi, phi / (2 pi) + 1/2).
//
// Parameters are FP32 masters in the layout W0 16x64, b0, W1 64x64, b1,
// W2 64x64, b2, W3 64x48, b3 (W[in][out]; the 5 inputs (13 with the scene
// bits of a batch) and 40 outputs are zero-padded to 16 and 48 for 16x16x16
// MFMA tiles).  The kernels use BF16
// copies of the weights (row-major for the backward pass, transposed for the
// forward pass) with FP32 accumulation.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace ar {
namespace mlp {

constexpr int kIn = 16, kHid = 64, kOut = 48, kUsedIn = 5, kUsedOut = 40, kLobes = 8;
// Batched renders train one network for all scenes: a vertex record holds
// the 5 features and the scene index, whose 8 bits (0 / 1) feed inputs
// 5..12 of the zero-padded first layer.  A scene-0 record therefore equals
// the plain 5-input network.
constexpr int kSceneBits = 8, kRec = 6;
constexpr uint32_t kMaxGuidedScenes = 1u << kSceneBits;

// input k of a record of stride xs (5: features only, scene 0; 6: + scene)
__host__ __device__ __forceinline__ float record_input(const float* __restrict__ r, int xs, int k) {
    if (k < kUsedIn) return r[k];
    if (xs == kRec && k < kUsedIn + kSceneBits) return (float)(((uint32_t)r[kUsedIn] >> (k - kUsedIn)) & 1u);
    return 0.0f;
}
constexpr int OW0 = 0;
constexpr int OB0 = OW0 + kIn * kHid;
constexpr int OW1 = OB0 + kHid;
constexpr int OB1 = OW1 + kHid * kHid;
constexpr int OW2 = OB1 + kHid;
constexpr int OB2 = OW2 + kHid * kHid;
constexpr int OW3 = OB2 + kHid;
constexpr int OB3 = OW3 + kHid * kOut;
constexpr int kParams = OB3 + kOut;
constexpr int kDecoded = 48;   // per vertex, evaluation form: 8 lobes x (pi, norm, kappa, mu xyz)
// per vertex as inference stores it: 8 lobes x (pi, kappa, mu as two 16-bit
// octahedral coordinates); unpack_mixture turns it into the evaluation form
constexpr int kPacked = 24;
// largest probability of sampling the learned guide in the one-sample
// mixture with the BSDF, reached after 64 samples per pixel as the voxel
// guide's 0.5 is (measured on the occluded-light room: 0.75 lowers the error
// at equal time against 0.5)
constexpr float kGuideProb = 0.75f;
constexpr float kTwoPi = 6.28318530717958647692f;

// network inputs of a vertex
__device__ __forceinline__ void features(float px, float py, float pz, float nx, float ny, float nz, const float lo[3],
                                         const float inv_ext[3], float x[5]) {
    x[0] = (px - lo[0]) * inv_ext[0];
    x[1] = (py - lo[1]) * inv_ext[1];
    x[2] = (pz - lo[2]) * inv_ext[2];
    x[3] = acosf(fminf(1.0f, fmaxf(-1.0f, nz))) * 0.318309886183790672f;
    x[4] = atan2f(ny, nx) * 0.159154943091895336f + 0.5f;
}

// decoded lobe k of a vertex: d[6k..6k+5] = (pi, pi * kappa / (2 pi (1 - e^-2kappa)), kappa, mu)
__device__ __forceinline__ float mixture_pdf(const float* __restrict__ d, float wx, float wy, float wz) {
    float q = 0.0f;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) {
        const float* l = d + 6 * k;
        q += l[1] * expf(l[2] * (l[3] * wx + l[4] * wy + l[5] * wz - 1.0f));
    }
    return q;
}

__device__ __forceinline__ void mixture_sample(const float* __restrict__ d, float u0, float u1, float u2, float& wx,
                                               float& wy, float& wz) {
    int k = 0;
    float c = 0.0f;
    for (; k < kLobes - 1; ++k) {
        c += d[6 * k];
        if (u0 < c) break;
    }
    const float* l = d + 6 * k;
    const float kap = l[2];
    // Jakob 2012: cos = 1 + log(u + (1 - u) e^{-2 kappa}) / kappa
    const float ct = fminf(1.0f, fmaxf(-1.0f, 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * kap)) / kap));
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float sp, cp;
    sincosf(kTwoPi * u2, &sp, &cp);
    const float mx = l[3], my = l[4], mz = l[5];
    const float sg = mz >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sg + mz), b = mx * my * a;
    const float tx = 1.0f + sg * mx * mx * a, ty = sg * b, tz = -sg * mx;
    const float bx = b, by = sg + my * my * a, bz = -my;
    wx = st * cp * tx + st * sp * bx + ct * mx;
    wy = st * cp * ty + st * sp * by + ct * my;
    wz = st * cp * tz + st * sp * bz + ct * mz;
}

// A direction as two signed-normalised 16-bit octahedral coordinates
// (Cigolle et al. 2014): the point (x, y) / (|x| + |y| + |z|), with the lower
// hemisphere folded over the square's diagonals.  w need not be unit length.
__device__ __forceinline__ uint32_t oct_pack(float x, float y, float z) {
    const float s = __builtin_amdgcn_rcpf(fabsf(x) + fabsf(y) + fabsf(z));
    float u = x * s, v = y * s;
    if (z < 0.0f) {
        const float fu = (1.0f - fabsf(v)) * (u >= 0.0f ? 1.0f : -1.0f);
        v = (1.0f - fabsf(u)) * (v >= 0.0f ? 1.0f : -1.0f);
        u = fu;
    }
    const int qu = __float2int_rn(fminf(1.0f, fmaxf(-1.0f, u)) * 32767.0f);
    const int qv = __float2int_rn(fminf(1.0f, fmaxf(-1.0f, v)) * 32767.0f);
    return ((uint32_t)qu & 0xFFFFu) | ((uint32_t)qv << 16);
}
// the unit direction of a packed pair
__device__ __forceinline__ void oct_unpack(uint32_t w, float& x, float& y, float& z) {
    x = (float)(int16_t)(w & 0xFFFFu) * (1.0f / 32767.0f);
    y = (float)(int16_t)(w >> 16) * (1.0f / 32767.0f);
    z = 1.0f - fabsf(x) - fabsf(y);
    const float t = fmaxf(-z, 0.0f);
    x += x >= 0.0f ? -t : t;
    y += y >= 0.0f ? -t : t;
    const float il = __builtin_amdgcn_rsqf(x * x + y * y + z * z);
    x *= il;
    y *= il;
    z *= il;
}
// evaluation form (kDecoded floats) of a stored mixture (kPacked words):
// the normaliser pi kappa / (2 pi (1 - e^-2kappa)) and unit mean directions
// are computed here, once per vertex, from the stored values that sampling
// also reads, so samples and densities describe the same distribution.
__device__ __forceinline__ void unpack_lobe(float pi, float kap, uint32_t oct, float* __restrict__ l) {
    constexpr float kLog2e = 1.4426950408889634f;
    l[0] = pi;
    l[1] = pi * kap * __builtin_amdgcn_rcpf(kTwoPi * (1.0f - __builtin_amdgcn_exp2f(-2.0f * kLog2e * kap)));
    l[2] = kap;
    oct_unpack(oct, l[3], l[4], l[5]);
}

// hemisphere fold of the mixture, as for the voxel probes (octahedron.cuh:
// 291-337): below-horizon samples are mirrored, pdf = q(w) + q(mirror(w)).
__device__ __forceinline__ void mirror(float wx, float wy, float wz, float nx, float ny, float nz, float& mx, float& my,
                                       float& mz) {
    const float d = nx * wx + ny * wy + nz * wz;
    mx = wx - 2.0f * d * nx;
    my = wy - 2.0f * d * ny;
    mz = wz - 2.0f * d * nz;
}

__device__ __forceinline__ float hemi_pdf(const float* __restrict__ d, float wx, float wy, float wz, float nx, float ny,
                                          float nz) {
    if (wx * nx + wy * ny + wz * nz <= 0.0f) return 0.0f;
    float mx, my, mz;
    mirror(wx, wy, wz, nx, ny, nz, mx, my, mz);
    return mixture_pdf(d, wx, wy, wz) + mixture_pdf(d, mx, my, mz);
}

__device__ __forceinline__ void hemi_sample(const float* __restrict__ d, float nx, float ny, float nz, const float u[3],
                                            float& wx, float& wy, float& wz, float& pdf) {
    mixture_sample(d, u[0], u[1], u[2], wx, wy, wz);
    float mx, my, mz;
    mirror(wx, wy, wz, nx, ny, nz, mx, my, mz);
    pdf = mixture_pdf(d, wx, wy, wz) + mixture_pdf(d, mx, my, mz);
    if (wx * nx + wy * ny + wz * nz < 0.0f) { wx = mx; wy = my; wz = mz; }
}

} // namespace mlp
} // namespace ar
