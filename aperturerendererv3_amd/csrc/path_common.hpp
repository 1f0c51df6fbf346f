// Per-path helpers shared by the integrator and the guiding update: the
// per-path spectral quantity, the wavelengths of a path, and wave-aggregated
// queue appends.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ar.h"

namespace ar {

// Per-path spectral quantity: RGB in channels 0..2 (RGB mode), one wavelength
// in channel 0 (single-wavelength mode) or four hero wavelengths.
struct S4 { float v[4]; };
__device__ __forceinline__ S4 s4(float a, float b, float c, float d) { return {{a, b, c, d}}; }
__device__ __forceinline__ S4 s4mul(const S4& a, const S4& b) {
    return {{a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]}};
}
__device__ __forceinline__ S4 s4add(const S4& a, const S4& b) {
    return {{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]}};
}
__device__ __forceinline__ S4 s4scale(const S4& a, float s) { return {{a.v[0] * s, a.v[1] * s, a.v[2] * s, a.v[3] * s}}; }

// Wavelengths of a path: spectral mode and the first wavelength.
struct Lam {
    uint32_t mode;   // AR_SPECTRAL_*
    float l[4];
};

__device__ __forceinline__ Lam make_lam(uint32_t mode, float l0) {
    Lam L;
    L.mode = mode;
    L.l[0] = l0;
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        float x = l0 - 380.0f + 100.0f * (float)k;
        if (x >= 400.0f) x -= 400.0f;
        L.l[k] = 380.0f + x;
    }
    return L;
}

// Appends `pred` lanes to a queue with one atomic per wave; returns the slot.
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t* counter) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return 0;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

} // namespace ar
