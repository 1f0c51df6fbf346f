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

// Appends the `pa` lanes of the block to queue a and the `pb` lanes to queue
// b with one atomic per queue per block (every thread of the block must
// call it): per-wave ballot counts, an exclusive scan over the block's waves
// in LDS, one atomicAdd on each counter.  Returns each lane's slots.  Same-
// address atomics from every wave of a large grid serialise at the L2, so
// fewer, larger reservations keep the appends off the kernel's critical path.
template <int kMaxWaves>
__device__ __forceinline__ void block_append2(bool pa, bool pb, uint32_t* ca, uint32_t* cb, uint32_t& sa,
                                              uint32_t& sb) {
    __shared__ uint32_t cnt[2][kMaxWaves];
    __shared__ uint32_t base[2];
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
    const uint64_t ma = __ballot(pa), mb = __ballot(pb);
    if (lane == 0) {
        cnt[0][w] = (uint32_t)__popcll(ma);
        cnt[1][w] = (uint32_t)__popcll(mb);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const uint32_t k = threadIdx.x;
        uint32_t run = 0;
        for (uint32_t i = 0; i < nw; ++i) {
            const uint32_t c = cnt[k][i];
            cnt[k][i] = run;
            run += c;
        }
        base[k] = run ? atomicAdd(k == 0 ? ca : cb, run) : 0u;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    sa = base[0] + cnt[0][w] + (uint32_t)__popcll(ma & below);
    sb = base[1] + cnt[1][w] + (uint32_t)__popcll(mb & below);
    __syncthreads();   // cnt and base are read before a later call rewrites them
}

// block_append2 for one queue: one atomic per block, every thread of the
// block must call it; returns the lane's slot.
template <int kMaxWaves>
__device__ __forceinline__ uint32_t block_append(bool pred, uint32_t* counter) {
    __shared__ uint32_t cnt[kMaxWaves];
    __shared__ uint32_t base;
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
    const uint64_t m = __ballot(pred);
    if (lane == 0) cnt[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t i = 0; i < nw; ++i) {
            const uint32_t c = cnt[i];
            cnt[i] = run;
            run += c;
        }
        base = run ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
    const uint32_t slot = base + cnt[w] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    __syncthreads();   // cnt and base are read before a later call rewrites them
    return slot;
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
