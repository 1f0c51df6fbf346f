// Learned guiding: BF16 MFMA forward and backward of the 5-64-64-64-40 MLP,
// Adam, and the guiding-MLP C ABI.
//
// One wave64 owns a tile of 64 samples and keeps the tile's activations in
// LDS.  Every layer is a product of 16x16x16 BF16 tiles
// (v_mfma_f32_16x16x16_bf16, FP32 accumulate), computed transposed so that
// the samples are the N dimension:
//   forward        Z^T[out x s]  = W^T[out x in] . H^T[in x s]
//   backward data  dH^T[in x s]  = W[in x out]   . dZ^T[out x s]
//   weight grads   dW[in x out] += H^T[in x s]   . dZ[s x out]   (K = samples)
// With lane l, i = l % 16, g = l / 16, the operand layouts are
//   A: a[r] = A[i][4g + r],  B: b[r] = B[4g + r][i],  D: d[r] = D[4g + r][i],
// so a sample's activations are contiguous in LDS for both the A (weights,
// from global memory) and B (activations) fetches of the forward and
// backward-data products.  The loss gradient of the vMF mixture is computed
// per sample (lane = sample) in FP32; weight gradients are summed with FP32
// atomics; Adam updates the FP32 masters and rewrites the BF16 copies.
#include "ar_internal.hpp"
#include "mlp.hpp"
#include "path_common.hpp"
#include "spectral.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace ar {
namespace {

using namespace mlp;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) { return *reinterpret_cast<const bf16x4*>(p); }
__device__ __forceinline__ bf16x4 ld4s(const uint16_t* p, int stride) {
    bf16x4 v;
    v[0] = (short)p[0];
    v[1] = (short)p[stride];
    v[2] = (short)p[2 * stride];
    v[3] = (short)p[3 * stride];
    return v;
}

// Z^T = W^T . H^T for 64 samples; H in LDS [64][IN] bf16, WT global [OUT][IN].
// RELU: store bf16(relu(z + b)) to Hout [64][OUT]; else store z + b as FP32 to Oout [64][OUT].
template <int IN, int OUT, bool RELU>
__device__ __forceinline__ void fwd_layer(const uint16_t* __restrict__ WT, const float* __restrict__ bias,
                                          const uint16_t* Hin, uint16_t* Hout, float* Oout, int lane) {
    const int i = lane & 15, g = lane >> 4;
#pragma unroll 1
    for (int mb = 0; mb < OUT / 16; ++mb) {
        float bb[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) bb[r] = bias[mb * 16 + 4 * g + r];
#pragma unroll 1
        for (int nb = 0; nb < 4; ++nb) {
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kb = 0; kb < IN / 16; ++kb)
                c = mfma(ld4(WT + (mb * 16 + i) * IN + kb * 16 + 4 * g), ld4(Hin + (nb * 16 + i) * IN + kb * 16 + 4 * g), c);
            const int s = nb * 16 + i, u0 = mb * 16 + 4 * g;
            if (RELU) {
                uint16_t* o = Hout + s * OUT + u0;
#pragma unroll
                for (int r = 0; r < 4; ++r) o[r] = f2bf(fmaxf(c[r] + bb[r], 0.0f));
            } else {
                float* o = Oout + s * OUT + u0;
#pragma unroll
                for (int r = 0; r < 4; ++r) o[r] = c[r] + bb[r];
            }
        }
    }
}

// dH^T = W . dZ^T, masked by the layer input's ReLU (Hin > 0); writes dZ of
// the layer below as bf16 [64][IN].
template <int IN, int OUT>
__device__ __forceinline__ void bwd_data(const uint16_t* __restrict__ W, const uint16_t* dZ, const uint16_t* Hin,
                                         uint16_t* dZin, int lane) {
    const int i = lane & 15, g = lane >> 4;
#pragma unroll 1
    for (int mb = 0; mb < IN / 16; ++mb) {
#pragma unroll 1
        for (int nb = 0; nb < 4; ++nb) {
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kb = 0; kb < OUT / 16; ++kb)
                c = mfma(ld4(W + (mb * 16 + i) * OUT + kb * 16 + 4 * g), ld4(dZ + (nb * 16 + i) * OUT + kb * 16 + 4 * g), c);
            const int s = nb * 16 + i, u0 = mb * 16 + 4 * g;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                dZin[s * IN + u0 + r] = f2bf(bf2f(Hin[s * IN + u0 + r]) > 0.0f ? c[r] : 0.0f);
        }
    }
}

// grad[W][in][out] += sum_s H[s][in] dZ[s][out]; grad[b][out] += sum_s dZ[s][out]
// (the bias sum reads dZf when given, else the BF16 dZ)
template <int IN, int OUT>
__device__ __forceinline__ void bwd_weights(const uint16_t* H, const uint16_t* dZ, const float* dZf, float* gW, float* gB,
                                            int lane) {
    const int i = lane & 15, g = lane >> 4;
#pragma unroll 1
    for (int mb = 0; mb < IN / 16; ++mb) {
#pragma unroll 1
        for (int nb = 0; nb < OUT / 16; ++nb) {
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kb = 0; kb < 4; ++kb)
                c = mfma(ld4s(H + (kb * 16 + 4 * g) * IN + mb * 16 + i, IN),
                         ld4s(dZ + (kb * 16 + 4 * g) * OUT + nb * 16 + i, OUT), c);
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (c[r] != 0.0f) atomicAdd(gW + (mb * 16 + 4 * g + r) * OUT + nb * 16 + i, c[r]);
        }
    }
    if (lane < OUT) {
        float s = 0.0f;
        for (int k = 0; k < 64; ++k) s += dZf ? dZf[k * OUT + lane] : bf2f(dZ[k * OUT + lane]);
        if (s != 0.0f) atomicAdd(gB + lane, s);
    }
}

struct Net {
    const float* p;        // FP32 masters (biases are read from here)
    const uint16_t* wb;    // BF16 weights, [in][out] per layer (same offsets as p)
    const uint16_t* wt;    // BF16 weights transposed, [out][in] per layer (same offsets)
};

// LDS of one forward tile
struct FwdLds {
    uint16_t x[64 * kIn];
    uint16_t h1[64 * kHid];
    uint16_t h2[64 * kHid];
    uint16_t h3[64 * kHid];
    float o[64 * kOut];
};

__device__ __forceinline__ void load_inputs(const float* __restrict__ x, int xs, size_t n, size_t s0, uint16_t* xl,
                                            int lane) {
    const size_t s = s0 + lane;
#pragma unroll
    for (int k = 0; k < kIn; ++k) xl[lane * kIn + k] = s < n ? f2bf(record_input(x + s * xs, xs, k)) : (uint16_t)0;
}

__device__ __forceinline__ void forward_tile(const Net& N, FwdLds& L, int lane) {
    fwd_layer<kIn, kHid, true>(N.wt + OW0, N.p + OB0, L.x, L.h1, nullptr, lane);
    __syncthreads();
    fwd_layer<kHid, kHid, true>(N.wt + OW1, N.p + OB1, L.h1, L.h2, nullptr, lane);
    __syncthreads();
    fwd_layer<kHid, kHid, true>(N.wt + OW2, N.p + OB2, L.h2, L.h3, nullptr, lane);
    __syncthreads();
    fwd_layer<kHid, kOut, false>(N.wt + OW3, N.p + OB3, L.h3, nullptr, L.o, lane);
    __syncthreads();
}

struct Lobes {
    float pi[kLobes], kappa[kLobes], mu[kLobes][3], clen[kLobes];
    bool clamped[kLobes];
};

__device__ __forceinline__ void decode(const float* o, Lobes& L) {
    float mx = -3.0e38f;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) mx = fmaxf(mx, o[5 * k]);
    float z = 0.0f;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) { L.pi[k] = expf(o[5 * k] - mx); z += L.pi[k]; }
#pragma unroll
    for (int k = 0; k < kLobes; ++k) {
        L.pi[k] /= z;
        float s = o[5 * k + 1];
        L.clamped[k] = s < -3.0f || s > 7.0f;
        s = fminf(7.0f, fmaxf(-3.0f, s));
        L.kappa[k] = expf(s);
        const float cx = o[5 * k + 2], cy = o[5 * k + 3], cz = o[5 * k + 4];
        const float l = sqrtf(cx * cx + cy * cy + cz * cz);
        L.clen[k] = l < 1e-12f ? 0.0f : l;
        if (l < 1e-12f) { L.mu[k][0] = 0.0f; L.mu[k][1] = 0.0f; L.mu[k][2] = 1.0f; }
        else { L.mu[k][0] = cx / l; L.mu[k][1] = cy / l; L.mu[k][2] = cz / l; }
    }
}

__device__ __forceinline__ float vmf_norm(float kap) { return kap / (kTwoPi * (1.0f - expf(-2.0f * kap))); }

// Inference: persistent one-wave blocks.  Each lane keeps its A fragments of
// all four layers in registers (96 VGPRs of BF16 weights) and each 64-sample
// tile's activations stay in registers too: with the samples as the N
// dimension, a layer's D tile (rows = units 4g..4g+3, column = sample i)
// after bias, ReLU and BF16 rounding is exactly the B operand of the next
// layer's K block, so hidden layers never touch LDS.  DECODE: the 40 outputs
// of a sample go through LDS to one lane, which writes the 48 decoded floats;
// else the 40 raw outputs are stored directly.
__device__ __forceinline__ bf16x4 relu_bf16(f32x4 c, const float b[4]) {
    bf16x4 h;
#pragma unroll
    for (int r = 0; r < 4; ++r) h[r] = (short)f2bf(fmaxf(c[r] + b[r], 0.0f));
    return h;
}

template <bool DECODE>
__global__ __launch_bounds__(64) void k_mlp_forward(Net N, const float* __restrict__ x5, int xs, size_t n,
                                                    float* __restrict__ out, const uint32_t* __restrict__ n_dev) {
    __shared__ float lo[DECODE ? 64 * kOut : 1];
    const int lane = threadIdx.x;
    const int i = lane & 15, g = lane >> 4;
    const size_t nn = n_dev ? std::min<size_t>(n, *n_dev) : n;
    if ((size_t)blockIdx.x * 64 >= nn) return;
    bf16x4 a0[4], a1[4][4], a2[4][4], a3[3][4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        a0[mb] = ld4(N.wt + OW0 + (mb * 16 + i) * kIn + 4 * g);
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            a1[mb][kb] = ld4(N.wt + OW1 + (mb * 16 + i) * kHid + kb * 16 + 4 * g);
            a2[mb][kb] = ld4(N.wt + OW2 + (mb * 16 + i) * kHid + kb * 16 + 4 * g);
        }
    }
#pragma unroll
    for (int mb = 0; mb < 3; ++mb)
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) a3[mb][kb] = ld4(N.wt + OW3 + (mb * 16 + i) * kHid + kb * 16 + 4 * g);
    for (size_t s0 = (size_t)blockIdx.x * 64; s0 < nn; s0 += (size_t)gridDim.x * 64) {
        // keep the per-tile bias loads inside the loop (hoisting them out would
        // cost 60 registers and halve the occupancy)
        asm volatile("" ::: "memory");
        bf16x4 h[4][4], h2[4][4];
        // layer 0: B[k = input 4g + r][sample i]
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const size_t s = s0 + nb * 16 + i;
            bf16x4 xb;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 4 * g + r;
                xb[r] = (short)(s < nn ? f2bf(record_input(x5 + s * xs, xs, k)) : (uint16_t)0);
            }
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                float b[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) b[r] = N.p[OB0 + mb * 16 + 4 * g + r];
                const f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
                h[mb][nb] = relu_bf16(mfma(a0[mb], xb, z), b);
            }
        }
        // layers 1 and 2
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            float b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) b[r] = N.p[OB1 + mb * 16 + 4 * g + r];
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) c = mfma(a1[mb][kb], h[kb][nb], c);
                h2[mb][nb] = relu_bf16(c, b);
            }
        }
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            float b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) b[r] = N.p[OB2 + mb * 16 + 4 * g + r];
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) c = mfma(a2[mb][kb], h2[kb][nb], c);
                h[mb][nb] = relu_bf16(c, b);
            }
        }
        // output layer
#pragma unroll
        for (int mb = 0; mb < 3; ++mb) {
            float b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) b[r] = N.p[OB3 + mb * 16 + 4 * g + r];
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) c = mfma(a3[mb][kb], h[kb][nb], c);
                const int u0 = mb * 16 + 4 * g, sl = nb * 16 + i;
                const float4 v = make_float4(c[0] + b[0], c[1] + b[1], c[2] + b[2], c[3] + b[3]);
                if (DECODE) {
                    *reinterpret_cast<float4*>(&lo[sl * kOut + u0]) = v;
                } else if (u0 < kUsedOut && s0 + sl < nn) {
                    *reinterpret_cast<float4*>(&out[(s0 + sl) * kUsedOut + u0]) = v;
                }
            }
        }
        if (DECODE) {
            __syncthreads();
            const size_t s = s0 + lane;
            if (s < nn) {
                Lobes D;
                decode(lo + lane * kOut, D);
                float* d = out + s * kDecoded;
#pragma unroll
                for (int k = 0; k < kLobes; ++k) {
                    d[6 * k + 0] = D.pi[k];
                    d[6 * k + 1] = D.pi[k] * vmf_norm(D.kappa[k]);
                    d[6 * k + 2] = D.kappa[k];
                    d[6 * k + 3] = D.mu[k][0];
                    d[6 * k + 4] = D.mu[k][1];
                    d[6 * k + 5] = D.mu[k][2];
                }
            }
            __syncthreads();
        }
    }
}

struct TrainLds {
    FwdLds f;
    uint16_t dob[64 * kOut];
    uint16_t dz[64 * kHid];
    uint16_t dz2[64 * kHid];
};

// loss gradient of one sample w.r.t. the 40 outputs (zero padded to 48)
__device__ __forceinline__ float loss_grad(const float* o, const float w[3], float wt, float* d) {
    Lobes L;
    decode(o, L);
    float f[kLobes], q = 0.0f;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) {
        const float mw = L.mu[k][0] * w[0] + L.mu[k][1] * w[1] + L.mu[k][2] * w[2];
        f[k] = vmf_norm(L.kappa[k]) * expf(L.kappa[k] * (mw - 1.0f));
        q += L.pi[k] * f[k];
    }
    q = fmaxf(q, 1e-30f);
#pragma unroll
    for (int k = 0; k < kLobes; ++k) {
        const float r = L.pi[k] * f[k] / q;
        const float kap = L.kappa[k];
        const float mw = L.mu[k][0] * w[0] + L.mu[k][1] * w[1] + L.mu[k][2] * w[2];
        d[5 * k] = wt * (L.pi[k] - r);
        const float e2 = expf(-2.0f * kap);
        const float dlk = 1.0f / kap - 2.0f * e2 / (1.0f - e2) + (mw - 1.0f);
        d[5 * k + 1] = L.clamped[k] ? 0.0f : -wt * r * dlk * kap;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            d[5 * k + 2 + c] = L.clen[k] > 0.0f ? -wt * r * kap * (w[c] - mw * L.mu[k][c]) / L.clen[k] : 0.0f;
    }
#pragma unroll
    for (int k = kUsedOut; k < kOut; ++k) d[k] = 0.0f;
    return wt == 0.0f ? 0.0f : -wt * logf(q);
}

// gradient (and loss) of sum_s weight_s * scale * (-log q(dir_s | x_s))
__global__ __launch_bounds__(64) void k_mlp_grad(Net N, const float* __restrict__ x5, int xs, const float* __restrict__ dir,
                                                 const float* __restrict__ weight, size_t n,
                                                 const uint32_t* __restrict__ n_dev, const float* __restrict__ scale,
                                                 float* __restrict__ grad, float* __restrict__ loss) {
    __shared__ TrainLds L;
    const int lane = threadIdx.x;
    const size_t nn = n_dev ? std::min<size_t>(n, *n_dev) : n;
    const float sc = *scale;
    for (size_t s0 = (size_t)blockIdx.x * 64; s0 < nn; s0 += (size_t)gridDim.x * 64) {
        load_inputs(x5, xs, nn, s0, L.f.x, lane);
        __syncthreads();
        forward_tile(N, L.f, lane);
        const size_t s = s0 + lane;
        float wv[3] = {0.0f, 0.0f, 1.0f}, wt = 0.0f;
        if (s < nn) {
            wv[0] = dir[3 * s]; wv[1] = dir[3 * s + 1]; wv[2] = dir[3 * s + 2];
            wt = weight[s] * sc;
        }
        float* d = L.f.o + lane * kOut;
        float l = loss_grad(d, wv, wt, d);
        for (int k = 0; k < kOut; ++k) L.dob[lane * kOut + k] = f2bf(d[k]);
        for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
        if (lane == 0 && l != 0.0f) atomicAdd(loss, l);
        __syncthreads();
        bwd_weights<kHid, kOut>(L.f.h3, L.dob, L.f.o, grad + OW3, grad + OB3, lane);
        bwd_data<kHid, kOut>(N.wb + OW3, L.dob, L.f.h3, L.dz, lane);
        __syncthreads();
        bwd_weights<kHid, kHid>(L.f.h2, L.dz, nullptr, grad + OW2, grad + OB2, lane);
        bwd_data<kHid, kHid>(N.wb + OW2, L.dz, L.f.h2, L.dz2, lane);
        __syncthreads();
        bwd_weights<kHid, kHid>(L.f.h1, L.dz2, nullptr, grad + OW1, grad + OB1, lane);
        bwd_data<kHid, kHid>(N.wb + OW1, L.dz2, L.f.h1, L.dz, lane);
        __syncthreads();
        bwd_weights<kIn, kHid>(L.f.x, L.dz, nullptr, grad + OW0, grad + OB0, lane);
        __syncthreads();
    }
}

// scale = 1 / max(sum of weights, tiny)
__global__ void k_weight_sum(const float* __restrict__ w, size_t n, const uint32_t* __restrict__ n_dev, float* sum) {
    const size_t nn = n_dev ? std::min<size_t>(n, *n_dev) : n;
    float s = 0.0f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x) s += w[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0 && s != 0.0f) atomicAdd(sum, s);
}
__global__ void k_inverse(float* v) { v[0] = 1.0f / fmaxf(v[0], 1e-30f); }
__global__ void k_set(float* v, float x) { v[0] = x; }

struct Adam {
    float* p; float* m; float* v; float* g;
    uint16_t* wb; uint16_t* wt;
    const float* t;                  // step count (device; advanced by k_adam_tick)
    float lr, b1, b2, eps;
};

// Advances the Adam step count kept in device memory, so that a captured
// training step replays with the right bias corrections.
__global__ void k_adam_tick(float* t) { t[0] += 1.0f; }

// maps a weight index to its transposed position ([out][in] per layer)
__device__ __forceinline__ int transposed(int i) {
    int base, nin, nout;
    if (i < OB0) { base = OW0; nin = kIn; nout = kHid; }
    else if (i >= OW1 && i < OB1) { base = OW1; nin = kHid; nout = kHid; }
    else if (i >= OW2 && i < OB2) { base = OW2; nin = kHid; nout = kHid; }
    else if (i >= OW3 && i < OB3) { base = OW3; nin = kHid; nout = kOut; }
    else return -1;
    const int k = i - base, in = k / nout, out = k % nout;
    return base + out * nin + in;
}

__global__ void k_adam(Adam A, int apply) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    if (apply) {
        const float t = A.t[0];
        const float c1 = 1.0f - powf(A.b1, t), c2 = 1.0f - powf(A.b2, t);
        const float g = A.g[i];
        const float m = A.b1 * A.m[i] + (1.0f - A.b1) * g;
        const float v = A.b2 * A.v[i] + (1.0f - A.b2) * g * g;
        A.m[i] = m;
        A.v[i] = v;
        A.p[i] -= A.lr * (m / c1) / (sqrtf(v / c2) + A.eps);
        A.g[i] = 0.0f;
    }
    const int t = transposed(i);
    if (t >= 0) {
        const uint16_t b = f2bf(A.p[i]);
        A.wb[i] = b;
        A.wt[t] = b;
    }
}

__global__ void k_zero(float* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0f;
}

uint32_t tiles(size_t n) { return (uint32_t)std::max<size_t>(1, std::min<size_t>((n + 63) / 64, 8192)); }

struct Slots {
    const float* gr;       // [5*4][stride] collected radiance
    const uint32_t* gflag;
    const float* gx;       // [5*kRec][stride]
    const float* gw;       // [5*3][stride]
    const float* gp;       // [5][stride]
    uint32_t stride;
};

// Training records of a frame: every slot that received radiance gives
// (inputs, direction, luminance(L) / pdf(direction)), the Monte-Carlo weight
// of the KL divergence from the incident radiance to the guide.
__global__ __launch_bounds__(256) void k_mlp_records(Slots S, uint32_t npaths, uint32_t spectral,
                                                     const float* __restrict__ lambda0, const float* __restrict__ cie,
                                                     float* rx, float* rd, float* rw, uint32_t* count, uint32_t cap) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    const bool live = p < npaths;
    const uint32_t flags = live ? S.gflag[p] : 0u;
    const Lam L = make_lam(spectral, live && spectral != AR_SPECTRAL_RGB ? lambda0[p] : 0.0f);
    for (int i = 0; i < 5; ++i) {
        float w = 0.0f;
        if ((flags >> i) & 1u) {
            const float* r = S.gr + (size_t)(4 * i) * S.stride + p;
            float lum;
            if (spectral == AR_SPECTRAL_RGB) {
                lum = r[0] * 0.2126f + r[S.stride] * 0.7152f + r[2 * (size_t)S.stride] * 0.0722f;
            } else if (spectral == AR_SPECTRAL_SINGLE) {
                float cx, cy, cz;
                spec::cie_xyz(cie, L.l[0], cx, cy, cz);
                lum = cy * r[0] / spec::lambda_pdf(L.l[0]);
            } else {
                lum = 0.0f;
                for (int k = 0; k < 4; ++k) {
                    float cx, cy, cz;
                    spec::cie_xyz(cie, L.l[k], cx, cy, cz);
                    lum += cy * r[(size_t)k * S.stride];
                }
                lum = lum / (4.0f / 400.0f);
            }
            const float pdf = S.gp[(size_t)i * S.stride + p];
            w = pdf > 0.0f ? lum / pdf : 0.0f;
        }
        const bool ok = w > 0.0f && w < 3.0e38f;
        const uint32_t slot = wave_append(ok, count);
        if (ok && slot < cap) {
            for (int k = 0; k < kRec; ++k) rx[kRec * (size_t)slot + k] = S.gx[(size_t)(kRec * i + k) * S.stride + p];
            for (int k = 0; k < 3; ++k) rd[3 * (size_t)slot + k] = S.gw[(size_t)(3 * i + k) * S.stride + p];
            rw[slot] = w;
        }
    }
}

__global__ void k_clamp_count(uint32_t* count, uint32_t cap) {
    if (count[0] > cap) count[0] = cap;
}

} // namespace

int mlp_ensure_buffers(ar_ctx* ctx, size_t paths) {
    MlpDev& M = ctx->mlp;
    if (M.cap >= paths) return AR_OK;
    invalidate_graphs(ctx);
    for (float** q : {&M.feat, &M.dec, &M.rx, &M.rd, &M.rw})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    if (M.rcount) { (void)hipFree(M.rcount); M.rcount = nullptr; }
    M.cap = M.rcap = 0;
    const size_t rcap = 5 * paths;
    hipError_t e = hipMalloc((void**)&M.feat, paths * kRec * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M.dec, paths * kDecoded * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M.rx, rcap * kRec * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M.rd, rcap * 3 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M.rw, rcap * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M.rcount, 4 * sizeof(uint32_t));
    if (e != hipSuccess) return set_error(ctx, AR_E_OOM, "mlp: device allocation failed");
    M.cap = paths;
    M.rcap = rcap;
    return AR_OK;
}

__global__ void k_zero_u32(uint32_t* p) { p[0] = 0u; }

int mlp_frame_update(ar_ctx* ctx, uint32_t spectral, const float* lambda0, const float* cie, size_t paths,
                     hipStream_t s) {
    MlpDev& M = ctx->mlp;
    const GuideDev& g = ctx->guide;
    Slots S{g.gr, g.gflag, g.gx, g.gw, g.gp, (uint32_t)g.pcap};
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, M.rcount);
    hipLaunchKernelGGL(k_mlp_records, dim3((uint32_t)((paths + 255) / 256)), dim3(256), 0, s, S, (uint32_t)paths,
                       spectral, lambda0, cie, M.rx, M.rd, M.rw, M.rcount, (uint32_t)M.rcap);
    hipLaunchKernelGGL(k_clamp_count, dim3(1), dim3(1), 0, s, M.rcount, (uint32_t)M.rcap);
    const size_t n = std::min<size_t>(M.rcap, M.batch);
    hipError_t e = hipSuccess;
    for (uint32_t it = 0; it < M.iters && e == hipSuccess; ++it) {
        e = mlp_gradient(ctx, M.rx, kRec, M.rd, M.rw, n, M.rcount, true, s);
        if (e == hipSuccess) e = mlp_adam_step(ctx, s);
    }
    return check_hip(ctx, e, "guiding network training");
}

void mlp_free(ar_ctx* ctx) {
    invalidate_graphs(ctx);
    MlpDev& M = ctx->mlp;
    for (float** q : {&M.p, &M.feat, &M.dec, &M.rx, &M.rd, &M.rw})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    if (M.rcount) { (void)hipFree(M.rcount); M.rcount = nullptr; }
    M.m = M.v = M.g = M.scal = nullptr;
    M.wb = M.wt = nullptr;
    M.cap = M.rcap = 0;
}

int mlp_ensure(ar_ctx* ctx) {
    MlpDev& M = ctx->mlp;
    if (M.p) return AR_OK;
    (void)hipSetDevice(ctx->device);
    const size_t n = kParams;
    hipError_t e = hipMalloc((void**)&M.p, 4 * n * sizeof(float) + 2 * n * sizeof(uint16_t) + 4 * sizeof(float));
    if (e != hipSuccess) return set_error(ctx, AR_E_OOM, "mlp: device allocation failed");
    M.m = M.p + n;
    M.v = M.m + n;
    M.g = M.v + n;
    M.scal = M.g + n;                // [0] weight scale, [1] loss, [2] Adam step count, [3] spare
    M.wb = (uint16_t*)(M.scal + 4);
    M.wt = M.wb + n;
    return mlp_init(ctx, M.seed);
}

int mlp_init(ar_ctx* ctx, uint64_t seed) {
    MlpDev& M = ctx->mlp;
    // He-normal weights on the used rows/columns, zero padding and biases
    std::vector<float> h(kParams, 0.0f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    const int off[4] = {OW0, OW1, OW2, OW3};
    const int din[4] = {kIn, kHid, kHid, kHid}, dout[4] = {kHid, kHid, kHid, kOut};
    const int uin[4] = {kUsedIn, kHid, kHid, kHid}, uout[4] = {kHid, kHid, kHid, kUsedOut};
    for (int l = 0; l < 4; ++l)
        for (int i = 0; i < uin[l]; ++i)
            for (int o = 0; o < uout[l]; ++o) h[off[l] + i * dout[l] + o] = nd(rng) * std::sqrt(2.0f / (float)uin[l]);
    (void)din;
    return mlp_set_params(ctx, h.data());
}

int mlp_set_params(ar_ctx* ctx, const float* host) {
    MlpDev& M = ctx->mlp;
    hipError_t e = hipMemcpy(M.p, host, kParams * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(M.m, 0, (3 * kParams + 4) * sizeof(float));   // m, v, g, scalars
    if (e != hipSuccess) return check_hip(ctx, e, "mlp parameters");
    Adam A{M.p, M.m, M.v, M.g, M.wb, M.wt, M.scal + 2, 0, 0, 0, 0};
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, nullptr, A, 0);
    return check_hip(ctx, hipDeviceSynchronize(), "mlp parameters");
}

hipError_t mlp_forward(ar_ctx* ctx, const float* x5, int xs, size_t n, const uint32_t* n_dev, float* out, bool decoded,
                       hipStream_t s) {
    MlpDev& M = ctx->mlp;
    Net N{M.p, M.wb, M.wt};
    // persistent: one wave per SIMD (the kernel keeps ~280 registers), each looping over 64-sample tiles
    const uint32_t grid = (uint32_t)std::min<size_t>((n + 63) / 64, (size_t)std::max(ctx->num_cus, 1) * 4);
    if (decoded)
        hipLaunchKernelGGL(k_mlp_forward<true>, dim3(std::max(grid, 1u)), dim3(64), 0, s, N, x5, xs, n, out, n_dev);
    else
        hipLaunchKernelGGL(k_mlp_forward<false>, dim3(std::max(grid, 1u)), dim3(64), 0, s, N, x5, xs, n, out, n_dev);
    return hipGetLastError();
}

// gradient into M.g (zeroed first unless accumulate) and loss into M.scal[1];
// normalize: scale = 1 / sum(weights), else 1 / n
hipError_t mlp_gradient(ar_ctx* ctx, const float* x5, int xs, const float* dir, const float* w, size_t n,
                        const uint32_t* n_dev, bool normalize, hipStream_t s) {
    MlpDev& M = ctx->mlp;
    Net N{M.p, M.wb, M.wt};
    hipLaunchKernelGGL(k_zero, dim3(64), dim3(256), 0, s, M.g, (size_t)kParams);
    hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, M.scal + 1, 0.0f);
    if (normalize) {
        hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, M.scal, 0.0f);
        hipLaunchKernelGGL(k_weight_sum, dim3(256), dim3(256), 0, s, w, n, n_dev, M.scal);
        hipLaunchKernelGGL(k_inverse, dim3(1), dim3(1), 0, s, M.scal);
    } else {
        hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, M.scal, 1.0f / (float)std::max<size_t>(n, 1));
    }
    hipLaunchKernelGGL(k_mlp_grad, dim3(tiles(n)), dim3(64), 0, s, N, x5, xs, dir, w, n, n_dev, M.scal, M.g,
                       M.scal + 1);
    return hipGetLastError();
}

hipError_t mlp_adam_step(ar_ctx* ctx, hipStream_t s) {
    MlpDev& M = ctx->mlp;
    Adam A{M.p, M.m, M.v, M.g, M.wb, M.wt, M.scal + 2, M.lr, 0.9f, 0.999f, 1e-8f};
    hipLaunchKernelGGL(k_adam_tick, dim3(1), dim3(1), 0, s, M.scal + 2);
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, A, 1);
    return hipGetLastError();
}

} // namespace ar

using namespace ar;

extern "C" {

int ar_mlp_configure(ar_ctx* ctx, const ar_mlp_opts* opts) {
    if (!ctx || !opts) return AR_E_INVALID;
    if (!(opts->learning_rate > 0.0f) || opts->batch == 0)
        return set_error(ctx, AR_E_INVALID, "mlp: learning_rate and batch must be positive");
    MlpDev& M = ctx->mlp;
    M.lr = opts->learning_rate;
    M.batch = opts->batch;
    M.iters = opts->iters_per_frame;
    M.seed = opts->seed;
    if (!M.p) return mlp_ensure(ctx);
    (void)hipDeviceSynchronize();
    return mlp_init(ctx, M.seed);
}

int ar_mlp_params_get(ar_ctx* ctx, float* out, size_t count) {
    if (!ctx || !out) return AR_E_INVALID;
    if (count < (size_t)kParams) return set_error(ctx, AR_E_INVALID, "mlp: output array too small");
    int rc = mlp_ensure(ctx);
    if (rc != AR_OK) return rc;
    (void)hipDeviceSynchronize();
    return check_hip(ctx, hipMemcpy(out, ctx->mlp.p, kParams * sizeof(float), hipMemcpyDeviceToHost), "mlp params");
}

int ar_mlp_params_set(ar_ctx* ctx, const float* in, size_t count) {
    if (!ctx || !in) return AR_E_INVALID;
    if (count != (size_t)kParams) return set_error(ctx, AR_E_INVALID, "mlp: wrong parameter count");
    int rc = mlp_ensure(ctx);
    if (rc != AR_OK) return rc;
    (void)hipDeviceSynchronize();
    return mlp_set_params(ctx, in);
}

int ar_mlp_forward(ar_ctx* ctx, const float* x5, size_t n, float* out40, void* stream) {
    if (!ctx || (n && (!x5 || !out40))) return AR_E_INVALID;
    int rc = mlp_ensure(ctx);
    if (rc != AR_OK) return rc;
    if (n == 0) return AR_OK;
    return check_hip(ctx, mlp_forward(ctx, x5, kUsedIn, n, nullptr, out40, false, (hipStream_t)stream), "mlp forward");
}

int ar_mlp_gradient(ar_ctx* ctx, const float* x5, const float* dir, const float* weight, size_t n, float* grad,
                    float* loss, void* stream) {
    if (!ctx || !grad || (n && (!x5 || !dir || !weight))) return AR_E_INVALID;
    int rc = mlp_ensure(ctx);
    if (rc != AR_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = mlp_gradient(ctx, x5, kUsedIn, dir, weight, n, nullptr, false, s);
    if (e == hipSuccess) e = hipMemcpyAsync(grad, ctx->mlp.g, kParams * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, ctx->mlp.scal + 1, sizeof(float), hipMemcpyDeviceToDevice, s);
    return check_hip(ctx, e, "mlp gradient");
}

int ar_mlp_train(ar_ctx* ctx, const float* x5, const float* dir, const float* weight, size_t n, uint32_t steps,
                 float* loss, void* stream) {
    if (!ctx || (n && (!x5 || !dir || !weight))) return AR_E_INVALID;
    int rc = mlp_ensure(ctx);
    if (rc != AR_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    for (uint32_t t = 0; t < steps && e == hipSuccess; ++t) {
        e = mlp_gradient(ctx, x5, kUsedIn, dir, weight, n, nullptr, true, s);
        if (e == hipSuccess) e = mlp_adam_step(ctx, s);
    }
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, ctx->mlp.scal + 1, sizeof(float), hipMemcpyDeviceToDevice, s);
    return check_hip(ctx, e, "mlp train");
}

} // extern "C"
