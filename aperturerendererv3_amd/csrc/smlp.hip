// Spectral basis network: the reference's RGB -> 4 and wavelength -> 4
// encoders (core/spectrum/spectrum_basis.cuh:2273-2470), their uplift tables
// (spectrum_lut.cuh:12-235) and their training (pytorch/spectrum/
// train_spectrum.py), on FP32 MFMA.
//
// Every layer is 32 wide.  One wave owns a tile of 32 samples (training
// batches up to 16384: 16 samples, for twice the waves) and computes
// each Linear as v_mfma_f32_16x16x4f32 products with the samples as the N
// dimension, lane l = 16 g + i:
//   A: a = A[i][k = g],  B: b = B[k = g][i],  D: d[r] = D[4 g + r][i].
// A layer's output in D layout (lane (i, g) holds units 16 mb + 4 g + r of
// sample 16 nb + i) is the next layer's B operand when K step kb takes, in
// lane g, the unit pi(kb, g) = 16 (kb / 4) + 4 g + kb % 4: a sum over K is
// the same in any order, so activations stay in registers and the weight
// fragments are read with the permuted column.  Weight gradients take the
// samples as K: the layer's inputs and output gradients go through LDS
// ([sample][unit]) to be read transposed.
//
// Training (train_spectrum.py:724-752) per step: k_smlp_train (forward,
// loss, backward; per-tile gradient partials), two fixed-order reductions,
// k_smlp_adamw (global-norm clipping, AdamW, EMA).  No atomics: the
// gradient, and so every trained parameter, is reproducible run to run.
#include "ar_internal.hpp"
#include "smlp.hpp"
#include "spectral.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace ar {
namespace {

constexpr int kW = 32;
constexpr int kBlocks = 7;
constexpr int kRinW = 0, kRinB = 1024;
__host__ __device__ constexpr int rblk(int k) { return 1056 + k * 1056; }
constexpr int kRout = 8448;
__host__ __device__ constexpr int lblk(int k) { return 8576 + k * 1056; }
constexpr int kLout = 15968;
constexpr int kParams = 16096;
constexpr int kNB = 2;                 // 16-sample column blocks per tile
constexpr int kTile = 16 * kNB;        // samples per wave
constexpr float kPi = 3.141592654f;    // CUDART_PI_F
constexpr int kChunk = 16;             // tiles summed per first-level reduction block
constexpr int kRedBlock = 256;
constexpr int kGradBlocks = (kParams + kRedBlock - 1) / kRedBlock;
constexpr uint32_t kMaxBatch = 1u << 18;
#ifndef AR_SMLP_SMALL_BATCH
#define AR_SMLP_SMALL_BATCH 16384
#endif

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// D-layout activations of one tile: v[mb][nb][r] = X[sample 16 nb + i][unit 16 mb + 4 g + r]
template <int kNB>
struct Act {
    f32x4 v[2][kNB];
};

// A fragments and bias of one 32 x 32 layer (the training kernel loads them a
// layer ahead of use, so that the L2 latency overlaps the previous layer's
// MFMAs)
struct Frag {
    float4 w0[2], w1[2];
    f32x4 b[2];
};

__device__ __forceinline__ void load_frag(const float* __restrict__ Wm, const float* __restrict__ b, Frag& f, int i,
                                          int g) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        f.w0[mb] = *reinterpret_cast<const float4*>(Wm + (mb * 16 + i) * kW + 4 * g);
        f.w1[mb] = *reinterpret_cast<const float4*>(Wm + (mb * 16 + i) * kW + 16 + 4 * g);
#pragma unroll
        for (int r = 0; r < 4; ++r) f.b[mb][r] = b ? b[mb * 16 + 4 * g + r] : 0.0f;
    }
}

// z = W x (+ b), W row-major [32 out][32 in]; the bias is the accumulator's start
template <int kNB>
__device__ __forceinline__ void linear_f(const Frag& f, const Act<kNB>& x, Act<kNB>& z) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        const float a[8] = {f.w0[mb].x, f.w0[mb].y, f.w0[mb].z, f.w0[mb].w,
                            f.w1[mb].x, f.w1[mb].y, f.w1[mb].z, f.w1[mb].w};
#pragma unroll
        for (int nb = 0; nb < kNB; ++nb) {
            f32x4 c = f.b[mb];
#pragma unroll
            for (int kb = 0; kb < 8; ++kb) c = mfma4(a[kb], x.v[kb >> 2][nb][kb & 3], c);
            z.v[mb][nb] = c;
        }
    }
}

template <int kNB>
__device__ __forceinline__ void linear(const float* __restrict__ Wm, const float* __restrict__ b, const Act<kNB>& x, Act<kNB>& z,
                                       int i, int g) {
    Frag f;
    load_frag(Wm, b, f, i, g);
    linear_f(f, x, z);
}

// 32 -> 4 without bias; valid in lanes g == 0: out[nb][r] = K[sample 16 nb + i][r]
template <int kNB>
__device__ __forceinline__ void linear_out(const float* __restrict__ Wo, const Act<kNB>& x, f32x4 out[kNB], int i, int g) {
    float a[8];
    if (i < 4) {
        const float4 w0 = *reinterpret_cast<const float4*>(Wo + i * kW + 4 * g);
        const float4 w1 = *reinterpret_cast<const float4*>(Wo + i * kW + 16 + 4 * g);
        a[0] = w0.x; a[1] = w0.y; a[2] = w0.z; a[3] = w0.w;
        a[4] = w1.x; a[5] = w1.y; a[6] = w1.z; a[7] = w1.w;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = 0.0f;
    }
#pragma unroll
    for (int nb = 0; nb < kNB; ++nb) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) c = mfma4(a[kb], x.v[kb >> 2][nb][kb & 3], c);
        out[nb] = c;
    }
}

// RGB embedding in B layout: entry k = 16 h + 4 g + r is channel 2 h + g / 2,
// sine for even g, cosine for odd g, harmonic r + 1 (RGBEncode's phase order)
template <int kNB>
__device__ __forceinline__ void embed_rgb(const float rgb[kNB][3], const float s[kNB], Act<kNB>& e, int g) {
    const bool cosine = g & 1;
#pragma unroll
    for (int nb = 0; nb < kNB; ++nb) {
        const float inv = 1.0f / (s[nb] + 0.001f);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int c = 2 * h + (g >> 1);
            const float v = c < 3 ? rgb[nb][c] * inv : s[nb];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float ph = v * kPi * (float)(r + 1);
                e.v[h][nb][r] = cosine ? cosf(ph) : sinf(ph);
            }
        }
    }
}

// wavelength embedding: entry 4 g + r is sin, 16 + 4 g + r cos, harmonic 4 g + r + 1
template <int kNB>
__device__ __forceinline__ void embed_lambda(const float x[kNB], Act<kNB>& e, int g) {
#pragma unroll
    for (int nb = 0; nb < kNB; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float ph = x[nb] * kPi * (float)(4 * g + r + 1);
            e.v[0][nb][r] = sinf(ph);
            e.v[1][nb][r] = cosf(ph);
        }
}

template <int kNB>
__device__ __forceinline__ void residual_f(const Frag& f, Act<kNB>& h) {
    Act<kNB> t;
    linear_f(f, h, t);
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < kNB; ++nb)
#pragma unroll
            for (int r = 0; r < 4; ++r) h.v[mb][nb][r] = h.v[mb][nb][r] + fmaxf(t.v[mb][nb][r], 0.0f);
}

// the 7 residual blocks starting at parameter offset first (inference: several
// waves per SIMD hide the fragment loads, which are not loaded ahead here)
template <int kNB>
__device__ __forceinline__ void blocks(const float* __restrict__ P, int first, Act<kNB>& h, int i, int g) {
#pragma unroll 1
    for (int k = 0; k < kBlocks; ++k) {
        Frag f;
        load_frag(P + first + k * 1056, P + first + k * 1056 + 1024, f, i, g);
        residual_f(f, h);
    }
}

template <int kNB>
__device__ __forceinline__ void rgb_forward(const float* __restrict__ P, const Act<kNB>& e, f32x4 K[kNB], int i, int g) {
    Act<kNB> h;
    linear(P + kRinW, P + kRinB, e, h, i, g);
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < kNB; ++nb)
#pragma unroll
            for (int r = 0; r < 4; ++r) h.v[mb][nb][r] = fmaxf(h.v[mb][nb][r], 0.0f);
    blocks(P, rblk(0), h, i, g);
    linear_out(P + kRout, h, K, i, g);
}

template <int kNB>
__device__ __forceinline__ void lambda_forward(const float* __restrict__ P, Act<kNB>& h, f32x4 K[kNB], int i, int g) {
    blocks(P, lblk(0), h, i, g);
    linear_out(P + kLout, h, K, i, g);
}

__device__ __forceinline__ float interact(const f32x4& a, const f32x4& b) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) acc = fmaf(a[k], b[k], acc);
    return fminf(fmaxf(acc, 0.0f), 1.0f);
}

// ---------------------------------------------------------------- inference

enum : int { kEval = 0, kRgbTable = 1, kLambdaTable = 2 };

// EVAL: x n x 4 (r, g, b, nm) -> out n;  RGB table: entry v of res^3 at the
// voxel centre -> out[4 v..];  wavelength table: entry v of res -> out[4 v..]
template <int MODE>
__global__ __launch_bounds__(64) void k_smlp_encode(const float* __restrict__ P, const float* __restrict__ x, size_t n,
                                                    uint32_t res, float* __restrict__ out) {
    const int lane = threadIdx.x, i = lane & 15, g = lane >> 4;
    for (size_t t0 = (size_t)blockIdx.x * kTile; t0 < n; t0 += (size_t)gridDim.x * kTile) {
        float rgb[kNB][3], s[kNB], lx[kNB];
#pragma unroll
        for (int nb = 0; nb < kNB; ++nb) {
            const size_t j = t0 + nb * 16 + i;
            const bool ok = j < n;
            float l = 360.0f;
            if (MODE == kEval) {
                const float4 v = ok ? *reinterpret_cast<const float4*>(x + 4 * j) : make_float4(0, 0, 0, 360.0f);
                rgb[nb][0] = v.x; rgb[nb][1] = v.y; rgb[nb][2] = v.z;
                l = v.w;
            } else if (MODE == kRgbTable) {
                const float inv = 1.0f / (float)res;
                const uint32_t v = ok ? (uint32_t)j : 0u;
                rgb[nb][0] = ((float)(v % res) + 0.5f) * inv;
                rgb[nb][1] = ((float)((v / res) % res) + 0.5f) * inv;
                rgb[nb][2] = ((float)(v / (res * res)) + 0.5f) * inv;
            } else {
                const float step = (spec::kLambdaMax - spec::kLambdaMin) / (float)res;
                l = spec::kLambdaMin + ((float)(ok ? j : 0) + 0.5f) * step;
            }
            s[nb] = fmaxf(fmaxf(rgb[nb][0], rgb[nb][1]), rgb[nb][2]);
            lx[nb] = fminf(fmaxf((l - 360.0f) / (780.0f - 360.0f), 0.0f), 1.0f);
        }
        f32x4 kr[kNB], kl[kNB];
        Act<kNB> e;
        if (MODE != kLambdaTable) {
            embed_rgb(rgb, s, e, g);
            rgb_forward(P, e, kr, i, g);
        }
        if (MODE != kRgbTable) {
            embed_lambda(lx, e, g);
            lambda_forward(P, e, kl, i, g);
        }
        if (g == 0) {
#pragma unroll
            for (int nb = 0; nb < kNB; ++nb) {
                const size_t j = t0 + nb * 16 + i;
                if (j >= n) continue;
                if (MODE == kEval) {
                    f32x4 a = kr[nb];
#pragma unroll
                    for (int k = 0; k < 4; ++k) a[k] *= s[nb];
                    out[j] = interact(a, kl[nb]);
                } else if (MODE == kRgbTable) {
                    const f32x4 a = kr[nb];
                    *reinterpret_cast<float4*>(out + 4 * j) =
                        make_float4(a[0] * s[nb], a[1] * s[nb], a[2] * s[nb], a[3] * s[nb]);
                } else {
                    const f32x4 b = kl[nb];
                    *reinterpret_cast<float4*>(out + 4 * j) = make_float4(b[0], b[1], b[2], b[3]);
                }
            }
        }
    }
}

}  // namespace

namespace {

__global__ __launch_bounds__(256) void k_smlp_query(const float* __restrict__ rgb_lut, int rgb_res,
                                                    const float* __restrict__ lam_lut, int lam_res,
                                                    const float* __restrict__ x, size_t n, float* __restrict__ out) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const float4 v = *reinterpret_cast<const float4*>(x + 4 * j);
    out[j] = sb::lut_interact(rgb_lut, rgb_res, sb::query_lambda(lam_lut, lam_res, v.w), v.x, v.y, v.z);
}

// ---------------------------------------------------------------- training

__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// standard normals of the slot pair (2 pair, 2 pair + 1) of (seed, step,
// sample): Box-Muller on two 24-bit uniforms, cosine for the even slot, sine
// for the odd one
__device__ __forceinline__ void normal2(uint32_t base, uint32_t pair, float& even, float& odd) {
    const uint32_t a = mix32(base + 2u * pair), b = mix32(base + 2u * pair + 1u);
    const float u1 = (float)((a >> 8) + 1u) * (1.0f / 16777216.0f);
    const float u2 = (float)(b >> 8) * (1.0f / 16777216.0f);
    const float r = sqrtf(-2.0f * logf(u1)), th = 6.283185307f * u2;
    float sn, cs;
    sincosf(th, &sn, &cs);
    even = r * cs;
    odd = r * sn;
}

__device__ __forceinline__ uint32_t noise_base(uint32_t seed, uint32_t step, uint32_t j) {
    return mix32(mix32(seed + step * 0x9E3779B9u) ^ j);
}

struct TrainArgs {
    const float* P;
    const float* x;      // dataset rows: r, g, b, nm
    const float* y;
    size_t n;
    uint32_t batch;
    uint32_t step;       // global step: selects the rows and the noise
    uint32_t seed;
    uint32_t noise;
    float* part;         // [tiles][kParams]
    float* loss;         // [tiles]
};

template <int kNB>
struct TrainLds {
    float x[8][16 * kNB * kW];    // block inputs 0..6 and the output layer's input, [sample][unit];
                               // slot 0 then takes the RGB features for the first Linear's gradient
    float dz[16 * kNB * kW];   // output gradient of the layer being differentiated
    float dk[16 * kNB * 4];    // gradient of the 4 encoder outputs
    float red[16 * kNB];
};

template <int kNB>
__device__ __forceinline__ void store_act(float* L, const Act<kNB>& a, int i, int g) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < kNB; ++nb)
            *reinterpret_cast<f32x4*>(L + (nb * 16 + i) * kW + mb * 16 + 4 * g) = a.v[mb][nb];
}

// block forward in training mode: the Linear sees h + noise, the residual h;
// records the block input in LDS (when given) and the ReLU mask
template <int kNB>
__device__ __forceinline__ void train_block(const Frag& f, Act<kNB>& h, float* L, uint32_t& mask,
                                            const uint32_t base[kNB], uint32_t slot0, bool noise, int i, int g) {
    Act<kNB> xin = h;
    if (noise) {                       // slot0 is even: units 4g + {0, 1} and 4g + {2, 3} are pairs
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < kNB; ++nb)
#pragma unroll
                for (int r = 0; r < 4; r += 2) {
                    float n0, n1;
                    normal2(base[nb], (slot0 + mb * 16 + 4 * g + r) >> 1, n0, n1);
                    xin.v[mb][nb][r] += 0.001f * n0;
                    xin.v[mb][nb][r + 1] += 0.001f * n1;
                }
    }
    if (L) store_act(L, xin, i, g);
    Act<kNB> z;
    linear_f(f, xin, z);
    mask = 0;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < kNB; ++nb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float zz = z.v[mb][nb][r];
                mask |= (zz > 0.0f ? 1u : 0u) << ((mb * kNB + nb) * 4 + r);
                h.v[mb][nb][r] = h.v[mb][nb][r] + fmaxf(zz, 0.0f);
            }
}

template <int kNB>
__device__ __forceinline__ void apply_mask(const Act<kNB>& dh, uint32_t mask, Act<kNB>& dz) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < kNB; ++nb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                dz.v[mb][nb][r] = (mask >> ((mb * kNB + nb) * 4 + r)) & 1u ? dh.v[mb][nb][r] : 0.0f;
}

// dW[u][v] = sum_s dZ[s][u] X[s][v] and db[u] = sum_s dZ[s][u] from LDS,
// written to this tile's partials at off (bias at off + 1024 when has_bias)
template <int kNB>
__device__ __forceinline__ void weight_grads(const float* dzl, const float* xl, float* __restrict__ part, int off,
                                             bool has_bias, int lane, int i, int g) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nv = 0; nv < 2; ++nv) {
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kk = 0; kk < (16 * kNB) / 4; ++kk)
                c = mfma4(dzl[(4 * kk + g) * kW + mb * 16 + i], xl[(4 * kk + g) * kW + nv * 16 + i], c);
#pragma unroll
            for (int r = 0; r < 4; ++r) part[off + (mb * 16 + 4 * g + r) * kW + nv * 16 + i] = c[r];
        }
    if (has_bias && lane < kW) {
        float s = 0.0f;
        for (int k = 0; k < (16 * kNB); ++k) s += dzl[k * kW + lane];
        part[off + 1024 + lane] = s;
    }
}

// transposed A fragments of one 32 x 32 layer for the backward-data product
struct BFrag {
    float a[2][8];
};

__device__ __forceinline__ void load_bfrag(const float* __restrict__ Wm, BFrag& f, int i, int g) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) f.a[mb][kb] = Wm[((kb >> 2) * 16 + 4 * g + (kb & 3)) * kW + mb * 16 + i];
}

// dh += W^T dz
template <int kNB>
__device__ __forceinline__ void backward_data_f(const BFrag& f, const Act<kNB>& dz, Act<kNB>& dh) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < kNB; ++nb) {
            f32x4 c = dh.v[mb][nb];
#pragma unroll
            for (int kb = 0; kb < 8; ++kb) c = mfma4(f.a[mb][kb], dz.v[kb >> 2][nb][kb & 3], c);
            dh.v[mb][nb] = c;
        }
}

// output layer backward: dh = Wo^T dK, dWo = dK^T h7 (dK in LDS [s][4], h7 in LDS)
template <int kNB>
__device__ __forceinline__ void out_backward(const float* __restrict__ Wo, const float* dk, const float* h7,
                                             float* __restrict__ part, int off, Act<kNB>& dh, int i, int g) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        const float a = Wo[g * kW + mb * 16 + i];
#pragma unroll
        for (int nb = 0; nb < kNB; ++nb) {
            const f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
            dh.v[mb][nb] = mfma4(a, dk[(nb * 16 + i) * 4 + g], z);
        }
    }
#pragma unroll
    for (int nv = 0; nv < 2; ++nv) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kk = 0; kk < (16 * kNB) / 4; ++kk)
            c = mfma4(i < 4 ? dk[(4 * kk + g) * 4 + i] : 0.0f, h7[(4 * kk + g) * kW + nv * 16 + i], c);
        if (g == 0)
#pragma unroll
            for (int r = 0; r < 4; ++r) part[off + r * kW + nv * 16 + i] = c[r];
    }
}

__device__ __forceinline__ float importance2(float l) {
    const float fx = 0.25f + fmaxf(spec::xbar_fit(l), 0.0f);
    const float fy = 0.25f + fmaxf(spec::ybar_fit(l), 0.0f);
    const float fz = 0.25f + fmaxf(spec::zbar_fit(l), 0.0f);
    return fx * fx + fy * fy + fz * fz;
}

template <int kNB>
__global__ __launch_bounds__(64) void k_smlp_train(TrainArgs A) {
    constexpr int kTile = 16 * kNB;
    __shared__ TrainLds<kNB> L;
    const int lane = threadIdx.x, i = lane & 15, g = lane >> 4;
    const uint32_t tile = blockIdx.x;
    const float* P = A.P;
    float* part = A.part + (size_t)tile * kParams;
    const bool noise = A.noise != 0;
    // samples of this lane
    float rgb[kNB][3], s[kNB], lx[kNB], lam[kNB], yv[kNB];
    uint32_t base[kNB];
#pragma unroll
    for (int nb = 0; nb < kNB; ++nb) {
        const uint32_t j = tile * kTile + nb * 16 + i;
        const size_t row = ((size_t)A.step * A.batch + j) % A.n;
        const float4 v = *reinterpret_cast<const float4*>(A.x + 4 * row);
        yv[nb] = A.y[row];
        base[nb] = noise_base(A.seed, A.step, j);
        rgb[nb][0] = v.x; rgb[nb][1] = v.y; rgb[nb][2] = v.z;
        lam[nb] = v.w;
        lx[nb] = (v.w - 360.0f) / 420.0f;
        if (noise && j < A.batch / 2) {
            float n0, n1, n2, n3;
            normal2(base[nb], 0, n0, n1);
            normal2(base[nb], 1, n2, n3);
            rgb[nb][0] += (0.025f / 256.0f) * n0;
            rgb[nb][1] += (0.025f / 256.0f) * n1;
            rgb[nb][2] += (0.025f / 256.0f) * n2;
            lx[nb] += (0.025f / 420.0f) * n3;
        }
        s[nb] = fmaxf(fmaxf(rgb[nb][0], rgb[nb][1]), rgb[nb][2]);
    }
    const uint32_t lam_slot = 4 + kBlocks * kW;
    // wavelength encoder forward (recomputed with records after the RGB backward)
    f32x4 kl[kNB];
    {
        Act<kNB> h;
        embed_lambda(lx, h, g);
        uint32_t mk;
        Frag f;
        load_frag(P + lblk(0), P + lblk(0) + 1024, f, i, g);
#pragma unroll 1
        for (int k = 0; k < kBlocks; ++k) {
            const int nx = lblk(k + 1 < kBlocks ? k + 1 : k);
            Frag nf;
            load_frag(P + nx, P + nx + 1024, nf, i, g);
            train_block(f, h, nullptr, mk, base, lam_slot + k * kW, noise, i, g);
            f = nf;
        }
        linear_out(P + kLout, h, kl, i, g);
    }
    // RGB encoder forward with records
    uint32_t mrgb[kBlocks + 1];
    f32x4 kr[kNB];
    {
        Act<kNB> e, h;
        embed_rgb(rgb, s, e, g);
        linear(P + kRinW, P + kRinB, e, h, i, g);
        uint32_t m = 0;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < kNB; ++nb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    m |= (h.v[mb][nb][r] > 0.0f ? 1u : 0u) << ((mb * kNB + nb) * 4 + r);
                    h.v[mb][nb][r] = fmaxf(h.v[mb][nb][r], 0.0f);
                }
        mrgb[0] = m;
        Frag f;
        load_frag(P + rblk(0), P + rblk(0) + 1024, f, i, g);
#pragma unroll 1
        for (int k = 0; k < kBlocks; ++k) {
            const int nx = rblk(k + 1 < kBlocks ? k + 1 : k);
            Frag nf;
            load_frag(P + nx, P + nx + 1024, nf, i, g);
            train_block(f, h, L.x[k], mrgb[1 + k], base, 4 + k * kW, noise, i, g);
            f = nf;
        }
        store_act(L.x[7], h, i, g);
        linear_out(P + kRout, h, kr, i, g);
    }
    // loss and the gradient of the encoder outputs (lanes g == 0 hold K)
    f32x4 dkl[kNB];
    if (g == 0) {
        float lsum = 0.0f;
#pragma unroll
        for (int nb = 0; nb < kNB; ++nb) {
            float res = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k) res = fmaf(kr[nb][k] * s[nb], kl[nb][k], res);
            const float pred = fminf(fmaxf(res, 0.0f), 1.0f);
            const float imp2 = importance2(lam[nb]);
            const float d = pred - yv[nb];
            lsum += d * d * imp2;
            const float gres = (res > 0.0f && res < 1.0f) ? 2.0f * d * imp2 / (3.0f * (float)A.batch) : 0.0f;
            f32x4 dr;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                dr[k] = gres * kl[nb][k] * s[nb];
                dkl[nb][k] = gres * (kr[nb][k] * s[nb]);
            }
            *reinterpret_cast<f32x4*>(L.dk + (nb * 16 + i) * 4) = dr;
        }
        L.red[i] = lsum;
    }
    __syncthreads();
    if (lane == 0) {
        float t = 0.0f;
        for (int k = 0; k < 16; ++k) t += L.red[k];
        A.loss[tile] = t;
    }
    // RGB backward
    {
        Act<kNB> dh, dz;
        out_backward(P + kRout, L.dk, L.x[7], part, kRout, dh, i, g);
        BFrag bf;
        load_bfrag(P + rblk(kBlocks - 1), bf, i, g);
#pragma unroll 1
        for (int k = kBlocks - 1; k >= 0; --k) {
            BFrag nbf;
            load_bfrag(P + rblk(k > 0 ? k - 1 : 0), nbf, i, g);
            apply_mask(dh, mrgb[1 + k], dz);
            store_act(L.dz, dz, i, g);
            __syncthreads();
            weight_grads<kNB>(L.dz, L.x[k], part, rblk(k), true, lane, i, g);
            backward_data_f(bf, dz, dh);
            __syncthreads();
            bf = nbf;
        }
        apply_mask(dh, mrgb[0], dz);
        store_act(L.dz, dz, i, g);
        {
            // the first block's input is no longer needed: its slot takes the
            // RGB features (recomputed) for the first Linear's weight gradient
            Act<kNB> e;
            embed_rgb(rgb, s, e, g);
            store_act(L.x[0], e, i, g);
        }
        __syncthreads();
        weight_grads<kNB>(L.dz, L.x[0], part, kRinW, true, lane, i, g);
        __syncthreads();
    }
    // wavelength encoder again, with records, then its backward
    uint32_t mlam[kBlocks];
    {
        Act<kNB> h;
        embed_lambda(lx, h, g);
        Frag f;
        load_frag(P + lblk(0), P + lblk(0) + 1024, f, i, g);
#pragma unroll 1
        for (int k = 0; k < kBlocks; ++k) {
            const int nx = lblk(k + 1 < kBlocks ? k + 1 : k);
            Frag nf;
            load_frag(P + nx, P + nx + 1024, nf, i, g);
            train_block(f, h, L.x[k], mlam[k], base, lam_slot + k * kW, noise, i, g);
            f = nf;
        }
        store_act(L.x[7], h, i, g);
        if (g == 0)
#pragma unroll
            for (int nb = 0; nb < kNB; ++nb) *reinterpret_cast<f32x4*>(L.dk + (nb * 16 + i) * 4) = dkl[nb];
        __syncthreads();
        Act<kNB> dh, dz;
        out_backward(P + kLout, L.dk, L.x[7], part, kLout, dh, i, g);
        BFrag bf;
        load_bfrag(P + lblk(kBlocks - 1), bf, i, g);
#pragma unroll 1
        for (int k = kBlocks - 1; k >= 0; --k) {
            BFrag nbf;
            load_bfrag(P + lblk(k > 0 ? k - 1 : 0), nbf, i, g);
            apply_mask(dh, mlam[k], dz);
            store_act(L.dz, dz, i, g);
            __syncthreads();
            weight_grads<kNB>(L.dz, L.x[k], part, lblk(k), true, lane, i, g);
            if (k > 0) backward_data_f(bf, dz, dh);
            __syncthreads();
            bf = nbf;
        }
    }
}

// first-level sums: part2[c][p] = sum over tiles [c kChunk, ...) of part[t][p]
__global__ __launch_bounds__(kRedBlock) void k_smlp_reduce1(const float* __restrict__ part, uint32_t tiles,
                                                             float* __restrict__ part2) {
    const uint32_t q = blockIdx.x * kRedBlock + threadIdx.x;   // float4 index
    if (q >= kParams / 4) return;
    const uint32_t t0 = blockIdx.y * kChunk, t1 = std::min<uint32_t>(tiles, t0 + kChunk);
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t t = t0; t < t1; ++t) {
        const float4 v = reinterpret_cast<const float4*>(part + (size_t)t * kParams)[q];
        acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
    }
    reinterpret_cast<float4*>(part2 + (size_t)blockIdx.y * kParams)[q] = acc;
}

// gradient, per-block squared norms, and the batch loss (block 0)
__global__ __launch_bounds__(kRedBlock) void k_smlp_reduce2(const float* __restrict__ part2, uint32_t chunks,
                                                             const float* __restrict__ tile_loss, uint32_t tiles,
                                                             float inv3b, float* __restrict__ grad,
                                                             float* __restrict__ sq, float* __restrict__ loss_out) {
    __shared__ float red[kRedBlock];
    const uint32_t p = blockIdx.x * kRedBlock + threadIdx.x;
    float gsum = 0.0f;
    if (p < kParams)
        for (uint32_t c = 0; c < chunks; ++c) gsum += part2[(size_t)c * kParams + p];
    if (p < kParams) grad[p] = gsum;
    red[threadIdx.x] = gsum * gsum;
    __syncthreads();
    for (int w = kRedBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) sq[blockIdx.x] = red[0];
    if (blockIdx.x == 0 && loss_out) {
        __syncthreads();
        float l = 0.0f;
        for (uint32_t t = threadIdx.x; t < tiles; t += kRedBlock) l += tile_loss[t];
        red[threadIdx.x] = l;
        __syncthreads();
        for (int w = kRedBlock / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) *loss_out = red[0] * inv3b;
    }
}

__host__ __device__ __forceinline__ bool decays(int p) {
    if (p >= kRinB && p < kRinB + kW) return false;
    if (p >= rblk(0) && p < kRout) return (p - rblk(0)) % 1056 < 1024;
    if (p >= lblk(0) && p < kLout) return (p - lblk(0)) % 1056 < 1024;
    return true;
}

struct AdamArgs {
    float lr, b1, b2, eps, wd, bc1, bc2_sqrt, clip, ema_decay;
};

// torch.nn.utils.clip_grad_norm_ then torch.optim.AdamW (decoupled decay on
// weight matrices), then ModelEMA.update
__global__ __launch_bounds__(kRedBlock) void k_smlp_adamw(float* __restrict__ P, float* __restrict__ m,
                                                           float* __restrict__ v, float* __restrict__ ema,
                                                           const float* __restrict__ grad, const float* __restrict__ sq,
                                                           AdamArgs a) {
    const int p = blockIdx.x * kRedBlock + threadIdx.x;
    if (p >= kParams) return;
    float ss = 0.0f;
    for (int b = 0; b < kGradBlocks; ++b) ss += sq[b];
    float coef = a.clip > 0.0f ? a.clip / (sqrtf(ss) + 1e-6f) : 1.0f;
    coef = fminf(coef, 1.0f);
    const float gr = grad[p] * coef;
    float w = P[p];
    if (decays(p)) w *= 1.0f - a.lr * a.wd;
    float mm = m[p];
    mm = mm + (gr - mm) * (1.0f - a.b1);
    const float vv = v[p] * a.b2 + (1.0f - a.b2) * gr * gr;
    const float denom = sqrtf(vv) / a.bc2_sqrt + a.eps;
    w = w - (a.lr / a.bc1) * (mm / denom);
    P[p] = w;
    m[p] = mm;
    v[p] = vv;
    ema[p] = ema[p] * a.ema_decay + (1.0f - a.ema_decay) * w;
}

// per material: the RGB bases of the albedo (divided by its largest channel
// above 1, which becomes ascale) and of the emission normalised by escale
__global__ __launch_bounds__(64) void k_material_bases(GpuMaterial* __restrict__ mats, uint32_t n,
                                                       const float* __restrict__ lut, int res) {
    const uint32_t k = blockIdx.x * 64 + threadIdx.x;
    if (k >= n) return;
    GpuMaterial& m = mats[k];
    float a[3] = {m.albedo[0], m.albedo[1], m.albedo[2]};
    const float ma = fmaxf(fmaxf(a[0], a[1]), a[2]);
    m.ascale = 1.0f;
    if (ma > 1.0f) {
        for (float& v : a) v /= ma;
        m.ascale = ma;
    }
    float e[3] = {m.emission[0], m.emission[1], m.emission[2]};
    const float me = fmaxf(fmaxf(e[0], e[1]), e[2]);
    m.escale = 1.0f;
    if (me > 1.0f) {
        for (float& v : e) v /= me;
        m.escale = me;
    }
    const float4 ka = sb::query_rgb(lut, res, a[0], a[1], a[2]);
    const float4 ke = sb::query_rgb(lut, res, e[0], e[1], e[2]);
    m.calb[0] = ka.x; m.calb[1] = ka.y; m.calb[2] = ka.z; m.calb[3] = ka.w;
    m.cem[0] = ke.x; m.cem[1] = ke.y; m.cem[2] = ke.z; m.cem[3] = ke.w;
}

int ensure_part(ar_ctx* ctx, uint32_t tiles) {
    SmlpDev& S = ctx->smlp;
    if (S.part && S.part_tiles >= tiles) return AR_OK;
    if (S.part) (void)hipFree(S.part);
    if (S.part2) (void)hipFree(S.part2);
    if (S.misc) (void)hipFree(S.misc);
    S.part = S.part2 = S.misc = nullptr;
    S.part_tiles = 0;
    const uint32_t chunks = (tiles + kChunk - 1) / kChunk;
    int rc = check_hip(ctx, hipMalloc((void**)&S.part, (size_t)tiles * kParams * sizeof(float)), "hipMalloc(smlp partials)");
    if (rc == AR_OK)
        rc = check_hip(ctx, hipMalloc((void**)&S.part2, (size_t)chunks * kParams * sizeof(float)), "hipMalloc(smlp sums)");
    if (rc == AR_OK)
        rc = check_hip(ctx, hipMalloc((void**)&S.misc, ((size_t)kGradBlocks + tiles) * sizeof(float)), "hipMalloc(smlp misc)");
    if (rc != AR_OK) return rc;
    S.part_tiles = tiles;
    return AR_OK;
}

int check_batch(ar_ctx* ctx, uint32_t batch, size_t n) {
    if (batch == 0 || batch % kTile != 0 || batch > kMaxBatch)
        return set_error(ctx, AR_E_INVALID, "batch must be a positive multiple of 32, at most 262144");
    if (n == 0) return set_error(ctx, AR_E_INVALID, "empty dataset");
    return AR_OK;
}

// gradient of one batch into S.g (and the batch loss into loss_out)
int gradient(ar_ctx* ctx, const float* x, const float* y, size_t n, uint32_t batch, uint32_t step, uint32_t seed,
             uint32_t noise, float* grad_out, float* loss_out, hipStream_t st) {
    SmlpDev& S = ctx->smlp;
    // small batches take 16-sample tiles: twice the waves, half the chain each
    const uint32_t nb = batch <= AR_SMLP_SMALL_BATCH ? 1u : 2u;
    const uint32_t tiles = batch / (16u * nb);
    int rc = ensure_part(ctx, tiles);
    if (rc != AR_OK) return rc;
    TrainArgs A{S.p, x, y, n, batch, step, seed, noise, S.part, S.misc + kGradBlocks};
    if (nb == 1)
        hipLaunchKernelGGL(k_smlp_train<1>, dim3(tiles), dim3(64), 0, st, A);
    else
        hipLaunchKernelGGL(k_smlp_train<2>, dim3(tiles), dim3(64), 0, st, A);
    const uint32_t chunks = (tiles + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(k_smlp_reduce1, dim3((kParams / 4 + kRedBlock - 1) / kRedBlock, chunks), dim3(kRedBlock), 0,
                       st, S.part, tiles, S.part2);
    hipLaunchKernelGGL(k_smlp_reduce2, dim3(kGradBlocks), dim3(kRedBlock), 0, st, S.part2, chunks,
                       S.misc + kGradBlocks, tiles, 1.0f / (3.0f * (float)batch), grad_out ? grad_out : S.g, S.misc,
                       loss_out);
    return check_hip(ctx, hipGetLastError(), "spectral network gradient launch");
}

double lr_factor(uint64_t step, uint64_t total, double warmup_ratio, double min_ratio) {
    int64_t warm = (int64_t)((double)total * warmup_ratio);
    if (warm < 1) warm = 1;
    if ((int64_t)step < warm) return (double)step / (double)warm;
    const int64_t span = (int64_t)total - warm > 1 ? (int64_t)total - warm : 1;
    const double prog = (double)((int64_t)step - warm) / (double)span;
    return min_ratio + (1.0 - min_ratio) * 0.5 * (1.0 + std::cos(M_PI * prog));
}

int need_params(ar_ctx* ctx) {
    return ctx->smlp.have ? AR_OK : set_error(ctx, AR_E_INVALID, "spectral network parameters not set");
}

}  // namespace

int smlp_material_bases(ar_ctx* ctx, GpuMaterial* mats, uint32_t n) {
    if (n == 0) return AR_OK;
    const SmlpDev& S = ctx->smlp;
    hipLaunchKernelGGL(k_material_bases, dim3((n + 63) / 64), dim3(64), 0, nullptr, mats, n, S.rgb_lut, (int)S.rgb_res);
    int rc = check_hip(ctx, hipGetLastError(), "material bases launch");
    if (rc == AR_OK) rc = check_hip(ctx, hipDeviceSynchronize(), "material bases");
    return rc;
}

void smlp_free(ar_ctx* ctx) {
    invalidate_graphs(ctx);
    SmlpDev& S = ctx->smlp;
    for (float* q : {S.p, S.part, S.part2, S.misc, S.rgb_lut, S.lam_lut})
        if (q) (void)hipFree(q);
    S = SmlpDev{};
}

}  // namespace ar

using namespace ar;

extern "C" int ar_spectral_mlp_set(ar_ctx* ctx, const float* params) {
    if (!ctx) return AR_E_INVALID;
    if (!params) return set_error(ctx, AR_E_INVALID, "params is NULL");
    for (int p = 0; p < kParams; ++p)
        if (!std::isfinite(params[p])) return set_error(ctx, AR_E_INVALID, "non-finite parameter");
    (void)hipSetDevice(ctx->device);
    SmlpDev& S = ctx->smlp;
    if (!S.p) {
        int rc = check_hip(ctx, hipMalloc((void**)&S.p, 5 * (size_t)kParams * sizeof(float)), "hipMalloc(smlp params)");
        if (rc != AR_OK) { S.p = nullptr; return rc; }
        S.m = S.p + kParams;
        S.v = S.m + kParams;
        S.ema = S.v + kParams;
        S.g = S.ema + kParams;
    }
    std::vector<float> h(5 * (size_t)kParams, 0.0f);
    std::memcpy(h.data(), params, kParams * sizeof(float));
    std::memcpy(h.data() + 3 * (size_t)kParams, params, kParams * sizeof(float));
    int rc = check_hip(ctx, hipMemcpy(S.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice),
                       "upload smlp params");
    if (rc != AR_OK) return rc;
    S.have = true;
    S.t = 0;
    return AR_OK;
}

extern "C" int ar_spectral_set_basis(ar_ctx* ctx, uint32_t basis) {
    if (!ctx) return AR_E_INVALID;
    if (basis != AR_BASIS_SIGMOID && basis != AR_BASIS_NETWORK) return set_error(ctx, AR_E_INVALID, "unknown basis");
    if (basis == AR_BASIS_NETWORK && (!ctx->smlp.rgb_lut || !ctx->smlp.lam_lut))
        return set_error(ctx, AR_E_INVALID, "basis network tables not built (ar_spectral_mlp_build_tables)");
    ctx->smlp.basis = basis;
    return AR_OK;
}

extern "C" int ar_spectral_mlp_get(ar_ctx* ctx, float* params, int ema) {
    if (!ctx) return AR_E_INVALID;
    if (!params) return set_error(ctx, AR_E_INVALID, "params is NULL");
    int rc = need_params(ctx);
    if (rc != AR_OK) return rc;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    return check_hip(ctx, hipMemcpy(params, ema ? ctx->smlp.ema : ctx->smlp.p, kParams * sizeof(float),
                                    hipMemcpyDeviceToHost), "read smlp params");
}

extern "C" int ar_spectral_mlp_eval(ar_ctx* ctx, const float* x, float* out, size_t n, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if ((!x || !out) && n) return set_error(ctx, AR_E_INVALID, "x / out is NULL");
    int rc = need_params(ctx);
    if (rc != AR_OK || n == 0) return rc;
    (void)hipSetDevice(ctx->device);
    const size_t tiles = (n + kTile - 1) / kTile;
    const uint32_t grid = (uint32_t)std::min<size_t>(tiles, (size_t)ctx->num_cus * 16);
    hipLaunchKernelGGL(k_smlp_encode<kEval>, dim3(grid), dim3(64), 0, (hipStream_t)stream, ctx->smlp.p, x, n, 0u, out);
    return check_hip(ctx, hipGetLastError(), "spectral network eval launch");
}

extern "C" int ar_spectral_mlp_build_tables(ar_ctx* ctx, uint32_t rgb_res, uint32_t lambda_res, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (rgb_res < 2 || rgb_res > 1024 || lambda_res < 2 || lambda_res > 65536)
        return set_error(ctx, AR_E_INVALID, "table resolutions must be 2..1024 (RGB) and 2..65536 (wavelength)");
    int rc = need_params(ctx);
    if (rc != AR_OK) return rc;
    (void)hipSetDevice(ctx->device);
    SmlpDev& S = ctx->smlp;
    hipStream_t st = (hipStream_t)stream;
    const size_t nrgb = (size_t)rgb_res * rgb_res * rgb_res;
    if (S.rgb_res != rgb_res) {
        (void)hipDeviceSynchronize();     // renders on any stream may still read the old table
        invalidate_graphs(ctx);
        if (S.rgb_lut) (void)hipFree(S.rgb_lut);
        S.rgb_lut = nullptr;
        S.rgb_res = 0;
        rc = check_hip(ctx, hipMalloc((void**)&S.rgb_lut, nrgb * 4 * sizeof(float)), "hipMalloc(RGB table)");
        if (rc != AR_OK) { S.rgb_lut = nullptr; return rc; }
        S.rgb_res = rgb_res;
    }
    if (S.lam_res != lambda_res) {
        (void)hipDeviceSynchronize();
        invalidate_graphs(ctx);
        if (S.lam_lut) (void)hipFree(S.lam_lut);
        S.lam_lut = nullptr;
        S.lam_res = 0;
        rc = check_hip(ctx, hipMalloc((void**)&S.lam_lut, (size_t)lambda_res * 4 * sizeof(float)),
                       "hipMalloc(wavelength table)");
        if (rc != AR_OK) { S.lam_lut = nullptr; return rc; }
        S.lam_res = lambda_res;
    }
    const uint32_t cap = (uint32_t)ctx->num_cus * 16;
    hipLaunchKernelGGL(k_smlp_encode<kRgbTable>, dim3((uint32_t)std::min<size_t>((nrgb + kTile - 1) / kTile, cap)),
                       dim3(64), 0, st, S.p, nullptr, nrgb, rgb_res, S.rgb_lut);
    hipLaunchKernelGGL(k_smlp_encode<kLambdaTable>, dim3((lambda_res + kTile - 1) / kTile), dim3(64), 0, st, S.p,
                       nullptr, (size_t)lambda_res, lambda_res, S.lam_lut);
    ctx->scene.spectral_ready = false;    // material bases come from the tables
    return check_hip(ctx, hipGetLastError(), "spectral table launch");
}

extern "C" int ar_spectral_mlp_read_tables(ar_ctx* ctx, float* rgb_table, float* lambda_table) {
    if (!ctx) return AR_E_INVALID;
    const SmlpDev& S = ctx->smlp;
    if (!S.rgb_lut || !S.lam_lut) return set_error(ctx, AR_E_INVALID, "tables not built");
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    int rc = AR_OK;
    if (rgb_table)
        rc = check_hip(ctx, hipMemcpy(rgb_table, S.rgb_lut, (size_t)S.rgb_res * S.rgb_res * S.rgb_res * 16,
                                      hipMemcpyDeviceToHost), "read RGB table");
    if (rc == AR_OK && lambda_table)
        rc = check_hip(ctx, hipMemcpy(lambda_table, S.lam_lut, (size_t)S.lam_res * 16, hipMemcpyDeviceToHost),
                       "read wavelength table");
    return rc;
}

extern "C" int ar_spectral_mlp_query(ar_ctx* ctx, const float* x, float* out, size_t n, void* stream) {
    if (!ctx) return AR_E_INVALID;
    const SmlpDev& S = ctx->smlp;
    if (!S.rgb_lut || !S.lam_lut) return set_error(ctx, AR_E_INVALID, "tables not built");
    if ((!x || !out) && n) return set_error(ctx, AR_E_INVALID, "x / out is NULL");
    if (n == 0) return AR_OK;
    (void)hipSetDevice(ctx->device);
    hipLaunchKernelGGL(k_smlp_query, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, S.rgb_lut,
                       (int)S.rgb_res, S.lam_lut, (int)S.lam_res, x, n, out);
    return check_hip(ctx, hipGetLastError(), "spectral query launch");
}

extern "C" int ar_spectral_mlp_gradient(ar_ctx* ctx, const float* x, const float* y, size_t n, uint32_t batch,
                                        uint32_t step, uint32_t seed, int noise, float* grad, float* loss,
                                        void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (!x || !y || !grad) return set_error(ctx, AR_E_INVALID, "x / y / grad is NULL");
    int rc = need_params(ctx);
    if (rc == AR_OK) rc = check_batch(ctx, batch, n);
    if (rc != AR_OK) return rc;
    (void)hipSetDevice(ctx->device);
    return gradient(ctx, x, y, n, batch, step, seed, noise ? 1u : 0u, grad, loss, (hipStream_t)stream);
}

extern "C" int ar_spectral_mlp_train(ar_ctx* ctx, const float* x, const float* y, size_t n,
                                     const ar_smlp_train_opts* o, float* loss_history, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if (!x || !y || !o) return set_error(ctx, AR_E_INVALID, "x / y / opts is NULL");
    int rc = need_params(ctx);
    if (rc == AR_OK) rc = check_batch(ctx, o->batch, n);
    if (rc != AR_OK) return rc;
    if (!(o->lr >= 0.0f) || !(o->beta1 >= 0.0f && o->beta1 < 1.0f) || !(o->beta2 >= 0.0f && o->beta2 < 1.0f) ||
        !(o->eps > 0.0f) || !(o->ema_decay >= 0.0f && o->ema_decay <= 1.0f))
        return set_error(ctx, AR_E_INVALID, "invalid optimiser settings");
    (void)hipSetDevice(ctx->device);
    SmlpDev& S = ctx->smlp;
    hipStream_t st = (hipStream_t)stream;
    const uint64_t total = o->total_steps ? o->total_steps : (uint64_t)o->first_step + o->steps;
    for (uint32_t k = 0; k < o->steps; ++k) {
        const uint64_t gs = (uint64_t)o->first_step + k;
        rc = gradient(ctx, x, y, n, o->batch, (uint32_t)gs, o->seed, o->noise ? 1u : 0u, nullptr,
                      loss_history ? loss_history + k : nullptr, st);
        if (rc != AR_OK) return rc;
        S.t += 1;
        AdamArgs a;
        a.lr = (float)((double)o->lr * lr_factor(gs, total, o->warmup_ratio, o->min_lr_ratio));
        a.b1 = o->beta1;
        a.b2 = o->beta2;
        a.eps = o->eps;
        a.wd = o->weight_decay;
        a.bc1 = (float)(1.0 - std::pow((double)o->beta1, (double)S.t));
        a.bc2_sqrt = (float)std::sqrt(1.0 - std::pow((double)o->beta2, (double)S.t));
        a.clip = o->clip_norm;
        a.ema_decay = o->ema_decay;
        hipLaunchKernelGGL(k_smlp_adamw, dim3(kGradBlocks), dim3(kRedBlock), 0, st, S.p, S.m, S.v, S.ema, S.g, S.misc, a);
        rc = check_hip(ctx, hipGetLastError(), "spectral network AdamW launch");
        if (rc != AR_OK) return rc;
    }
    return AR_OK;
}

// Training set of the uplift this library ships: RGB drawn as the reference's
// synthetic branch draws it (spectrum_dataset.py:337-341: u 1.01 - 0.005
// clipped, to the power 0.75), wavelengths 75 % uniform on [360, 780] nm and
// 25 % from the CIE sampling mixture with +-2 nm jitter (:350-361), the
// target the Jakob-Hanika sigmoid spectrum of the RGB at that wavelength.
extern "C" int ar_spectral_uplift_dataset(ar_ctx* ctx, size_t n, uint32_t seed, float* x, float* y, void* stream) {
    if (!ctx) return AR_E_INVALID;
    if ((!x || !y) && n) return set_error(ctx, AR_E_INVALID, "x / y is NULL");
    if (n == 0) return AR_OK;
    const std::vector<float>& lut = spec::uplift_lut(spec::cie_xyz_table(), spec::d65_table());
    std::vector<float> hx(4 * n), hy(n);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    for (size_t k = 0; k < n; ++k) {
        float rgb[3];
        for (int c = 0; c < 3; ++c) {
            double v = std::min(1.0, std::max(0.0, U(rng) * 1.01 - 0.005));
            rgb[c] = (float)std::min(1.0, std::max(0.0, std::pow(v, 0.75)));
        }
        double l;
        if (U(rng) < 0.75) {
            l = 360.0 + U(rng) * 420.0;
        } else {
            const float a = (float)U(rng), b = (float)U(rng), c = (float)U(rng);
            l = spec::sample_lambda(a, b, c) + (U(rng) - 0.5) * 4.0;
        }
        l += (U(rng) - 0.5) * 0.01;
        l = std::min(780.0, std::max(360.0, l));
        float cf[3];
        spec::uplift(lut.data(), rgb, cf);
        hx[4 * k] = rgb[0]; hx[4 * k + 1] = rgb[1]; hx[4 * k + 2] = rgb[2];
        hx[4 * k + 3] = (float)l;
        hy[k] = spec::sigmoid_poly(cf[0], cf[1], cf[2], (float)l);
    }
    (void)hipSetDevice(ctx->device);
    hipStream_t st = (hipStream_t)stream;
    int rc = check_hip(ctx, hipMemcpyAsync(x, hx.data(), hx.size() * sizeof(float), hipMemcpyHostToDevice, st),
                       "upload dataset");
    if (rc == AR_OK)
        rc = check_hip(ctx, hipMemcpyAsync(y, hy.data(), hy.size() * sizeof(float), hipMemcpyHostToDevice, st),
                       "upload targets");
    if (rc == AR_OK) rc = check_hip(ctx, hipStreamSynchronize(st), "dataset upload sync");
    return rc;
}
