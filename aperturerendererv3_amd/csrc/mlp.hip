// Learned guiding: BF16 MFMA forward and backward of the 5-64-64-64-40 MLP,
// Adam, and the guiding-MLP C ABI.
//
// Both kernels run on v_mfma_f32_16x16x32_bf16 with FP32 accumulation, the
// samples as N, the weights staged in LDS and each wave's 32-sample tile
// chained through the layers in registers.  Training (k_mlp_train, 8-wave
// blocks) accumulates weight gradients per block from LDS-transposed
// activations and reduces them in a fixed order (no atomics: two runs give
// the same bits); inference (k_mlp_infer, 4-wave blocks) optionally decodes
// the vMF mixture.  The loss gradient of the mixture is computed per sample
// in FP32; Adam updates the FP32 masters and rewrites the BF16 copies.
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

// FP32 -> BF16, round to nearest even (gfx950's v_cvt_pk_bf16_f32, which
// converts two values per instruction; for finite values the same bits as
// adding 0x7FFF + lsb and truncating)
__device__ __forceinline__ uint16_t f2bf(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
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
struct Net {
    const float* p;        // FP32 masters (biases are read from here)
    const uint16_t* wb;    // BF16 weights, [in][out] per layer (same offsets as p)
    const uint16_t* wt;    // BF16 weights transposed, [out][in] per layer (same offsets)
};

__device__ __forceinline__ float vmf_norm(float kap) { return kap / (kTwoPi * (1.0f - expf(-2.0f * kap))); }

// ---------------------------------------------------------------- training
//
// One training step's gradient, deterministic (no atomics): blocks of eight
// waves, each wave a tile of 32 samples per round (256 per block round).
//  * Forward and backward-data products run on v_mfma_f32_16x16x32_bf16 with
//    the samples as N.  A layer's D fragment (units 4g..4g+3 of sample i)
//    after bias, ReLU and BF16 rounding is half of the next layer's B
//    fragment: two D blocks (units 32j + 4g.. and 32j + 16 + 4g..) form the
//    8 K values of lane group g, and the weight fragments are loaded in that
//    same K order, so activations never leave registers.
//  * Weight gradients take the samples as K: each layer's activations and
//    output gradients are written transposed (unit-major, padded rows) to
//    LDS, and every wave accumulates its share of the layer's 16x16
//    gradient blocks (blocks w and w + 8) over the block's 256 samples in
//    registers, across all rounds.  Bias gradients are the same product with an all-ones A on
//    the BF16 output gradients dZ.
//  * Each block writes its partial gradient once; k_mlp_reduce1/2 add the
//    partials in block order.
constexpr int kTT = 32;              // samples per wave tile
constexpr int kTW = 8;               // waves per block (two per SIMD)
constexpr int kTB = kTW * 64;        // threads per block
constexpr int kTR = kTT * kTW;       // samples per block round
constexpr int kTS = kTR + 8;         // transposed-row stride (shorts): 16 B of padding per row
constexpr int kRS = 52;              // loss-stage row stride (floats)
constexpr int kDS = 56;              // output-gradient row stride (shorts)

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}
__device__ __forceinline__ s16x8 cat(bf16x4 lo, bf16x4 hi) {
    s16x8 r;
#pragma unroll
    for (int k = 0; k < 4; ++k) { r[k] = lo[k]; r[4 + k] = hi[k]; }
    return r;
}
// 8 consecutive BF16 values in LDS (16-byte aligned)
__device__ __forceinline__ s16x8 lds8(const uint16_t* p) { return *reinterpret_cast<const s16x8*>(p); }
// the K-permuted weight fragment of a 64-wide K block j: units 32j + 4g .. + 3 and 32j + 16 + 4g .. + 3
__device__ __forceinline__ s16x8 wfrag(const uint16_t* row, int j, int g) {
    return cat(ld4(row + 32 * j + 4 * g), ld4(row + 32 * j + 16 + 4 * g));
}
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ s16x2 bf16_pair(float a, float b) {
    return __builtin_bit_cast(s16x2, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
// ReLU and BF16 rounding of a D fragment (whose accumulator started at the
// bias): the next layer's B half.  BF16 of max(c, 0) for every c but NaN:
// each pair is rounded to BF16 first, then kept or zeroed by a max with 0 as
// signed 16-bit integers (a negative value, -0 included, has its sign bit
// set and so is a negative integer), so the result is +0 or positive.
__device__ __forceinline__ bf16x4 relu_pk(f32x4 c) {
    const s16x2 z = {0, 0};
    const s16x2 lo = __builtin_elementwise_max(bf16_pair(c[0], c[1]), z);
    const s16x2 hi = __builtin_elementwise_max(bf16_pair(c[2], c[3]), z);
    return (bf16x4){lo[0], lo[1], hi[0], hi[1]};
}
// BF16 of c where the layer's ReLU output h (from relu_pk: +0 or positive)
// is non-zero, else +0: the mask is 0 - min(h, 1) per 16-bit half.
__device__ __forceinline__ bf16x4 relu_mask(bf16x4 h, f32x4 c) {
    const u16x2 one = {1, 1}, zero = {0, 0};
    const u16x2 mlo = zero - __builtin_elementwise_min(__builtin_bit_cast(u16x2, (s16x2){h[0], h[1]}), one);
    const u16x2 mhi = zero - __builtin_elementwise_min(__builtin_bit_cast(u16x2, (s16x2){h[2], h[3]}), one);
    const u16x2 lo = __builtin_bit_cast(u16x2, bf16_pair(c[0], c[1])) & mlo;
    const u16x2 hi = __builtin_bit_cast(u16x2, bf16_pair(c[2], c[3])) & mhi;
    return (bf16x4){(short)lo[0], (short)lo[1], (short)hi[0], (short)hi[1]};
}
// a layer's four biases of lane group g, the initial accumulator of its D fragment
__device__ __forceinline__ f32x4 bias4(const float* b) { return *reinterpret_cast<const f32x4*>(b); }
// Keeps the loads of one layer next to its products: the scheduler would
// otherwise issue every layer's weight fragments at once (about 130
// registers) and spill.
__device__ __forceinline__ void layer_fence() { asm volatile("" ::: "memory"); }

// orders a wave's LDS writes before its later LDS reads of other lanes' data
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// BF16 weights staged in LDS per block, rows padded by 4 values so that the
// 16 rows of a fragment fall in different banks: forward W^T [out][in],
// backward W [in][out] (layers 1-3), then the FP32 biases.
constexpr int kS16 = kIn + 4, kS64 = kHid + 4, kS48 = kOut + 4;
constexpr int LT0 = 0, LT1 = LT0 + kHid * kS16, LT2 = LT1 + kHid * kS64, LT3 = LT2 + kHid * kS64;
constexpr int LB3 = LT3 + kOut * kS64, LB2 = LB3 + kHid * kS48, LB1 = LB2 + kHid * kS64;
constexpr int kLdsW = LB1 + kHid * kS64;   // BF16 values
constexpr int kBias = 3 * kHid + kOut;     // b0, b1, b2, b3

struct TrainShared {
    uint16_t wgt[kLdsW];
    alignas(16) float bias[kBias];   // read four at a time (bias4)
    union {
        struct { uint16_t ht[kHid * kTS]; uint16_t zt[kHid * kTS]; } t;   // transposed activations / gradients
        float r[kTW][kTT * kRS];                                          // loss stage (per wave)
    } u;
    uint16_t dor[kTW][kTT * kDS];   // BF16 output gradient, [sample][unit] per wave
    float red[kTW];
};

// transposed store of a layer's D-layout fragments (units 4g + r of sample i)
__device__ __forceinline__ void store_t(uint16_t* t, const bf16x4 (&v)[4][2], int col0, int i, int g) {
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int r = 0; r < 4; ++r) t[(mb * 16 + 4 * g + r) * kTS + col0 + nb * 16 + i] = (uint16_t)v[mb][nb][r];
}

// the next layer's activations from a 64-unit layer: D = b + W^T . H over two
// K blocks.  BIAS_INIT: the bias is the accumulator's initial value (the
// inference kernel); else it is added to the product (the training kernel,
// whose gradient accumulators leave no registers for the bias fragments
// across the products).
template <bool BIAS_INIT>
__device__ __forceinline__ void fwd_hidden(const uint16_t* wt, const float* bias, const bf16x4 (&hin)[4][2],
                                           bf16x4 (&hout)[4][2], int i, int g) {
    layer_fence();
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        const uint16_t* row = wt + (mb * 16 + i) * kS64;
        const s16x8 a0 = wfrag(row, 0, g), a1 = wfrag(row, 1, g);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            f32x4 c = mfma32(a0, cat(hin[0][nb], hin[1][nb]), BIAS_INIT ? bias4(bias + mb * 16 + 4 * g) : zero);
            c = mfma32(a1, cat(hin[2][nb], hin[3][nb]), c);
            if (!BIAS_INIT) c += bias4(bias + mb * 16 + 4 * g);
            hout[mb][nb] = relu_pk(c);
        }
    }
}

// dZ of the layer below: D = W . dZ^T over two K blocks, masked by the layer input's ReLU
__device__ __forceinline__ void bwd_hidden(const uint16_t* wb, const bf16x4 (&dz)[4][2], const bf16x4 (&hin)[4][2],
                                           bf16x4 (&dzin)[4][2], int i, int g) {
    layer_fence();
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        const uint16_t* row = wb + (mb * 16 + i) * kS64;
        const s16x8 a0 = wfrag(row, 0, g), a1 = wfrag(row, 1, g);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
            c = mfma32(a0, cat(dz[0][nb], dz[1][nb]), c);
            c = mfma32(a1, cat(dz[2][nb], dz[3][nb]), c);
            dzin[mb][nb] = relu_mask(hin[mb][nb], c);
        }
    }
}

// accumulates gradient block (rows mb, columns nb) of H^T . dZ over the block's 128 samples
__device__ __forceinline__ f32x4 wgrad(const uint16_t* ht, const uint16_t* zt, int mb, int nb, int i, int g, f32x4 acc) {
#pragma unroll 2
    for (int kb = 0; kb < kTR / 32; ++kb)
        acc = mfma32(lds8(ht + (mb * 16 + i) * kTS + 32 * kb + 8 * g), lds8(zt + (nb * 16 + i) * kTS + 32 * kb + 8 * g),
                     acc);
    return acc;
}
__device__ __forceinline__ f32x4 bsum(const uint16_t* zt, int nb, int i, int g, f32x4 acc) {
    s16x8 ones;
#pragma unroll
    for (int k = 0; k < 8; ++k) ones[k] = (short)0x3F80;   // BF16 1.0
#pragma unroll 2
    for (int kb = 0; kb < kTR / 32; ++kb) acc = mfma32(ones, lds8(zt + (nb * 16 + i) * kTS + 32 * kb + 8 * g), acc);
    return acc;
}

// Loss gradient of one sample from its outputs o (40 floats, LDS), shared by
// two lanes: lane half h = 0 / 1 (lanes s and s + 32) takes lobes 4h .. 4h+3;
// the maximum logit, the softmax normaliser and the mixture density are
// combined across the halves with one exchange each.  Hardware exp and
// reciprocal square root (about 1 ulp); the result is a training signal
// compared with the FP64 network at the BF16 tolerance.  Writes the BF16
// gradient of this half's 20 outputs (half 1 also zeroes the padding) to
// gout and returns the loss (half 0; half 1 returns 0).
__device__ __forceinline__ float loss_half(const float* o, const float w[3], float wt, uint16_t* gout, int h) {
    constexpr float kLog2e = 1.4426950408889634f;
    float lg[4], pu[4], f[4], kap[4], e2[4], mu[4][3], clen[4];
    bool cl[4];
    float mx = -3.0e38f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        lg[j] = o[5 * (4 * h + j)];
        mx = fmaxf(mx, lg[j]);
    }
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    float z = 0.0f, qz = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float* ok = o + 5 * (4 * h + j);
        pu[j] = exp2f((lg[j] - mx) * kLog2e);
        const float sk = ok[1];
        cl[j] = sk < -3.0f || sk > 7.0f;
        kap[j] = exp2f(fminf(7.0f, fmaxf(-3.0f, sk)) * kLog2e);
        const float cx = ok[2], cy = ok[3], cz = ok[4];
        const float l2 = cx * cx + cy * cy + cz * cz;
        const bool deg = l2 < 1e-24f;
        const float rl = deg ? 0.0f : __builtin_amdgcn_rsqf(l2);
        clen[j] = deg ? 0.0f : l2 * rl;
        mu[j][0] = deg ? 0.0f : cx * rl;
        mu[j][1] = deg ? 0.0f : cy * rl;
        mu[j][2] = deg ? 1.0f : cz * rl;
        const float mw = mu[j][0] * w[0] + mu[j][1] * w[1] + mu[j][2] * w[2];
        e2[j] = exp2f(-2.0f * kap[j] * kLog2e);
        f[j] = kap[j] * __builtin_amdgcn_rcpf(kTwoPi * (1.0f - e2[j])) * exp2f(kap[j] * (mw - 1.0f) * kLog2e);
        z += pu[j];
        qz += pu[j] * f[j];
    }
    z += __shfl_xor(z, 32);
    qz += __shfl_xor(qz, 32);
    const float rz = __builtin_amdgcn_rcpf(z);
    const float q = fmaxf(qz * rz, 1e-30f);
    const float rq = __builtin_amdgcn_rcpf(q);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k = 4 * h + j;
        const float pi = pu[j] * rz;
        const float r = pi * f[j] * rq;
        const float mw = mu[j][0] * w[0] + mu[j][1] * w[1] + mu[j][2] * w[2];
        gout[5 * k] = f2bf(wt * (pi - r));
        const float dlk = __builtin_amdgcn_rcpf(kap[j]) - 2.0f * e2[j] * __builtin_amdgcn_rcpf(1.0f - e2[j]) + (mw - 1.0f);
        gout[5 * k + 1] = f2bf(cl[j] ? 0.0f : -wt * r * dlk * kap[j]);
        const float rc = clen[j] > 0.0f ? -wt * r * kap[j] * __builtin_amdgcn_rcpf(clen[j]) : 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) gout[5 * k + 2 + c] = f2bf(rc * (w[c] - mw * mu[j][c]));
    }
    if (h == 1) {
#pragma unroll
        for (int k = kUsedOut; k < kOut; ++k) gout[k] = 0;
    }
    return (h == 1 || wt == 0.0f) ? 0.0f : -wt * logf(q);
}

// gradient of sum_s weight_s * scale * (-log q(dir_s | x_s)): block partials
// part[block][kParams] and part_loss[block]
__global__ __launch_bounds__(kTB) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_mlp_train(Net N, const float* __restrict__ x, int xs,
                                                   const float* __restrict__ dir, const float* __restrict__ weight,
                                                   size_t n, const uint32_t* __restrict__ n_dev,
                                                   const float* __restrict__ scale, float* __restrict__ part,
                                                   float* __restrict__ part_loss) {
    __shared__ TrainShared S;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, i = lane & 15, g = lane >> 4;
    const size_t nn = n_dev ? std::min<size_t>(n, *n_dev) : n;
    const float sc = *scale;
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    // gradient blocks of this wave: q = w and q = w + 8 of each layer's
    // 16x16 blocks (W3: 12, W2 and W1: 16, W0: 4), biases on waves 4..7
    f32x4 aw3[2] = {zero, zero}, aw2[2] = {zero, zero}, aw1[2] = {zero, zero};
    f32x4 aw0 = zero, ab3 = zero, ab2 = zero, ab1 = zero, ab0 = zero;
    float aloss = 0.0f;
    uint16_t* ht = S.u.t.ht;
    uint16_t* zt = S.u.t.zt;
    const int col0 = w * kTT;
    // ---- stage the BF16 weights (padded rows) and the biases in LDS
    {
        struct Mat { int src, rows, cols, dst, stride; const uint16_t* base; };
        const Mat mats[7] = {{OW0, kHid, kIn, LT0, kS16, N.wt}, {OW1, kHid, kHid, LT1, kS64, N.wt},
                             {OW2, kHid, kHid, LT2, kS64, N.wt}, {OW3, kOut, kHid, LT3, kS64, N.wt},
                             {OW3, kHid, kOut, LB3, kS48, N.wb}, {OW2, kHid, kHid, LB2, kS64, N.wb},
                             {OW1, kHid, kHid, LB1, kS64, N.wb}};
#pragma unroll
        for (int m = 0; m < 7; ++m) {   // unrolled: a loop over the table would move it to LDS
            const Mat& M = mats[m];
            const int quads = M.rows * M.cols / 4;
            // every thread's loads of a matrix are issued before its stores
            // (at most 2 per thread), so the copy waits on memory once
            bf16x4 v[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int e = tid + kTB * u;
                if (e < quads) v[u] = *reinterpret_cast<const bf16x4*>(M.base + M.src + 4 * e);
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int e = tid + kTB * u;
                if (e < quads) {
                    const int row = (4 * e) / M.cols, col = (4 * e) % M.cols;
                    *reinterpret_cast<bf16x4*>(&S.wgt[M.dst + row * M.stride + col]) = v[u];
                }
            }
        }
        if (tid < kHid) {
            S.bias[tid] = N.p[OB0 + tid];
            S.bias[kHid + tid] = N.p[OB1 + tid];
            S.bias[2 * kHid + tid] = N.p[OB2 + tid];
            if (tid < kOut) S.bias[3 * kHid + tid] = N.p[OB3 + tid];
        }
        __syncthreads();
    }
    for (size_t r0 = (size_t)blockIdx.x * kTR; r0 < nn; r0 += (size_t)gridDim.x * kTR) {
        const size_t t0 = r0 + (size_t)col0;
        // ---- forward: inputs (K = 16, 16x16x16 tiles), three hidden layers, outputs
        bf16x4 xb[2];
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const size_t s = t0 + nb * 16 + i;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                xb[nb][r] = (short)(s < nn ? f2bf(record_input(x + s * xs, xs, 4 * g + r)) : (uint16_t)0);
        }
        bf16x4 h1[4][2], h2[4][2], h3[4][2];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const bf16x4 a = ld4(S.wgt + LT0 + (mb * 16 + i) * kS16 + 4 * g);
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) h1[mb][nb] = relu_pk(mfma(a, xb[nb], zero) + bias4(S.bias + mb * 16 + 4 * g));
        }
        fwd_hidden<false>(S.wgt + LT1, S.bias + kHid, h1, h2, i, g);
        fwd_hidden<false>(S.wgt + LT2, S.bias + 2 * kHid, h2, h3, i, g);
        float* R = S.u.r[w];
        layer_fence();
#pragma unroll
        for (int mb = 0; mb < 3; ++mb) {
            const uint16_t* row = S.wgt + LT3 + (mb * 16 + i) * kS64;
            const s16x8 a0 = wfrag(row, 0, g), a1 = wfrag(row, 1, g);
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                f32x4 c = mfma32(a0, cat(h3[0][nb], h3[1][nb]), zero);
                c = mfma32(a1, cat(h3[2][nb], h3[3][nb]), c);
                c += bias4(S.bias + 3 * kHid + mb * 16 + 4 * g);
                *reinterpret_cast<float4*>(&R[(nb * 16 + i) * kRS + mb * 16 + 4 * g]) = make_float4(c[0], c[1], c[2], c[3]);
            }
        }
        // ---- loss gradient, two lanes per sample (lanes s and s + 32)
        wave_sync();
        {
            const int sl = lane & (kTT - 1);
            const size_t s = t0 + sl;
            float wv[3] = {0.0f, 0.0f, 1.0f}, wt = 0.0f;
            if (s < nn) {
                wv[0] = dir[3 * s]; wv[1] = dir[3 * s + 1]; wv[2] = dir[3 * s + 2];
                wt = weight[s] * sc;
            }
            aloss += loss_half(&R[sl * kRS], wv, wt, &S.dor[w][sl * kDS], lane >> 5);
        }
        __syncthreads();                           // the loss stage's LDS (aliasing ht / zt) is free
        // ---- output layer: dW3 = H3^T . dO over the block's samples; dZ2 = W3 . dO^T masked by H3
        store_t(ht, h3, col0, i, g);
        if (lane < kTT) {
#pragma unroll 4
            for (int k = 0; k < kOut; ++k) zt[k * kTS + col0 + lane] = S.dor[w][lane * kDS + k];
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int q = w + kTW * t;
            if (q < 12) aw3[t] = wgrad(ht, zt, q / 3, q % 3, i, g, aw3[t]);
        }
        if (w >= 4 && w < 7) ab3 = bsum(zt, w - 4, i, g, ab3);
        bf16x4 dz[4][2];
        layer_fence();
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            // K = 48 outputs: units 0..31, then units 32..47 as a 16x16x32
            // product whose upper K half is zero (a 16x16x16 MFMA accumulating
            // onto this chain's 16x16x32 result loses rows 4g and 4g + 1)
            const uint16_t* row = S.wgt + LB3 + (mb * 16 + i) * kS48;
            const bf16x4 z4 = {0, 0, 0, 0};
            const s16x8 a0 = cat(ld4(row + 8 * g), ld4(row + 8 * g + 4));
            const s16x8 a1 = cat(ld4(row + 32 + 4 * g), z4);
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                const uint16_t* dr = &S.dor[w][(nb * 16 + i) * kDS];
                f32x4 c = mfma32(a0, lds8(dr + 8 * g), zero);
                c = mfma32(a1, cat(*reinterpret_cast<const bf16x4*>(dr + 32 + 4 * g), z4), c);
                dz[mb][nb] = relu_mask(h3[mb][nb], c);
            }
        }
        __syncthreads();
        // ---- hidden layer 2: dW2, b2; dZ1
        store_t(ht, h2, col0, i, g);
        store_t(zt, dz, col0, i, g);
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int q = w + kTW * t;
            aw2[t] = wgrad(ht, zt, q / 4, q % 4, i, g, aw2[t]);
        }
        if (w >= 4) ab2 = bsum(zt, w - 4, i, g, ab2);
        bf16x4 dz1[4][2];
        bwd_hidden(S.wgt + LB2, dz, h2, dz1, i, g);
        __syncthreads();
        // ---- hidden layer 1: dW1, b1; dZ0
        store_t(ht, h1, col0, i, g);
        store_t(zt, dz1, col0, i, g);
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int q = w + kTW * t;
            aw1[t] = wgrad(ht, zt, q / 4, q % 4, i, g, aw1[t]);
        }
        if (w >= 4) ab1 = bsum(zt, w - 4, i, g, ab1);
        bwd_hidden(S.wgt + LB1, dz1, h1, dz, i, g);
        __syncthreads();
        // ---- input layer: dW0 = X^T . dZ0 (16 input rows), b0
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int r = 0; r < 4; ++r) ht[(4 * g + r) * kTS + col0 + nb * 16 + i] = (uint16_t)xb[nb][r];
        store_t(zt, dz, col0, i, g);
        __syncthreads();
        if (w < 4) aw0 = wgrad(ht, zt, 0, w, i, g, aw0);
        else ab0 = bsum(zt, w - 4, i, g, ab0);
        __syncthreads();
    }
    // ---- this block's partial gradient and loss
    float* P = part + (size_t)blockIdx.x * kParams;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = 4 * g + r;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int q = w + kTW * t;
            if (q < 12) P[OW3 + ((q / 3) * 16 + row) * kOut + (q % 3) * 16 + i] = aw3[t][r];
            P[OW2 + ((q / 4) * 16 + row) * kHid + (q % 4) * 16 + i] = aw2[t][r];
            P[OW1 + ((q / 4) * 16 + row) * kHid + (q % 4) * 16 + i] = aw1[t][r];
        }
        if (w < 4) P[OW0 + row * kHid + w * 16 + i] = aw0[r];
    }
    if (g == 0 && w >= 4) {
        const int nb = w - 4;
        if (nb < 3) P[OB3 + nb * 16 + i] = ab3[0];
        P[OB2 + nb * 16 + i] = ab2[0];
        P[OB1 + nb * 16 + i] = ab1[0];
        P[OB0 + nb * 16 + i] = ab0[0];
    }
    for (int off = 32; off > 0; off >>= 1) aloss += __shfl_xor(aloss, off);
    if (lane == 0) S.red[w] = aloss;
    __syncthreads();
    if (tid == 0) {
        float l = 0.0f;
        for (int k = 0; k < kTW; ++k) l += S.red[k];
        part_loss[blockIdx.x] = l;
    }
}

// ---------------------------------------------------------------- inference
//
// k_mlp_infer: the training kernel's forward on 4-wave blocks (two blocks per
// CU), forward weights and biases staged in LDS, a 32-sample tile per wave
// chained through the layers in registers on the 16x16x32 BF16 MFMA.  Each
// layer's bias is the accumulator's initial value and its ReLU acts on the
// BF16 pairs (relu_pk).  Raw outputs are stored from the D fragments; DECODE
// sends each tile's outputs through LDS to two lanes per vertex (lobes 0..3 /
// 4..7), which write kPacked words per vertex: per lobe (pi, kappa, mu as
// 16-bit octahedral coordinates), with hardware exp and reciprocal (about
// 1 ulp each; sampling and densities both read these stored values, so such
// roundings change the guide, not the agreement of its samples with its
// densities).  Offsets are 32-bit: one launch covers at most kInferMax
// vertices.
constexpr int kIW = 4;
constexpr size_t kInferMax = (size_t)1 << 26;
struct InferShared {
    uint16_t wgt[LB3];                 // forward weights W^T, padded rows (LT0 .. LT3)
    alignas(16) float bias[kBias];   // read four at a time (bias4)
    float r[kIW][kTT * kRS];           // output tile per wave (DECODE)
};

template <bool DECODE, int XS>
__global__ __launch_bounds__(kIW * 64) void k_mlp_infer(Net N, const float* __restrict__ x, uint32_t n,
                                                        float* __restrict__ out, const uint32_t* __restrict__ n_dev) {
    __shared__ InferShared S;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, i = lane & 15, g = lane >> 4;
    const uint32_t nn = n_dev ? std::min(n, *n_dev) : n;
    {
        struct Mat { int src, rows, cols, dst, stride; };
        const Mat mats[4] = {{OW0, kHid, kIn, LT0, kS16}, {OW1, kHid, kHid, LT1, kS64}, {OW2, kHid, kHid, LT2, kS64},
                             {OW3, kOut, kHid, LT3, kS64}};
#pragma unroll
        for (int m = 0; m < 4; ++m) {   // unrolled: a loop over the table would move it to LDS
            const Mat& M = mats[m];
            const int quads = M.rows * M.cols / 4;
            bf16x4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = tid + kIW * 64 * u;
                if (e < quads) v[u] = *reinterpret_cast<const bf16x4*>(N.wt + M.src + 4 * e);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = tid + kIW * 64 * u;
                if (e < quads) {
                    const int row = (4 * e) / M.cols, col = (4 * e) % M.cols;
                    *reinterpret_cast<bf16x4*>(&S.wgt[M.dst + row * M.stride + col]) = v[u];
                }
            }
        }
        if (tid < kHid) {
            S.bias[tid] = N.p[OB0 + tid];
            S.bias[kHid + tid] = N.p[OB1 + tid];
            S.bias[2 * kHid + tid] = N.p[OB2 + tid];
            if (tid < kOut) S.bias[3 * kHid + tid] = N.p[OB3 + tid];
        }
        __syncthreads();
    }
    float* R = S.r[w];
    // The inputs of a tile are loaded one tile ahead, so their latency
    // overlaps the previous tile's products: raw[nb][r] is the record value
    // behind input 4g + r (a feature, or the scene index for its bits).
    const uint32_t tstride = gridDim.x * kIW * kTT;
    auto load_raw = [&](uint32_t t, float raw[2][4]) {
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const uint32_t s = t + nb * 16 + i;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 4 * g + r;
                const bool used = k < kUsedIn || (XS == kRec && k < kUsedIn + kSceneBits);
                raw[nb][r] = (s < nn && used) ? x[s * XS + (k < kUsedIn ? k : kUsedIn)] : 0.0f;
            }
        }
    };
    float cur[2][4];
    load_raw((blockIdx.x * kIW + w) * kTT, cur);
    for (uint32_t t0 = (blockIdx.x * kIW + w) * kTT; t0 < nn; t0 += tstride) {
        float nxt[2][4];
        load_raw(t0 + tstride, nxt);
        bf16x4 xb[2];
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 4 * g + r;
                const float v = k < kUsedIn ? cur[nb][r]
                                            : (float)(((uint32_t)cur[nb][r] >> (k < kUsedIn + kSceneBits ? k - kUsedIn : 0)) & 1u);
                xb[nb][r] = (short)f2bf(k < kUsedIn || (XS == kRec && k < kUsedIn + kSceneBits) ? v : 0.0f);
            }
        }
        bf16x4 h1[4][2], h2[4][2];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const bf16x4 a = ld4(S.wgt + LT0 + (mb * 16 + i) * kS16 + 4 * g);
            const f32x4 b = bias4(S.bias + mb * 16 + 4 * g);
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) h1[mb][nb] = relu_pk(mfma(a, xb[nb], b));
        }
        fwd_hidden<true>(S.wgt + LT1, S.bias + kHid, h1, h2, i, g);
        fwd_hidden<true>(S.wgt + LT2, S.bias + 2 * kHid, h2, h1, i, g);   // h1 now holds the third layer
        layer_fence();
        if (DECODE) wave_sync();                                        // the previous tile's decode has read R
#pragma unroll
        for (int mb = 0; mb < 3; ++mb) {
            const uint16_t* row = S.wgt + LT3 + (mb * 16 + i) * kS64;
            const s16x8 a0 = wfrag(row, 0, g), a1 = wfrag(row, 1, g);
            const f32x4 b = bias4(S.bias + 3 * kHid + mb * 16 + 4 * g);
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                f32x4 c = mfma32(a0, cat(h1[0][nb], h1[1][nb]), b);
                c = mfma32(a1, cat(h1[2][nb], h1[3][nb]), c);
                const float4 v = make_float4(c[0], c[1], c[2], c[3]);
                const uint32_t u0 = mb * 16 + 4 * g, sl = nb * 16 + i;
                if (DECODE) *reinterpret_cast<float4*>(&R[sl * kRS + u0]) = v;
                else if (u0 < kUsedOut && t0 + sl < nn) *reinterpret_cast<float4*>(&out[(t0 + sl) * kUsedOut + u0]) = v;
            }
        }
        if (DECODE) {
            constexpr float kLog2e = 1.4426950408889634f;
            wave_sync();
            const int sl = lane & (kTT - 1), h = lane >> 5;
            const float* o = &R[sl * kRS];
            float mx = -3.0e38f;
#pragma unroll
            for (int j = 0; j < 4; ++j) mx = fmaxf(mx, o[5 * (4 * h + j)]);
            mx = fmaxf(mx, __shfl_xor(mx, 32));
            float pu[4], z = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) { pu[j] = __builtin_amdgcn_exp2f((o[5 * (4 * h + j)] - mx) * kLog2e); z += pu[j]; }
            z += __shfl_xor(z, 32);
            const float iz = __builtin_amdgcn_rcpf(z);
            float dv[12];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = 4 * h + j;
                const float cx = o[5 * k + 2], cy = o[5 * k + 3], cz = o[5 * k + 4];
                const bool deg = !(cx * cx + cy * cy + cz * cz >= 1e-24f);
                dv[3 * j + 0] = pu[j] * iz;
                dv[3 * j + 1] = __builtin_amdgcn_exp2f(fminf(7.0f, fmaxf(-3.0f, o[5 * k + 1])) * kLog2e);
                dv[3 * j + 2] = __builtin_bit_cast(float, deg ? oct_pack(0.0f, 0.0f, 1.0f) : oct_pack(cx, cy, cz));
            }
            // the stored tile (32 vertices x 24 words) goes back through R,
            // then out in whole 16-byte pieces of consecutive vertices
            wave_sync();                                            // every lane has read its raw outputs
#pragma unroll
            for (int q = 0; q < 3; ++q)
                *reinterpret_cast<float4*>(&R[sl * kRS + 12 * h + 4 * q]) =
                    make_float4(dv[4 * q], dv[4 * q + 1], dv[4 * q + 2], dv[4 * q + 3]);
            wave_sync();
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const uint32_t c = lane + 64 * q, v = c / 6, part = c - 6 * v;
                if (t0 + v < nn)
                    *reinterpret_cast<float4*>(&out[(t0 + v) * kPacked + 4 * part]) =
                        *reinterpret_cast<const float4*>(&R[v * kRS + 4 * part]);
            }
        }
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int r = 0; r < 4; ++r) cur[nb][r] = nxt[nb][r];
    }
}

// Fixed-order sums of the block partials: chunks of kRedChunk blocks, then the chunk sums.
constexpr int kRedChunk = 16;
// tick (optional): the Adam step count, advanced here when an Adam step
// follows (k_mlp_reduce2_adam reads it after this kernel has finished)
__global__ void k_mlp_reduce1(const float* __restrict__ part, const float* __restrict__ part_loss, int blocks,
                              float* __restrict__ part2, float* __restrict__ loss2, float* __restrict__ tick) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (tick && p == 0 && c == 0) tick[0] += 1.0f;
    const int b0 = c * kRedChunk, b1 = std::min(blocks, b0 + kRedChunk);
    if (p < kParams) {
        float s = 0.0f;
        for (int b = b0; b < b1; ++b) s += part[(size_t)b * kParams + p];
        part2[(size_t)c * kParams + p] = s;
    }
    if (p == kParams) {
        float s = 0.0f;
        for (int b = b0; b < b1; ++b) s += part_loss[b];
        loss2[c] = s;
    }
}
__global__ void k_mlp_reduce2(const float* __restrict__ part2, const float* __restrict__ loss2, int chunks,
                              float* __restrict__ g, float* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kParams) {
        float s = 0.0f;
        for (int c = 0; c < chunks; ++c) s += part2[(size_t)c * kParams + p];
        g[p] = s;
    }
    if (p == kParams) {
        float s = 0.0f;
        for (int c = 0; c < chunks; ++c) s += loss2[c];
        *loss = s;
    }
}

// Per-block sums of the sample weights in a fixed order (strided thread
// sums, then an LDS tree); k_inverse adds the block sums in block order and
// writes scale = 1 / max(sum, tiny).
constexpr int kWsBlocks = 256;
__global__ __launch_bounds__(256) void k_weight_sum(const float* __restrict__ w, size_t n,
                                                    const uint32_t* __restrict__ n_dev, float* __restrict__ partial) {
    __shared__ float t[256];
    const size_t nn = n_dev ? std::min<size_t>(n, *n_dev) : n;
    float s = 0.0f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x) s += w[i];
    t[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) t[threadIdx.x] += t[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = t[0];
}
__global__ __launch_bounds__(256) void k_inverse(const float* __restrict__ partial, int blocks, float* v) {
    __shared__ float t[256];
    t[threadIdx.x] = (int)threadIdx.x < blocks ? partial[threadIdx.x] : 0.0f;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) t[threadIdx.x] += t[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) v[0] = 1.0f / fmaxf(t[0], 1e-30f);
}
__global__ void k_set(float* v, float x) { v[0] = x; }

struct Adam {
    float* p; float* m; float* v; float* g;
    uint16_t* wb; uint16_t* wt;
    const float* t;                  // step count (device; advanced by k_mlp_reduce1)
    float lr, b1, b2, eps;
};



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

// k_mlp_reduce2 and k_adam in one pass: the gradient of each parameter is
// summed over the chunks in the same order and applied at once (the same
// arithmetic as the two kernels; the gradient buffer is left zero as k_adam
// leaves it)
__global__ void k_mlp_reduce2_adam(const float* __restrict__ part2, const float* __restrict__ loss2, int chunks,
                                   float* __restrict__ loss, Adam A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == kParams) {
        float s = 0.0f;
        for (int c = 0; c < chunks; ++c) s += loss2[c];
        *loss = s;
    }
    if (i >= kParams) return;
    float g = 0.0f;
    for (int c = 0; c < chunks; ++c) g += part2[(size_t)c * kParams + i];
    const float t = A.t[0];
    const float c1 = 1.0f - powf(A.b1, t), c2 = 1.0f - powf(A.b2, t);
    const float m = A.b1 * A.m[i] + (1.0f - A.b1) * g;
    const float v = A.b2 * A.v[i] + (1.0f - A.b2) * g * g;
    A.m[i] = m;
    A.v[i] = v;
    A.p[i] -= A.lr * (m / c1) / (sqrtf(v / c2) + A.eps);
    A.g[i] = 0.0f;
    const int tr = transposed(i);
    if (tr >= 0) {
        const uint16_t b = f2bf(A.p[i]);
        A.wb[i] = b;
        A.wt[tr] = b;
    }
}

__global__ void k_zero(float* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0f;
}

// blocks of a training launch over n samples (at most the resident blocks the partials are sized for)
uint32_t train_grid(const MlpDev& M, size_t n) {
    return (uint32_t)std::max<size_t>(1, std::min<size_t>((n + kTR - 1) / kTR, M.part_blocks));
}

struct Slots {
    const float* gr;       // [5*4][stride] path radiance where the slot was completed, less that vertex's emission
    const float* gm;       // [5*4][stride] path throughput there
    const uint32_t* gflag; // bit i: slot i complete
    const float* gx;       // [5*kRec][stride]
    const float* gw;       // [5*3][stride]
    const float* gp;       // [5][stride]
    const float* rad[4];   // the paths' radiance at the end of the frame
    uint32_t stride;
};

// Training records of a frame: every complete slot gives (inputs, direction,
// luminance(L) / pdf(direction)), with L = (radiance at the end of the path -
// gr) / gm the radiance that arrived along the slot's direction (the next
// vertex's unweighted emission, its NEE and every later bounce): the
// Monte-Carlo weight of the KL divergence from the incident radiance to the
// guide (Muller et al. 2019).  Weighting by the cosine as well (the
// integrand of a diffuse vertex) trained a worse guide for the one-sample
// mixture with the BSDF, which supplies the cosine.  The records
// are numbered in (path, slot) order by a block count, a scan and a write
// pass, so a frame's records and the trained network do not depend on
// scheduling.  When a frame has T records and the training batch is B < T,
// the batch is records floor(j T / B), j < B: evenly strided over the frame
// (over the image), not its first B.
__device__ __forceinline__ uint32_t record_weights(const Slots& S, uint32_t p, uint32_t npaths, uint32_t spectral,
                                                   const float* __restrict__ lambda0, const float* __restrict__ cie,
                                                   float w[5]) {
    const bool live = p < npaths;
    const uint32_t flags = live ? S.gflag[p] : 0u;
    const Lam L = make_lam(spectral, live && spectral != AR_SPECTRAL_RGB ? lambda0[p] : 0.0f);
    uint32_t mask = 0;
    for (int i = 0; i < 5; ++i) {
        w[i] = 0.0f;
        if ((flags >> i) & 1u) {
            const float* base = S.gr + (size_t)(4 * i) * S.stride + p;
            const float* thr = S.gm + (size_t)(4 * i) * S.stride + p;
            float li[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float m = thr[(size_t)k * S.stride];
                li[k] = m > 0.0f ? fmaxf(0.0f, (S.rad[k][p] - base[(size_t)k * S.stride]) / m) : 0.0f;
            }
            const float* r = li;
            const size_t rs = 1;
            float lum;
            if (spectral == AR_SPECTRAL_RGB) {
                lum = r[0] * 0.2126f + r[rs] * 0.7152f + r[2 * rs] * 0.0722f;
            } else if (spectral == AR_SPECTRAL_SINGLE) {
                float cx, cy, cz;
                spec::cie_xyz(cie, L.l[0], cx, cy, cz);
                lum = cy * r[0] / spec::lambda_pdf(L.l[0]);
            } else {
                lum = 0.0f;
                for (int k = 0; k < 4; ++k) {
                    float cx, cy, cz;
                    spec::cie_xyz(cie, L.l[k], cx, cy, cz);
                    lum += cy * r[k * rs];
                }
                lum = lum / (4.0f / 400.0f);
            }
            const float pdf = S.gp[(size_t)i * S.stride + p];
            w[i] = pdf > 0.0f ? lum / pdf : 0.0f;
        }
        if (w[i] > 0.0f && w[i] < 3.0e38f) mask |= 1u << i;
    }
    return mask;
}

constexpr uint32_t kRecBlock = 256;

// The paths a frame's records come from: `quota` per segment (scene) of the
// pass, so that every scene of a guided batch trains the network alike
// whatever its size; in a segment of m paths, path k < min(quota, m) is one
// per stratum of m / min(quota, m) consecutive paths, at a hashed offset
// inside it (a fixed function of k, so the records stay a deterministic
// function of the frame).  ranges: (first path, path count) per segment.
// Returns npaths (no path) for unused selection slots.
__device__ __forceinline__ uint32_t record_path(uint32_t j, uint32_t quota, const uint32_t* __restrict__ ranges,
                                               uint32_t nseg, uint32_t npaths) {
    const uint32_t sg = j / quota, k = j - sg * quota;
    if (sg >= nseg) return npaths;
    const uint32_t base = ranges[2 * sg], m = ranges[2 * sg + 1];
    const uint32_t sel = min(quota, m);
    if (k >= sel) return npaths;
    const uint64_t a = (uint64_t)k * m / sel, b = (uint64_t)(k + 1) * m / sel;
    uint32_t h = j * 0x9E3779B1u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    return base + (uint32_t)(a + h % (uint32_t)(b - a));
}

// in-block exclusive scan of c (Hillis-Steele over LDS); returns the prefix, *total the block sum
__device__ __forceinline__ uint32_t block_exclusive(uint32_t* sc, uint32_t c, uint32_t* total) {
    const uint32_t t = threadIdx.x;
    sc[t] = c;
    __syncthreads();
    for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
        const uint32_t v = t >= off ? sc[t - off] : 0u;
        __syncthreads();
        sc[t] += v;
        __syncthreads();
    }
    *total = sc[blockDim.x - 1];
    return sc[t] - c;
}

__global__ __launch_bounds__(kRecBlock) void k_rec_count(Slots S, uint32_t npaths, uint32_t quota,
                                                         const uint32_t* __restrict__ ranges, uint32_t nseg,
                                                         uint32_t spectral, const float* __restrict__ lambda0,
                                                         const float* __restrict__ cie, uint32_t* __restrict__ blk) {
    __shared__ uint32_t sc[kRecBlock];
    float w[5];
    const uint32_t j = blockIdx.x * kRecBlock + threadIdx.x;
    const uint32_t m = record_weights(S, record_path(j, quota, ranges, nseg, npaths), npaths, spectral, lambda0, cie,
                                      w);
    uint32_t total;
    (void)block_exclusive(sc, (uint32_t)__popc(m), &total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// exclusive offsets of the block counts in place; count[0] = min(T, batch), count[1] = T
__global__ __launch_bounds__(1024) void k_rec_scan(uint32_t* __restrict__ blk, uint32_t nblocks, uint32_t batch,
                                                   uint32_t* __restrict__ count) {
    __shared__ uint32_t sc[1024];
    const uint32_t t = threadIdx.x, per = (nblocks + 1023u) / 1024u;
    const uint32_t b0 = min(nblocks, t * per), b1 = min(nblocks, b0 + per);
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; ++b) sum += blk[b];
    uint32_t total;
    uint32_t run = block_exclusive(sc, sum, &total);
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t c = blk[b];
        blk[b] = run;
        run += c;
    }
    if (t == 0) {
        count[0] = min(total, batch);
        count[1] = total;
    }
}

__global__ __launch_bounds__(kRecBlock) void k_mlp_records(Slots S, uint32_t npaths, uint32_t quota,
                                                           const uint32_t* __restrict__ ranges, uint32_t nseg,
                                                           uint32_t spectral, const float* __restrict__ lambda0,
                                                           const float* __restrict__ cie,
                                                           const uint32_t* __restrict__ blk,
                                                           const uint32_t* __restrict__ count, float* rx, float* rd,
                                                           float* rw) {
    __shared__ uint32_t sc[kRecBlock];
    const uint32_t p = record_path(blockIdx.x * kRecBlock + threadIdx.x, quota, ranges, nseg, npaths);
    float w[5];
    const uint32_t m = record_weights(S, p, npaths, spectral, lambda0, cie, w);
    uint32_t total;
    uint32_t r = blk[blockIdx.x] + block_exclusive(sc, (uint32_t)__popc(m), &total);
    const uint32_t B = count[0], T = count[1];
    for (int i = 0; i < 5; ++i) {
        if (!((m >> i) & 1u)) continue;
        // record r is batch slot j = ceil(r B / T) when floor(j T / B) = r
        uint64_t j = r;
        bool keep = true;
        if (T > B) {
            j = ((uint64_t)r * B + T - 1) / T;
            keep = j < B && (j * T) / B == r;
        }
        if (keep) {
            for (int k = 0; k < kRec; ++k) rx[kRec * j + k] = S.gx[(size_t)(kRec * i + k) * S.stride + p];
            for (int k = 0; k < 3; ++k) rd[3 * j + k] = S.gw[(size_t)(3 * i + k) * S.stride + p];
            rw[j] = w[i];
        }
        ++r;
    }
}

} // namespace

int mlp_ensure_buffers(ar_ctx* ctx, size_t paths) {
    MlpDev& M = ctx->mlp;
    // a frame's records fill at most one training batch
    const size_t rcap = std::min<size_t>(5 * paths, M.batch);
    if (M.cap >= paths && M.rcap >= rcap) return AR_OK;
    paths = std::max(paths, M.cap);
    invalidate_graphs(ctx);
    for (float** q : {&M.feat, &M.dec, &M.rx, &M.rd, &M.rw})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    if (M.rcount) { (void)hipFree(M.rcount); M.rcount = nullptr; }
    if (M.rblk) { (void)hipFree(M.rblk); M.rblk = nullptr; }
    M.cap = M.rcap = M.rblocks = 0;
    hipError_t e = hipMalloc((void**)&M.feat, paths * kRec * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M.dec, paths * kPacked * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M.rx, rcap * kRec * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M.rd, rcap * 3 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M.rw, rcap * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M.rcount, 4 * sizeof(uint32_t));
    // record selection: at most batch + kMaxGuidedScenes threads (one per slot of each scene's quota)
    const size_t rblocks = (std::max<size_t>(paths, M.batch) + kMaxGuidedScenes + kRecBlock - 1) / kRecBlock;
    if (e == hipSuccess) e = hipMalloc((void**)&M.rblk, rblocks * sizeof(uint32_t));
    if (e != hipSuccess) return set_error(ctx, AR_E_OOM, "mlp: device allocation failed");
    M.cap = paths;
    M.rcap = rcap;
    M.rblocks = rblocks;
    return AR_OK;
}

int mlp_frame_update(ar_ctx* ctx, uint32_t spectral, const float* lambda0, const float* cie, size_t paths,
                     const uint32_t* ranges, uint32_t nseg, hipStream_t s) {
    MlpDev& M = ctx->mlp;
    const GuideDev& g = ctx->guide;
    const RenderBuffers& b = ctx->rb;
    Slots S{g.gr, g.gm, g.gflag, g.gx, g.gw, g.gp, {b.rad[0], b.rad[1], b.rad[2], b.rad[3]}, (uint32_t)g.pcap};
    const size_t n = std::min<size_t>(M.rcap, M.batch);
    // records come from one path per batch slot (each gives up to 5), the
    // slots shared equally by the frame's scenes and spread evenly over each
    // scene's paths
    nseg = std::max(1u, nseg);
    uint32_t quota = (uint32_t)std::max<size_t>(1, (n + nseg - 1) / nseg);
    quota = (uint32_t)std::min<size_t>(quota, M.rblocks * kRecBlock / nseg);   // fits the block counts
    const uint32_t nb = (uint32_t)(((size_t)quota * nseg + kRecBlock - 1) / kRecBlock);
    hipLaunchKernelGGL(k_rec_count, dim3(nb), dim3(kRecBlock), 0, s, S, (uint32_t)paths, quota, ranges, nseg,
                       spectral, lambda0, cie, M.rblk);
    hipLaunchKernelGGL(k_rec_scan, dim3(1), dim3(1024), 0, s, M.rblk, nb, (uint32_t)n, M.rcount);
    hipLaunchKernelGGL(k_mlp_records, dim3(nb), dim3(kRecBlock), 0, s, S, (uint32_t)paths, quota, ranges, nseg,
                       spectral, lambda0, cie, M.rblk, M.rcount, M.rx, M.rd, M.rw);
    hipError_t e = hipSuccess;
    // the weight normaliser depends only on the records: computed by the first step
    for (uint32_t it = 0; it < M.iters && e == hipSuccess; ++it)
        e = mlp_gradient(ctx, M.rx, kRec, M.rd, M.rw, n, M.rcount, it == 0 ? kNormSum : kNormKeep, s, true);
    return check_hip(ctx, e, "guiding network training");
}

void mlp_free(ar_ctx* ctx) {
    invalidate_graphs(ctx);
    MlpDev& M = ctx->mlp;
    for (float** q : {&M.p, &M.feat, &M.dec, &M.rx, &M.rd, &M.rw, &M.part})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    if (M.rcount) { (void)hipFree(M.rcount); M.rcount = nullptr; }
    if (M.rblk) { (void)hipFree(M.rblk); M.rblk = nullptr; }
    M.m = M.v = M.g = M.scal = nullptr;
    M.part2 = M.ploss = M.ploss2 = M.wpart = nullptr;
    M.part_blocks = 0;
    M.wb = M.wt = nullptr;
    M.cap = M.rcap = M.rblocks = 0;
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
    // training partials: one gradient per resident block of k_mlp_train, chunk sums, losses, weight sums
    int bpc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_mlp_train, kTB, 0) != hipSuccess || bpc < 1) bpc = 1;
    M.part_blocks = (uint32_t)std::max(1, ctx->num_cus) * (uint32_t)std::min(bpc, 4);
    const size_t chunks = (M.part_blocks + kRedChunk - 1) / kRedChunk;
    const size_t pf = ((size_t)M.part_blocks + chunks) * kParams + M.part_blocks + chunks + kWsBlocks;
    e = hipMalloc((void**)&M.part, pf * sizeof(float));
    if (e != hipSuccess) { mlp_free(ctx); return set_error(ctx, AR_E_OOM, "mlp: device allocation failed"); }
    M.part2 = M.part + (size_t)M.part_blocks * kParams;
    M.ploss = M.part2 + chunks * kParams;
    M.ploss2 = M.ploss + M.part_blocks;
    M.wpart = M.ploss2 + chunks;
    return mlp_init(ctx, M.seed);
}

int mlp_init(ar_ctx* ctx, uint64_t seed) {
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
    // a device count bounds one launch (the integrator's queues hold at most
    // kInferMax vertices); a host count is split into launches of kInferMax
    if (n_dev && n > kInferMax) return hipErrorInvalidValue;
    const int ow = decoded ? kPacked : kUsedOut;
    for (size_t first = 0; first < n; first += kInferMax) {
        const uint32_t m = (uint32_t)std::min(n - first, kInferMax);
        const float* xf = x5 + first * xs;
        float* of = out + first * ow;
        // persistent: two 4-wave blocks per CU, each wave looping over 32-sample tiles
        const uint32_t grid = (uint32_t)std::min<size_t>((m + kIW * kTT - 1) / (kIW * kTT),
                                                         (size_t)std::max(ctx->num_cus, 1) * 2);
        const dim3 g(std::max(grid, 1u)), b(kIW * 64);
        if (decoded && xs == kRec) hipLaunchKernelGGL((k_mlp_infer<true, kRec>), g, b, 0, s, N, xf, m, of, n_dev);
        else if (decoded) hipLaunchKernelGGL((k_mlp_infer<true, kUsedIn>), g, b, 0, s, N, xf, m, of, n_dev);
        else if (xs == kRec) hipLaunchKernelGGL((k_mlp_infer<false, kRec>), g, b, 0, s, N, xf, m, of, n_dev);
        else hipLaunchKernelGGL((k_mlp_infer<false, kUsedIn>), g, b, 0, s, N, xf, m, of, n_dev);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// gradient into M.g (zeroed first unless accumulate) and loss into M.scal[1];
// normalize: scale = 1 / sum(weights), else 1 / n
hipError_t mlp_gradient(ar_ctx* ctx, const float* x5, int xs, const float* dir, const float* w, size_t n,
                        const uint32_t* n_dev, int norm, hipStream_t s, bool adam) {
    MlpDev& M = ctx->mlp;
    Net N{M.p, M.wb, M.wt};
    if (norm == kNormSum) {
        hipLaunchKernelGGL(k_weight_sum, dim3(kWsBlocks), dim3(256), 0, s, w, n, n_dev, M.wpart);
        hipLaunchKernelGGL(k_inverse, dim3(1), dim3(256), 0, s, M.wpart, kWsBlocks, M.scal);
    } else if (norm == kNormCount) {
        hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, M.scal, 1.0f / (float)std::max<size_t>(n, 1));
    }
    const uint32_t grid = train_grid(M, n);
    hipLaunchKernelGGL(k_mlp_train, dim3(grid), dim3(kTB), 0, s, N, x5, xs, dir, w, n, n_dev, M.scal, M.part, M.ploss);
    const uint32_t chunks = (grid + kRedChunk - 1) / kRedChunk;
    const uint32_t pblocks = (uint32_t)((kParams + 2 + 255) / 256);
    hipLaunchKernelGGL(k_mlp_reduce1, dim3(pblocks, chunks), dim3(256), 0, s, M.part, M.ploss, (int)grid, M.part2,
                       M.ploss2, adam ? M.scal + 2 : nullptr);
    if (adam) {
        Adam A{M.p, M.m, M.v, M.g, M.wb, M.wt, M.scal + 2, M.lr, 0.9f, 0.999f, 1e-8f};
        hipLaunchKernelGGL(k_mlp_reduce2_adam, dim3(pblocks), dim3(256), 0, s, M.part2, M.ploss2, (int)chunks,
                           M.scal + 1, A);
    } else {
        hipLaunchKernelGGL(k_mlp_reduce2, dim3(pblocks), dim3(256), 0, s, M.part2, M.ploss2, (int)chunks, M.g,
                           M.scal + 1);
    }
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

int ar_mlp_records_get(ar_ctx* ctx, float* x, float* dir, float* weight, size_t cap, uint32_t* count,
                       uint32_t* total) {
    if (!ctx) return AR_E_INVALID;
    const MlpDev& M = ctx->mlp;
    uint32_t c[2] = {0, 0};
    int rc = check_hip(ctx, hipDeviceSynchronize(), "mlp records");
    if (rc == AR_OK && M.rcount)
        rc = check_hip(ctx, hipMemcpy(c, M.rcount, sizeof c, hipMemcpyDeviceToHost), "mlp records");
    if (rc != AR_OK) return rc;
    if (count) *count = c[0];
    if (total) *total = c[1];
    const size_t n = std::min<size_t>(c[0], cap);
    if (n && x) rc = check_hip(ctx, hipMemcpy(x, M.rx, n * kRec * sizeof(float), hipMemcpyDeviceToHost), "mlp records");
    if (rc == AR_OK && n && dir)
        rc = check_hip(ctx, hipMemcpy(dir, M.rd, n * 3 * sizeof(float), hipMemcpyDeviceToHost), "mlp records");
    if (rc == AR_OK && n && weight)
        rc = check_hip(ctx, hipMemcpy(weight, M.rw, n * sizeof(float), hipMemcpyDeviceToHost), "mlp records");
    return rc;
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
    hipError_t e = mlp_gradient(ctx, x5, kUsedIn, dir, weight, n, nullptr, kNormCount, s, false);
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
    for (uint32_t t = 0; t < steps && e == hipSuccess; ++t)
        e = mlp_gradient(ctx, x5, kUsedIn, dir, weight, n, nullptr, t == 0 ? kNormSum : kNormKeep, s, true);
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, ctx->mlp.scal + 1, sizeof(float), hipMemcpyDeviceToDevice, s);
    return check_hip(ctx, e, "mlp train");
}

} // extern "C"
