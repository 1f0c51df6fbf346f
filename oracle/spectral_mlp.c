/*
 * Spectral basis network of the reference (TEST INFRASTRUCTURE ONLY: the
 * library never links or calls this file).
 *
 * Inference, FP32, in the reference's operation order
 * (core/spectrum/spectrum_basis.cuh:2273-2470, core/tensor/tensor.cuh:57-160):
 *   RGBEncode     s = max(r, g, b); sin / cos of pi k c / (s + 0.001) for
 *                 k = 1..4 and c = r, g, b, and of pi k s; Linear 32 -> 32
 *                 + ReLU; 7 blocks h += ReLU(W h + b); Linear 32 -> 4 without
 *                 bias; times s
 *   LambdaEncode  x = clamp((lambda - 360) / 420, 0, 1); sin / cos of pi k x
 *                 for k = 1..16; 7 blocks; Linear 32 -> 4 without bias
 *   Interact      clamp(sum_k K_rgb[k] K_lambda[k], 0, 1), an FMA chain
 *   LUTs          spectrum_lut.cuh:12-235: RGB table at voxel centres
 *                 (x + 0.5) / res, [z][y][x][4]; wavelength table at
 *                 360 + (i + 0.5) 420 / res nm; trilinear / linear queries
 *                 with coordinates c res - 0.5 clamped to [0, res - 1];
 *                 Interact divides an RGB above 1 by its largest channel
 *                 and scales the result back
 *
 * Training, FP64 (pytorch/spectrum/train_spectrum.py):
 *   SpectralMLP.forward in training mode (:103-190): Gaussian noise of
 *   standard deviation 0.025 / 256 on the RGB and 0.025 / 420 on the
 *   normalised wavelength of the first half of the batch, and 0.001 on the
 *   input of every block's Linear; no clamp of the normalised wavelength;
 *   loss (:733-739) mean over the batch and the three CIE channels of
 *   ((pred - target) (0.25 + max(cie_c(lambda), 0)))^2 with the analytic CIE
 *   fits of :476-494; clamp01 = relu(1 - relu(1 - x)) passes the gradient
 *   only for 0 < x < 1; gradient clipping to a global norm (:744-745);
 *   AdamW with decoupled weight decay on the weight matrices only
 *   (:28-50), torch's update order; warmup-cosine learning rate (:52-60);
 *   EMA of the parameters (:62-73).
 *   The noise is not torch.randn: normals come in Box-Muller pairs drawn from
 *   a counter hash of (seed, step, sample in batch, slot pair), the same on
 *   the GPU.
 *
 * Parity pins (tests/test_oracle_spectral_mlp.py): the reference's golden
 * vectors (inputTest / gtTest) through the FP32 forward with the reference's
 * weights; the FP64 gradient against central differences and against
 * torch.autograd of the same forward; clipping against
 * torch.nn.utils.clip_grad_norm_; the AdamW step against torch.optim.AdamW.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

#define W 32
#define KOUT 4
#define NBLK 7
#define PI_F 3.141592654f
#define PI_D 3.14159265358979323846
#define M_PI_F32 3.14159265358979323846f

/* flat parameter layout (the header's order) */
#define O_RIN_W 0
#define O_RIN_B 1024
#define O_RBLK(k) (1056 + (k) * 1056)
#define O_ROUT 8448
#define O_LBLK(k) (8576 + (k) * 1056)
#define O_LOUT 15968
#define N_PARAMS 16096

int orc_smlp_param_count(void) { return N_PARAMS; }

/* 1 for entries that take weight decay (weight matrices), 0 for biases */
int orc_smlp_decays(int p) {
    if (p >= O_RIN_B && p < O_RIN_B + W) return 0;
    for (int k = 0; k < NBLK; ++k) {
        if (p >= O_RBLK(k) + 1024 && p < O_RBLK(k) + 1056) return 0;
        if (p >= O_LBLK(k) + 1024 && p < O_LBLK(k) + 1056) return 0;
    }
    return 1;
}

/* ---------------- FP32 inference ---------------- */

/* TF32 operand rounding (10-bit mantissa, round to nearest, ties away) */
static float tf32(float v) {
    uint32_t u;
    memcpy(&u, &v, 4);
    u = (u + 0x1000u) & 0xFFFFE000u;
    memcpy(&v, &u, 4);
    return v;
}

/* y = W x (+ b) (+ ReLU), W row-major [to][from]; an FMA chain over j from the
 * bias (LinearRowMajor); with t32 the operands are rounded to TF32 first */
static void lin_x(const float* w, const float* b, const float* x, float* y, int from, int to, int relu, int t32) {
    for (int i = 0; i < to; ++i) {
        float acc = b ? b[i] : 0.0f;
        for (int j = 0; j < from; ++j)
            acc = t32 ? fmaf(tf32(x[j]), tf32(w[i * from + j]), acc) : fmaf(x[j], w[i * from + j], acc);
        y[i] = relu ? fmaxf(acc, 0.0f) : acc;
    }
}

static void lin_f(const float* w, const float* b, const float* x, float* y, int from, int to, int relu) {
    lin_x(w, b, x, y, from, to, relu, 0);
}

static void blocks_x(const float* w, int first, float* h, int t32) {
    float t[W];
    for (int k = 0; k < NBLK; ++k) {
        const int o = first + k * 1056;
        lin_x(w + o, w + o + 1024, h, t, W, W, 1, t32);
        for (int i = 0; i < W; ++i) h[i] = h[i] + t[i];
    }
}

static void blocks_f(const float* w, int first, float* h) { blocks_x(w, first, h, 0); }

void orc_smlp_rgb_encode(const float* w, const float rgb[3], float out[4]) {
    const float s = fmaxf(fmaxf(rgb[0], rgb[1]), rgb[2]);
    const float inv = 1.0f / (s + 0.001f);
    float e[W], h[W];
    for (int c = 0; c < 4; ++c) {
        for (int k = 0; k < 4; ++k) {
            const float ph = c < 3 ? rgb[c] * inv * PI_F * (float)(k + 1) : s * PI_F * (float)(k + 1);
            e[8 * c + k] = sinf(ph);
            e[8 * c + 4 + k] = cosf(ph);
        }
    }
    lin_f(w + O_RIN_W, w + O_RIN_B, e, h, W, W, 1);
    blocks_f(w, O_RBLK(0), h);
    lin_f(w + O_ROUT, 0, h, out, W, KOUT, 0);
    for (int k = 0; k < KOUT; ++k) out[k] *= s;
}

void orc_smlp_lambda_encode(const float* w, float lam, float out[4]) {
    const float x = fminf(fmaxf((lam - 360.0f) / (780.0f - 360.0f), 0.0f), 1.0f);
    float h[W];
    for (int k = 0; k < 16; ++k) {
        const float ph = x * PI_F * (float)(k + 1);
        h[k] = sinf(ph);
        h[16 + k] = cosf(ph);
    }
    blocks_f(w, O_LBLK(0), h);
    lin_f(w + O_LOUT, 0, h, out, W, KOUT, 0);
}

float orc_smlp_interact(const float a[4], const float b[4]) {
    float acc = 0.0f;
    for (int k = 0; k < KOUT; ++k) acc = fmaf(a[k], b[k], acc);
    return fminf(fmaxf(acc, 0.0f), 1.0f);
}

/* x: n x 4 (r, g, b, lambda nm) -> out: n reflectances */
void orc_smlp_eval(const float* w, size_t n, const float* x, float* out) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; ++i) {
        float a[4], b[4];
        orc_smlp_rgb_encode(w, x + 4 * i, a);
        orc_smlp_lambda_encode(w, x[4 * i + 3], b);
        out[i] = orc_smlp_interact(a, b);
    }
}

/* The evaluation the reference's golden vectors come from
 * (train_spectrum.py:417-421: model.forward in eval mode on the GPU with
 * torch.backends.cuda.matmul.allow_tf32, :20): SinCosEmbedding's phases
 * (c / (s + 0.001)) * float(pi k) and x * float(pi k), every Linear's operands
 * rounded to TF32, FP32 accumulation.  x holds the wavelength in nm. */
void orc_smlp_eval_torch_tf32(const float* w, size_t n, const float* x, float* out) {
    for (size_t i = 0; i < n; ++i) {
        const float* r = x + 4 * i;
        const float s = fmaxf(fmaxf(r[0], r[1]), r[2]);
        const float se = s + 0.001f;
        float e[W], h[W], a[4], b[4];
        for (int c = 0; c < 4; ++c)
            for (int k = 0; k < 4; ++k) {
                const float om = (float)M_PI_F32 * (float)(k + 1);
                const float ph = (c < 3 ? r[c] / se : s) * om;
                e[8 * c + k] = sinf(ph);
                e[8 * c + 4 + k] = cosf(ph);
            }
        lin_x(w + O_RIN_W, w + O_RIN_B, e, h, W, W, 1, 1);
        blocks_x(w, O_RBLK(0), h, 1);
        lin_x(w + O_ROUT, 0, h, a, W, KOUT, 0, 1);
        for (int k = 0; k < KOUT; ++k) a[k] *= s;
        const float lx = (r[3] - 360.0f) / 420.0f;
        for (int k = 0; k < 16; ++k) {
            const float ph = lx * ((float)M_PI_F32 * (float)(k + 1));
            h[k] = sinf(ph);
            h[16 + k] = cosf(ph);
        }
        blocks_x(w, O_LBLK(0), h, 1);
        lin_x(w + O_LOUT, 0, h, b, W, KOUT, 0, 1);
        float acc = 0.0f;
        for (int k = 0; k < KOUT; ++k) acc += a[k] * b[k];
        out[i] = fminf(fmaxf(acc, 0.0f), 1.0f);
    }
}

/* floorf-based coordinate of __linear_coord */
static void lin_coord(float g, int res, int* i0, int* i1, float* t) {
    const float f = floorf(g);
    *i0 = (int)f;
    *i1 = *i0 + 1 >= res ? res - 1 : *i0 + 1;
    *t = *i0 == *i1 ? 0.0f : g - f;
}

/* RGB basis of QueryRGBBasis at table resolution res, each corner evaluated
 * directly at its voxel centre (the table entry) */
void orc_smlp_query_rgb(const float* w, int res, const float rgb[3], float out[4]) {
    int i0[3], i1[3];
    float t[3];
    for (int c = 0; c < 3; ++c) {
        float g = rgb[c] * (float)res - 0.5f;
        g = fminf(fmaxf(g, 0.0f), (float)res - 1.0f);
        lin_coord(g, res, &i0[c], &i1[c], &t[c]);
    }
    for (int k = 0; k < 4; ++k) out[k] = 0.0f;
    const float inv = 1.0f / (float)res;
    /* corner order of QueryRGBBasis: 000 100 010 110 001 101 011 111 */
    for (int corner = 0; corner < 8; ++corner) {
        const int bx = corner & 1, by = (corner >> 1) & 1, bz = corner >> 2;
        const int ix = bx ? i1[0] : i0[0], iy = by ? i1[1] : i0[1], iz = bz ? i1[2] : i0[2];
        const float wx = bx ? t[0] : 1.0f - t[0], wy = by ? t[1] : 1.0f - t[1], wz = bz ? t[2] : 1.0f - t[2];
        const float wgt = wx * wy * wz;
        const float c3[3] = {((float)ix + 0.5f) * inv, ((float)iy + 0.5f) * inv, ((float)iz + 0.5f) * inv};
        float v[4];
        orc_smlp_rgb_encode(w, c3, v);
        for (int k = 0; k < 4; ++k) out[k] = fmaf(wgt, v[k], out[k]);
    }
}

/* wavelength-table entry i of a res-entry table */
void orc_smlp_lambda_entry(const float* w, int res, int i, float out[4]) {
    const float step = (780.0f - 360.0f) / (float)res;
    orc_smlp_lambda_encode(w, 360.0f + ((float)i + 0.5f) * step, out);
}

/* QueryLambda at table resolution res */
void orc_smlp_query_lambda(const float* w, int res, float lam, float out[4]) {
    float u = (lam - 360.0f) / (780.0f - 360.0f);
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    float g = u * (float)res - 0.5f;
    g = fminf(fmaxf(g, 0.0f), (float)res - 1.0f);
    int i0, i1;
    float t;
    lin_coord(g, res, &i0, &i1, &t);
    float a[4], b[4];
    orc_smlp_lambda_entry(w, res, i0, a);
    orc_smlp_lambda_entry(w, res, i1, b);
    for (int k = 0; k < 4; ++k) out[k] = a[k] * (1.0f - t) + b[k] * t;
}

/* Interact: an RGB above 1 is divided by its largest channel first */
float orc_smlp_lut_interact(const float* w, int rgb_res, int lam_res, const float rgb[3], float lam) {
    float kl[4], kr[4];
    orc_smlp_query_lambda(w, lam_res, lam, kl);
    const float mx = fmaxf(fmaxf(rgb[0], rgb[1]), rgb[2]);
    if (mx > 1.0f) {
        const float n3[3] = {rgb[0] / mx, rgb[1] / mx, rgb[2] / mx};
        orc_smlp_query_rgb(w, rgb_res, n3, kr);
        return orc_smlp_interact(kr, kl) * mx;
    }
    orc_smlp_query_rgb(w, rgb_res, rgb, kr);
    return orc_smlp_interact(kr, kl);
}

/* ---------------- training noise ---------------- */

static uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

/* standard normal for (seed, step, sample, slot): Box-Muller on two 24-bit
 * uniforms of the slot pair slot / 2, u1 in (0, 1]; an even slot takes the
 * cosine, an odd slot the sine of the pair */
double orc_smlp_normal(uint32_t seed, uint32_t step, uint32_t sample, uint32_t slot) {
    const uint32_t base = mix32(mix32(seed + step * 0x9E3779B9u) ^ sample);
    const uint32_t pair = slot >> 1;
    const uint32_t a = mix32(base + 2u * pair), b = mix32(base + 2u * pair + 1u);
    const float u1 = (float)((a >> 8) + 1u) * (1.0f / 16777216.0f);
    const float u2 = (float)(b >> 8) * (1.0f / 16777216.0f);
    const float r = sqrtf(-2.0f * logf(u1)), th = 6.283185307f * u2;
    return (double)(r * ((slot & 1u) ? sinf(th) : cosf(th)));
}

/* ---------------- FP64 training step ---------------- */

static double cie_fit(int c, double l) {
    if (c == 0) {
        const double k1 = (l - 595.8) / 33.33, k2 = (l - 446.8) / 19.44;
        return 1.065 * exp(-0.5 * k1 * k1) + 0.366 * exp(-0.5 * k2 * k2);
    }
    if (c == 1) {
        const double k = (log(l) - 6.3213077) / 0.075;
        return 1.014 * exp(-0.5 * k * k);
    }
    const double k = (log(l) - 6.1088030) / 0.051;
    return 1.839 * exp(-0.5 * k * k);
}

/* loss weight sum_c (0.25 + max(cie_c, 0))^2 of a wavelength in nm */
double orc_smlp_importance2(double lam) {
    double s = 0.0;
    for (int c = 0; c < 3; ++c) {
        const double v = 0.25 + fmax(cie_fit(c, lam), 0.0);
        s += v * v;
    }
    return s;
}

typedef struct {
    double e[W];               /* embedding (RGB: input of the first Linear; lambda: h0) */
    double zin[W];             /* RGB first Linear pre-activation */
    double h[NBLK + 1][W];     /* block inputs h0..h6 and h7 */
    double xin[NBLK][W];       /* block Linear inputs (h + noise) */
    double z[NBLK][W];         /* block pre-activations */
    double k[KOUT];            /* output before the RGB scale */
} enc_state;

static void lin_d(const float* w, const float* b, const double* x, double* y, int from, int to) {
    for (int i = 0; i < to; ++i) {
        double acc = b ? (double)b[i] : 0.0;
        for (int j = 0; j < from; ++j) acc += x[j] * (double)w[i * from + j];
        y[i] = acc;
    }
}

static void blocks_fwd_d(const float* w, int first, enc_state* S, double bnoise, uint32_t seed, uint32_t step,
                         uint32_t j, uint32_t slot0) {
    for (int k = 0; k < NBLK; ++k) {
        const int o = first + k * 1056;
        for (int u = 0; u < W; ++u)
            S->xin[k][u] = S->h[k][u] + (bnoise > 0.0 ? bnoise * orc_smlp_normal(seed, step, j, slot0 + k * W + u) : 0.0);
        lin_d(w + o, w + o + 1024, S->xin[k], S->z[k], W, W);
        for (int u = 0; u < W; ++u) S->h[k + 1][u] = S->h[k][u] + fmax(S->z[k][u], 0.0);
    }
}

static void blocks_bwd_d(const float* w, int first, const enc_state* S, double* dh, double* g, int need_dh0) {
    for (int k = NBLK - 1; k >= 0; --k) {
        const int o = first + k * 1056;
        double dz[W];
        for (int u = 0; u < W; ++u) dz[u] = S->z[k][u] > 0.0 ? dh[u] : 0.0;
        for (int u = 0; u < W; ++u) {
            for (int i = 0; i < W; ++i) g[o + u * W + i] += dz[u] * S->xin[k][i];
            g[o + 1024 + u] += dz[u];
        }
        if (k > 0 || need_dh0)
            for (int i = 0; i < W; ++i) {
                double s = 0.0;
                for (int u = 0; u < W; ++u) s += (double)w[o + u * W + i] * dz[u];
                dh[i] += s;
            }
    }
}

/* One batch (B rows of x: r, g, b, lambda nm; targets y): loss (returned)
 * and its gradient added into grad (N_PARAMS doubles, zeroed by the caller).
 * noise != 0 adds the training-mode noise for (seed, step). */
double orc_smlp_loss_grad(const float* w, uint32_t B, const float* x, const float* y, int noise, uint32_t seed,
                          uint32_t step, double* grad) {
    double loss = 0.0;
    const double n_rgb = noise ? 0.025 / 256.0 : 0.0, n_lam = noise ? 0.025 / 420.0 : 0.0;
    const double bnoise = noise ? 0.001 : 0.0;
    for (uint32_t j = 0; j < B; ++j) {
        const float* xi = x + 4 * (size_t)j;
        const int noisy_in = j < B / 2;
        double rgb[3], lamn = ((double)xi[3] - 360.0) / 420.0;
        for (int c = 0; c < 3; ++c) rgb[c] = (double)xi[c] + (noisy_in ? n_rgb * orc_smlp_normal(seed, step, j, c) : 0.0);
        if (noisy_in) lamn += n_lam * orc_smlp_normal(seed, step, j, 3);
        enc_state R, L;
        /* RGB encoder */
        const double s = fmax(fmax(rgb[0], rgb[1]), rgb[2]);
        const double inv = 1.0 / (s + 0.001);
        for (int c = 0; c < 4; ++c)
            for (int k = 0; k < 4; ++k) {
                const double ph = (c < 3 ? rgb[c] * inv : s) * PI_D * (double)(k + 1);
                R.e[8 * c + k] = sin(ph);
                R.e[8 * c + 4 + k] = cos(ph);
            }
        lin_d(w + O_RIN_W, w + O_RIN_B, R.e, R.zin, W, W);
        for (int u = 0; u < W; ++u) R.h[0][u] = fmax(R.zin[u], 0.0);
        blocks_fwd_d(w, O_RBLK(0), &R, bnoise, seed, step, j, 4);
        lin_d(w + O_ROUT, 0, R.h[NBLK], R.k, W, KOUT);
        /* wavelength encoder */
        for (int k = 0; k < 16; ++k) {
            const double ph = lamn * PI_D * (double)(k + 1);
            L.h[0][k] = sin(ph);
            L.h[0][16 + k] = cos(ph);
        }
        blocks_fwd_d(w, O_LBLK(0), &L, bnoise, seed, step, j, 4 + NBLK * W);
        lin_d(w + O_LOUT, 0, L.h[NBLK], L.k, W, KOUT);
        double res = 0.0;
        for (int k = 0; k < KOUT; ++k) res += R.k[k] * s * L.k[k];
        const double pred = fmin(fmax(res, 0.0), 1.0);
        const double imp2 = orc_smlp_importance2((double)xi[3]);
        const double d = pred - (double)y[j];
        loss += d * d * imp2;
        const double gres = (res > 0.0 && res < 1.0) ? 2.0 * d * imp2 / (3.0 * (double)B) : 0.0;
        if (gres == 0.0) continue;
        /* output layers */
        double dkr[KOUT], dkl[KOUT], dh[W];
        for (int k = 0; k < KOUT; ++k) {
            dkr[k] = gres * L.k[k] * s;
            dkl[k] = gres * R.k[k] * s;
        }
        /* RGB */
        for (int i = 0; i < W; ++i) {
            dh[i] = 0.0;
            for (int k = 0; k < KOUT; ++k) {
                grad[O_ROUT + k * W + i] += dkr[k] * R.h[NBLK][i];
                dh[i] += (double)w[O_ROUT + k * W + i] * dkr[k];
            }
        }
        blocks_bwd_d(w, O_RBLK(0), &R, dh, grad, 1);
        for (int u = 0; u < W; ++u) {
            const double dz = R.zin[u] > 0.0 ? dh[u] : 0.0;
            for (int i = 0; i < W; ++i) grad[O_RIN_W + u * W + i] += dz * R.e[i];
            grad[O_RIN_B + u] += dz;
        }
        /* wavelength */
        for (int i = 0; i < W; ++i) {
            dh[i] = 0.0;
            for (int k = 0; k < KOUT; ++k) {
                grad[O_LOUT + k * W + i] += dkl[k] * L.h[NBLK][i];
                dh[i] += (double)w[O_LOUT + k * W + i] * dkl[k];
            }
        }
        blocks_bwd_d(w, O_LBLK(0), &L, dh, grad, 0);
    }
    return loss / (3.0 * (double)B);
}

/* torch.nn.utils.clip_grad_norm_: scales grad when its norm exceeds max_norm;
 * returns the norm before clipping */
double orc_smlp_clip(double* grad, double max_norm) {
    double ss = 0.0;
    for (int p = 0; p < N_PARAMS; ++p) ss += grad[p] * grad[p];
    const double norm = sqrt(ss);
    const double coef = max_norm / (norm + 1e-6);
    if (coef < 1.0)
        for (int p = 0; p < N_PARAMS; ++p) grad[p] *= coef;
    return norm;
}

/* torch.optim.AdamW step t (1-based), decoupled decay on the weight matrices;
 * then ema = d ema + (1 - d) p when ema is given */
void orc_smlp_adamw(double* p, double* m, double* v, const double* g, int t, double lr, double b1, double b2,
                    double eps, double wd, double* ema, double ema_decay) {
    const double bc1 = 1.0 - pow(b1, t), bc2 = 1.0 - pow(b2, t);
    for (int i = 0; i < N_PARAMS; ++i) {
        if (orc_smlp_decays(i)) p[i] *= 1.0 - lr * wd;
        m[i] += (g[i] - m[i]) * (1.0 - b1);
        v[i] = v[i] * b2 + (1.0 - b2) * g[i] * g[i];
        const double denom = sqrt(v[i]) / sqrt(bc2) + eps;
        p[i] -= lr / bc1 * m[i] / denom;
        if (ema) ema[i] = ema[i] * ema_decay + (1.0 - ema_decay) * p[i];
    }
}

/* learning-rate factor of build_warmup_cosine_scheduler at a step */
double orc_smlp_lr_factor(int step, int total, double warmup_ratio, double min_ratio) {
    int warm = (int)(total * warmup_ratio);
    if (warm < 1) warm = 1;
    if (step < warm) return (double)step / (double)warm;
    const double prog = (double)(step - warm) / (double)(total - warm > 1 ? total - warm : 1);
    return min_ratio + (1.0 - min_ratio) * 0.5 * (1.0 + cos(PI_D * prog));
}
