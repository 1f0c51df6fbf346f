/*
 * Oracle of the learned guiding distribution.  TEST INFRASTRUCTURE ONLY.
 *
 * The network: 5 inputs (position normalised to the scene box, 3; the
 * shading normal as theta/pi and phi/(2 pi) + 1/2, 2), three hidden layers of
 * 64 ReLU units, 40 outputs = 8 von Mises-Fisher lobes x (logit, log kappa,
 * mean direction 3).  Parameters are stored as W[in][out] with the input
 * padded to 16 rows and the output to 48 columns (zero padding), biases per
 * layer: W0 16x64, b0 64, W1 64x64, b1 64, W2 64x64, b2 64, W3 64x48, b3 48.
 *
 * Distribution: q(w) = sum_k pi_k vMF(w; mu_k, kappa_k) with
 * pi = softmax(logits), kappa_k = exp(clamp(s_k, -3, 7)), mu_k = c_k/|c_k|,
 * vMF(w; mu, kappa) = kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1)).
 * Loss of a sample (x, w, weight): -weight * log q(w | x), the Monte-Carlo
 * form of the KL divergence from the sampled radiance distribution to q
 * (Mueller et al., "Neural Importance Sampling", 2019).
 *
 * Everything here is FP64.  orc_mlp_*: no rounding, the mathematical network;
 * orc_mlp_*_bf16: the same with values rounded to BF16 exactly where the GPU
 * kernels store BF16, so that the GPU's MFMA data flow can be checked to FP32
 * accumulation accuracy.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

#define NI 16
#define NH 64
#define NO 48
#define NLOBE 8

static const double kPi = 3.14159265358979323846;

/* BF16 emulation: when set, values are rounded where the GPU kernels store
 * BF16 (inputs, weights, post-ReLU activations, the loss gradient and the
 * masked backward activations), through FP32 with round-to-nearest-even. */
static int g_bf16 = 0;

static double bf(double x)
{
    if (!g_bf16) return x;
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

/* parameter offsets (floats) */
enum {
    OW0 = 0,
    OB0 = OW0 + NI * NH,
    OW1 = OB0 + NH,
    OB1 = OW1 + NH * NH,
    OW2 = OB1 + NH,
    OB2 = OW2 + NH * NH,
    OW3 = OB2 + NH,
    OB3 = OW3 + NH * NO,
    NPARAM = OB3 + NO
};

int orc_mlp_param_count(void) { return NPARAM; }

static void dense(const double* W, const double* b, const double* x, int nin, int nout, double* y, int relu)
{
    for (int o = 0; o < nout; ++o) {
        double s = 0;
        for (int i = 0; i < nin; ++i) s += x[i] * bf(W[i * nout + o]);
        s += b[o];
        y[o] = relu ? bf(s > 0 ? s : 0) : s;
    }
}

static void forward(const double* P, const double* x16, double* h1, double* h2, double* h3, double* o)
{
    dense(P + OW0, P + OB0, x16, NI, NH, h1, 1);
    dense(P + OW1, P + OB1, h1, NH, NH, h2, 1);
    dense(P + OW2, P + OB2, h2, NH, NH, h3, 1);
    dense(P + OW3, P + OB3, h3, NH, NO, o, 0);
}

void orc_mlp_forward(const double* P, const float* x5, double* out40)
{
    double x[NI] = { 0 }, h1[NH], h2[NH], h3[NH], o[NO];
    for (int i = 0; i < 5; ++i) x[i] = bf(x5[i]);
    forward(P, x, h1, h2, h3, o);
    for (int i = 0; i < 40; ++i) out40[i] = o[i];
}

typedef struct { double pi[NLOBE], kappa[NLOBE], mu[NLOBE][3], clen[NLOBE]; int kclamped[NLOBE]; } lobes;

static void decode(const double* o, lobes* L)
{
    double mx = -1e300;
    for (int k = 0; k < NLOBE; ++k) mx = o[5 * k] > mx ? o[5 * k] : mx;
    double z = 0;
    for (int k = 0; k < NLOBE; ++k) { L->pi[k] = exp(o[5 * k] - mx); z += L->pi[k]; }
    for (int k = 0; k < NLOBE; ++k) {
        L->pi[k] /= z;
        double s = o[5 * k + 1];
        L->kclamped[k] = s < -3.0 || s > 7.0;
        s = s < -3.0 ? -3.0 : (s > 7.0 ? 7.0 : s);
        L->kappa[k] = exp(s);
        const double cx = o[5 * k + 2], cy = o[5 * k + 3], cz = o[5 * k + 4];
        double l = sqrt(cx * cx + cy * cy + cz * cz);
        if (l < 1e-12) { L->mu[k][0] = 0; L->mu[k][1] = 0; L->mu[k][2] = 1; L->clen[k] = 0; }
        else { L->mu[k][0] = cx / l; L->mu[k][1] = cy / l; L->mu[k][2] = cz / l; L->clen[k] = l; }
    }
}

static double vmf(double kappa, double cosang)
{
    return kappa / (2.0 * kPi * (1.0 - exp(-2.0 * kappa))) * exp(kappa * (cosang - 1.0));
}

double orc_vmf_mixture_pdf(const double* out40, const double* w)
{
    lobes L;
    decode(out40, &L);
    double q = 0;
    for (int k = 0; k < NLOBE; ++k)
        q += L.pi[k] * vmf(L.kappa[k], L.mu[k][0] * w[0] + L.mu[k][1] * w[1] + L.mu[k][2] * w[2]);
    return q;
}

/* sample: u[0] lobe, u[1..2] direction (Jakob, "Numerically stable sampling
 * of the von Mises-Fisher distribution on S^2", 2012) */
void orc_vmf_mixture_sample(const double* out40, const double* u, double* w)
{
    lobes L;
    decode(out40, &L);
    int k = 0;
    double c = 0;
    for (k = 0; k < NLOBE - 1; ++k) {
        c += L.pi[k];
        if (u[0] < c) break;
    }
    const double kap = L.kappa[k];
    const double ct = 1.0 + log(u[1] + (1.0 - u[1]) * exp(-2.0 * kap)) / kap;
    const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    const double ph = 2.0 * kPi * u[2];
    const double* m = L.mu[k];
    /* orthonormal basis around m (Duff et al. 2017) */
    const double sg = m[2] >= 0 ? 1.0 : -1.0;
    const double a = -1.0 / (sg + m[2]), b = m[0] * m[1] * a;
    const double t[3] = { 1.0 + sg * m[0] * m[0] * a, sg * b, -sg * m[0] };
    const double bt[3] = { b, sg + m[1] * m[1] * a, -m[1] };
    for (int i = 0; i < 3; ++i) w[i] = st * cos(ph) * t[i] + st * sin(ph) * bt[i] + ct * m[i];
}

/* loss of a batch and its gradient (grad may be NULL); returns the mean of
 * -weight * log q over the batch.  x: n x 5, w: n x 3 unit directions. */
double orc_mlp_loss_grad(const double* P, int n, const float* x5, const float* dir, const float* weight, double* grad)
{
    if (grad) memset(grad, 0, sizeof(double) * NPARAM);
    double loss = 0;
    for (int s = 0; s < n; ++s) {
        double x[NI] = { 0 }, h1[NH], h2[NH], h3[NH], o[NO];
        for (int i = 0; i < 5; ++i) x[i] = bf(x5[5 * s + i]);
        forward(P, x, h1, h2, h3, o);
        lobes L;
        decode(o, &L);
        const double w[3] = { dir[3 * s], dir[3 * s + 1], dir[3 * s + 2] };
        double f[NLOBE], q = 0;
        for (int k = 0; k < NLOBE; ++k) {
            f[k] = vmf(L.kappa[k], L.mu[k][0] * w[0] + L.mu[k][1] * w[1] + L.mu[k][2] * w[2]);
            q += L.pi[k] * f[k];
        }
        const double wt = weight[s] / n;
        loss += -wt * log(fmax(q, 1e-300));
        if (!grad) continue;
        double d_o[NO] = { 0 };
        for (int k = 0; k < NLOBE; ++k) {
            const double r = L.pi[k] * f[k] / q;
            const double kap = L.kappa[k];
            const double mw = L.mu[k][0] * w[0] + L.mu[k][1] * w[1] + L.mu[k][2] * w[2];
            d_o[5 * k] = wt * (L.pi[k] - r);
            /* d log f / d kappa = 1/kappa - 2 e^{-2k}/(1 - e^{-2k}) + (mu.w - 1); kappa = exp(s) */
            const double e2 = exp(-2.0 * kap);
            const double dlk = 1.0 / kap - 2.0 * e2 / (1.0 - e2) + (mw - 1.0);
            d_o[5 * k + 1] = L.kclamped[k] ? 0.0 : -wt * r * dlk * kap;
            if (L.clen[k] > 0)
                for (int i = 0; i < 3; ++i)
                    d_o[5 * k + 2 + i] = -wt * r * kap * (w[i] - mw * L.mu[k][i]) / L.clen[k];
        }
        /* backward */
        double d3[NH], d2[NH], d1[NH], dob[NO];
        for (int o2 = 0; o2 < NO; ++o2) dob[o2] = bf(d_o[o2]);
        for (int i = 0; i < NH; ++i) {
            double s3 = 0;
            for (int o2 = 0; o2 < NO; ++o2) {
                s3 += bf(P[OW3 + i * NO + o2]) * dob[o2];
                grad[OW3 + i * NO + o2] += h3[i] * dob[o2];
            }
            d3[i] = bf(h3[i] > 0 ? s3 : 0);
        }
        for (int o2 = 0; o2 < NO; ++o2) grad[OB3 + o2] += dob[o2];
        for (int i = 0; i < NH; ++i) {
            double s2 = 0;
            for (int o2 = 0; o2 < NH; ++o2) {
                s2 += bf(P[OW2 + i * NH + o2]) * d3[o2];
                grad[OW2 + i * NH + o2] += h2[i] * d3[o2];
            }
            d2[i] = bf(h2[i] > 0 ? s2 : 0);
        }
        for (int o2 = 0; o2 < NH; ++o2) grad[OB2 + o2] += d3[o2];
        for (int i = 0; i < NH; ++i) {
            double s1 = 0;
            for (int o2 = 0; o2 < NH; ++o2) {
                s1 += bf(P[OW1 + i * NH + o2]) * d2[o2];
                grad[OW1 + i * NH + o2] += h1[i] * d2[o2];
            }
            d1[i] = bf(h1[i] > 0 ? s1 : 0);
        }
        for (int o2 = 0; o2 < NH; ++o2) grad[OB1 + o2] += d2[o2];
        for (int i = 0; i < NI; ++i)
            for (int o2 = 0; o2 < NH; ++o2) grad[OW0 + i * NH + o2] += x[i] * d1[o2];
        for (int o2 = 0; o2 < NH; ++o2) grad[OB0 + o2] += d1[o2];
    }
    return loss;
}

/* the same with the GPU kernels' BF16 roundings emulated */
void orc_mlp_forward_bf16(const double* P, const float* x5, double* out40)
{
    g_bf16 = 1;
    orc_mlp_forward(P, x5, out40);
    g_bf16 = 0;
}

double orc_mlp_loss_grad_bf16(const double* P, int n, const float* x5, const float* dir, const float* weight,
                              double* grad)
{
    g_bf16 = 1;
    const double l = orc_mlp_loss_grad(P, n, x5, dir, weight, grad);
    g_bf16 = 0;
    return l;
}
