// RGB -> reflectance-spectrum uplift table (Jakob & Hanika, "A Low-Dimensional
// Function Space for Efficient Spectral Upsampling", EG 2019), built on the
// host once per process.
//
// For every largest channel l and grid point (x, y, z) (z = that channel's
// value on the smoothstep^2 scale, x and y = the other two channels / z), the
// three coefficients c of s(lambda) = sigmoid(c0 t^2 + c1 t + c2), t = (lambda -
// 360) / 420, are found by damped Gauss-Newton so that the linear-sRGB colour
// of s under D65 (CIE 1931 2-degree, 1 nm sums, the XYZ->sRGB matrix of
// core/tonemap.cuh:116-126) matches the target.  Solves march along z from
// z ~ 0.2 outward, each warm-started from its neighbour.
#include "ar_internal.hpp"
#include "spectral.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>

namespace ar {
namespace spec {
namespace {

struct Weights {
    double w[3][kTableCount];   // linear sRGB weight of each 1 nm sample under D65
};

void make_weights(const float* xyz, const float* d65, Weights& W) {
    double norm = 0.0;
    for (int i = 0; i < kTableCount; ++i) norm += (double)d65[i] * xyz[3 * i + 1];
    const double M[3][3] = {{3.2406, -1.5372, -0.4986}, {-0.9689, 1.8758, 0.0415}, {0.0557, -0.204, 1.057}};
    for (int i = 0; i < kTableCount; ++i) {
        const double e = (double)d65[i] / norm;
        const double X = xyz[3 * i] * e, Y = xyz[3 * i + 1] * e, Z = xyz[3 * i + 2] * e;
        for (int c = 0; c < 3; ++c) W.w[c][i] = M[c][0] * X + M[c][1] * Y + M[c][2] * Z;
    }
}

double sig(double q) { return 0.5 + q / (2.0 * std::sqrt(1.0 + q * q)); }
double dsig(double q) { return 0.5 / std::pow(1.0 + q * q, 1.5); }

// rgb(c) and its Jacobian
void eval(const Weights& W, const double c[3], double rgb[3], double J[3][3]) {
    for (int k = 0; k < 3; ++k) { rgb[k] = 0.0; J[k][0] = J[k][1] = J[k][2] = 0.0; }
    for (int i = 0; i < kTableCount; ++i) {
        const double t = (double)i / 420.0;
        const double q = (c[0] * t + c[1]) * t + c[2];
        const double s = sig(q), ds = dsig(q);
        const double g[3] = {ds * t * t, ds * t, ds};
        for (int k = 0; k < 3; ++k) {
            rgb[k] += W.w[k][i] * s;
            for (int m = 0; m < 3; ++m) J[k][m] += W.w[k][i] * g[m];
        }
    }
}

bool solve3(double A[3][3], const double b[3], double x[3]) {
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                       A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                       A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    if (!(std::fabs(det) > 1e-300)) return false;
    for (int k = 0; k < 3; ++k) {
        double B[3][3];
        std::memcpy(B, A, sizeof B);
        for (int r = 0; r < 3; ++r) B[r][k] = b[r];
        x[k] = (B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) - B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0])) / det;
    }
    return true;
}

void gauss_newton(const Weights& W, const double target[3], double c[3]) {
    double rgb[3], J[3][3];
    eval(W, c, rgb, J);
    double err = 0.0;
    for (int k = 0; k < 3; ++k) err += (rgb[k] - target[k]) * (rgb[k] - target[k]);
    for (int it = 0; it < 40 && err > 1e-14; ++it) {
        const double r[3] = {target[0] - rgb[0], target[1] - rgb[1], target[2] - rgb[2]};
        double d[3];
        if (!solve3(J, r, d)) break;
        double step = 1.0;
        bool improved = false;
        for (int ls = 0; ls < 12; ++ls) {
            const double cn[3] = {c[0] + step * d[0], c[1] + step * d[1], c[2] + step * d[2]};
            double rn[3], Jn[3][3];
            eval(W, cn, rn, Jn);
            double en = 0.0;
            for (int k = 0; k < 3; ++k) en += (rn[k] - target[k]) * (rn[k] - target[k]);
            if (en < err) {
                std::memcpy(c, cn, sizeof cn);
                std::memcpy(rgb, rn, sizeof rn);
                std::memcpy(J, Jn, sizeof Jn);
                err = en;
                improved = true;
                break;
            }
            step *= 0.5;
        }
        if (!improved) break;
    }
}

std::mutex g_mu;
std::vector<float> g_lut;   // [l][k][j][i][3]

} // namespace

size_t lut_floats() { return (size_t)3 * kLutRes * kLutRes * kLutRes * 3; }

const std::vector<float>& uplift_lut(const float* xyz, const float* d65) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_lut.empty()) return g_lut;
    Weights W;
    make_weights(xyz, d65, W);
    std::vector<float> lut(lut_floats(), 0.0f);
    const int R = kLutRes;
    const int k0 = R / 5;
    auto work = [&](int task_begin, int task_end) {
        for (int task = task_begin; task < task_end; ++task) {
            const int l = task / (R * R), ji = task % (R * R), j = ji / R, i = ji % R;
            const double x = (double)i / (R - 1), y = (double)j / (R - 1);
            auto solve_at = [&](int k, double c[3]) {
                const double z = lut_scale(k);
                double t[3];
                t[l] = z;
                t[(l + 1) % 3] = x * z;
                t[(l + 2) % 3] = y * z;
                float* dst = &lut[((((size_t)l * R + k) * R + j) * R + i) * 3];
                if (z <= 0.0) {
                    dst[0] = 0.0f; dst[1] = 0.0f; dst[2] = -1e30f;
                    return;
                }
                gauss_newton(W, t, c);
                dst[0] = (float)c[0]; dst[1] = (float)c[1]; dst[2] = (float)c[2];
            };
            double c[3] = {0.0, 0.0, 0.0};
            solve_at(k0, c);
            double cs[3] = {c[0], c[1], c[2]};
            for (int k = k0 + 1; k < R; ++k) solve_at(k, c);
            std::memcpy(c, cs, sizeof cs);
            for (int k = k0 - 1; k >= 0; --k) solve_at(k, c);
        }
    };
    const int tasks = 3 * R * R;
    const unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back(work, (int)((size_t)tasks * t / nt), (int)((size_t)tasks * (t + 1) / nt));
    for (auto& t : th) t.join();
    g_lut.swap(lut);
    return g_lut;
}

// Coefficients for an RGB colour with channels in [0, 1] (trilinear in the
// coefficient table of the largest channel).
void uplift(const float* lut, const float rgb_in[3], float c[3]) {
    float rgb[3] = {spec::clampf(rgb_in[0], 0.0f, 1.0f), spec::clampf(rgb_in[1], 0.0f, 1.0f),
                    spec::clampf(rgb_in[2], 0.0f, 1.0f)};
    int l = 0;
    if (rgb[1] > rgb[l]) l = 1;
    if (rgb[2] > rgb[l]) l = 2;
    const float z = rgb[l];
    if (!(z > 0.0f)) { c[0] = 0.0f; c[1] = 0.0f; c[2] = -1e30f; return; }
    const float x = rgb[(l + 1) % 3] / z, y = rgb[(l + 2) % 3] / z;
    const int R = kLutRes;
    int k = 0;
    while (k < R - 2 && lut_scale(k + 1) <= z) ++k;
    const float z0 = lut_scale(k), z1 = lut_scale(k + 1);
    const float tz = spec::clampf((z - z0) / (z1 - z0), 0.0f, 1.0f);
    const float fx = x * (R - 1), fy = y * (R - 1);
    const int i = std::min((int)fx, R - 2), j = std::min((int)fy, R - 2);
    const float tx = fx - (float)i, ty = fy - (float)j;
    for (int m = 0; m < 3; ++m) {
        float acc = 0.0f;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const float w = (dx ? tx : 1.0f - tx) * (dy ? ty : 1.0f - ty) * (dz ? tz : 1.0f - tz);
                    const float v = lut[((((size_t)l * R + (k + dz)) * R + (j + dy)) * R + (i + dx)) * 3 + m];
                    acc += w * v;
                }
        c[m] = acc;
    }
}

} // namespace spec
} // namespace ar

namespace ar {
namespace spec {
#include "cie_tables.inc"
const float* cie_xyz_table() { return kCieXyz1nm; }
const float* d65_table() { return kD65_1nm; }
} // namespace spec
} // namespace ar

extern "C" int ar_spectrum_lut_build(float* out, size_t count) {
    if (!out || count < ar::spec::lut_floats()) return AR_E_INVALID;
    const std::vector<float>& lut = ar::spec::uplift_lut(ar::spec::cie_xyz_table(), ar::spec::d65_table());
    std::memcpy(out, lut.data(), lut.size() * sizeof(float));
    return AR_OK;
}

extern "C" int ar_spectrum_uplift(const float* rgb, size_t n, float* coeffs) {
    if (n && (!rgb || !coeffs)) return AR_E_INVALID;
    const std::vector<float>& lut = ar::spec::uplift_lut(ar::spec::cie_xyz_table(), ar::spec::d65_table());
    for (size_t i = 0; i < n; ++i) ar::spec::uplift(lut.data(), rgb + 3 * i, coeffs + 3 * i);
    return AR_OK;
}
