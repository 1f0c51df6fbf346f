"""Oracle denoiser (core/denoise) pinned against numpy, scipy and the
reference's own quantile table.

  - the t quantile table equals the reference's table (golden fixture, bit
    for bit) and scipy's t.ppf(1 - 0.005 / 2, nu);
  - Welford statistics equal numpy's mean and central moments of the Box-Cox
    transformed samples (a dropped n0 factor or a wrong sign in the M3
    update fails);
  - the G-buffer means equal numpy means over the hit samples and the model
    histogram the frequencies over all samples;
  - with equal variances and counts and no skew, the gate is scipy's Welch
    t-test at 0.5 % (two-sided): the asymmetric statistic reduces to
    |mean_a - mean_b| / sqrt(s_a^2 / n_a + s_b^2 / n_b) there;
  - the filter keeps a constant image, keeps a two-level image whose levels
    differ in their statistics (the gate separates them), and averages a
    noisy region whose statistics agree."""
import math
from pathlib import Path

import numpy as np
import pytest
from scipy import stats as sps

import oracle_lib as orc

GOLDEN = Path(__file__).resolve().parent / "golden" / "t005_quantiles_reference.npy"


def _u(v):
    return np.array([v], np.uint32).view(np.float32)[0]


def test_t_quantile_table():
    ref = np.load(GOLDEN)
    ours = np.array([orc.t_quantile(nu) for nu in range(1, 1025)], np.float32)
    np.testing.assert_array_equal(ours, ref)
    np.testing.assert_allclose(ours, sps.t.ppf(1 - 0.0025, np.arange(1, 1025)), rtol=1e-6)
    assert orc.t_quantile(0) == ref[0] and orc.t_quantile(-5) == ref[0] and orc.t_quantile(5000) == ref[-1]


def test_welford_matches_numpy_moments():
    rng = np.random.default_rng(1)
    for n in (1, 2, 7, 300):
        xs = (rng.gamma(0.7, 2.0, (n, 3)) * np.array([1.0, 0.3, 3.0])).astype(np.float32)
        xs[rng.random(n) < 0.1] = 0.0
        st = np.zeros(12, np.float32)
        for x in xs:
            orc.stats_record(st, x)
        assert st[:1].view(np.uint32)[0] == n
        bc = (np.sqrt(np.maximum(xs.astype(np.float64), 0)) - 1.0) * 2.0
        mean = bc.mean(axis=0)
        m2 = ((bc - mean) ** 2).mean(axis=0)
        m3 = ((bc - mean) ** 3).mean(axis=0)
        np.testing.assert_allclose(st[1:4], mean, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(st[4:7], m2, rtol=1e-4, atol=1e-6)
        scale = np.maximum(m2, 1e-3) ** 1.5
        assert np.all(np.abs(st[7:10] - m3) <= 1e-3 * scale + 1e-6), (st[7:10], m3)


def test_gbuffer_record_means_and_histogram():
    rng = np.random.default_rng(2)
    g = np.zeros(16, np.float32)
    hits, models, alb, nrm, dep, met, rou = [], [], [], [], [], [], []
    for k in range(50):
        hit = rng.random() < 0.7
        a, nv = rng.random(3), rng.normal(size=3)
        d, m, r = rng.uniform(0.5, 3), rng.random(), rng.random()
        model = int(rng.integers(0, 4))
        orc.gbuffer_record(g, hit, a, nv, d, m, r, model)
        hits.append(hit); models.append(model if hit else 4)
        if hit:
            alb.append(a); nrm.append(nv); dep.append(d); met.append(m); rou.append(r)
    np.testing.assert_allclose(g[0:3], np.mean(alb, axis=0), rtol=1e-5)
    np.testing.assert_allclose(g[3:6], np.mean(nrm, axis=0), rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(g[6:9], [np.mean(dep), np.mean(met), np.mean(rou)], rtol=1e-5)
    np.testing.assert_allclose(g[9:14], np.bincount(models, minlength=5) / 50.0, atol=1e-5)
    assert g[14:16].view(np.uint32).tolist() == [sum(hits), 50]


def _stat(n, mean, var_pop, m3=0.0):
    st = np.zeros(12, np.float32)
    st[0] = _u(n)
    st[1:4] = mean
    st[4:7] = var_pop
    st[7:10] = m3
    return st


def test_gate_is_welch_t_test_without_skew():
    rng = np.random.default_rng(3)
    agree = checked = 0
    for _ in range(400):
        n = int(rng.integers(4, 200))
        s2 = float(rng.uniform(0.1, 2.0))            # Bessel-corrected variance of both
        ma = float(rng.normal())
        mb = ma + float(rng.normal()) * math.sqrt(2 * s2 / n) * rng.uniform(0.5, 6.0)
        var_pop = s2 * (n - 1) / n
        a, b = _stat(n, ma, var_pop), _stat(n, mb, var_pop)
        gate = orc.tstat_gate(a, b)
        sa = float(np.float32(np.float32(var_pop) * n / (n - 1)))
        t = abs(np.float32(ma) - np.float32(mb)) / math.sqrt(2 * sa / n)
        nu = 2 * (n - 1)
        q = sps.t.ppf(1 - 0.0025, nu)
        if abs(t - q) < 1e-3 * q:
            continue                                  # too close to the threshold for float rounding
        p = sps.ttest_ind_from_stats(ma, math.sqrt(sa), n, mb, math.sqrt(sa), n, equal_var=False).pvalue
        expect = 1.0 if p > 0.005 else 0.0
        checked += 1
        agree += gate == expect
    assert checked > 350 and agree == checked


def test_gate_skew_correction_and_degenerate_records():
    # Johnson's correction moves the mean by M3 / (6 s^2 n) per sample
    a = _stat(50, 0.0, 1.0, m3=0.0)
    b = _stat(50, 0.0, 1.0, m3=0.0)
    assert orc.tstat_gate(a, b) == 1.0
    far = _stat(50, 10.0, 1.0)
    assert orc.tstat_gate(a, far) == 0.0
    # empty records: NaN statistics fail the gate, as in the reference
    empty = np.zeros(12, np.float32)
    assert orc.tstat_gate(empty, empty) == 0.0


def _records(h, w, albedo, n, mean_bc, var_pop):
    g = np.zeros((h, w, 16), np.float32)
    g[..., 0:3] = albedo
    g[..., 3:6] = (0.0, 0.0, 1.0)
    g[..., 6] = 2.0
    g[..., 9] = 1.0
    g[..., 14] = _u(n)
    g[..., 15] = _u(n)
    st = np.zeros((h, w, 12), np.float32)
    st[..., 0] = _u(n)
    st[..., 1:4] = mean_bc
    st[..., 4:7] = var_pop
    return g, st


def test_bilateral_constant_and_two_level_images():
    h, w = 20, 24
    g, st = _records(h, w, 0.5, 64, 0.3, 0.2)
    hdr = np.full((h, w, 3), 0.7, np.float32)
    np.testing.assert_allclose(orc.bilateral(hdr, g, st), hdr, rtol=2e-6)
    # two levels whose statistics differ: the gate keeps them apart, every
    # pixel keeps its own level although its window crosses the edge
    hdr[:, 12:] = 3.0
    st[:, 12:, 1:4] = 5.0
    out = orc.bilateral(hdr, g, st)
    np.testing.assert_allclose(out, hdr, rtol=2e-6)
    # without the statistics difference the edge blurs
    st[:, 12:, 1:4] = 0.3
    out = orc.bilateral(hdr, g, st)
    assert out[:, 11].mean() > 0.9 and out[:, 12].mean() < 2.8


def test_bilateral_smooths_noise_with_agreeing_statistics():
    rng = np.random.default_rng(4)
    h, w = 32, 32
    g, st = _records(h, w, 0.5, 64, 0.3, 0.2)
    hdr = (1.0 + 0.3 * rng.normal(size=(h, w, 3))).astype(np.float32)
    out = orc.bilateral(hdr, g, st)
    core = (slice(8, 24), slice(8, 24))
    assert out[core].std() < 0.25 * hdr[core].std()
    # the log-domain mean of a symmetric noise field stays within a few percent
    assert abs(out[core].mean() - hdr[core].mean()) < 0.05
    # a pixel with an opposite normal and another shading model is left alone
    g2 = g.copy()
    g2[16, 16, 3:6] = (0.0, 0.0, -1.0)
    g2[16, 16, 9], g2[16, 16, 11] = 0.0, 1.0
    out2 = orc.bilateral(hdr, g2, st)
    assert abs(out2[16, 16] - hdr[16, 16]).max() < 0.02 * abs(hdr[16, 16]).max()
