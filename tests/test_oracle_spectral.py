"""Spectral pieces of the oracle, pinned by textbook facts and closed forms.

* CIE 1931 2-degree / D65 tables: the Y normaliser sum(ybar) = 106.857, the
  D65 white point (X/Y, Z/Y) = (0.95047, 1.08883), and the D65 * ybar
  integral 10567.065 that core/spectrum/ciexyz.cuh normalises D65 by.
* The normal quantile against scipy.special.ndtri.
* Wavelength sampling (core/spectrum/sample_ciexyz.cuh:159-326): the mixture
  density integrates to one and a histogram of the sampler follows it (a
  wrong mixture weight, mean or width fails here).
* The sigmoid-polynomial uplift table: reflectance of uplift(rgb) under D65,
  projected on the CMFs and mapped by XYZ->linear sRGB, returns rgb.
* Spectral furnace: a gray Lambert floor under a white constant sky renders
  the closed-form integral of the uplifted spectra (single-wavelength and
  four-hero-wavelength estimators, XYZ and sRGB outputs).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as orc

XYZ2SRGB = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.204, 1.057]])
LAM = np.arange(360.0, 781.0)


def _sig(c, lam):
    x = (lam - 360.0) / 420.0
    q = (c[..., 0:1] * x + c[..., 1:2]) * x + c[..., 2:3]
    return 0.5 + q / (2.0 * np.sqrt(1.0 + q * q))


def test_cie_tables_textbook_values():
    xyz, d65 = orc.cie_tables()
    xyz = xyz.astype(np.float64)
    d65 = d65.astype(np.float64)
    assert xyz.shape == (421, 3) and d65.shape == (421,)
    assert abs(xyz[:, 1].sum() - 106.857) < 0.01
    assert abs(xyz[LAM == 555.0, 1][0] - 1.0) < 1e-4            # ybar peaks at 555 nm
    assert abs(d65[LAM == 560.0][0] - 100.0) < 1e-6             # D65 is normalised to 100 at 560 nm
    w = d65 @ xyz
    assert abs(w[1] - 10567.065) < 0.05
    assert abs(w[0] / w[1] - 0.95047) < 2e-4 and abs(w[2] / w[1] - 1.08883) < 2e-4
    assert np.allclose(XYZ2SRGB @ (w / w[1]), 1.0, atol=2e-3)   # D65 white maps to sRGB white


def test_ndtri_matches_scipy():
    from scipy.special import ndtri
    f = orc.fn("orc_ndtri_f", C.c_float, [C.c_float])
    u = np.concatenate([np.logspace(-7, -1, 200), np.linspace(0.01, 0.99, 400), 1.0 - np.logspace(-1, -6, 200)])
    got = np.array([f(float(x)) for x in u.astype(np.float32)])
    want = ndtri(u.astype(np.float32).astype(np.float64))
    assert np.all(np.abs(got - want) <= 2e-5 * np.abs(want) + 2e-6), np.max(np.abs(got - want))


def test_lambda_pdf_integrates_to_one():
    pdf = orc.fn("orc_lambda_pdf_f", C.c_float, [C.c_float])
    lam = np.linspace(360.0, 780.0, 8401)
    p = np.array([pdf(float(x)) for x in lam])
    assert p.min() > 0
    assert abs(np.trapezoid(p, lam) - 1.0) < 2e-3


def test_lambda_sampler_follows_pdf():
    samp = orc.fn("orc_sample_lambda_f", C.c_float, [C.c_float, C.c_float, C.c_float])
    pdf = orc.fn("orc_lambda_pdf_f", C.c_float, [C.c_float])
    rng = np.random.default_rng(7)
    u = rng.random((200_000, 3), dtype=np.float32)
    lam = np.array([samp(float(a), float(b), float(c)) for a, b, c in u])
    assert lam.min() >= 360.0 and lam.max() <= 780.0
    edges = np.linspace(360.0, 780.0, 43)
    hist, _ = np.histogram(lam, edges)
    fine = np.linspace(360.0, 780.0, 4201)
    pf = np.array([pdf(float(x)) for x in fine])
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (pf[1:] + pf[:-1]) * np.diff(fine))])
    expect = np.diff(np.interp(edges, fine, cdf)) / cdf[-1] * len(lam)
    chi2 = np.sum((hist - expect) ** 2 / expect)
    assert chi2 < 90.0, chi2      # 41 dof: p ~ 1e-5


def test_sigmoid_limits():
    s = orc.fn("orc_sigmoid_f", C.c_float, [C.c_void_p, C.c_float])
    for c, want in (((0, 0, 0), 0.5), ((0, 0, 1e30), 1.0), ((0, 0, -1e30), 0.0), ((0, 0, 1.0), 0.5 + 0.5 / np.sqrt(2))):
        c = np.array(c, np.float32)
        assert abs(s(c.ctypes.data, 500.0) - want) < 1e-6


def _reflect_to_srgb(coef, lam=LAM):
    xyz, d65 = orc.cie_tables()
    xyz = xyz.astype(np.float64)
    d65 = d65.astype(np.float64)
    refl = _sig(coef.astype(np.float64), lam)
    XYZ = (refl * d65) @ xyz / (d65 @ xyz[:, 1])
    return XYZ @ XYZ2SRGB.T


def test_uplift_round_trip():
    rng = np.random.default_rng(3)
    special = np.array([[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.9, 0.1, 0.3],
                        [0.02, 0.02, 0.02], [0.2, 0.8, 0.6]], np.float32)
    rgb = np.concatenate([special, rng.random((1500, 3), dtype=np.float32)])
    coef = orc.uplift(rgb)
    assert np.isfinite(coef).all()
    err = np.abs(_reflect_to_srgb(coef) - rgb).max(axis=1)
    assert err.max() < 3e-3 and err.mean() < 5e-4, (err.max(), err.mean())
    # through the oracle's own sigmoid as well
    s = orc.fn("orc_sigmoid_f", C.c_float, [C.c_void_p, C.c_float])
    c0 = np.ascontiguousarray(coef[6])
    refl = np.array([s(c0.ctypes.data, float(l)) for l in LAM])
    assert np.allclose(refl, _sig(c0.astype(np.float64)[None], LAM)[0], atol=2e-6)


def test_library_uplift_matches_oracle():
    """ar_spectrum_uplift (host code in the library, no device) against the
    oracle's trilinear lookup of the same table."""
    from aperturerendererv3_amd import _lib as L
    lib = L.load()
    rng = np.random.default_rng(5)
    rgb = np.ascontiguousarray(rng.random((512, 3), dtype=np.float32) * np.float32(1.0))
    rgb[:4] = [[0, 0, 0], [1, 1, 1], [1, 0, 0], [0.25, 0.5, 0.75]]
    out = np.zeros_like(rgb)
    assert lib.ar_spectrum_uplift(rgb.ctypes.data, rgb.shape[0], out.ctypes.data) == 0
    ref = orc.uplift(rgb)
    assert np.allclose(out, ref, rtol=1e-5, atol=1e-4), np.abs(out - ref).max()


def _furnace_expect(albedo, mode, output):
    """Closed form of the spectral furnace: every path hits the floor once and
    escapes into the constant sky, so a sample's value is
    refl(l) * sky(l) * D65n(l), weighted by cmf(l)/pdf(l)."""
    lut_c = orc.uplift(np.array([[albedo] * 3, [1.0, 1.0, 1.0]], np.float32))
    lam = np.linspace(360.0, 780.0, 42001) if mode == 1 else np.linspace(380.0, 780.0, 40001)
    xyz, d65 = orc.cie_tables()
    cmf = np.stack([np.interp(lam, LAM, xyz[:, k].astype(np.float64)) for k in range(3)], 1)
    d = np.interp(lam, LAM, d65.astype(np.float64)) / np.float64(np.float32(10567.065))
    v = _sig(lut_c[0:1].astype(np.float64), lam)[0] * _sig(lut_c[1:2].astype(np.float64), lam)[0] * d
    XYZ = np.trapezoid(v[:, None] * cmf, lam, axis=0)
    return XYZ if output == 0 else np.maximum(XYZ2SRGB @ XYZ, 0)


@pytest.mark.parametrize("mode,output", [(1, 0), (1, 1), (4, 0), (4, 1)])
def test_spectral_furnace_closed_form(mode, output):
    import geom
    V, F, M, mats, _ = geom.floor_and_light(albedo=0.6)
    F, M, mats = F[:2], M[:2], mats[:1]
    img = orc.render(V, F, M, mats, 32, 32, 64, max_depth=4, nee=0, sky=2, env=(1.0, 1.0, 1.0),
                     cam_pos=(0, 0.5, 0.01), cam_dir=(0, -1, 1e-3), vfov=20.0, rr_min_depth=100, precision=64,
                     spectral=mode, output=output)
    got = img.reshape(-1, 3)
    mean = got.mean(axis=0)
    sem = got.std(axis=0) / np.sqrt(got.shape[0])
    expect = _furnace_expect(0.6, mode, output)
    assert np.all(np.abs(mean - expect) < 4 * sem + 1e-4), (mean, expect, sem)
    if output == 1:
        assert np.allclose(expect, 0.6, atol=6e-3)


def test_spectral_hero_has_less_noise_than_single():
    import geom
    V, F, M, mats, _ = geom.floor_and_light(albedo=0.6)
    F, M, mats = F[:2], M[:2], mats[:1]
    kw = dict(max_depth=4, nee=0, sky=2, env=(1.0, 1.0, 1.0), cam_pos=(0, 0.5, 0.01), cam_dir=(0, -1, 1e-3),
              vfov=20.0, rr_min_depth=100, output=1)
    s = orc.render(V, F, M, mats, 16, 16, 16, spectral=1, **kw)
    h = orc.render(V, F, M, mats, 16, 16, 16, spectral=4, **kw)
    assert h.std() < 0.5 * s.std()


def test_spectral_fp32_and_fp64_trace_the_same_paths():
    from aperturerendererv3_amd.workloads import CORNELL_CAMERA, cornell_box
    V, F, M, mats = cornell_box()
    cam = dict(cam_pos=CORNELL_CAMERA["pos"], cam_dir=CORNELL_CAMERA["dir"], vfov=CORNELL_CAMERA["vfov"])
    for mode in (1, 4):
        a = orc.render(V, F, M, mats, 24, 24, 4, seed=9, precision=32, spectral=mode, output=1, **cam)
        b = orc.render(V, F, M, mats, 24, 24, 4, seed=9, precision=64, spectral=mode, output=1, **cam)
        assert orc.rel_mse(a, b) < 1e-6, (mode, orc.rel_mse(a, b))


def test_spectral_cornell_color_agrees_with_rgb_render():
    """Spectral and RGB transport differ only through interreflection of
    non-flat spectra; the white-lit Cornell box mean colour agrees to a few %."""
    from aperturerendererv3_amd.workloads import CORNELL_CAMERA, cornell_box
    V, F, M, mats = cornell_box()
    cam = dict(cam_pos=CORNELL_CAMERA["pos"], cam_dir=CORNELL_CAMERA["dir"], vfov=CORNELL_CAMERA["vfov"])
    rgb = orc.render(V, F, M, mats, 16, 16, 256, seed=2, **cam).reshape(-1, 3).mean(0)
    spec = orc.render(V, F, M, mats, 16, 16, 256, seed=2, spectral=4, output=1, **cam).reshape(-1, 3).mean(0)
    assert np.all(np.abs(spec - rgb) / rgb < 0.08), (rgb, spec)
