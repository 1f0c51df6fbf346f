"""The oracle's learned guiding distribution (FP64), pinned by mathematics:

* each vMF lobe and the 8-lobe mixture integrate to one over the sphere
  (quadrature on Clarberg's equal-area grid);
* the mixture sampler's frequencies follow the mixture density (chi^2 over
  equal-area cells);
* the analytic gradient of the loss equals central finite differences
  (a dropped term, a wrong sign or a transposed weight fails here);
* gradient descent on the loss fits a known target distribution.
"""
import ctypes as C
import math

import numpy as np

import oracle_lib as orc
from test_oracle_guiding import uv_to_dir

NP = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 * 48 + 48


def _f(name, res, args):
    return orc.fn(name, res, args)


def init_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    P = np.zeros(NP)
    off = 0
    for nin, nout, used_in, used_out in ((16, 64, 5, 64), (64, 64, 64, 64), (64, 64, 64, 64), (64, 48, 64, 40)):
        W = np.zeros((nin, nout))
        W[:used_in, :used_out] = rng.normal(size=(used_in, used_out)) * scale * math.sqrt(2.0 / used_in)
        P[off:off + nin * nout] = W.ravel()
        off += nin * nout
        b = np.zeros(nout)
        b[:used_out] = rng.normal(size=used_out) * 0.1
        P[off:off + nout] = b
        off += nout
    assert off == NP
    return P


def test_param_count():
    assert _f("orc_mlp_param_count", C.c_int, [])() == NP


def forward(P, x):
    f = _f("orc_mlp_forward", None, [C.c_void_p, C.c_void_p, C.c_void_p])
    out = np.zeros(40)
    x = np.ascontiguousarray(x, np.float32)
    f(P.ctypes.data, x.ctypes.data, out.ctypes.data)
    return out


def pdf(out, w):
    f = _f("orc_vmf_mixture_pdf", C.c_double, [C.c_void_p, C.c_void_p])
    return f(out.ctypes.data, np.ascontiguousarray(w, np.float64).ctypes.data)


def test_forward_matches_numpy_restatement_of_layers():
    P = init_params(1)
    x = np.array([0.2, 0.7, 0.4, 0.3, 0.9], np.float32)
    h = np.zeros(16)
    h[:5] = x
    off = 0
    for nin, nout, relu in ((16, 64, True), (64, 64, True), (64, 64, True), (64, 48, False)):
        W = P[off:off + nin * nout].reshape(nin, nout)
        off += nin * nout
        b = P[off:off + nout]
        off += nout
        h = h @ W + b
        if relu:
            h = np.maximum(h, 0)
    assert np.allclose(forward(P, x), h[:40], rtol=1e-12, atol=1e-12)


def test_mixture_integrates_to_one():
    out = forward(init_params(2, scale=2.0), np.array([0.1, 0.5, 0.9, 0.2, 0.4], np.float32))
    out[1::5] = np.linspace(-2.0, 5.0, 8)          # kappa from 0.14 to 148
    n = 400
    g = (np.arange(n) + 0.5) / n
    vals = np.array([pdf(out, uv_to_dir(u, v).astype(np.float64)) for u in g for v in g])
    assert abs(vals.mean() * 4 * math.pi - 1.0) < 2e-3


def test_mixture_sampler_follows_pdf():
    out = forward(init_params(3, scale=2.0), np.array([0.3, 0.3, 0.6, 0.5, 0.5], np.float32))
    out[1::5] = np.linspace(-1.0, 3.0, 8)
    samp = _f("orc_vmf_mixture_sample", None, [C.c_void_p, C.c_void_p, C.c_void_p])
    from test_oracle_guiding import dir_to_uv
    rng = np.random.default_rng(4)
    n = 60000
    res = 12
    counts = np.zeros((res, res))
    w = np.zeros(3)
    for _ in range(n):
        u = rng.random(3)
        samp(out.ctypes.data, u.ctypes.data, w.ctypes.data)
        assert abs(np.linalg.norm(w) - 1) < 1e-9
        uv = dir_to_uv(w.astype(np.float32))
        counts[min(res - 1, int(uv[0] * res)), min(res - 1, int(uv[1] * res))] += 1
    # expected mass per equal-area cell by quadrature
    sub = 6
    expect = np.zeros((res, res))
    for i in range(res):
        for j in range(res):
            g = [((i + (a + 0.5) / sub) / res, (j + (b + 0.5) / sub) / res) for a in range(sub) for b in range(sub)]
            expect[i, j] = np.mean([pdf(out, uv_to_dir(u, v).astype(np.float64)) for u, v in g]) * 4 * math.pi / res ** 2
    e = expect * n
    m = e > 5
    chi2 = np.sum((counts[m] - e[m]) ** 2 / e[m])
    assert chi2 < m.sum() + 5 * math.sqrt(2 * m.sum()), (chi2, m.sum())


def loss_grad(P, x, d, wt, grad=True):
    f = _f("orc_mlp_loss_grad", C.c_double, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p])
    g = np.zeros(NP) if grad else None
    x = np.ascontiguousarray(x, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    wt = np.ascontiguousarray(wt, np.float32)
    loss = f(P.ctypes.data, len(wt), x.ctypes.data, d.ctypes.data, wt.ctypes.data, g.ctypes.data if grad else None)
    return loss, g


def batch(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 5)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    wt = rng.gamma(1.0, 1.0, n).astype(np.float32)
    return x, d, wt


def test_gradient_matches_finite_differences():
    P = init_params(5)
    x, d, wt = batch(6, 24)
    _, g = loss_grad(P, x, d, wt)
    rng = np.random.default_rng(7)
    live = np.nonzero(g)[0]
    assert len(live) > 1000
    idx = rng.choice(live, 300, replace=False)
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (loss_grad(Pp, x, d, wt, False)[0] - loss_grad(Pm, x, d, wt, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])) + 1e-7, (i, fd, g[i])
    # padding rows/columns get no gradient
    W0 = g[:16 * 64].reshape(16, 64)
    assert np.all(W0[5:] == 0)
    W3 = g[16 * 64 + 64 + 2 * (64 * 64 + 64):][:64 * 48].reshape(64, 48)
    assert np.all(W3[:, 40:] == 0)


def test_descent_fits_a_target_lobe():
    """Samples from one vMF lobe around a position-dependent mean; plain
    gradient descent must cut the loss well below its starting value."""
    rng = np.random.default_rng(8)
    n = 256
    x = rng.random((n, 5)).astype(np.float32)
    mu = np.stack([np.cos(3 * x[:, 0]), np.sin(3 * x[:, 0]), 0.3 + 0 * x[:, 0]], 1)
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    d = mu + 0.15 * rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    wt = np.ones(n, np.float32)
    P = init_params(9, scale=0.5)
    l0, _ = loss_grad(P, x, d, wt, False)
    m = np.zeros(NP); v = np.zeros(NP)
    for t in range(1, 301):
        l, g = loss_grad(P, x, d, wt)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        P -= 3e-3 * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
    print("loss", l0, "->", l)
    assert l < l0 - 1.5
