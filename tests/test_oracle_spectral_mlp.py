"""Oracle spectral basis network (core/spectrum/spectrum_basis.cuh and
pytorch/spectrum/train_spectrum.py) pinned against the reference's golden
vectors, torch.autograd, torch's clipping and torch's AdamW.

  - the FP32 forward with the reference's shipped weights reproduces its 100
    golden vectors (gtTest, the reflectances its trained model returned);
  - the FP64 loss gradient equals torch.autograd's gradient of the same
    network and loss, with and without the training noise, and central
    differences of the oracle's own loss;
  - clipping equals torch.nn.utils.clip_grad_norm_, the AdamW step equals
    torch.optim.AdamW with the reference's parameter groups (two steps);
  - the table queries equal direct evaluation at table nodes and interpolate
    linearly between them; Interact scales an RGB above 1;
  - the noise normals are standard normal (Kolmogorov-Smirnov)."""
import math
from pathlib import Path

import numpy as np
import pytest
import torch
from scipy import stats as sps

import oracle_lib as orc

FIX = Path(__file__).resolve().parent / "golden" / "spectral_mlp_reference.npz"


def _ref():
    d = np.load(FIX)
    return d["weights"], d["inputs"], d["expected"]


def _random_weights(seed=0):
    rng = np.random.default_rng(seed)
    w = np.zeros(orc.SMLP_PARAMS, np.float32)
    mask = orc.smlp_decay_mask()
    w[mask] = rng.normal(0.0, math.sqrt(2.0 / 32.0) * 0.5, mask.sum())
    w[~mask] = rng.normal(0.0, 0.05, (~mask).sum())
    return w


def _batch(n, seed=1):
    rng = np.random.default_rng(seed)
    x = np.empty((n, 4), np.float32)
    x[:, :3] = rng.random((n, 3)) ** 0.75
    x[:, 3] = rng.uniform(360.0, 780.0, n)
    y = rng.random(n).astype(np.float32)
    return x, y


def test_golden_vectors_of_the_reference():
    w, x, gt = _ref()
    # the golden vectors come from PyTorch on a GPU with TF32 matmuls: that
    # evaluation reproduces them to FP32 accumulation-order noise, so every
    # layer, embedding entry and weight index is pinned
    t = orc.smlp_eval_torch_tf32(w, x)
    et = np.abs(t - gt)
    # the reference's own FP32 CUDA forward is off by TF32's rounding; its self
    # test (spectrum_basis.cuh:2491-2512) reports that MSE
    out = orc.smlp_eval(w, x)
    err = np.abs(out - gt)
    print(f"TF32 evaluation: max {et.max():.3e}, within 1e-6 {np.mean(et <= 1e-6):.2f}; "
          f"FP32 forward: max {err.max():.3e}, MSE {np.mean(err ** 2):.3e}")
    assert et.max() < 5e-5 and np.mean(et <= 1e-6) >= 0.8
    assert err.max() < 1e-3 and np.mean(err ** 2) < 3e-8
    # a mistake in the FP32 forward shows against the TF32 one
    assert np.abs(out - t).max() < 1e-3


# ---- torch restatement of the training forward (autograd supplies the backward)

def _torch_loss(w, x, y, noise=None):
    W = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    xt = torch.tensor(x, dtype=torch.float64)
    rgb = xt[:, :3].clone()
    lam = (xt[:, 3:4] - 360.0) / 420.0
    if noise is not None:
        rgb = rgb + noise["rgb"]
        lam = lam + noise["lam"]

    def lin(o, fin, fout, h, bias=True):
        Wm = W[o:o + fin * fout].view(fout, fin)
        z = h @ Wm.T
        return z + W[o + fin * fout:o + fin * fout + fout] if bias else z

    s = rgb.max(dim=1, keepdim=True).values
    se = s + 0.001
    k = math.pi * torch.arange(1, 5, dtype=torch.float64)
    emb = torch.cat([torch.cat([torch.sin(v * k), torch.cos(v * k)], 1)
                     for v in (rgb[:, 0:1] / se, rgb[:, 1:2] / se, rgb[:, 2:3] / se, s)], 1)
    h = torch.relu(lin(0, 32, 32, emb))
    for b in range(7):
        xin = h + (noise["blk"][:, b] if noise is not None else 0.0)
        h = h + torch.relu(lin(1056 + b * 1056, 32, 32, xin))
    kr = lin(8448, 32, 4, h, bias=False) * s
    kk = math.pi * torch.arange(1, 17, dtype=torch.float64)
    h = torch.cat([torch.sin(lam * kk), torch.cos(lam * kk)], 1)
    for b in range(7):
        xin = h + (noise["blk"][:, 7 + b] if noise is not None else 0.0)
        h = h + torch.relu(lin(8576 + b * 1056, 32, 32, xin))
    kl = lin(15968, 32, 4, h, bias=False)
    res = (kr * kl).sum(dim=1, keepdim=True)
    pred = torch.relu(1.0 - torch.relu(1.0 - res))
    ln = xt[:, 3:4]
    fits = torch.cat([1.065 * torch.exp(-0.5 * ((ln - 595.8) / 33.33) ** 2)
                      + 0.366 * torch.exp(-0.5 * ((ln - 446.8) / 19.44) ** 2),
                      1.014 * torch.exp(-0.5 * ((torch.log(ln) - 6.3213077) / 0.075) ** 2),
                      1.839 * torch.exp(-0.5 * ((torch.log(ln) - 6.1088030) / 0.051) ** 2)], 1)
    imp = 0.25 + torch.clamp(fits, min=0.0)
    loss = torch.square((pred - torch.tensor(y, dtype=torch.float64)[:, None]) * imp).mean()
    loss.backward()
    return loss.item(), W.grad.numpy()


def _noise(n, seed, step):
    rgb = np.zeros((n, 3))
    lam = np.zeros((n, 1))
    blk = np.zeros((n, 14, 32))
    for j in range(n):
        if j < n // 2:
            rgb[j] = [orc.smlp_normal(seed, step, j, c) * 0.025 / 256.0 for c in range(3)]
            lam[j] = orc.smlp_normal(seed, step, j, 3) * 0.025 / 420.0
        for e in range(14):
            blk[j, e] = [orc.smlp_normal(seed, step, j, 4 + e * 32 + u) * 0.001 for u in range(32)]
    return {"rgb": torch.tensor(rgb), "lam": torch.tensor(lam), "blk": torch.tensor(blk)}


@pytest.mark.parametrize("noise", [False, True])
def test_gradient_equals_autograd(noise):
    w = _random_weights(3)
    x, y = _batch(24, 4)
    loss, g = orc.smlp_loss_grad(w, x, y, noise=noise, seed=11, step=5)
    tl, tg = _torch_loss(w, x, y, _noise(len(x), 11, 5) if noise else None)
    assert abs(loss - tl) <= 1e-12 * max(1.0, abs(tl))
    assert np.abs(g).max() > 0
    np.testing.assert_allclose(g, tg, rtol=1e-8, atol=1e-12 * np.abs(tg).max())


def test_gradient_central_differences():
    w = _random_weights(5)
    x, y = _batch(16, 6)
    _, g = orc.smlp_loss_grad(w, x, y)
    rng = np.random.default_rng(0)
    picks = list(rng.choice(np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max()), 12, replace=False))
    for p in picks:
        wp, wm = w.copy(), w.copy()
        h = np.float32(1e-3) * max(1.0, abs(float(w[p])))
        wp[p] += h
        wm[p] -= h
        lp, _ = orc.smlp_loss_grad(wp, x, y)
        lm, _ = orc.smlp_loss_grad(wm, x, y)
        fd = (lp - lm) / (float(wp[p]) - float(wm[p]))
        assert abs(fd - g[p]) <= 2e-3 * abs(g[p]) + 1e-9, (p, fd, g[p])


def test_clip_and_adamw_equal_torch():
    rng = np.random.default_rng(7)
    g = rng.normal(size=orc.SMLP_PARAMS)
    norm, gc = orc.smlp_clip(g.copy(), 1.0)
    t = torch.tensor(g, requires_grad=True)
    t.grad = torch.tensor(g)
    tn = torch.nn.utils.clip_grad_norm_([t], 1.0)
    assert abs(norm - tn.item()) < 1e-9 * norm
    np.testing.assert_allclose(gc, t.grad.numpy(), rtol=1e-12)

    mask = orc.smlp_decay_mask()
    p0 = rng.normal(size=orc.SMLP_PARAMS)
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    ema = p0.copy()
    tw = torch.tensor(p0[mask], requires_grad=True)
    tb = torch.tensor(p0[~mask], requires_grad=True)
    opt = torch.optim.AdamW([{"params": [tw], "weight_decay": 1e-3}, {"params": [tb], "weight_decay": 0.0}],
                            lr=5e-4, betas=(0.9, 0.95), eps=1e-8)
    after = []
    for step in (1, 2):
        gs = rng.normal(size=orc.SMLP_PARAMS)
        lr = 5e-4 * (0.7 if step == 2 else 1.0)
        for group in opt.param_groups:
            group["lr"] = lr
        tw.grad = torch.tensor(gs[mask])
        tb.grad = torch.tensor(gs[~mask])
        opt.step()
        orc.smlp_adamw(p, m, v, gs, step, lr, ema=ema)
        after.append(p.copy())
        np.testing.assert_allclose(p[mask], tw.detach().numpy(), rtol=0, atol=1e-15)
        np.testing.assert_allclose(p[~mask], tb.detach().numpy(), rtol=0, atol=1e-15)
    # ModelEMA.update after each step (decay 0.999)
    np.testing.assert_allclose(ema, 0.999 * (0.999 * p0 + 0.001 * after[0]) + 0.001 * after[1], rtol=1e-14)


def test_learning_rate_schedule():
    total = 10000
    assert orc.smlp_lr_factor(0, total) == 0.0
    assert orc.smlp_lr_factor(5, total) == pytest.approx(0.5)
    assert orc.smlp_lr_factor(10, total) == pytest.approx(1.0)
    assert orc.smlp_lr_factor(total, total, min_ratio=0.01) == pytest.approx(0.01)
    mid = 10 + (total - 10) // 2
    assert orc.smlp_lr_factor(mid, total, min_ratio=0.0) == pytest.approx(0.5, abs=1e-4)


def test_table_queries():
    w, _, _ = _ref()
    res = 16
    # at a voxel centre the trilinear query returns the node's value
    node = (np.array([3, 7, 12]) + 0.5) / res
    np.testing.assert_allclose(orc.smlp_query_rgb(w, res, node), orc.smlp_encode(w, rgb=node.astype(np.float32)),
                               rtol=1e-6, atol=1e-7)
    # halfway between two nodes along x: the mean of both
    a = (np.array([3, 7, 12]) + 0.5) / res
    b = (np.array([4, 7, 12]) + 0.5) / res
    mid = ((a + b) / 2).astype(np.float32)
    expect = 0.5 * (orc.smlp_encode(w, rgb=a.astype(np.float32)) + orc.smlp_encode(w, rgb=b.astype(np.float32)))
    np.testing.assert_allclose(orc.smlp_query_rgb(w, res, mid), expect, rtol=1e-5, atol=1e-6)
    # below the first centre the coordinate clamps to node 0
    np.testing.assert_allclose(orc.smlp_query_rgb(w, res, np.zeros(3, np.float32)),
                               orc.smlp_encode(w, rgb=np.full(3, 0.5 / res, np.float32)), rtol=1e-6, atol=1e-7)
    lres = 840
    i = 123
    lam_i = 360.0 + (i + 0.5) * 420.0 / lres
    np.testing.assert_allclose(orc.smlp_query_lambda(w, lres, lam_i), orc.smlp_lambda_entry(w, lres, i), rtol=1e-5,
                               atol=1e-6)
    np.testing.assert_allclose(orc.smlp_lambda_entry(w, lres, i), orc.smlp_encode(w, lam=np.float32(lam_i)),
                               rtol=1e-6, atol=1e-7)
    # Interact divides an RGB above 1 by its largest channel and scales back
    rgb = np.array([0.8, 0.4, 0.2], np.float32)
    r1 = orc.smlp_lut_interact(w, res, lres, rgb, 550.0)
    r3 = orc.smlp_lut_interact(w, res, lres, rgb * np.float32(3.0 / 0.8), 550.0)
    rn = orc.smlp_lut_interact(w, res, lres, rgb / np.float32(0.8), 550.0)
    assert r3 == pytest.approx(rn * 3.0, rel=1e-6)
    assert 0.0 <= r1 <= 1.0


def test_noise_is_standard_normal():
    z = np.array([orc.smlp_normal(9, s, j, slot) for s in range(4) for j in range(50) for slot in range(25)])
    assert abs(z.mean()) < 0.05 and abs(z.std() - 1.0) < 0.05
    assert sps.kstest(z, "norm").pvalue > 1e-3
