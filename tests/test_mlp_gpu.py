"""BF16 MFMA forward/backward of the guiding MLP (ar_mlp_* through the C ABI)
against the FP64 oracle network on the same FP32 parameters.

Two references: the oracle with the kernels' BF16 roundings emulated (the
same values rounded at the same places, FP64 sums), which the GPU must match
to FP32-accumulation accuracy (outputs within 1e-3 of the output scale; per
layer gradient cosine > 0.9999 and relative L2 error < 1 %: a rare BF16
rounding flip moves one element by 0.4 %); and the unrounded FP64 network,
with the BF16 bars (outputs within 3 % of the output scale; gradient cosine
> 0.995, relative L2 < 10 %, the error of 8-bit mantissas compounded over
three backward layers).
"""
import ctypes as C
import numpy as np
import pytest

import oracle_lib as orc
from test_oracle_mlp import NP, batch, forward, init_params, loss_grad


def forward_bf16(P, x):
    f = orc.fn("orc_mlp_forward_bf16", None, [C.c_void_p, C.c_void_p, C.c_void_p])
    out = np.zeros(40)
    x = np.ascontiguousarray(x, np.float32)
    f(P.ctypes.data, x.ctypes.data, out.ctypes.data)
    return out


def loss_grad_bf16(P, x, d, wt):
    f = orc.fn("orc_mlp_loss_grad_bf16", C.c_double,
               [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p])
    g = np.zeros(NP)
    x, d, wt = (np.ascontiguousarray(a, np.float32) for a in (x, d, wt))
    loss = f(P.ctypes.data, len(wt), x.ctypes.data, d.ctypes.data, wt.ctypes.data, g.ctypes.data)
    return loss, g

pytestmark = pytest.mark.gpu

LAYERS = {"W0": (0, 16 * 64), "b0": (1024, 64), "W1": (1088, 4096), "b1": (5184, 64), "W2": (5248, 4096),
          "b2": (9344, 64), "W3": (9408, 3072), "b3": (12480, 48)}


@pytest.fixture(scope="module")
def ctx():
    from aperturerendererv3_amd import Context
    c = Context(0)
    yield c
    c.close()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def test_param_roundtrip(ctx):
    P = init_params(11).astype(np.float32)
    ctx.mlp_set_params(P)
    assert np.array_equal(ctx.mlp_params(), P)


def test_forward_matches_fp64_oracle(ctx):
    P = init_params(12).astype(np.float32)
    ctx.mlp_set_params(P)
    x, _, _ = batch(13, 1000)
    g = ctx.mlp_forward(_dev(x)).cpu().numpy()
    o = np.stack([forward(P.astype(np.float64), xi) for xi in x])
    ob = np.stack([forward_bf16(P.astype(np.float64), xi) for xi in x])
    scale = np.sqrt(np.mean(o ** 2))
    err = np.abs(g - o).max()
    errb = np.abs(g - ob).max()
    print(f"forward: max |gpu - fp64| = {err:.4f}, max |gpu - bf16 oracle| = {errb:.2e}, output rms {scale:.4f}",
          flush=True)
    assert err < 0.03 * scale
    assert errb < 1e-3 * scale


def test_forward_splits_host_counts_above_2_26(ctx):
    """A host count above 2^26 vertices runs as launches of at most 2^26 (the
    kernel's offsets are 32-bit): vertices on both sides of the split get
    the same bits as a call on them alone (each vertex's outputs depend on
    its own inputs only)."""
    import torch
    P = init_params(14).astype(np.float32)
    ctx.mlp_set_params(P)
    n = (1 << 26) + 1000
    gen = torch.Generator(device="cuda").manual_seed(5)
    x = torch.rand((n, 5), device="cuda", generator=gen)
    out = ctx.mlp_forward(x)
    lo = (1 << 26) - 1000
    part = ctx.mlp_forward(x[lo:].contiguous())
    torch.cuda.synchronize()
    a, b = out[lo:].cpu().numpy(), part.cpu().numpy()
    del out, x
    torch.cuda.empty_cache()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.all(np.isfinite(b)) and np.abs(b).max() > 0


def test_gradient_matches_fp64_oracle(ctx):
    P = init_params(14).astype(np.float32)
    ctx.mlp_set_params(P)
    x, d, wt = batch(15, 777)      # not a multiple of the 64-sample tile
    g, loss = ctx.mlp_gradient(_dev(x), _dev(d), _dev(wt))
    g = g.cpu().numpy().astype(np.float64)
    lo, go = loss_grad(P.astype(np.float64), x, d, wt)
    lb, gb = loss_grad_bf16(P.astype(np.float64), x, d, wt)
    print(f"loss gpu {loss.item():.6f} fp64 {lo:.6f} bf16 oracle {lb:.6f}", flush=True)
    assert abs(loss.item() - lb) < 1e-4 * abs(lb) + 1e-5
    assert abs(loss.item() - lo) < 0.01 * abs(lo) + 1e-3
    for name, (off, n) in LAYERS.items():
        for ref, cmin, rmax, tag in ((gb, 0.9999, 0.01, "bf16 oracle"), (go, 0.995, 0.10, "fp64")):
            a, b = g[off:off + n], ref[off:off + n]
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
            rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
            print(f"{name} vs {tag}: cos {cos:.6f}, rel L2 {rel:.4f}", flush=True)
            assert cos > cmin and rel < rmax, (name, tag)
    # the padding never receives gradient
    assert np.all(g[0:1024].reshape(16, 64)[5:] == 0)
    assert np.all(g[9408:9408 + 3072].reshape(64, 48)[:, 40:] == 0)


def test_training_fits_a_target_lobe(ctx):
    """Same target as the oracle descent test; 300 Adam steps on the GPU."""
    rng = np.random.default_rng(8)
    n = 4096
    x = rng.random((n, 5)).astype(np.float32)
    mu = np.stack([np.cos(3 * x[:, 0]), np.sin(3 * x[:, 0]), 0.3 + 0 * x[:, 0]], 1)
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    d = mu + 0.15 * rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    wt = np.ones(n, np.float32)
    ctx.mlp_configure(learning_rate=3e-3, seed=3)
    P0 = ctx.mlp_params().astype(np.float64)
    l0, _ = loss_grad(P0, x[:512], d[:512], wt[:512], False)
    ctx.mlp_train(_dev(x), _dev(d), _dev(wt), 300)
    P1 = ctx.mlp_params().astype(np.float64)
    l1, _ = loss_grad(P1, x[:512], d[:512], wt[:512], False)
    print(f"fp64 loss of the GPU-trained network: {l0:.4f} -> {l1:.4f}", flush=True)
    assert l1 < l0 - 1.5


def test_gradient_over_many_rounds_matches_bf16_oracle(ctx):
    """60,000 samples: every block of the training kernel runs two rounds of
    128 samples, so the register accumulators carry across rounds."""
    P = init_params(21).astype(np.float32)
    ctx.mlp_set_params(P)
    x, d, wt = batch(22, 60_000)
    g, loss = ctx.mlp_gradient(_dev(x), _dev(d), _dev(wt))
    g = g.cpu().numpy().astype(np.float64)
    lb, gb = loss_grad_bf16(P.astype(np.float64), x, d, wt)
    print(f"60k: loss gpu {loss.item():.6f} bf16 oracle {lb:.6f}", flush=True)
    assert abs(loss.item() - lb) < 1e-4 * abs(lb) + 1e-5
    for name, (off, n) in LAYERS.items():
        a, b = g[off:off + n], gb[off:off + n]
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
        print(f"60k {name}: cos {cos:.6f}, rel L2 {rel:.4f}", flush=True)
        assert cos > 0.9999 and rel < 0.01, name


def test_gradient_and_training_are_bit_reproducible(ctx):
    """No atomics: the same batch gives the same gradient bits, and the same
    Adam steps give the same parameters in another context."""
    from aperturerendererv3_amd import Context
    P = init_params(31).astype(np.float32)
    ctx.mlp_set_params(P)
    x, d, wt = batch(32, 1 << 17)
    xd, dd, wd = _dev(x), _dev(d), _dev(wt)
    g1, l1 = ctx.mlp_gradient(xd, dd, wd)
    g2, l2 = ctx.mlp_gradient(xd, dd, wd)
    assert np.array_equal(g1.cpu().numpy().view(np.uint32), g2.cpu().numpy().view(np.uint32))
    assert l1.item() == l2.item()
    params = []
    for c in (ctx, Context(0)):
        c.mlp_configure(learning_rate=3e-3, batch=1 << 16, seed=5)
        c.mlp_train(xd, dd, wd, 20)
        params.append(c.mlp_params())
    assert np.array_equal(params[0].view(np.uint32), params[1].view(np.uint32))
