"""Learned guiding in the integrator (AR_GUIDING_MLP): the network is trained
online on each frame's path vertices and sampled in the BSDF mixture.

* Unbiased: guided and plain renders of the Cornell box converge to the same
  mean (per-render spread ~0.1 %, bar 0.5 %).
* Gain: on the occluded-light room (light only through a door gap) the
  guided render's relMSE against a 8192-spp reference is below 0.8x the plain
  render's at equal spp, averaged over two seeds.
* Reproducible: a frame's training records are numbered in path order (no
  atomics) and gradients are reduced in a fixed order, so two guided renders
  with the same seeds give the same image and network, bit for bit; also with
  growing frames.
* Records: on a floor lit by one area light (no sky), a frame rendered before
  the guide is sampled (cosine-sampled directions) gives records only for
  directions that reach the light, each of weight lum(Le) pi / cos: the
  radiance that arrived (the light's emission without MIS weight, without the
  floor's own NEE sample or BSDF factor) over the direction's pdf.
"""
import numpy as np
import pytest

import oracle_lib as orc
from aperturerendererv3_amd.workloads import CORNELL_CAMERA, OCCLUDED_CAMERA, cornell_box, occluded_room

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from aperturerendererv3_amd import Context
    c = Context(0)
    yield c
    c.close()


def _cam(c):
    return dict(cam_pos=c["pos"], cam_dir=c["dir"], vfov=c["vfov"])


def _load(ctx, scene):
    V, F, M, mats = scene
    ctx.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
    ctx.build_bvh()


def test_learned_guiding_is_unbiased(ctx):
    import torch
    _load(ctx, cornell_box())
    ctx.mlp_configure(learning_rate=5e-3, iters_per_frame=2, seed=2)
    kw = dict(_cam(CORNELL_CAMERA), max_depth=5)
    g = ctx.render(128, 128, 256, guiding=2, seed=31, **kw)
    p = ctx.render(128, 128, 256, seed=32, **kw)
    torch.cuda.synchronize()
    a = g.reshape(-1, 3).mean(0).cpu().numpy()
    b = p.reshape(-1, 3).mean(0).cpu().numpy()
    print("learned", a, "plain", b, flush=True)
    assert np.isfinite(g.cpu().numpy()).all()
    assert np.all(np.abs(a - b) / b < 0.005), (a, b)


def test_learned_guiding_reduces_error_on_occluded_light(ctx):
    import torch
    _load(ctx, occluded_room())
    kw = dict(_cam(OCCLUDED_CAMERA), max_depth=8)
    ref = ctx.render(96, 96, 8192, seed=99, **kw).cpu().numpy()
    rp, rl = 0.0, 0.0
    for seed in (5, 6):
        plain = ctx.render(96, 96, 64, seed=seed, **kw).cpu().numpy()
        ctx.mlp_configure(learning_rate=5e-3, iters_per_frame=4, seed=seed)
        learned = ctx.render(96, 96, 64, seed=seed, guiding=2, **kw)
        torch.cuda.synchronize()
        rp += orc.rel_mse(plain, ref) / 2
        rl += orc.rel_mse(learned.cpu().numpy(), ref) / 2
    print(f"relMSE at 64 spp (2 seeds): plain {rp:.4e}, learned guiding {rl:.4e} ({rl / rp:.2f}x)", flush=True)
    assert rl < 0.8 * rp
    params = ctx.mlp_params()
    assert np.isfinite(params).all()
    ctx.mlp_configure(learning_rate=5e-3, iters_per_frame=4, seed=6)
    again = ctx.render(96, 96, 64, seed=6, guiding=2, **kw)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(again.cpu().numpy().view(np.uint32), learned.cpu().numpy().view(np.uint32))
    np.testing.assert_array_equal(ctx.mlp_params().view(np.uint32), params.view(np.uint32))


@pytest.mark.parametrize("growth", [1, 2])
def test_learned_guiding_reproducible_with_frame_growth(ctx, growth):
    import torch
    _load(ctx, cornell_box())
    kw = dict(_cam(CORNELL_CAMERA), max_depth=5)
    out = []
    for _ in range(2):
        ctx.mlp_configure(learning_rate=1e-2, iters_per_frame=8, batch=1 << 16, seed=3)
        g = ctx.render(96, 96, 120, guiding=2, seed=9, frame_spp=8, frame_growth=growth, **kw)
        torch.cuda.synchronize()
        out.append((g.cpu().numpy(), ctx.mlp_params()))
    np.testing.assert_array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    np.testing.assert_array_equal(out[0][1].view(np.uint32), out[1][1].view(np.uint32))
    p = ctx.render(96, 96, 1024, seed=10, **kw).reshape(-1, 3).mean(0).cpu().numpy()
    a = out[0][0].reshape(-1, 3).mean(0)
    print(f"growth {growth}: guided mean {a}, plain mean {p}", flush=True)
    assert np.all(np.abs(a - p) / p < 0.01), (a, p)


def test_learned_records_are_incident_radiance(ctx):
    le = (4.0, 2.0, 1.0)
    V = np.array([[-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1],
                  [-0.3, 1, -0.3], [0.3, 1, -0.3], [0.3, 1, 0.3], [-0.3, 1, 0.3]], np.float32)
    F = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]], np.uint32)
    M = np.array([0, 0, 1, 1], np.uint32)
    mats = [{"albedo": (0.5, 0.5, 0.5)}, {"albedo": (0.0, 0.0, 0.0), "emission": le}]
    ctx.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
    ctx.build_bvh()
    ctx.mlp_configure(learning_rate=1e-2, iters_per_frame=1, batch=1 << 18, seed=1)
    # one frame of 8 samples: the guide is not sampled yet (probability 0 before sample 1)
    ctx.render(64, 64, 8, cam_pos=(0.0, 0.6, 1.6), cam_dir=(0.0, -0.6, -1.6), vfov=60.0, seed=5, max_depth=5,
               guiding=2, frame_spp=8)
    r = ctx.mlp_records()
    n = r["weight"].shape[0]
    print(f"{n} records of {r['total']}", flush=True)
    assert n == r["total"] and n > 1000
    th = r["x"][:, 3] * np.pi
    ph = (r["x"][:, 4] - 0.5) * 2.0 * np.pi
    nrm = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    np.testing.assert_allclose(nrm, np.broadcast_to([0.0, 1.0, 0.0], nrm.shape), atol=1e-5)
    d = r["dir"].astype(np.float64)
    assert (d[:, 1] > 0).all()
    cos = d[:, 1]
    lum = 0.2126 * le[0] + 0.7152 * le[1] + 0.0722 * le[2]
    np.testing.assert_allclose(r["weight"] * cos / np.pi, lum, rtol=2e-4)
