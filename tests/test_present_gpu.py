"""The output stage and host-side render of the interactive / offline
SceneRenderer::Render on the GPU, through the C ABI:

- ar_present: the debug views (albedo, normal, depth / metallic / roughness
  heat maps, statistics) equal the oracle's views of the same records (the
  sample-count view goes through a double log: >= 99.9 %); the image view is
  ar_tonemap; the denoised view is ar_denoise then ar_tonemap; missing
  records and unknown views are errors.
- ar_render_host: the same image as ar_render (bit for bit) and its
  ar_tonemap, in host memory.
- Guided progressive frames (voxel probes, learned network): each frame
  updates the guide; the running mean stays an unbiased estimate (its mean
  within 3 % of a plain 4096-spp render)."""
import numpy as np
import pytest

import oracle_lib as orc
from aperturerendererv3_amd.workloads import CORNELL_CAMERA, cornell_box

pytestmark = pytest.mark.gpu

W, H = 64, 48
CAM = dict(cam_pos=CORNELL_CAMERA["pos"], cam_dir=CORNELL_CAMERA["dir"], vfov=CORNELL_CAMERA["vfov"])


@pytest.fixture(scope="module")
def ctx():
    from aperturerendererv3_amd import Context
    c = Context(0)
    V, F, M, mats = cornell_box()
    mats = [dict(m, roughness=0.2 + 0.2 * k, metallic=0.5 * (k % 2), model=1 if k == 1 else 0)
            for k, m in enumerate(mats)]
    c.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
    c.build_bvh()
    yield c
    c.close()


def _frames(ctx, n=4, **kw):
    import torch
    acc = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    gb = torch.empty((H, W, 16), dtype=torch.float32, device="cuda")
    st = torch.empty((H, W, 12), dtype=torch.float32, device="cuda")
    for f in range(n):
        ctx.render_progressive(f, acc, spp=2, seed=5, max_depth=5, gbuffer=gb, stats=st, **CAM, **kw)
    return acc, gb, st


def test_debug_views_match_oracle(ctx):
    import torch
    acc, gb, st = _frames(ctx)
    torch.cuda.synchronize()
    g, s = gb.cpu().numpy(), st.cpu().numpy()
    assert g[..., 8].max() > g[..., 8].min() and g[..., 7].max() > 0.0
    for view in range(2, 11):
        print(f"view {view}", flush=True)
        got = ctx.present(gbuffer=gb, stats=st, view=view).cpu().numpy().view(np.uint32).reshape(-1)
        ref = orc.view(g, s, view, W * H)
        same = np.mean(got == ref)
        assert same >= (0.999 if view == 7 else 1.0), (view, same)
        if view == 7:   # every pixel has its 8 samples: log(9) / 100 in red
            assert np.all(got == (0xFF000000 | (int(255.0 * np.float32(np.log(9.0) * 0.01)) << 16))), view
        else:
            assert len(np.unique(got)) > 1, view


def test_image_and_denoised_views(ctx):
    import torch
    acc, gb, st = _frames(ctx)
    img = ctx.present(hdr=acc, view=0, exposure=0.8, tone=2)
    ref = ctx.tonemap(acc, exposure=0.8, tone=2)
    den = ctx.present(hdr=acc, gbuffer=gb, stats=st, view=1, exposure=0.8, tone=2)
    ref_den = ctx.tonemap(ctx.denoise(acc, gb, st), exposure=0.8, tone=2)
    torch.cuda.synchronize()
    assert torch.equal(img, ref) and torch.equal(den, ref_den)
    assert not torch.equal(img, den)


def test_present_rejects_missing_records_and_unknown_views(ctx):
    import torch
    from aperturerendererv3_amd import ArError, _lib
    acc, gb, st = _frames(ctx, n=1)
    out = torch.empty((H, W), dtype=torch.int32, device="cuda")
    for kw in (dict(hdr=acc, view=11), dict(hdr=acc, view=2), dict(hdr=acc, gbuffer=gb, view=1),
               dict(gbuffer=gb, view=8, width=W, height=H)):
        with pytest.raises(ArError) as e:
            ctx.present(out=out, **kw)
        assert e.value.code == _lib.AR_E_INVALID, kw


def test_render_host_equals_render(ctx):
    import torch
    kw = dict(seed=9, max_depth=5, **CAM)
    hdr, ldr = ctx.render_host(W, H, 8, exposure=1.3, tone=2, **kw)
    dev = ctx.render(W, H, 8, **kw)
    ref_ldr = ctx.tonemap(dev, exposure=1.3, tone=2)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(hdr.view(np.uint32), dev.cpu().numpy().view(np.uint32))
    np.testing.assert_array_equal(ldr, ref_ldr.cpu().numpy().view(np.uint32))
    only_ldr = ctx.render_host(W, H, 8, exposure=1.3, tone=2, hdr=False, **kw)[1]
    np.testing.assert_array_equal(only_ldr, ldr)


@pytest.mark.parametrize("guiding", [1, 2], ids=["voxel", "mlp"])
def test_guided_progressive_frames(ctx, guiding):
    import torch
    kw = dict(seed=21, max_depth=8, **CAM)
    ref = ctx.render(W, H, 4096, seed=101, max_depth=8, **CAM)
    ctx.guiding_reset()
    if guiding == 2:
        ctx.mlp_configure(learning_rate=5e-3, batch=1 << 14, iters_per_frame=1, seed=3)
    acc = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    odd = torch.empty_like(acc)
    even = torch.empty_like(acc)
    params = []
    spp, frames = (1, 96) if guiding == 1 else (4, 24)
    for f in range(frames):
        ctx.render_progressive(f * spp, acc, odd, even, spp=spp, guiding=guiding, frame_spp=spp, **kw)
        if guiding == 2 and f in (2, frames - 1):
            params.append(ctx.mlp_params())
    torch.cuda.synchronize()
    a, r = acc.cpu().numpy().astype(np.float64), ref.cpu().numpy().astype(np.float64)
    print(f"guided progressive ({'voxel' if guiding == 1 else 'mlp'}): mean {a.mean():.5f} vs {r.mean():.5f}",
          flush=True)
    assert np.isfinite(a).all() and abs(a.mean() - r.mean()) / r.mean() < 0.03
    half = 0.5 * (odd + even).cpu().numpy()
    assert abs(half.mean() - a.mean()) / a.mean() < 1e-3
    if guiding == 1:
        assert ctx.guiding_info()["probes"] > 0
    else:
        assert not np.array_equal(params[0], params[1])
