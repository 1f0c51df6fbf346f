"""Guided batches, captured guided renders and the context's safety rules,
through the C ABI.

- A guided batch (voxel probes or the learned network) renders every scene
  frame by frame with one guide update per frame over all scenes; each
  scene's image is an unbiased estimate, so its mean matches a plain render
  of the same scene within the noise, and its error against a high-spp plain
  reference stays at the plain render's level.
- Voxel probes of a batch are keyed by scene (bits 48..55 of the key); the
  network takes the scene index as input bits (first-layer rows 5..12), so
  with two scenes only the row of bit 0 learns.
- Guided renders can be captured into a HIP graph and replayed: each replay
  runs further frames of the same guide (the network's Adam step count lives
  on the device, so its parameters keep moving between replays).
- A captured render is refused once the memory it reads was freed (BVH
  rebuild) or its context destroyed; calls of one context on two streams are
  ordered (hits equal the one-stream result).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as orc
from aperturerendererv3_amd.workloads import (CORNELL_CAMERA, OCCLUDED_CAMERA, cornell_box, occluded_room,
                                              traversal_mesh, traversal_rays)

pytestmark = pytest.mark.gpu


def _cam(c):
    return dict(cam_pos=c["pos"], cam_dir=c["dir"], vfov=c["vfov"])


@pytest.fixture(scope="module")
def scenes():
    from aperturerendererv3_amd import Context
    specs = [(cornell_box(), _cam(CORNELL_CAMERA), 48, 48), (occluded_room(target_tris=20_000), _cam(OCCLUDED_CAMERA), 64, 36)]
    ctxs, views = [], []
    for (V, F, M, mats), cam, w, h in specs:
        c = Context(0)
        c.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
        c.build_bvh()
        ctxs.append(c)
        views.append(dict(width=w, height=h, **cam))
    yield ctxs, views
    for c in ctxs:
        c.close()


def _np(imgs):
    import torch
    torch.cuda.synchronize()
    return [i.cpu().numpy().astype(np.float64) for i in imgs]


def _plain(ctxs, views, spp, seed):
    return [_np(c.render_views([dict(v, spp=spp)], max_depth=8, seed=seed))[0] for c, v in zip(ctxs, views)]


@pytest.mark.parametrize("guiding", [1, 2], ids=["voxel", "mlp"])
def test_guided_batch_is_unbiased(scenes, guiding):
    from aperturerendererv3_amd import make_batch
    ctxs, views = scenes
    ref = _plain(ctxs, views, 4096, seed=101)
    plain = _plain(ctxs, views, 256, seed=7)
    b = make_batch(ctxs)
    if guiding == 2:
        b.mlp_configure(learning_rate=5e-3, batch=1 << 16, iters_per_frame=1, seed=3)
    v2 = [dict(v, spp=256) for v in views]
    print(f"guided batch ({'voxel' if guiding == 1 else 'mlp'}): 2 scenes, 256 spp, frames of 16", flush=True)
    got = _np(b.render_views(v2, max_depth=8, seed=7, guiding=guiding, frame_spp=16))
    for k in range(2):
        assert np.isfinite(got[k]).all()
        m_ref, m_got = ref[k].mean(), got[k].mean()
        assert abs(m_got - m_ref) / m_ref < 0.03, (k, m_got, m_ref)
        e_plain, e_got = orc.rel_mse(plain[k], ref[k]), orc.rel_mse(got[k], ref[k])
        print(f"scene {k}: relMSE plain {e_plain:.4g}, guided {e_got:.4g}", flush=True)
        assert e_got < 2.0 * e_plain, (k, e_got, e_plain)
    if guiding == 1:
        keys, _, counts = b.guiding_snapshot()
        scene_of = (keys >> np.uint64(48)) & np.uint64(0xFF)
        assert set(np.unique(scene_of).tolist()) == {0, 1}
        assert (counts > 0).all()
    else:
        p = b.mlp_params()
        w0 = p[:16 * 64].reshape(16, 64)
        assert np.abs(w0[5]).max() > 0.0           # scene bit 0 learned
        assert np.abs(w0[6:16]).max() == 0.0       # bits 1..7 of two scenes stay zero, rows 13..15 are padding
    b.close()


def test_single_scene_guided_render_unchanged_by_batch_support(scenes):
    ctxs, views = scenes
    c, v = ctxs[1], views[1]
    ref = _plain([c], [v], 4096, seed=101)[0]
    c.mlp_configure(learning_rate=5e-3, batch=1 << 16, iters_per_frame=1, seed=3)
    got = _np(c.render_views([dict(v, spp=256)], max_depth=8, seed=7, guiding=2, frame_spp=16))[0]
    assert abs(got.mean() - ref.mean()) / ref.mean() < 0.03
    p = c.mlp_params()
    assert np.abs(p[5 * 64:16 * 64]).max() == 0.0  # a plain context never feeds scene bits


@pytest.mark.parametrize("guiding", [1, 2], ids=["voxel", "mlp"])
def test_guided_render_graph_replays(scenes, guiding):
    ctxs, views = scenes
    c, v = ctxs[0], dict(views[0], spp=64)
    ref = _plain([c], [views[0]], 4096, seed=101)[0]
    c.guiding_reset()
    if guiding == 2:
        c.mlp_configure(learning_rate=5e-3, batch=1 << 16, iters_per_frame=1, seed=3)
    g = c.capture([v], max_depth=8, seed=11, guiding=guiding, frame_spp=8)
    params = []
    for r in range(3):
        print(f"guided graph ({'voxel' if guiding == 1 else 'mlp'}) replay {r}", flush=True)
        g.launch()
        img = _np(g.outputs)[0]
        assert np.isfinite(img).all()
        assert abs(img.mean() - ref.mean()) / ref.mean() < 0.05, (r, img.mean(), ref.mean())
        if guiding == 2:
            params.append(c.mlp_params())
        else:
            assert c.guiding_info()["probes"] > 0
    if guiding == 2:
        assert not np.array_equal(params[0], params[1]) and not np.array_equal(params[1], params[2])
    g.close()


def test_graph_refused_after_bvh_rebuild_and_destroy():
    from aperturerendererv3_amd import ArError, Context, _lib
    V, F, M, mats = cornell_box()
    c = Context(0)
    c.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
    c.build_bvh()
    g = c.capture([dict(width=16, height=16, spp=2, **_cam(CORNELL_CAMERA))], max_depth=4, seed=1)
    g.launch()
    c.build_bvh()                      # frees the nodes the graph reads
    with pytest.raises(ArError) as e:
        g.launch()
    assert e.value.code == _lib.AR_E_INVALID
    g.close()
    g2 = c.capture([dict(width=16, height=16, spp=2, **_cam(CORNELL_CAMERA))], max_depth=4, seed=1)
    g2.launch()
    c.close()                          # the graph's context is gone
    assert c._lib.ar_graph_launch(g2._g, None) == _lib.AR_E_INVALID
    g2.close()


def test_one_context_on_two_streams_is_ordered():
    import torch
    from aperturerendererv3_amd import Context
    V, F = traversal_mesh(levels=7)
    rays = traversal_rays(V, F, res=512)
    with Context(0) as c:
        c.set_meshes([{"vertices": V, "indices": F}])
        c.build_bvh()
        r = torch.from_numpy(rays).cuda()
        ref = c.trace_closest(r)
        torch.cuda.synchronize()
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        h1 = c.alloc_hits(r.shape[1])
        h2 = c.alloc_hits(r.shape[1])
        torch.cuda.synchronize()
        for _ in range(3):
            c.trace_closest(r, hits=h1, stream=s1)     # both calls share the context's refill counter
            c.trace_closest(r, hits=h2, stream=s2)
        torch.cuda.synchronize()
        for h in (h1, h2):
            assert torch.equal(h["prim"], ref["prim"]) and torch.equal(h["tuv"], ref["tuv"])
