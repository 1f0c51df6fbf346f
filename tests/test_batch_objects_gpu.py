"""Batches of scenes with analytic objects (ar_batch_create / ar_render_batch):
the object Cornell box (box light), triangles with sphere and box lights, a
glass sphere around its camera (the path starts inside the medium), two
scenes with signed-distance tori and voxel volumes (each its own voxel
pool), and a triangle-only scene, merged into one batch.

A batched image must equal the scene's own render bit for bit, at BVH widths
2, 4 and 8, in RGB and with 4 hero wavelengths, and on a replay of the
captured graph: the merge rebases object ids of leaf slots, material and
light ids, the objects' own leaf slots and voxel offsets, and object prim
ids in the light table, and each view starts in its own scene's medium."""
import numpy as np
import pytest

import oracle_lib as orc
from aperturerendererv3_amd.workloads import (CORNELL_CAMERA, SPECTRAL_CAMERA, cornell_box, cornell_objects,
                                              spectral_scene)

pytestmark = pytest.mark.gpu


def _cam(c):
    return dict(cam_pos=c["pos"], cam_dir=c["dir"], vfov=c["vfov"])


def _mixed():
    V, F, M, mats = cornell_box()
    objs = [
        {"type": "sphere", "center": (0.2, 0.25, -0.2), "extent": (0.06, 0.06, 0.06),
         "material": {"albedo": (0, 0, 0), "emission": (30.0, 25.0, 20.0)}},
        {"type": "box", "center": (-0.25, 0.2, 0.1), "extent": (0.05, 0.04, 0.08), "rotation": (10, 30, 0),
         "material": {"albedo": (0, 0, 0), "emission": (4.0, 8.0, 12.0)}},
        {"type": "box", "center": (0.05, -0.05, 0.0), "extent": (0.1, 0.12, 0.08), "rotation": (25, -40, 15),
         "material": {"albedo": (0.7, 0.5, 0.3), "model": 1, "roughness": 0.3, "specular": 0.5}},
    ]
    return (V, F, M, mats), objs


def _glass_around_camera():
    V, F, M, mats = cornell_box()
    glass = {"albedo": (0.9, 0.8, 0.7), "model": 2, "ior": 1.4, "sigma_a": 0.3}
    objs = [{"type": "sphere", "center": tuple(CORNELL_CAMERA["pos"]), "extent": (0.2, 0.2, 0.2), "material": glass},
            {"type": "box", "center": (0.0, -0.3, 0.0), "extent": (0.1, 0.1, 0.1), "rotation": (0, 20, 0),
             "material": glass}]
    return (V, F, M, mats), objs


def _sdf(ctx, seed):
    rng = np.random.default_rng(seed)
    V, F, M, mats = cornell_box()
    vol = orc.sdf_bake("torus", (0.35, 0.12), (24 + 8 * seed, 24, 24))
    if seed == 2:   # a second, unused volume first: this scene's volume is not the pool's first
        ctx.sdf_volume_add(np.ones((4, 4, 4), np.float32))
    vidx = ctx.sdf_volume_add(vol)
    objs = []
    for k in range(6):
        kind = "torus" if k % 2 == 0 else "volume"
        ob = {"type": kind, "center": (float(rng.uniform(-0.3, 0.3)), float(rng.uniform(-0.35, 0.1)),
                                       float(rng.uniform(-0.3, 0.3))),
              "extent": (0.12, 0.12, 0.12), "rotation": tuple(float(x) for x in rng.uniform(-90, 90, 3)),
              "material": {"albedo": tuple(float(x) for x in rng.uniform(0.2, 0.9, 3))}}
        if kind == "torus":
            ob["sdf"] = (0.35, 0.1)
        else:
            ob["volume"] = vidx
        objs.append(ob)
    return (V, F, M, mats), objs


@pytest.fixture(scope="module")
def scenes():
    from aperturerendererv3_amd import Context
    ctxs, views = [], []

    def add(tri, objs, view, setup=None):
        c = Context(0)
        if setup is not None:
            tri, objs = setup(c)
        if tri is None:
            c.set_meshes([])
        else:
            V, F, M, mats = tri
            c.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
        c.set_objects(objs)
        ctxs.append(c)
        views.append(view)

    add(None, cornell_objects(), dict(width=48, height=48, spp=8, **_cam(CORNELL_CAMERA)))
    add(*_mixed(), dict(width=40, height=32, spp=8, **_cam(CORNELL_CAMERA)))
    add(None, None, dict(width=32, height=32, spp=16, **_cam(CORNELL_CAMERA)), setup=lambda c: _sdf(c, 1))
    add(*_glass_around_camera(), dict(width=32, height=24, spp=8, **_cam(CORNELL_CAMERA)))
    V, F, M, mats = spectral_scene(target_tris=3000, seed=5)
    add((V, F, M, mats), [], dict(width=48, height=27, spp=4, **_cam(SPECTRAL_CAMERA)))
    add(None, None, dict(width=24, height=24, spp=8, **_cam(CORNELL_CAMERA)), setup=lambda c: _sdf(c, 2))
    yield ctxs, views
    for c in ctxs:
        c.close()


def _np(imgs):
    import torch
    torch.cuda.synchronize()
    return [i.cpu().numpy() for i in imgs]


@pytest.mark.parametrize("width", [2, 4, 8])
@pytest.mark.parametrize("shared", [dict(max_depth=6, seed=3), dict(max_depth=6, seed=4, spectral=4, output=1)],
                         ids=["rgb", "spectral"])
def test_object_scenes_batch_equals_solo(scenes, width, shared):
    from aperturerendererv3_amd import make_batch
    ctxs, views = scenes
    for c in ctxs:
        c.build_bvh(width=width)
    solo = _np([c.render_views([v], **shared)[0] for c, v in zip(ctxs, views)])
    b = make_batch(ctxs)
    print(f"width {width} {shared}: {len(ctxs)} scenes batched", flush=True)
    together = _np(b.render_views(views, **shared))
    for k, (a, t) in enumerate(zip(solo, together)):
        assert a.mean() > 0, k
        assert np.array_equal(a.view(np.uint32), t.view(np.uint32)), (k, float(np.abs(a - t).max()))
    g = b.capture(views, **shared)
    g.launch()
    for k, (a, t) in enumerate(zip(solo, _np(g.outputs))):
        assert np.array_equal(a.view(np.uint32), t.view(np.uint32)), ("graph", k)
    g.close()
    b.close()


def test_camera_inside_glass_starts_in_the_medium(scenes):
    """The scene whose camera sits inside a glass sphere (its paths start in
    the medium, as the oracle's) matches the FP32 oracle; the batch equals
    it (test above), so the batched view started in its own scene's medium."""
    ctxs, views = scenes
    img = _np([ctxs[3].render_views([views[3]], max_depth=6, seed=3)[0]])[0]
    (V, F, M, mats), objs = _glass_around_camera()
    o = orc.render(V, F, M, mats, views[3]["width"], views[3]["height"], views[3]["spp"], seed=3, max_depth=6,
                   precision=32, objects=objs, **_cam(CORNELL_CAMERA))
    rel = orc.rel_mse(img, o)
    print(f"glass around the camera: relMSE(gpu, oracle) {rel:.3e}", flush=True)
    assert rel <= 1e-4
