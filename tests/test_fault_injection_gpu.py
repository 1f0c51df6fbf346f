"""Fault injection on the GPU, through the C ABI: the cases of
test_fault_injection.py (which pins the oracle on them) run on the HIP
library and are compared with the oracle at equal seed.

- Non-finite material values are refused (AR_E_INVALID) before anything
  reaches the device; the context stays usable.
- Radiance clamp and scrub (core/render/render.cuh:966-969): a light of
  radiance 4e3 gives exactly 1e3 per directly seen sample; at 3e38 some FP32
  paths overflow and count 0, so no pixel is brighter than at 1e30.  GPU
  equals the FP32 oracle with the integrator's bar (relMSE <= 1e-6, 99.5 %
  of pixels within 1e-5 relative) at 4e3; at 3e38 a path whose sum lands
  next to FLT_MAX can overflow on one side only (the GPU fuses multiply-adds
  that the oracle rounds twice), so the bar there is 99 % of pixels and
  relMSE <= 1e-4 (one flipped path moves a pixel by 1e3/8).
- A scene without lights under NEE: black without the sky; with the sky
  NEE on equals NEE off bit for bit and matches the oracle.
- Guiding capacity overflow: with room for 64 probes the voxel guide keeps
  64, counts the samples it cannot place in `dropped`, and the render stays
  finite and unbiased (mean within 3 % of a plain render).
- Coincident and zero-area triangles at widths 2, 4 and 8, leaf sizes 1, 3
  and 15, and a depth bound of 4 (far below what 1,300 triangles need): hits
  and occlusion equal brute force bit for bit (ties to the smallest index)."""
import numpy as np
import pytest

import oracle_lib as orc
from aperturerendererv3_amd.workloads import CORNELL_CAMERA, cornell_box
from test_fault_injection import bad_materials, coincident_mesh, coincident_rays, first_copies

pytestmark = pytest.mark.gpu

CAM = dict(cam_pos=CORNELL_CAMERA["pos"], cam_dir=CORNELL_CAMERA["dir"], vfov=CORNELL_CAMERA["vfov"])
W, H = 48, 36


@pytest.fixture(scope="module")
def ctx():
    from aperturerendererv3_amd import Context
    c = Context(0)
    yield c
    c.close()


def render(ctx, le, spp=8, seed=3, max_depth=5, **kw):
    import torch
    V, F, M, mats = cornell_box(light_le=(le, le, le))
    ctx.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
    ctx.build_bvh()
    img = ctx.render(W, H, spp, seed=seed, max_depth=max_depth, **CAM, **kw)
    torch.cuda.synchronize()
    return img.cpu().numpy()


def oracle(le, spp=8, seed=3, max_depth=5, **kw):
    V, F, M, mats = cornell_box(light_le=(le, le, le))
    return orc.render(V, F, M, mats, W, H, spp, seed=seed, max_depth=max_depth, precision=32, **CAM, **kw)


def close_fraction(g, o, rtol=1e-5, atol=1e-6):
    return (np.abs(g - o) <= rtol * np.abs(o) + atol).all(axis=2).mean()


@pytest.mark.parametrize("field,value", bad_materials())
def test_non_finite_materials_are_refused(ctx, field, value):
    from aperturerendererv3_amd import ArError, _lib
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    F = np.array([[0, 1, 2]], np.uint32)
    mat = {"albedo": (0.5, 0.5, 0.5), "roughness": 1.0}
    if field in ("albedo", "emission"):
        mat[field] = (0.5, value, 0.5)
    else:
        mat[field] = value
    with pytest.raises(ArError) as e:
        ctx.set_meshes([{"vertices": V, "indices": F, "tri_material": np.zeros(1, np.uint32), "materials": [mat]}])
    assert e.value.code == _lib.AR_E_INVALID
    with pytest.raises(ArError) as e:
        ctx.set_objects([{"type": _lib.AR_OBJ_SPHERE, "center": (0, 0, 0), "extent": (0.5, 0.5, 0.5),
                          "material": mat}])
    assert e.value.code == _lib.AR_E_INVALID
    ctx.set_objects([])
    ctx.set_meshes([{"vertices": V, "indices": F}])   # still usable
    ctx.build_bvh()


def test_radiance_clamp_matches_oracle(ctx):
    g = render(ctx, 4.0e3)
    o = oracle(4.0e3)
    lit = (o == np.float32(1.0e3)).all(axis=2)
    print(f"clamp: {lit.sum()} pixels see only the light; relMSE {orc.rel_mse(g, o):.3e}, "
          f"close {close_fraction(g, o):.5f}", flush=True)
    assert np.isfinite(g).all() and g.min() >= 0.0 and g.max() <= 1.0e3
    assert lit.sum() >= 4 and (g[lit] == np.float32(1.0e3)).all()
    assert orc.rel_mse(g, o) <= 1e-6 and close_fraction(g, o) >= 0.995


def test_overflowing_paths_are_scrubbed(ctx):
    g = render(ctx, 3.0e38)
    big = render(ctx, 1.0e30)
    o = oracle(3.0e38)
    print(f"overflow: relMSE {orc.rel_mse(g, o):.3e}, close {close_fraction(g, o):.5f}, "
          f"darker pixels {int(((big - g) >= 124.0).any(axis=2).sum())}", flush=True)
    assert np.isfinite(g).all() and g.min() >= 0.0 and g.max() <= 1.0e3
    assert (g <= big).all()
    assert orc.rel_mse(g, o) <= 1e-4 and close_fraction(g, o) >= 0.99


def test_scene_without_lights_under_nee(ctx):
    from aperturerendererv3_amd import _lib
    black = render(ctx, 0.0, spp=4, seed=2, max_depth=6, nee=True)
    assert (black == 0.0).all()
    sky = dict(sky=_lib.AR_SKY_GRADIENT)
    on = render(ctx, 0.0, spp=4, seed=2, max_depth=6, nee=True, **sky)
    off = render(ctx, 0.0, spp=4, seed=2, max_depth=6, nee=False, **sky)
    np.testing.assert_array_equal(on.view(np.uint32), off.view(np.uint32))
    o = oracle(0.0, spp=4, seed=2, max_depth=6, nee=1, **sky)
    print(f"no lights: relMSE {orc.rel_mse(on, o):.3e}, close {close_fraction(on, o, 1e-3, 1e-5):.5f}", flush=True)
    assert on.max() > 0.0
    assert orc.rel_mse(on, o) <= 1e-6 and close_fraction(on, o, 1e-3, 1e-5) >= 0.98


def test_guiding_capacity_overflow_drops_samples(ctx):
    import torch
    V, F, M, mats = cornell_box()
    ctx.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
    ctx.build_bvh()
    plain = ctx.render(W, H, 1024, seed=11, max_depth=6, **CAM).cpu().numpy().astype(np.float64)
    ctx.guiding_configure(64)
    try:
        ctx.guiding_reset()
        img = ctx.render(W, H, 256, seed=5, max_depth=6, guiding=1, frame_spp=16, **CAM)
        torch.cuda.synchronize()
        info = ctx.guiding_info()
        g = img.cpu().numpy().astype(np.float64)
        print(f"capacity 64: probes {info['probes']}, samples {info['samples']}, dropped {info['dropped']}, "
              f"mean {g.mean():.5f} vs plain {plain.mean():.5f}", flush=True)
        assert info["capacity"] == 64 and info["probes"] == 64 and info["dropped"] > 0
        assert np.isfinite(g).all() and abs(g.mean() - plain.mean()) / plain.mean() < 0.03
    finally:
        ctx.guiding_configure(1 << 20)
        ctx.guiding_reset()
    ctx.render(W, H, 16, seed=5, max_depth=6, guiding=1, frame_spp=16, **CAM)
    torch.cuda.synchronize()
    assert ctx.guiding_info()["dropped"] == 0


@pytest.mark.parametrize("width", [2, 4, 8])
@pytest.mark.parametrize("leaf_size,max_depth", [(1, 64), (3, 4), (15, 64)])
def test_coincident_triangles_match_brute_force(ctx, width, leaf_size, max_depth):
    import torch
    V, F = coincident_mesh()
    rays = coincident_rays()
    ctx.set_meshes([{"vertices": V, "indices": F}])
    ctx.build_bvh(width=width, leaf_size=leaf_size, max_depth=max_depth)
    info = ctx.bvh_info()
    print(f"width {width} leaf {leaf_size} depth bound {max_depth}: {F.shape[0]} tris, built depth {info['depth']}",
          flush=True)
    r = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    h = ctx.trace_closest(r)
    occ = ctx.trace_any(r)
    torch.cuda.synchronize()
    got = {"prim": h["prim"].cpu().numpy().view(np.uint32), "t": h["tuv"][0].cpu().numpy(),
           "u": h["tuv"][1].cpu().numpy(), "v": h["tuv"][2].cpu().numpy(),
           "flags": h["flags"].cpu().numpy().view(np.uint32)}
    ref = orc.brute_closest(V, F, rays)
    for k in ("prim", "t", "u", "v", "flags"):
        np.testing.assert_array_equal(got[k].view(np.uint32), ref[k].view(np.uint32), err_msg=k)
    np.testing.assert_array_equal(occ.cpu().numpy().view(np.uint32), orc.brute_any(V, F, rays))
    prim = got["prim"]
    hit = prim != 0xFFFFFFFF
    assert (first_copies(V, F)[prim[hit]] == prim[hit]).all()
