"""Parity of the HIP traversal with analytic objects (boxes, spheres) mixed
into the triangle BVH, through the C ABI, against the oracle's brute force
(itself pinned to FP64 closed forms in test_oracle_objects.py).

Bar: bit-exact prim, t, u, v and flags (AR_HIT_INSIDE for hits from inside
an object) for closest hit and exact occlusion bits for any hit, at BVH
widths 2 and 8, including rays that start inside objects, a duplicated
object (ties go to the smaller prim id) and tmin / tmax windows."""
import numpy as np
import pytest

import oracle_lib as orc
from aperturerendererv3_amd.workloads import cornell_objects, random_object_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from aperturerendererv3_amd import Context
    c = Context(0)
    yield c
    c.close()


def _rays(rng, n, spread=1.3):
    o = rng.uniform(-spread, spread, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((8, n), np.float32)
    r[0:3] = o.T
    r[4:7] = d.T
    r[7] = np.inf
    return r


def _gpu(ctx, V, F, objs, rays, width):
    import torch
    ctx.set_meshes([{"vertices": V, "indices": F}] if F.shape[0] else [])
    ctx.set_objects(objs)
    ctx.build_bvh(width=width, leaf_size=3)
    r = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    h = ctx.trace_closest(r)
    occ = ctx.trace_any(r)
    torch.cuda.synchronize()
    out = {"prim": h["prim"].cpu().numpy().view(np.uint32), "t": h["tuv"][0].cpu().numpy(),
           "u": h["tuv"][1].cpu().numpy(), "v": h["tuv"][2].cpu().numpy(),
           "flags": h["flags"].cpu().numpy().view(np.uint32)}
    return out, orc.unpack_bits(occ.cpu().numpy(), rays.shape[1])


@pytest.mark.parametrize("width", [2, 4, 8])
def test_mixed_scene_matches_oracle(ctx, width):
    V, F, objs = random_object_scene(n_tris=1500, n_boxes=25, n_spheres=25, seed=4)
    objs.append(dict(objs[-1]))
    rng = np.random.default_rng(21)
    rays = _rays(rng, 30000)
    rays[3, ::7] = 0.05
    rays[7, ::5] = rng.uniform(0.2, 2.0, rays[7, ::5].shape[0]).astype(np.float32)
    print(f"mixed scene width {width}: {F.shape[0]} tris, {len(objs)} objects, {rays.shape[1]} rays", flush=True)
    got, occ = _gpu(ctx, V, F, objs, rays, width)
    g = orc.Geom(V, F, objs)
    ref = g.brute(rays)
    for k in ("prim", "t", "u", "v", "flags"):
        a, b = got[k].view(np.uint32), ref[k].view(np.uint32)
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"{k} differs on {bad.size} rays, first {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"
    np.testing.assert_array_equal(occ, g.any_brute(rays))
    assert np.sum(ref["prim"] >= F.shape[0]) > 3000 and np.any(ref["flags"] & orc.HIT_INSIDE)


def test_objects_only_cornell(ctx):
    """Object-only scene (no triangles): rays from inside the box hit walls."""
    objs = cornell_objects()
    rng = np.random.default_rng(3)
    rays = _rays(rng, 8000, spread=0.45)
    V = np.zeros((0, 3), np.float32)
    F = np.zeros((0, 3), np.uint32)
    got, occ = _gpu(ctx, V, F, objs, rays, 8)
    ref = orc.Geom(V, F, objs).brute(rays)
    for k in ("prim", "t", "flags"):
        np.testing.assert_array_equal(got[k].view(np.uint32), ref[k].view(np.uint32))
    back = rays[6] < 0
    assert np.all(got["prim"][back] != orc.MISS)
