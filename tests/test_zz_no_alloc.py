"""ar_trace_closest_noalloc against ar_trace_closest, through the C ABI.

The no-allocation entry point runs the same traversal kernels with their
scratch (refill counter and stack spill) in the library's fixed per-device
array.  Bar: bit-exact prim, t, u, v and flags against ar_trace_closest on
the same rays, for two calls in a row (the second with fewer rays, so stale
scratch from the first call would show), at BVH widths 2 and 8, on a
triangle mesh and on a scene with analytic objects.  Also: the call rejects
ray counts beyond its 32-bit refill counter before launching anything.
"""
import ctypes as C

import numpy as np
import pytest

from aperturerendererv3_amd.workloads import random_object_scene, traversal_mesh, traversal_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from aperturerendererv3_amd import Context
    c = Context(0)
    yield c
    c.close()


def _host(h):
    return {"prim": h["prim"].cpu().numpy().view(np.uint32), "t": h["tuv"][0].cpu().numpy().view(np.uint32),
            "u": h["tuv"][1].cpu().numpy().view(np.uint32), "v": h["tuv"][2].cpu().numpy().view(np.uint32),
            "flags": h["flags"].cpu().numpy().view(np.uint32)}


def _assert_same(got, ref, name):
    for k in ("prim", "t", "u", "v", "flags"):
        bad = np.nonzero(got[k] != ref[k])[0]
        assert bad.size == 0, f"{name}: {k} differs on {bad.size} rays, first {bad[:5]}"


def _two_calls(ctx, rays_np, name):
    import torch
    r = torch.from_numpy(np.ascontiguousarray(rays_np)).cuda()
    n = r.shape[1]
    small = r[:, : n // 3 + 17].contiguous()
    for rr, tag in ((r, "first call"), (small, "second call, fewer rays")):
        print(f"{name}: {tag}, {rr.shape[1]} rays", flush=True)
        ref = ctx.trace_closest(rr)
        got = ctx.trace_closest_noalloc(rr)
        torch.cuda.synchronize()
        hr, hg = _host(ref), _host(got)
        assert (hr["prim"] != 0xFFFFFFFF).sum() > rr.shape[1] // 10
        _assert_same(hg, hr, f"{name}: {tag}")


@pytest.mark.parametrize("width", [2, 4, 8])
def test_noalloc_equals_trace_closest_mesh(ctx, width):
    V, F = traversal_mesh(levels=7)
    rays = traversal_rays(V, F, res=512)
    ctx.set_meshes([{"vertices": V, "indices": F}])
    ctx.build_bvh(width=width)
    _two_calls(ctx, rays, f"mesh width {width}")


@pytest.mark.parametrize("width", [2, 4, 8])
def test_noalloc_equals_trace_closest_objects(ctx, width):
    V, F, objs = random_object_scene(n_tris=1500, n_boxes=25, n_spheres=25, seed=4)
    ctx.set_meshes([{"vertices": V, "indices": F}])
    ctx.set_objects(objs)
    ctx.build_bvh(width=width, leaf_size=3)
    rng = np.random.default_rng(5)
    n = 40000
    o = rng.uniform(-1.3, 1.3, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((8, n), np.float32)
    rays[0:3] = o.T
    rays[4:7] = d.T
    rays[7] = np.inf
    _two_calls(ctx, rays, f"objects width {width}")
    ctx.set_objects([])


def test_noalloc_rejects_too_many_rays(ctx):
    import torch
    from aperturerendererv3_amd import _lib
    V, F = traversal_mesh(levels=3)
    ctx.set_meshes([{"vertices": V, "indices": F}])
    ctx.build_bvh()
    r = torch.zeros((8, 64), dtype=torch.float32, device="cuda")
    h = ctx.alloc_hits(64)
    rc = ctx._lib.ar_trace_closest_noalloc(ctx._h, C.byref(ctx._rays_c(r)), (1 << 31) + 1,
                                           C.byref(ctx._hits_c(h)), None)
    assert rc == _lib.AR_E_CAPACITY
    assert b"2^31" in ctx._lib.ar_last_error(ctx._h)
    # the context still traces afterwards
    got = ctx.trace_closest_noalloc(r)
    torch.cuda.synchronize()
    assert (got["prim"].cpu().numpy().view(np.uint32) == 0xFFFFFFFF).all()   # zero directions never hit
