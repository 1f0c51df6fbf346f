"""Parity of the HIP traversal (through the C ABI) with the CPU oracle.

Bar: bit-exact prim, t, u, v and flags for closest hit, bit-exact occlusion
words for any hit, on identical rays.  Small cases are checked against the
oracle's brute force (every triangle), larger ones against the oracle's BVH
traversal in its canonical mode (itself pinned to brute force in
test_oracle_trace.py), and the full-size traversal workload (16,777,216
rays and their shadow segments) on every ray, plus size-independent
properties.
"""
import numpy as np
import pytest

import geom
import oracle_lib as orc

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


@pytest.fixture(scope="module")
def ctx():
    from aperturerendererv3_amd import Context
    c = Context(0)
    yield c
    c.close()


def gpu_closest(ctx, V, F, rays, leaf_size=4, bins=16, width=8):
    torch = _torch()
    ctx.set_meshes([{"vertices": V, "indices": F}])
    ctx.build_bvh(width=width, leaf_size=leaf_size, bin_count=bins)
    r = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    h = ctx.trace_closest(r)
    occ = ctx.trace_any(r)
    torch.cuda.synchronize()
    out = {"prim": h["prim"].cpu().numpy().view(np.uint32), "t": h["tuv"][0].cpu().numpy(),
           "u": h["tuv"][1].cpu().numpy(), "v": h["tuv"][2].cpu().numpy(),
           "flags": h["flags"].cpu().numpy().view(np.uint32)}
    return out, occ.cpu().numpy().view(np.uint32)


def assert_hits_equal(got, ref, name=""):
    for k in ("prim", "t", "u", "v", "flags"):
        a, b = got[k].view(np.uint32), ref[k].view(np.uint32)
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"{name}: {k} differs on {bad.size} rays, first {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"


CASES = [
    ("sphere-random", lambda: geom.sphere_mesh(4), lambda V, F: geom.random_rays(20000, seed=1)),
    ("sphere-vertex", lambda: geom.sphere_mesh(4), lambda V, F: geom.aimed_rays(V, 8000, mode="vertex")),
    ("sphere-edge", lambda: geom.sphere_mesh(4), lambda V, F: geom.aimed_rays(V, 8000, mode="edge", F=F)),
    ("sphere-axis", lambda: geom.sphere_mesh(4), lambda V, F: geom.axis_rays(5000)),
    ("soup-random", lambda: geom.soup(3000, seed=3), lambda V, F: geom.random_rays(20000, seed=2)),
    ("soup-window", lambda: geom.soup(3000, seed=3),
     lambda V, F: geom.random_rays(10000, seed=4, tmin=0.5, tmax=3.0)),
    ("grid-vertex", lambda: geom.grid_mesh(6), lambda V, F: geom.aimed_rays(V, 5000, mode="vertex")),
    ("grid-edge-cases", lambda: geom.grid_mesh(4), lambda V, F: geom.edge_case_rays()),
    ("one-triangle", lambda: (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32)),
     lambda V, F: geom.random_rays(3000, seed=5, extent=1.0, target_extent=0.6)),
]


@pytest.mark.parametrize("width", [2, 4, 8])
@pytest.mark.parametrize("case", range(len(CASES)), ids=[c[0] for c in CASES])
def test_closest_and_any_match_brute_force(ctx, case, width):
    name, mk_mesh, mk_rays = CASES[case]
    V, F = mk_mesh()
    rays = mk_rays(V, F)
    print(f"case {name} width {width}: {F.shape[0]} tris, {rays.shape[1]} rays", flush=True)
    got, occ = gpu_closest(ctx, V, F, rays, width=width)
    ref = orc.brute_closest(V, F, rays)
    assert_hits_equal(got, ref, name)
    ref_occ = orc.brute_any(V, F, rays)
    assert np.array_equal(occ, ref_occ), f"{name}: occlusion words differ"


@pytest.mark.parametrize("width,leaf_size,bins", [(2, 1, 4), (2, 8, 12), (2, 15, 64), (8, 1, 16), (8, 3, 8),
                                                  (8, 15, 64)])
def test_bvh_parameters_do_not_change_results(ctx, width, leaf_size, bins):
    V, F = geom.sphere_mesh(5)
    rays = geom.random_rays(20000, seed=7)
    got, occ = gpu_closest(ctx, V, F, rays, leaf_size=leaf_size, bins=bins, width=width)
    ref = orc.OracleBVH(V, F).closest(rays)
    assert_hits_equal(got, ref, f"leaf={leaf_size} bins={bins}")


def test_empty_scene_all_miss(ctx):
    torch = _torch()
    ctx.set_meshes([])
    ctx.build_bvh()
    r = torch.from_numpy(geom.random_rays(1000)).cuda()
    h = ctx.trace_closest(r)
    occ = ctx.trace_any(r)
    torch.cuda.synchronize()
    assert (h["prim"].cpu().numpy().view(np.uint32) == 0xFFFFFFFF).all()
    assert np.isinf(h["tuv"][0].cpu().numpy()).all()
    assert (occ.cpu().numpy() == 0).all()


def test_trace_without_bvh_raises():
    torch = _torch()
    from aperturerendererv3_amd import ArError, Context
    with Context(0) as c:
        r = torch.from_numpy(geom.random_rays(10)).cuda()
        with pytest.raises(ArError) as e:
            c.trace_closest(r)
        assert e.value.code == -3


def test_scaled_traversal_workload(ctx):
    """The traversal workload at 1/64 of its triangle count (16,384 tris) and
    512^2 ray slots: every ray checked against the oracle's BVH traversal,
    a subsample against brute force, plus the shadow any-hit pass."""
    torch = _torch()
    from aperturerendererv3_amd.workloads import traversal_mesh, traversal_rays
    V, F = traversal_mesh(levels=6)
    rays = traversal_rays(V, F, res=512)
    for width in (2, 4, 8):
        got, _ = gpu_closest(ctx, V, F, rays, width=width)
        obvh = orc.OracleBVH(V, F)
        ref = obvh.closest(rays)
        assert_hits_equal(got, ref, f"scaled workload width {width}")
    sub = np.random.default_rng(0).choice(rays.shape[1], 2000, replace=False)
    bref = orc.brute_closest(V, F, rays[:, sub].copy())
    assert np.array_equal(ref["prim"][sub], bref["prim"])
    # shadow segments generated on the GPU, traced by both
    r = torch.from_numpy(rays).cuda()
    h = ctx.trace_closest(r)
    sh = ctx.make_shadow_rays(r, h, (2.0, 2.0, 2.0))
    occ = ctx.trace_any(sh)
    torch.cuda.synchronize()
    shn = sh.cpu().numpy()
    assert np.array_equal(occ.cpu().numpy().view(np.uint32), obvh.any(shn))
    hit = ref["prim"] != orc.MISS
    assert (shn[7][~hit] == 0).all() and (shn[7][hit] > 0).all()


def test_full_size_traversal_workload(ctx):
    """1,048,576 triangles, 4096^2 ray slots: a 262,144-ray subsample checked
    against the oracle's BVH traversal; over all rays, any-hit on the same
    rays must equal (closest hit exists), and rays with a hit at t must be
    unoccluded on [0, t) and occluded on [0, t*(1+2^-20)]."""
    torch = _torch()
    from aperturerendererv3_amd.workloads import traversal_mesh, traversal_rays
    V, F = traversal_mesh(levels=9)
    rays = traversal_rays(V, F, res=4096)
    print(f"full workload: {F.shape[0]} tris, {rays.shape[1]} rays", flush=True)
    ctx.set_meshes([{"vertices": V, "indices": F}])
    ctx.build_bvh()
    r = torch.from_numpy(rays).cuda()
    h = ctx.trace_closest(r)
    occ = ctx.trace_any(r)
    torch.cuda.synchronize()
    prim = h["prim"].cpu().numpy().view(np.uint32)
    t = h["tuv"][0].cpu().numpy()
    assert np.array_equal(orc.unpack_bits(occ.cpu().numpy(), rays.shape[1]), prim != orc.MISS)
    hit = np.nonzero(prim != orc.MISS)[0]
    assert hit.size > rays.shape[1] // 4
    # [0, t) must be free, [0, t + eps] must be blocked
    r2 = r[:, torch.from_numpy(hit).cuda()].contiguous()
    tt = torch.from_numpy(t[hit]).cuda()
    r2[7] = tt
    free = orc.unpack_bits(ctx.trace_any(r2).cpu().numpy(), hit.size)
    free2 = orc.unpack_bits(ctx.trace_any(r2).cpu().numpy(), hit.size)
    if free.any() or free2.any():
        bad = hit[free | free2][:8]
        rs = rays[:, bad].copy()
        oref = orc.OracleBVH(V, F).closest(rs)
        print("violations", int(free.sum()), int(free2.sum()), "rays", bad, "gpu prim", prim[bad], "t", t[bad],
              "oracle prim", oref["prim"], "t", oref["t"], flush=True)
    assert not free.any() and not free2.any()
    r2[7] = tt * (1 + 2.0 ** -20) + 1e-30
    blocked = orc.unpack_bits(ctx.trace_any(r2).cpu().numpy(), hit.size)
    assert blocked.all()
    # determinism: a second closest-hit pass is bit-identical
    h2 = ctx.trace_closest(r)
    torch.cuda.synchronize()
    assert torch.equal(h2["prim"], h["prim"]) and torch.equal(h2["tuv"], h["tuv"])
    # every ray against the oracle's BVH traversal (OpenMP over the host cores)
    obvh = orc.OracleBVH(V, F)
    print("oracle closest hit over all rays", flush=True)
    ref = obvh.closest(rays)
    got = {"prim": prim, "t": t, "u": h["tuv"][1].cpu().numpy(), "v": h["tuv"][2].cpu().numpy(),
           "flags": h["flags"].cpu().numpy().view(np.uint32)}
    assert_hits_equal(got, ref, "full workload, all rays")
    # the benchmark's shadow segments toward (2, 2, 2): any-hit words against the oracle on every segment
    sh = ctx.make_shadow_rays(r, h, (2.0, 2.0, 2.0))
    occ_sh = ctx.trace_any(sh)
    torch.cuda.synchronize()
    print("oracle any hit over all shadow segments", flush=True)
    ref_occ = obvh.any(sh.cpu().numpy())
    got_occ = occ_sh.cpu().numpy().view(np.uint32)
    bad = np.nonzero(got_occ != ref_occ)[0]
    assert bad.size == 0, f"shadow occlusion words differ at {bad[:5]}"
    assert 0 < orc.unpack_bits(got_occ, rays.shape[1]).sum() < hit.size
