"""Pins the CPU oracle's traversal functions (no GPU needed).

Each check is chosen so that a plausible mistake fails it: closed-form FP64
hits catch a wrong sign/index in the triangle test, watertightness catches
a dropped zero-edge case, brute force catches any culling or tie error in
the BVH traversal, a numpy re-derivation of the SAH split catches a wrong
cost, and the structural invariants catch a broken tree.
"""
import numpy as np
import pytest

import geom
import oracle_lib as orc


def fp64_hit(o, d, a, b, c):
    """Moller-Trumbore in float64: (t, w0, w1, w2)."""
    o, d, a, b, c = (np.asarray(x, np.float64) for x in (o, d, a, b, c))
    e1, e2 = b - a, c - a
    p = np.cross(d, e2)
    det = e1 @ p
    if det == 0:
        return None
    s = o - a
    w1 = (s @ p) / det
    q = np.cross(s, e1)
    w2 = (d @ q) / det
    t = (e2 @ q) / det
    return t, 1 - w1 - w2, w1, w2


def test_tri_known_answers():
    v0, v1, v2 = (0, 0, 0), (1, 0, 0), (0, 1, 0)
    hit, t, u, v, f = orc.tri_test((0.25, 0.25, -1), (0, 0, 1), v0, v1, v2)
    assert hit and t == 1.0 and u == 0.25 and v == 0.25 and (f & 7) == 0
    back = orc.tri_test((0.25, 0.25, -1), (0, 0, 1), v0, v2, v1)
    assert back[0] and back[1] == 1.0 and back[2] == 0.25 and back[3] == 0.25
    assert (f & 8) != (back[4] & 8)          # exactly one of the windings is a backface
    # through v0: weights of v1 and v2 exactly zero
    hit, t, u, v, f = orc.tri_test((0, 0, 2), (0, 0, -1), v0, v1, v2)
    assert hit and t == 2.0 and u == 0.0 and v == 0.0 and (f & 6) == 6
    # on the edge v1-v2: weight of v0 exactly zero
    hit, t, u, v, f = orc.tri_test((0.5, 0.5, 2), (0, 0, -1), v0, v1, v2)
    assert hit and (f & 1) == 1 and u == 0.5 and v == 0.5
    # outside, parallel, zero direction
    assert not orc.tri_test((1.0, 1.0, 2), (0, 0, -1), v0, v1, v2)[0]
    assert not orc.tri_test((0.2, 0.2, 1), (1, 0, 0), v0, v1, v2)[0]
    assert not orc.tri_test((0.2, 0.2, 1), (0, 0, 0), v0, v1, v2)[0]
    # t is reported for any sign (callers range-check)
    hit, t, *_ = orc.tri_test((0.25, 0.25, 1), (0, 0, 1), v0, v1, v2)
    assert hit and t == -1.0


def test_tri_matches_fp64_closed_form():
    rng = np.random.default_rng(7)
    n_in = n_out = 0
    for _ in range(4000):
        a, b, c = rng.normal(size=(3, 3))
        o = rng.normal(size=3) * 3
        tgt = a * rng.random() + b * rng.random() + c * rng.random()
        tgt /= 1.5
        d = tgt - o
        d /= np.linalg.norm(d)
        ref = fp64_hit(o, d, a, b, c)
        hit, t, u, v, _ = orc.tri_test(o, d, a, b, c)
        if ref is None:
            continue
        t64, w0, w1, w2 = ref
        if min(w0, w1, w2) > 1e-4:
            n_in += 1
            assert hit
            assert abs(t - t64) <= 1e-5 * max(1.0, abs(t64))
            assert abs(u - w1) <= 1e-4 and abs(v - w2) <= 1e-4
        elif min(w0, w1, w2) < -1e-4:
            n_out += 1
            assert not hit
    assert n_in > 500 and n_out > 500


def test_watertight_closed_sphere():
    V, F = geom.sphere_mesh(4)
    rng = np.random.default_rng(11)
    n = 20000
    o = rng.uniform(-0.3, 0.3, (n, 3)).astype(np.float32)
    tgt = np.concatenate([V[rng.integers(0, V.shape[0], n // 2)],              # vertices
                          0.5 * (V[F[rng.integers(0, F.shape[0], n - n // 2), 0]]
                                 + V[F[rng.integers(0, F.shape[0], n - n // 2), 0]])])
    d = (tgt - o).astype(np.float32)
    r = np.zeros((8, n), np.float32)
    r[0:3], r[4:7], r[7] = o.T, d.T, np.inf
    h = orc.brute_closest(V, F, r)
    assert (h["prim"] != orc.MISS).all(), "a ray escaped a closed mesh"


def test_shared_edges_min_id_tie_rule():
    V, F = geom.grid_mesh(8)
    xs = np.arange(0, 9) / 8.0
    pts = [(x, y) for x in xs for y in xs] + [(x + 1 / 16, y) for x in xs[:-1] for y in xs]
    r = np.zeros((8, len(pts)), np.float32)
    r[0], r[1], r[2] = [p[0] for p in pts], [p[1] for p in pts], 2.0
    r[6], r[7] = -1.0, np.inf
    h = orc.brute_closest(V, F, r)
    for k, (x, y) in enumerate(pts):
        # triangles whose closed region holds (x, y), by exact rational tests
        holders = []
        for i, f in enumerate(F):
            a, b, c = V[f].astype(np.float64)
            ref = fp64_hit((x, y, 2.0), (0, 0, -1.0), a, b, c)
            if ref and min(ref[1:]) >= -1e-12:
                holders.append(i)
        assert h["prim"][k] == min(holders)
        assert h["t"][k] == 2.0


def _bvh_invariants(bvh, n, leaf_size):
    nodes = bvh.nodes
    seen = np.zeros(n, np.int64)
    leaves = 0
    for i, nd in enumerate(nodes):
        if nd["count"] > 0:
            leaves += 1
            ids = bvh.indices[nd["first"]:nd["first"] + nd["count"]]
            seen[ids] += 1
            assert (bvh.pmin[ids] >= nd["bmin"]).all() and (bvh.pmax[ids] <= nd["bmax"]).all()
        else:
            assert nd["left"] == i + 1, "pre-order: left child follows its parent"
            for c in (nd["left"], nd["right"]):
                assert (nodes[c]["bmin"] >= nd["bmin"]).all() and (nodes[c]["bmax"] <= nd["bmax"]).all()
    assert (seen == 1).all(), "every primitive in exactly one leaf"
    assert len(nodes) == 2 * leaves - 1
    return leaves


@pytest.mark.parametrize("levels", [2, 4, 6])
def test_bvh_build_invariants(levels):
    V, F = geom.sphere_mesh(levels)
    bvh = orc.OracleBVH(V, F, leaf_size=8, bins=12)
    _bvh_invariants(bvh, F.shape[0], 8)
    big = [nd for nd in bvh.nodes if nd["count"] > 8]
    assert not big                           # distinct centroids: leaves stay within leaf_size


def test_bvh_root_split_matches_numpy_sah():
    V, F = geom.soup(3000, seed=5)
    bvh = orc.OracleBVH(V, F, leaf_size=8, bins=12)
    cen, pmin, pmax = bvh.pcen.astype(np.float64), bvh.pmin.astype(np.float64), bvh.pmax.astype(np.float64)
    ext = cen.max(0) - cen.min(0)
    axis = 0
    if ext[1] > ext[0] and ext[1] >= ext[2]:
        axis = 1
    elif ext[2] > ext[0] and ext[2] >= ext[1]:
        axis = 2
    c0 = np.float32(cen[:, axis].min())
    invw = np.float32(12) / (np.float32(cen[:, axis].max()) - c0)
    b = np.clip(((bvh.pcen[:, axis] - c0) * invw).astype(np.int64), 0, 11)

    def area(lo, hi):
        d = np.maximum(0, hi - lo)
        return 2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])
    costs = []
    for s in range(11):
        L, R = b <= s, b > s
        if L.sum() == 0 or R.sum() == 0:
            costs.append(np.inf)
            continue
        costs.append(area(pmin[L].min(0), pmax[L].max(0)) * L.sum() + area(pmin[R].min(0), pmax[R].max(0)) * R.sum())
    best = int(np.argmin(costs))
    left = bvh.nodes[bvh.nodes[0]["left"]]
    right = bvh.nodes[bvh.nodes[0]["right"]]
    L = b <= best
    assert np.allclose(left["bmin"], pmin[L].min(0)) and np.allclose(left["bmax"], pmax[L].max(0))
    assert np.allclose(right["bmin"], pmin[~L].min(0)) and np.allclose(right["bmax"], pmax[~L].max(0))


def _ray_sets():
    Vs, Fs = geom.sphere_mesh(4)
    Vsoup, Fsoup = geom.soup(1500, seed=3)
    Vg, Fg = geom.grid_mesh(6)
    return [
        ("sphere-random", Vs, Fs, geom.random_rays(6000, seed=1)),
        ("sphere-vertex", Vs, Fs, geom.aimed_rays(Vs, 4000, mode="vertex")),
        ("sphere-edge", Vs, Fs, geom.aimed_rays(Vs, 4000, mode="edge", F=Fs)),
        ("sphere-axis", Vs, Fs, geom.axis_rays(3000)),
        ("soup-random", Vsoup, Fsoup, geom.random_rays(6000, seed=2)),
        ("soup-window", Vsoup, Fsoup, geom.random_rays(4000, seed=4, tmin=0.5, tmax=3.0)),
        ("grid-vertex", Vg, Fg, geom.aimed_rays(Vg, 3000, mode="vertex")),
        ("grid-edges", Vg, Fg, geom.edge_case_rays()),
    ]


@pytest.mark.parametrize("case", range(8))
def test_bvh_traversal_equals_brute_force(case):
    name, V, F, r = _ray_sets()[case]
    bvh = orc.OracleBVH(V, F)
    got = bvh.closest(r, mode=0)
    ref = orc.brute_closest(V, F, r)
    for k in ("prim", "t", "u", "v", "flags"):
        assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), f"{name}: {k}"
    occ = orc.unpack_bits(bvh.any(r, mode=0), r.shape[1])
    assert np.array_equal(occ, ref["prim"] != orc.MISS), f"{name}: any-hit != (closest hit exists)"
    assert np.array_equal(occ, orc.unpack_bits(orc.brute_any(V, F, r), r.shape[1]))


def test_reference_faithful_mode_agrees_without_ties():
    V, F = geom.sphere_mesh(5)
    r = geom.random_rays(20000, seed=9)
    bvh = orc.OracleBVH(V, F)
    a = bvh.closest(r, mode=0)
    b = bvh.closest(r, mode=1)
    assert np.array_equal(a["prim"], b["prim"])
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))


def test_edge_case_rays_on_grid():
    V, F = geom.grid_mesh(4)
    r = geom.edge_case_rays()
    h = orc.brute_closest(V, F, r)
    miss = h["prim"] == orc.MISS
    assert miss[0] and miss[1] and miss[2] and miss[3]   # zero dir, NaN, empty window, tmax < tmin
    assert miss[4]                                        # hit exactly at tmax is excluded
    assert not miss[5] and h["t"][5] == 2.0               # hit exactly at tmin is included
    assert not miss[6] and not miss[7] and not miss[8] and not miss[9]
    assert miss[10]                                       # ray in the triangles' plane


def test_empty_scene():
    V = np.zeros((0, 3), np.float32)
    F = np.zeros((0, 3), np.uint32)
    bvh = orc.OracleBVH(V, F)
    assert bvh.count == 0
    h = bvh.closest(geom.random_rays(10))
    assert (h["prim"] == orc.MISS).all()
