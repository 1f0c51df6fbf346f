"""Oracle signed-distance objects (torus, baked distance volume) pinned
against closed forms and library routines.

The oracle restates core/object/sdf (RaymarchLocal, IntersectSDFObject,
sdfNormalUnit, SdfTorus, SampleSDFAdaptive01, FillSDFVolume, SdfDiamondUnit).
Each check would fail for a plausible slip:
  - the march against the torus's exact ray intersection (roots of its
    quartic in FP64): a wrong unit-space scale, a dropped sMin factor or a
    wrong inside/outside rule moves or loses hits;
  - the normal against the torus's analytic gradient: a swapped axis or a
    missing UnitNormalToLocal scale tilts it;
  - the baked grid against the FP64 torus at the grid points, and the volume
    lookup against scipy's trilinear map_coordinates: x/z transposed or an
    off-by-one grid mapping shifts the values;
  - the volume object's hits against the exact torus: the whole volume path;
  - the diamond's 8-fold symmetry and its table plane value;
  - BVH traversal of tori, volumes and triangles against brute force."""
import math

import numpy as np
import pytest

import oracle_lib as orc


def _w2o(obj):
    return np.array(obj.w2o[:], np.float64).reshape(3, 3)


def _torus_roots(lo, ld, e, R, r):
    """Real roots t of the torus (|u.xz| - R)^2 + u.y^2 = r^2 along u(t) =
    (lo + ld t) / (2 e), in FP64, ascending."""
    P = np.polynomial.Polynomial
    ux, uy, uz = (P([lo[k] / (2 * e[k]), ld[k] / (2 * e[k])]) for k in range(3))
    s = ux * ux + uy * uy + uz * uz + (R * R - r * r)
    q = s * s - 4 * R * R * (ux * ux + uz * uz)
    rts = q.roots()
    real = np.sort(rts[np.abs(rts.imag) < 1e-9].real)
    return real


def _torus_f(u, R, r):
    return math.hypot(math.hypot(u[0], u[2]) - R, u[1]) - r


def test_torus_march_matches_exact_intersection():
    rng = np.random.default_rng(11)
    R, r, ext = 0.375, 0.125, 0.4
    checked = inside_checked = misses = 0
    for k in range(300):
        rot = rng.uniform(-180, 180, 3)
        c = rng.uniform(-1, 1, 3)
        obj = {"type": "torus", "center": c, "extent": (ext, ext, ext), "rotation": rot}
        arr = orc.make_objects([obj])
        W = _w2o(arr[0])
        start_inside = k % 4 == 0
        if start_inside:
            # origin on the core circle of the torus, in object space
            phi = rng.uniform(0, 2 * math.pi)
            lo0 = np.array([math.cos(phi) * R, 0.0, math.sin(phi) * R]) * 2 * ext
            o = W.T @ lo0 + c
        else:
            o = c + rng.normal(size=3) * 1.5
        target = c + W.T @ (rng.uniform(-0.5, 0.5, 3) * 2 * ext)
        d = target - o
        d /= np.linalg.norm(d)
        meps = 5e-4
        hit, t, ins = orc.object_test(arr[0], o.astype(np.float32), d.astype(np.float32), meps)
        lo = W @ (o.astype(np.float32).astype(np.float64) - c.astype(np.float32))
        ld = W @ d.astype(np.float32)
        ld /= np.linalg.norm(ld)
        roots = _torus_roots(lo, ld, (ext, ext, ext), R, r)
        fwd = roots[roots > 1e-6]
        u0 = lo / (2 * ext)
        assert bool(ins) == (_torus_f(u0, R, r) <= 0) or abs(_torus_f(u0, R, r)) < 1e-6
        if len(fwd) == 0:
            assert not hit, (k, t)
            misses += 1
            continue
        t_ex = fwd[0]
        # angle between the ray and the surface normal at the exact hit
        p = lo + ld * t_ex
        rho = math.hypot(p[0], p[2])
        g = np.array([p[0] * (1 - R * 2 * ext / rho), p[1], p[2] * (1 - R * 2 * ext / rho)])
        cos = abs(g @ ld) / np.linalg.norm(g)
        if cos < 0.3:
            continue   # grazing: a shallow crossing can be stepped over (min step eps / 2)
        assert hit, (k, t_ex)
        # the march stops just past the surface: depth below eps, so t within eps / cos
        assert t_ex - 1e-5 <= t <= t_ex + meps / cos + 1e-5, (k, t, t_ex, cos)
        checked += 1
        inside_checked += start_inside
    assert checked > 120 and inside_checked > 20 and misses > 10


def test_torus_normal_matches_gradient():
    rng = np.random.default_rng(12)
    R, r = 0.375, 0.125
    n_ok = 0
    for k in range(200):
        ext = rng.uniform(0.2, 0.6)
        obj = {"type": "torus", "center": rng.uniform(-1, 1, 3), "extent": (ext, ext, ext),
               "rotation": rng.uniform(-180, 180, 3)}
        arr = orc.make_objects([obj])
        W = _w2o(arr[0])
        c = np.asarray(obj["center"], np.float32).astype(np.float64)
        o = (c + rng.normal(size=3) * 2).astype(np.float32)
        d = (c + W.T @ (rng.uniform(-0.3, 0.3, 3) * 2 * ext) - o).astype(np.float32)
        d /= np.linalg.norm(d)
        hit, t, ins = orc.object_test(arr[0], o, d)
        if not hit:
            continue
        pos, n = orc.object_surface(arr[0], o, d, t, ins)
        lp = W @ (pos.astype(np.float64) - c)
        rho = math.hypot(lp[0], lp[2])
        g = np.array([lp[0] * (1 - R * 2 * ext / rho), lp[1], lp[2] * (1 - R * 2 * ext / rho)])
        gw = W.T @ (g / np.linalg.norm(g))
        assert abs(np.linalg.norm(n) - 1) < 1e-5
        assert gw @ n > 1 - 2e-5, (k, gw, n)
        n_ok += 1
    assert n_ok > 100


def _fp64_torus_grid(dims, R, r):
    nx, ny, nz = dims
    x = np.arange(nx) / (nx - 1) - 0.5
    y = np.arange(ny) / (ny - 1) - 0.5
    z = np.arange(nz) / (nz - 1) - 0.5
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return np.hypot(np.hypot(X, Z) - R, Y) - r


def test_bake_matches_fp64_torus_and_lookup_matches_scipy():
    from scipy.ndimage import map_coordinates
    dims = (40, 33, 27)           # distinct sizes expose a transposed index
    vol = orc.sdf_bake("torus", (0.3, 0.1), dims)
    assert vol.shape == (27, 33, 40)
    np.testing.assert_allclose(vol, _fp64_torus_grid(dims, 0.3, 0.1), atol=2e-7)
    # a deliberately asymmetric volume for the lookup
    rng = np.random.default_rng(13)
    vol = rng.normal(size=(9, 7, 5)).astype(np.float32)
    arr = orc.make_objects([{"type": "volume", "center": (0, 0, 0), "extent": (1, 1, 1), "volume_data": vol}])
    vmax = math.sqrt(1 / 25 + 1 / 49 + 1 / 81) * 2.5
    assert arr[0].vmax == pytest.approx(vmax, rel=1e-6)
    n_lin = n_near = 0
    for _ in range(3000):
        u = rng.uniform(-0.55, 0.55, 3).astype(np.float32)
        got = orc.sdf_value(arr[0], u)
        g = (u.astype(np.float64) + 0.5) * (np.array([5, 7, 9]) - 1)
        gi = np.clip(np.floor(g + 0.5).astype(int), 0, [4, 6, 8])
        near = vol[gi[2], gi[1], gi[0]]
        if abs(near) > vmax:
            assert got == near
            n_near += 1
        else:
            gc = np.clip(g, 0, [4, 6, 8])
            lin = map_coordinates(vol.astype(np.float64), [[gc[2]], [gc[1]], [gc[0]]], order=1, mode="nearest")[0]
            assert got == pytest.approx(lin, abs=2e-6)
            n_lin += 1
    assert n_lin > 500 and n_near > 500


def test_volume_object_hits_the_torus():
    R, r, ext = 0.375, 0.125, 0.5
    vol = orc.sdf_bake("torus", (R, r), (64, 64, 64))
    rng = np.random.default_rng(14)
    objs = [{"type": "volume", "center": (0.1, -0.2, 0.3), "extent": (ext, ext, ext), "rotation": (20, -35, 50),
             "volume_data": vol}]
    arr = orc.make_objects(objs)
    W = _w2o(arr[0])
    c = np.array([0.1, -0.2, 0.3], np.float32).astype(np.float64)
    checked = 0
    for _ in range(300):
        o = (c + rng.normal(size=3) * 2).astype(np.float32)
        d = (c + W.T @ (rng.uniform(-0.4, 0.4, 3) * 2 * ext) - o).astype(np.float32)
        d /= np.linalg.norm(d)
        hit, t, _ = orc.object_test(arr[0], o, d)
        lo = W @ (o.astype(np.float64) - c)
        ld = W @ d.astype(np.float64)
        ld /= np.linalg.norm(ld)
        fwd = _torus_roots(lo, ld, (ext, ext, ext), R, r)
        fwd = fwd[fwd > 0]
        if len(fwd) == 0:
            continue
        p = lo + ld * fwd[0]
        rho = math.hypot(p[0], p[2])
        g = np.array([p[0] * (1 - R * 2 * ext / rho), p[1], p[2] * (1 - R * 2 * ext / rho)])
        if abs(g @ ld) / np.linalg.norm(g) < 0.3:
            continue
        assert hit
        # trilinear error of a 64^3 grid: well under a voxel (1/63 of the box side)
        assert abs(t - fwd[0]) < 0.25 * (2 * ext) / 63, (t, fwd[0])
        checked += 1
    assert checked > 100


def test_diamond_bake_symmetry_and_table():
    dims = (48, 48, 48)
    vol = orc.sdf_bake("diamond", (0.0, 0.0, 0.0, 0.5), dims)
    # CreateDiamond(scale 1) = (d 0.94, bias 0.5, table 0.3, scale 0.5): inside at the centre, outside at a corner
    assert vol[24, 24, 24] < 0 and vol[0, 0, 0] > 0
    # 8-fold symmetry about z: a quarter turn of the grid (a multiple of pi / 4) maps it onto itself
    rot = np.rot90(vol, k=1, axes=(1, 2))
    assert (np.abs(rot - vol) < 1e-5).mean() > 0.99
    # no mirror symmetry z -> -z (crown vs pavilion)
    assert np.abs(vol[::-1] - vol).max() > 0.05
    # a column along the axis: p01 = (0.5, 0.5, z) with dims (3, 3, n); above the
    # crown the table plane dominates, value (2 u.z - 0.3) * 0.5
    n = 101
    col = orc.sdf_bake("diamond", (0.0, 0.0, 0.0, 0.5), (3, 3, n))[:, 1, 1]
    z = np.arange(n) / (n - 1) - 0.5
    on_table = (z > 0.16) & (z < 0.45)
    np.testing.assert_allclose(col[on_table], (2 * z[on_table] - 0.3) * 0.5, atol=2e-6)


def test_overlap_of_sdf_objects():
    ext = 0.5
    arr = orc.make_objects([{"type": "torus", "center": (1, 2, 3), "extent": (ext, ext, ext),
                             "rotation": (0, 0, 90)}])
    ov = orc.fn("orc_object_overlap", orc.C.c_int, [orc.C.c_void_p, orc.C.c_void_p])
    W = _w2o(arr[0])
    core = np.array([0.375, 0.0, 0.0]) * 2 * ext
    for p_local, expect in ((np.zeros(3), 0), (core, 1), (core * 1.2, 1), (core * 2, 0)):
        p = (W.T @ p_local + np.array([1, 2, 3])).astype(np.float32)
        assert ov(orc.C.byref(arr[0]), p.ctypes.data) == expect, p_local


def test_traversal_with_sdf_objects_equals_brute_force():
    rng = np.random.default_rng(15)
    vol = orc.sdf_bake("torus", (0.35, 0.12), (32, 32, 32))
    objs = []
    for k in range(24):
        kind = ("torus", "volume", "box", "sphere")[k % 4]
        ob = {"type": kind, "center": rng.uniform(-1.5, 1.5, 3), "extent": rng.uniform(0.1, 0.4, 3),
              "rotation": rng.uniform(-180, 180, 3)}
        if kind == "volume":
            ob["volume_data"] = vol
        if kind == "torus":
            ob["sdf"] = (rng.uniform(0.25, 0.4), rng.uniform(0.05, 0.1))
        objs.append(ob)
    V = rng.uniform(-2, 2, (60, 3)).astype(np.float32)
    F = rng.integers(0, 60, (40, 3)).astype(np.uint32)
    g = orc.Geom(V, F, objs, leaf_size=3)
    n = 8000
    o = rng.uniform(-2.5, 2.5, (n, 3))
    tgt = np.array([ob["center"] for ob in objs])[rng.integers(0, len(objs), n)] + rng.normal(size=(n, 3)) * 0.2
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((9, n), np.float32)
    rays[0:3] = o.T
    rays[3] = np.where(np.arange(n) % 3 == 0, 0.0, 1e-3)
    rays[4:7] = d.T
    rays[7] = np.inf
    rays[8] = rng.choice([5e-4, 1e-3, 4e-3], n)
    a, b = g.closest(rays), g.brute(rays)
    for k in ("prim", "t", "flags"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    hit = a["prim"] != 0xFFFFFFFF
    prim_obj = a["prim"][hit].astype(np.int64) - F.shape[0]
    kinds = np.array([k % 4 for k in range(len(objs))])
    hit_kinds = kinds[prim_obj[prim_obj >= 0]]
    assert (hit_kinds == 0).sum() > 50 and (hit_kinds == 1).sum() > 50
    rays[7] = np.where(a["prim"] < 0xFFFFFFFF, a["t"], 1e3)
    np.testing.assert_array_equal(g.any(rays), g.any_brute(rays))
