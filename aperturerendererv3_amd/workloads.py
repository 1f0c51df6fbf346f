"""Deterministic synthetic workloads (numpy, seeded).

traversal_mesh / traversal_rays build the traversal microbenchmark: a
tetrahedron subdivided `levels` times (4 * 4**levels triangles; 9 levels =
1,048,576), projected to the unit sphere and displaced radially by
1 + amp * n(v), n = seeded 3-D value noise in [-1, 1] at frequency 8; and a
square grid of rays where slots with x+y even are pinhole primary rays from
(0, 0, 3) toward the origin (vertical fov 60 degrees, pixel centres) and slots
with x+y odd are incoherent rays: a uniform point on an area-weighted random
triangle, offset 1e-4 along the inward geometric normal, with a
cosine-weighted direction about that inward normal (so the ray crosses the
sphere's interior and traverses the whole tree).
"""
from __future__ import annotations

import math

import numpy as np

_TETRA_V = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=np.float64)
_TETRA_F = np.array([[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]], dtype=np.int64)


def _hash_u32(x: np.ndarray) -> np.ndarray:
    """32-bit avalanche hash (lowbias32 by C. Wellons), vectorised."""
    x = x.astype(np.uint32)
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7FEB352D)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846CA68B)
    x ^= x >> np.uint32(16)
    return x


def _lattice(ix, iy, iz, seed):
    h = _hash_u32(ix.astype(np.uint32) * np.uint32(73856093) ^ iy.astype(np.uint32) * np.uint32(19349663)
                  ^ iz.astype(np.uint32) * np.uint32(83492791) ^ np.uint32(seed * 0x9E3779B9 & 0xFFFFFFFF))
    return (h >> np.uint32(8)).astype(np.float64) * (2.0 / (1 << 24)) - 1.0


def value_noise(p: np.ndarray, seed: int) -> np.ndarray:
    """Trilinear value noise with smoothstep fade; p (n,3) in lattice units."""
    f = np.floor(p)
    i = f.astype(np.int64)
    t = p - f
    t = t * t * (3.0 - 2.0 * t)
    out = np.zeros(p.shape[0])
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = ((t[:, 0] if dx else 1 - t[:, 0]) * (t[:, 1] if dy else 1 - t[:, 1])
                     * (t[:, 2] if dz else 1 - t[:, 2]))
                out += w * _lattice(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
    return out


def subdivided_tetra(levels: int):
    """Unit-sphere geodesic tetrahedron: (V (nv,3) f64, F (nf,3) i64), outward winding."""
    V = _TETRA_V / np.sqrt(3.0)
    F = _TETRA_F.copy()
    for _ in range(levels):
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        lo, hi = np.minimum(e[:, 0], e[:, 1]), np.maximum(e[:, 0], e[:, 1])
        key = lo * (V.shape[0] + 1) + hi
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // (V.shape[0] + 1), uk % (V.shape[0] + 1)
        mid = V[a] + V[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = V.shape[0]
        V = np.concatenate([V, mid])
        m = F.shape[0]
        ab, bc, ca = base + inv[:m], base + inv[m:2 * m], base + inv[2 * m:]
        F = np.concatenate([
            np.stack([F[:, 0], ab, ca], 1), np.stack([F[:, 1], bc, ab], 1),
            np.stack([F[:, 2], ca, bc], 1), np.stack([ab, bc, ca], 1)])
    return V, F


def traversal_mesh(levels: int = 9, amp: float = 0.05, freq: float = 8.0, seed: int = 1):
    """Displaced subdivided tetrahedron: vertices (nv,3) f32, indices (nf,3) u32."""
    V, F = subdivided_tetra(levels)
    r = 1.0 + amp * value_noise(V * freq, seed)
    Vd = (V * r[:, None]).astype(np.float32)
    return Vd, F.astype(np.uint32)


def camera_dirs(width, height, fwd, vfov_deg=60.0, jitter=None):
    """Pinhole directions with the reference camera convention
    (core/scene.cu:145-164): right = norm(fwd x +Y), up = norm(fwd x right),
    ndc.y = (1-v)*2-1.  jitter: (h,w,2) offsets in [0,1) or None (centres)."""
    fwd = np.asarray(fwd, dtype=np.float64)
    fwd = fwd / np.linalg.norm(fwd)
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(fwd, right)
    up /= np.linalg.norm(up)
    jj, ii = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
    ox = 0.5 if jitter is None else jitter[..., 0]
    oy = 0.5 if jitter is None else jitter[..., 1]
    u = (ii + ox) / width
    v = (jj + oy) / height
    th = math.tan(math.radians(vfov_deg) * 0.5)
    sx = (u * 2 - 1) * th * (width / max(1.0, height))
    sy = ((1 - v) * 2 - 1) * th
    d = fwd[None, None, :] + right[None, None, :] * sx[..., None] + up[None, None, :] * sy[..., None]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return d


def _onb(n):
    """Branchless orthonormal basis (Duff et al., JCGT 6(1) 2017)."""
    s = np.where(n[:, 2] >= 0, 1.0, -1.0)
    a = -1.0 / (s + n[:, 2])
    b = n[:, 0] * n[:, 1] * a
    t = np.stack([1 + s * n[:, 0] ** 2 * a, s * b, -s * n[:, 0]], 1)
    bt = np.stack([b, s + n[:, 1] ** 2 * a, -n[:, 1]], 1)
    return t, bt


def traversal_rays(V, F, res: int = 4096, seed: int = 1, cam_pos=(0.0, 0.0, 3.0)):
    """Ray slots of a res x res grid as an (8, res*res) float32 array
    (rows ox, oy, oz, tmin, dx, dy, dz, tmax)."""
    n = res * res
    rng = np.random.Generator(np.random.PCG64(seed))
    rays = np.empty((8, n), dtype=np.float32)
    yy, xx = np.divmod(np.arange(n), res)
    primary = ((xx + yy) & 1) == 0
    cam = np.asarray(cam_pos, dtype=np.float64)
    d = camera_dirs(res, res, -cam).reshape(n, 3)
    pi = np.nonzero(primary)[0]
    rays[0:3, pi] = cam[:, None]
    rays[4:7, pi] = d[pi].T
    qi = np.nonzero(~primary)[0]
    m = qi.size
    a = V[F[:, 0]].astype(np.float64)
    e1 = V[F[:, 1]].astype(np.float64) - a
    e2 = V[F[:, 2]].astype(np.float64) - a
    cr = np.cross(e1, e2)
    area = np.linalg.norm(cr, axis=1)
    cdf = np.cumsum(area)
    cdf /= cdf[-1]
    tri = np.minimum(np.searchsorted(cdf, rng.random(m), side="right"), F.shape[0] - 1)
    r1, r2 = rng.random(m), rng.random(m)
    s1 = np.sqrt(r1)
    p = a[tri] + e1[tri] * (s1 * (1 - r2))[:, None] + e2[tri] * (s1 * r2)[:, None]
    nin = -cr[tri] / area[tri, None]
    p = p + 1e-4 * nin
    u1, u2 = rng.random(m), rng.random(m)
    rr = np.sqrt(u1)
    phi = 2 * math.pi * u2
    lx, ly, lz = rr * np.cos(phi), rr * np.sin(phi), np.sqrt(np.maximum(0.0, 1 - u1))
    t, b = _onb(nin)
    dd = t * lx[:, None] + b * ly[:, None] + nin * lz[:, None]
    dd /= np.linalg.norm(dd, axis=1, keepdims=True)
    rays[0:3, qi] = p.T
    rays[4:7, qi] = dd.T
    rays[3] = 0.0
    rays[7] = np.inf
    return rays


# --------------------------------------------------------------------------
# Cornell-box-like scene of 32 triangles (the plain-integrator workload)
# --------------------------------------------------------------------------

CORNELL_CAMERA = {"pos": (0.0, 0.0, 1.1), "dir": (0.0, 0.0, -1.0), "vfov": 60.0}


def _quad(V, F, M, corners, mat):
    """Append a quad (4 corners, counter-clockwise seen from the side the
    normal (b-a)x(c-a) points to) as two triangles."""
    base = len(V)
    V.extend(corners)
    F.extend([[base, base + 1, base + 2], [base, base + 2, base + 3]])
    M.extend([mat, mat])


def _box(V, F, M, center, size, yrot_deg, mat):
    """Five faces (no bottom) of an axis-aligned box rotated about +Y,
    outward winding."""
    hx, hy, hz = (s * 0.5 for s in size)
    c, s = math.cos(math.radians(yrot_deg)), math.sin(math.radians(yrot_deg))

    def P(x, y, z):
        return (center[0] + c * x + s * z, center[1] + y, center[2] - s * x + c * z)
    # top (+y), then +x, -x, +z, -z sides, each counter-clockwise from outside
    _quad(V, F, M, [P(-hx, hy, -hz), P(-hx, hy, hz), P(hx, hy, hz), P(hx, hy, -hz)], mat)
    _quad(V, F, M, [P(hx, -hy, -hz), P(hx, hy, -hz), P(hx, hy, hz), P(hx, -hy, hz)], mat)
    _quad(V, F, M, [P(-hx, -hy, -hz), P(-hx, -hy, hz), P(-hx, hy, hz), P(-hx, hy, -hz)], mat)
    _quad(V, F, M, [P(-hx, -hy, hz), P(hx, -hy, hz), P(hx, hy, hz), P(-hx, hy, hz)], mat)
    _quad(V, F, M, [P(-hx, -hy, -hz), P(-hx, hy, -hz), P(hx, hy, -hz), P(hx, -hy, -hz)], mat)


def cornell_box(light_le=(17.0, 12.0, 4.0)):
    """Unit box [-0.5, 0.5]^3 open at +z: floor, ceiling and back wall grey 0.75,
    left wall red (0.63, 0.065, 0.05), right wall green (0.14, 0.45, 0.09); a tall
    0.3x0.6x0.3 box turned +17 degrees and a 0.3^3 box turned -17 degrees (five
    faces each, grey 0.75); a 0.25 x 0.25 light quad facing down just below the
    ceiling.  All Lambert.  Returns (V f32 (n,3), F u32 (32,3), tri_material u32,
    materials as dicts of ar_material fields)."""
    V, F, M = [], [], []
    g, r, gr, lt = 0, 1, 2, 3
    h = 0.5
    _quad(V, F, M, [(-h, -h, -h), (-h, -h, h), (h, -h, h), (h, -h, -h)], g)      # floor, +y
    _quad(V, F, M, [(-h, h, -h), (h, h, -h), (h, h, h), (-h, h, h)], g)          # ceiling, -y
    _quad(V, F, M, [(-h, -h, -h), (h, -h, -h), (h, h, -h), (-h, h, -h)], g)      # back, +z
    _quad(V, F, M, [(-h, -h, -h), (-h, h, -h), (-h, h, h), (-h, -h, h)], r)      # left, +x
    _quad(V, F, M, [(h, -h, -h), (h, -h, h), (h, h, h), (h, h, -h)], gr)         # right, -x
    _box(V, F, M, (-0.16, -0.2, -0.14), (0.3, 0.6, 0.3), 17.0, g)
    _box(V, F, M, (0.17, -0.35, 0.12), (0.3, 0.3, 0.3), -17.0, g)
    ly, l = h - 1e-3, 0.125
    _quad(V, F, M, [(-l, ly, -l), (l, ly, -l), (l, ly, l), (-l, ly, l)], lt)      # light, -y
    mats = [
        {"albedo": (0.75, 0.75, 0.75)},
        {"albedo": (0.63, 0.065, 0.05)},
        {"albedo": (0.14, 0.45, 0.09)},
        {"albedo": (0.0, 0.0, 0.0), "emission": tuple(light_le)},
    ]
    return (np.asarray(V, np.float32), np.asarray(F, np.uint32), np.asarray(M, np.uint32), mats)


CORNELL_THICK = 0.01


def cornell_objects(light_le=(17.0, 12.0, 4.0)):
    """The Cornell box in analytic objects, as the reference builds it
    (CreateDiffuseCornellBox, core/helper/scene_helper.h:92-142): five slabs
    of half-thickness 0.01 outside the unit box (floor, ceiling and back grey
    0.75, left red (0.63, 0.065, 0.05), right green (0.14, 0.45, 0.09)), a box
    light of 0.25 x 0.25 footprint just under the ceiling, and the two boxes
    of cornell_box as rotated OBJ_CUBEs.  Returns a list of object dicts
    (Context.set_objects)."""
    t = CORNELL_THICK
    grey = {"albedo": (0.75, 0.75, 0.75)}
    return [
        {"type": "box", "center": (0.0, -0.5 - t, 0.0), "extent": (0.5, t, 0.5), "material": grey},
        {"type": "box", "center": (0.0, 0.5 + t, 0.0), "extent": (0.5, t, 0.5), "material": grey},
        {"type": "box", "center": (0.0, 0.5 - t, 0.0), "extent": (0.125, t, 0.125),
         "material": {"albedo": (0.0, 0.0, 0.0), "emission": tuple(light_le)}},
        {"type": "box", "center": (0.0, 0.0, -0.5 - t), "extent": (0.5, 0.5, t), "material": grey},
        {"type": "box", "center": (-0.5 - t, 0.0, 0.0), "extent": (t, 0.5, 0.5),
         "material": {"albedo": (0.63, 0.065, 0.05)}},
        {"type": "box", "center": (0.5 + t, 0.0, 0.0), "extent": (t, 0.5, 0.5),
         "material": {"albedo": (0.14, 0.45, 0.09)}},
        {"type": "box", "center": (-0.16, -0.2, -0.14), "extent": (0.15, 0.3, 0.15), "rotation": (0.0, 17.0, 0.0),
         "material": grey},
        {"type": "box", "center": (0.17, -0.35, 0.12), "extent": (0.15, 0.15, 0.15), "rotation": (0.0, -17.0, 0.0),
         "material": grey},
    ]


def random_object_scene(n_tris: int = 2000, n_boxes: int = 30, n_spheres: int = 30, seed: int = 1):
    """Random triangles (edges ~0.1) and rotated boxes / spheres (non-uniform
    extents) scattered in [-1, 1]^3: the traversal parity workload for
    analytic objects.  Returns (V f32, F u32, objects)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.uniform(-1.0, 1.0, (n_tris, 1, 3))
    V = (c + rng.normal(0.0, 0.06, (n_tris, 3, 3))).reshape(-1, 3).astype(np.float32)
    F = np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3)
    objects = []
    for k in range(n_boxes + n_spheres):
        objects.append({"type": "box" if k < n_boxes else "sphere",
                        "center": tuple(float(x) for x in rng.uniform(-1.0, 1.0, 3)),
                        "extent": tuple(float(x) for x in rng.uniform(0.03, 0.2, 3)),
                        "rotation": tuple(float(x) for x in rng.uniform(-180.0, 180.0, 3))})
    return V, F, objects


def _hsv(h, s=1.0, v=1.0):
    """HSV -> RGB (h in [0, 1))."""
    import colorsys
    return colorsys.hsv_to_rgb(float(h) % 1.0, s, v)


def glass_sdf_grid(kind="torus", half=2, scale=1.0, object_scale=1.0, position=(0.0, 0.0, 0.0), volume=-1,
                   seed=1):
    """The glass stress grids of the reference: CreateGlassTorus
    (core/helper/scene_helper.h:376-410, kind 'torus') and CreateGlassSDFVolume
    (:412-445, kind 'volume', `volume` = an SDF volume index): (2 half + 1)^3
    glass SDF objects on a lattice of spacing 0.25 scale / half around
    `position`, half extent 0.05 scale object_scale, random rotations, hue and
    the reference's IOR / absorption / dispersion ramps over the lattice
    coordinates.  The reference draws its random numbers with Hash11f; here a
    seeded numpy generator does.  Returns object dicts."""
    rng = np.random.Generator(np.random.PCG64(seed))
    objs = []
    for x in range(-half, half + 1):
        for y in range(-half, half + 1):
            for z in range(-half, half + 1):
                nx, ny, nz = ((v + half) / half * 0.5 for v in (x, y, z))
                loc = tuple(float(position[k] + v * 0.25 * scale / half) for k, v in enumerate((x, y, z)))
                rot = tuple(float(a) for a in rng.uniform(0.0, 360.0, 3))
                col = np.asarray(_hsv(rng.random()), np.float64)
                if kind == "torus":
                    col = col * (1 - nx * 0.7) + nx * 0.7
                    mat = {"albedo": tuple(col), "model": 2, "ior": 1.3 + 2.0 * ny, "sigma_a": 5.0 + nz * 50.0,
                           "abbe": 1.0, "tilt": 0.1}
                    ob = {"type": "torus", "sdf": (0.375, 0.125)}
                else:
                    col = col * (1 - nz * 0.7) + nz * 0.7
                    mat = {"albedo": tuple(col), "model": 2, "ior": 1.6 + 2.0 * ny, "sigma_a": 5.0 + nx * 25.0,
                           "abbe": 1.0, "tilt": 0.1}
                    ob = {"type": "volume", "volume": int(volume)}
                e = 0.05 * scale * object_scale
                ob.update({"center": loc, "extent": (e, e, e), "rotation": rot, "material": mat})
                objs.append(ob)
    return objs


def sdf_cornell(kind="torus", volume=-1, glass=True, half=1):
    """The object Cornell box (cornell_objects without its two boxes) holding
    a glass_sdf_grid of tori or volumes at 1.6x the reference's object size
    (so that the objects are several pixels wide at 32-64 px); glass=False
    gives the same grid in white Lambert."""
    objs = cornell_objects()[:6]
    grid = glass_sdf_grid(kind, half=half, scale=1.2, object_scale=1.6, position=(0.0, -0.1, 0.0), volume=volume)
    if not glass:
        for ob in grid:
            ob["material"] = {"albedo": (0.8, 0.8, 0.8)}
    return objs + grid


# --------------------------------------------------------------------------
# Spectral instanced scene (~500k triangles, 64 triangle lights, Lambert/GGX)
# --------------------------------------------------------------------------

SPECTRAL_CAMERA = {"pos": (0.0, 0.0, 30.0), "dir": (0.0, 0.0, -1.0), "vfov": 60.0}


def _subdivide_sphere(V, F, levels):
    """Loop-style midpoint subdivision projected to the unit sphere."""
    for _ in range(levels):
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        lo, hi = np.minimum(e[:, 0], e[:, 1]), np.maximum(e[:, 0], e[:, 1])
        key = lo * (V.shape[0] + 1) + hi
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // (V.shape[0] + 1), uk % (V.shape[0] + 1)
        mid = V[a] + V[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = V.shape[0]
        V = np.concatenate([V, mid])
        m = F.shape[0]
        ab, bc, ca = base + inv[:m], base + inv[m:2 * m], base + inv[2 * m:]
        F = np.concatenate([
            np.stack([F[:, 0], ab, ca], 1), np.stack([F[:, 1], bc, ab], 1),
            np.stack([F[:, 2], ca, bc], 1), np.stack([ab, bc, ca], 1)])
    return V, F


def icosphere(levels: int = 3):
    """Unit icosphere with outward winding: 20 * 4**levels triangles
    (1,280 at 3 levels)."""
    phi = (1.0 + math.sqrt(5.0)) / 2.0
    V = []
    for a in (-1.0, 1.0):
        for b in (-phi, phi):
            V += [(a, b, 0.0), (0.0, a, b), (b, 0.0, a)]
    V = np.asarray(V)
    n = len(V)
    d = np.linalg.norm(V[:, None] - V[None], axis=-1)
    F = []
    for i in range(n):
        for j in range(i + 1, n):
            for k in range(j + 1, n):
                if abs(d[i, j] - 2) < 1e-9 and abs(d[j, k] - 2) < 1e-9 and abs(d[i, k] - 2) < 1e-9:
                    nrm = np.cross(V[j] - V[i], V[k] - V[i])
                    F.append((i, j, k) if nrm @ (V[i] + V[j] + V[k]) > 0 else (i, k, j))
    V = V / np.linalg.norm(V, axis=1, keepdims=True)
    return _subdivide_sphere(V, np.asarray(F, np.int64), levels)


_BOX_V = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)])
_BOX_F = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1],
                   [2, 3, 7], [2, 7, 6], [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]], np.int64)


def spectral_scene(target_tris: int = 500_000, nlights: int = 64, seed: int = 1, sphere_levels: int = 3):
    """Instances of displaced icospheres (radial 1 + 0.1 n(p), value noise at
    frequency 4 with a per-instance offset) and boxes (random turn about +Y),
    scale U[0.2, 1], centres U[-10, 10]^3, one kind or the other with equal
    probability, added until `target_tris` is reached (within one sphere);
    above a 100 x 100 two-triangle floor at y = -10 (grey 0.5 Lambert).
    Each instance has its own material: Lambert or GGX with equal probability,
    albedo U[0.05, 0.95]^3, roughness U[0.05, 1], metallic 0, specular 0.5.
    `nlights` distinct instance triangles become white emitters with
    Le = min(U^(-1/1.5), 1000), each its own material (its instance's
    material plus the emission).
    Returns (V f32 (n,3), F u32 (m,3), tri_material u32 (m,), materials)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    SV, SF = icosphere(sphere_levels)
    BV = _BOX_V
    parts_v, parts_f, parts_m = [], [], []
    mats = []
    nv = 0
    ntri = 0
    inst = 0
    while ntri < target_tris:
        sphere = rng.random() < 0.5
        s = 0.2 + 0.8 * rng.random()
        c = -10.0 + 20.0 * rng.random(3)
        if sphere:
            off = rng.random(3) * 1000.0
            r = 1.0 + 0.1 * value_noise(SV * 4.0 + off, seed + 7)
            v = SV * (r * s)[:, None] + c
            f = SF
        else:
            a = 2.0 * math.pi * rng.random()
            ca, sa = math.cos(a), math.sin(a)
            R = np.array([[ca, 0.0, sa], [0.0, 1.0, 0.0], [-sa, 0.0, ca]])
            v = (BV * s) @ R.T + c
            f = _BOX_F
        ggx = rng.random() < 0.5
        alb = tuple(float(x) for x in 0.05 + 0.9 * rng.random(3))
        rough = float(0.05 + 0.95 * rng.random())
        mats.append({"albedo": alb, "model": 1 if ggx else 0, "roughness": rough, "metallic": 0.0, "specular": 0.5})
        parts_v.append(v)
        parts_f.append(f + nv)
        parts_m.append(np.full(f.shape[0], inst, np.uint32))
        nv += v.shape[0]
        ntri += f.shape[0]
        inst += 1
    L = 50.0
    fl = len(mats)
    mats.append({"albedo": (0.5, 0.5, 0.5)})
    parts_v.append(np.array([[-L, -10.0, -L], [-L, -10.0, L], [L, -10.0, L], [L, -10.0, -L]]))
    parts_f.append(np.array([[0, 1, 2], [0, 2, 3]], np.int64) + nv)
    parts_m.append(np.full(2, fl, np.uint32))
    V = np.concatenate(parts_v).astype(np.float32)
    F = np.concatenate(parts_f).astype(np.uint32)
    M = np.concatenate(parts_m)
    lights = rng.choice(ntri, size=nlights, replace=False)
    le = np.minimum(rng.random(nlights) ** (-1.0 / 1.5), 1000.0)
    for k, t in enumerate(lights):
        m = dict(mats[M[t]])
        m["emission"] = (float(le[k]),) * 3
        M[t] = len(mats)
        mats.append(m)
    return V, F, M, mats


# --------------------------------------------------------------------------
# Occluded-light interior: a closed room lit only through a door gap
# --------------------------------------------------------------------------

OCCLUDED_CAMERA = {"pos": (1.2, 1.4, 1.5), "dir": (-0.8, -0.25, -0.6), "vfov": 70.0}


def _grid_quad(origin, eu, ev, nu, nv, keep=None):
    """Quad origin + s*eu + t*ev split into nu x nv cells of two triangles,
    wound so the normal is eu x ev; keep(s, t) -> bool drops cells (holes)."""
    o, eu, ev = (np.asarray(a, np.float64) for a in (origin, eu, ev))
    s = np.linspace(0, 1, nu + 1)
    t = np.linspace(0, 1, nv + 1)
    P = o + s[:, None, None] * eu + t[None, :, None] * ev
    V = P.reshape(-1, 3)
    idx = lambda i, j: i * (nv + 1) + j
    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    if keep is not None:
        m = keep((ii + 0.5) / nu, (jj + 0.5) / nv)
        ii, jj = ii[m], jj[m]
    a, b, c, d = idx(ii, jj), idx(ii + 1, jj), idx(ii + 1, jj + 1), idx(ii, jj + 1)
    F = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    return V, F


def occluded_room(target_tris: int = 200_000, le: float = 40.0):
    """A 4 x 3 x 4 m room ([-2, 2] x [0, 3] x [-2, 2], all walls facing
    inward, grey 0.6 Lambert, the floor a GGX 0.3) whose +x wall has a
    0.5 m wide, 2 m high door gap (z in [-0.25, 0.25]); the only light is a
    2 x 2 m emitter (radiance `le`, white) outside at x = 3.5 facing the door,
    so the room is lit by light through the gap.  Walls are tessellated so
    that the scene has about `target_tris` triangles."""
    per = max(1, int(math.sqrt(target_tris / 12.0)))
    parts = []
    W, H = 4.0, 3.0
    # floor (+y), ceiling (-y), walls facing inward
    parts.append((_grid_quad((-2, 0, -2), (0, 0, W), (W, 0, 0), per, per), 1))
    parts.append((_grid_quad((-2, H, -2), (W, 0, 0), (0, 0, W), per, per), 0))
    parts.append((_grid_quad((-2, 0, -2), (0, H, 0), (0, 0, W), per, per), 0))            # x = -2, normal +x
    parts.append((_grid_quad((-2, 0, -2), (W, 0, 0), (0, H, 0), per, per), 0))            # z = -2, normal +z
    parts.append((_grid_quad((-2, 0, 2), (0, H, 0), (W, 0, 0), per, per), 0))             # z = +2, normal -z
    door = lambda s, t: ~((np.abs(s * W - 2.0) < 0.25) & (t * H < 2.0))
    parts.append((_grid_quad((2, 0, -2), (0, 0, W), (0, H, 0), per, per, keep=door), 0))  # x = +2, normal -x
    parts.append((_grid_quad((3.5, 0.0, -1.0), (0, 0, 2), (0, 2, 0), 1, 1), 2))           # light at x = 3.5, normal -x
    Vs, Fs, Ms = [], [], []
    nv = 0
    for (V, F), m in parts:
        Vs.append(V)
        Fs.append(F + nv)
        Ms.append(np.full(F.shape[0], m, np.uint32))
        nv += V.shape[0]
    mats = [{"albedo": (0.6, 0.6, 0.6)},
            {"albedo": (0.5, 0.45, 0.4), "model": 1, "roughness": 0.3, "specular": 0.5, "metallic": 0.0},
            {"albedo": (0.0, 0.0, 0.0), "emission": (le, le, le)}]
    return (np.concatenate(Vs).astype(np.float32), np.concatenate(Fs).astype(np.uint32), np.concatenate(Ms), mats)


# --------------------------------------------------------------------------
# Batched workload: many scenes of uneven size, resolution and spp
# --------------------------------------------------------------------------

BATCH_RESOLUTIONS = ((256, 256), (512, 512), (1024, 1024), (1920, 1080))
BATCH_SPP = (4, 16, 64)


def batch_specs(count: int = 48, seed: int = 1, tri_range=(1_000, 2_000_000)):
    """Per scene k (seed + k): triangle count log-uniform in tri_range,
    resolution uniform in BATCH_RESOLUTIONS, spp uniform in BATCH_SPP.
    Returns a list of dicts (seed, tris, width, height, spp)."""
    out = []
    lo, hi = math.log(tri_range[0]), math.log(tri_range[1])
    for k in range(count):
        rng = np.random.Generator(np.random.PCG64(seed + k))
        tris = int(round(math.exp(lo + (hi - lo) * rng.random())))
        w, h = BATCH_RESOLUTIONS[int(rng.integers(len(BATCH_RESOLUTIONS)))]
        spp = BATCH_SPP[int(rng.integers(len(BATCH_SPP)))]
        out.append({"seed": seed + k, "tris": tris, "width": w, "height": h, "spp": spp})
    return out


def batch_scene(spec):
    """The instanced spectral-scene generator at the spec's size and seed."""
    return spectral_scene(target_tris=max(spec["tris"], 200), seed=spec["seed"])
