"""Aperture MI355X path-tracing core: Python host mirror of include/ar.h.

The compute path is libar_hip.so (hand-written HIP kernels for gfx950 behind
a C ABI).  This module only marshals arguments: device memory comes from
torch tensors on a ROCm device, and every call goes through the C ABI.  If
the library is not built, or no GPU is present, calls raise; nothing here
computes a result on the CPU.

Naming follows the reference's host API (SceneRenderer / Camera in
core/scene.hpp:18-135 and core/camera.hpp:14-99): a Context owns the scene,
its BVH and the device state.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (AR_OBJ_SDF_TORUS, AR_OBJ_SDF_VOLUME, AR_TONE_ACES, AR_TONE_GAMMA, AR_TONE_LINEAR, AR_HIT_INSIDE, AR_MISS, AR_MODEL_GGX, AR_MODEL_GLASS, AR_MODEL_LAMBERT, AR_OBJ_BOX, AR_OBJ_SPHERE, AR_OK, AR_SKY_CONSTANT, AR_SKY_GRADIENT, AR_SKY_NONE,
                   ar_bvh_info, ar_bvh_opts, ar_camera, ar_device_opts, ar_hits, ar_material, ar_mesh, ar_rays,
                   ar_render_opts)

__all__ = ["Context", "ArError", "write_bmp", "AR_TONE_LINEAR", "AR_TONE_GAMMA", "AR_TONE_ACES", "AR_MISS", "AR_MODEL_LAMBERT", "AR_MODEL_GGX", "AR_MODEL_GLASS", "AR_OBJ_BOX",
           "AR_OBJ_SPHERE", "AR_HIT_INSIDE", "AR_SKY_NONE", "AR_SKY_GRADIENT",
           "AR_SKY_CONSTANT", "Material", "RAY_FIELDS"]

RAY_FIELDS = ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")


class ArError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Material:
    """Material of a triangle or analytic object (the reference GBuffer);
    glass parameters default to the reference's (IOR 1.5, sigmaA 100, Abbe
    20, tilt 0)."""

    def __init__(self, albedo=(0.75, 0.75, 0.75), emission=(0.0, 0.0, 0.0), specular=0.5, metallic=0.0,
                 roughness=1.0, model=AR_MODEL_LAMBERT, ior=1.5, sigma_a=100.0, abbe=20.0, tilt=0.0):
        self.albedo = tuple(float(x) for x in albedo)
        self.emission = tuple(float(x) for x in emission)
        self.specular = float(specular)
        self.metallic = float(metallic)
        self.roughness = float(roughness)
        self.model = int(model)
        self.ior, self.sigma_a, self.abbe, self.tilt = float(ior), float(sigma_a), float(abbe), float(tilt)

    def to_c(self) -> ar_material:
        m = ar_material()
        m.albedo[:] = self.albedo
        m.emission[:] = self.emission
        m.specular, m.metallic, m.roughness, m.model = self.specular, self.metallic, self.roughness, self.model
        m.ior, m.sigma_a, m.abbe, m.tilt = self.ior, self.sigma_a, self.abbe, self.tilt
        return m


_OBJ_TYPES = {"box": AR_OBJ_BOX, "sphere": AR_OBJ_SPHERE, "torus": _lib.AR_OBJ_SDF_TORUS,
              "sdf_torus": _lib.AR_OBJ_SDF_TORUS, "volume": _lib.AR_OBJ_SDF_VOLUME,
              "sdf_volume": _lib.AR_OBJ_SDF_VOLUME}
_OBJ_TYPES.update({v: v for v in list(_OBJ_TYPES.values())})
TORUS_DEFAULT = (0.375, 0.125)   # CreateTorus(scale 1): ring and tube radius in unit space


def object_array(objects):
    """ctypes ar_object array from dicts: type ('box' / 'sphere' / 'torus' /
    'volume'), center, extent (half extents; sphere radius = max), rotation
    (Euler degrees, objectToWorld = Rz Ry Rx), material (Material or its
    keyword dict); a torus's 'sdf' radii (default 0.375, 0.125), a volume's
    'volume' index."""
    arr = (_lib.ar_object * max(1, len(objects)))()
    for k, o in enumerate(objects):
        arr[k].type = _OBJ_TYPES[o.get("type", "box")]
        sdf = list(o.get("sdf", TORUS_DEFAULT if arr[k].type == _lib.AR_OBJ_SDF_TORUS else ()))
        arr[k].sdf[:] = [float(x) for x in (sdf + [0.0] * 4)[:4]]
        arr[k].sdf_volume = int(o.get("volume", -1))
        arr[k].center[:] = [float(x) for x in o["center"]]
        arr[k].extent[:] = [float(x) for x in o["extent"]]
        arr[k].rotation[:] = [float(x) for x in o.get("rotation", (0.0, 0.0, 0.0))]
        m = o.get("material") or Material()
        arr[k].material = (m if isinstance(m, Material) else Material(**m)).to_c()
    return arr


def _ptr(t) -> int:
    return int(t.data_ptr())


def _stream_ptr(stream) -> int | None:
    if stream is None:
        import torch
        return int(torch.cuda.current_stream().cuda_stream)
    return int(stream.cuda_stream)


def _camera_opts(width, height, spp, cam_pos=(0.0, 0.0, 1.1), cam_dir=(0.0, 0.0, -1.0), vfov=60.0, seed=1,
                 max_depth=5, nee=True, sky=AR_SKY_NONE, env=(0.0, 0.0, 0.0), rr_min_depth=3, sample_offset=0,
                 max_paths_per_pass=0, spectral=0, output=0, guiding=0, frame_spp=0, frame_growth=0):
    cam = ar_camera()
    cam.pos[:] = [float(x) for x in cam_pos]
    cam.dir[:] = [float(x) for x in cam_dir]
    cam.vfov_deg = float(vfov)
    o = ar_render_opts()
    o.width, o.height, o.spp, o.sample_offset = width, height, spp, sample_offset
    o.seed, o.max_depth, o.rr_min_depth, o.nee, o.sky = seed, max_depth, rr_min_depth, int(bool(nee)), sky
    o.env[:] = [float(x) for x in env]
    o.max_paths_per_pass = max_paths_per_pass
    o.spectral, o.output, o.guiding, o.frame_spp = spectral, output, guiding, frame_spp
    o.frame_growth = frame_growth
    return cam, o


def _view_arrays(views, shared):
    """views: list of dicts (width, height, spp, cam_pos, cam_dir, vfov, sample_offset)."""
    n = len(views)
    cams = (ar_camera * n)()
    opts = (ar_render_opts * n)()
    for i, v in enumerate(views):
        kw = dict(shared)
        kw.update(v)
        c, o = _camera_opts(**kw)
        cams[i] = c
        opts[i] = o
    return cams, opts


class Graph:
    """A captured render (ar_graph): launch() replays it, writing `outputs`."""

    def __init__(self, lib, handle, outputs):
        self._lib, self._g, self.outputs = lib, handle, outputs

    def launch(self, stream=None):
        rc = self._lib.ar_graph_launch(self._g, _stream_ptr(stream))
        if rc != AR_OK:
            raise ArError(rc, "ar_graph_launch")

    def close(self):
        if getattr(self, "_g", None):
            self._lib.ar_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spectral_mlp_init(seed=0):
    """Initial parameters of the spectral basis network as the reference's
    trainer sets them (pytorch/spectrum/train_spectrum.py:103-160): Kaiming
    normal (fan in, ReLU gain: std sqrt(2 / 32)) for the RGB input layer and
    every block, zero biases, and torch's default for the two output layers
    (uniform on +-1 / sqrt(32)).  AR_SMLP_PARAMS float32 values."""
    import numpy as np
    rng = np.random.default_rng(seed)
    w = np.zeros(_lib.SMLP_PARAMS, np.float32)
    std = np.sqrt(2.0 / 32.0)
    w[0:1024] = rng.normal(0.0, std, 1024)
    for k in range(7):
        w[1056 + k * 1056:1056 + k * 1056 + 1024] = rng.normal(0.0, std, 1024)
        w[8576 + k * 1056:8576 + k * 1056 + 1024] = rng.normal(0.0, std, 1024)
    bound = 1.0 / np.sqrt(32.0)
    w[8448:8576] = rng.uniform(-bound, bound, 128)
    w[15968:16096] = rng.uniform(-bound, bound, 128)
    return w


class Context:
    """One device context (ar_ctx).  Thread-compatible, not thread-safe."""

    @classmethod
    def _adopt(cls, handle, device):
        c = cls.__new__(cls)
        c._lib = _lib.load()
        c._h = handle
        c.device = device
        c._keep = []
        return c

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        h = C.c_void_p()
        opts = ar_device_opts(device, 0)
        rc = self._lib.ar_create(C.byref(h), C.byref(opts))
        if rc != AR_OK:
            raise ArError(rc, f"ar_create(device={device}) failed")
        self._h = h
        self.device = device
        self._keep = []

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.ar_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int, what: str):
        if rc != AR_OK:
            msg = self._lib.ar_last_error(self._h)
            raise ArError(rc, f"{what}: {msg.decode() if msg else ''}")

    # -- scene -------------------------------------------------------------
    def set_meshes(self, meshes):
        """meshes: list of dicts with 'vertices' (nv,3) f32, 'indices' (ntri,3) u32,
        optional 'tri_material' (ntri,) u32 and 'materials' (list of Material)."""
        cms = (ar_mesh * max(1, len(meshes)))()
        keep = []
        for i, m in enumerate(meshes):
            v = np.ascontiguousarray(m["vertices"], dtype=np.float32).reshape(-1, 3)
            ix = np.ascontiguousarray(m["indices"], dtype=np.uint32).reshape(-1, 3)
            keep += [v, ix]
            cms[i].vertices = v.ctypes.data
            cms[i].nv = v.shape[0]
            cms[i].indices = ix.ctypes.data
            cms[i].ntri = ix.shape[0]
            tm = m.get("tri_material")
            if tm is not None:
                tm = np.ascontiguousarray(tm, dtype=np.uint32)
                keep.append(tm)
                cms[i].tri_material = tm.ctypes.data
            mats = [mm if isinstance(mm, Material) else Material(**mm) for mm in (m.get("materials") or [])]
            if mats:
                arr = (ar_material * len(mats))(*[mm.to_c() for mm in mats])
                keep.append(arr)
                cms[i].materials = C.cast(arr, C.c_void_p)
                cms[i].nmaterials = len(mats)
        self._check(self._lib.ar_scene_set(self._h, cms, len(meshes)), "ar_scene_set")

    def set_objects(self, objects):
        """Analytic boxes and spheres (ar_scene_set_objects); object k is
        primitive ntri + k.  See object_array for the dict keys."""
        arr = object_array(objects)
        self._check(self._lib.ar_scene_set_objects(self._h, arr, len(objects)), "ar_scene_set_objects")

    def build_bvh(self, width=4, leaf_size=2, bin_count=16, max_depth=60) -> dict:
        o = ar_bvh_opts(width, leaf_size, bin_count, max_depth)
        self._check(self._lib.ar_bvh_build(self._h, C.byref(o)), "ar_bvh_build")
        return self.bvh_info()

    # ---------------------------------------------------------------- batches
    def set_tile(self, threads: int):
        self._check(self._lib.ar_set_tile(self._h, int(threads)), "ar_set_tile")

    def _outputs(self, views, shared):
        import torch
        return [torch.empty((v["height"], v["width"], 3), dtype=torch.float32, device=f"cuda:{self.device}")
                for v in views]

    def render_views(self, views, outs=None, stream=None, **shared):
        """Render every view (batch contexts: one per scene; plain: one) in one
        call (ar_render_batch).  Returns the list of (h, w, 3) images."""
        cams, opts = _view_arrays(views, shared)
        outs = outs or self._outputs(views, shared)
        ptrs = (C.c_void_p * len(views))(*[_ptr(o) for o in outs])
        self._check(self._lib.ar_render_batch(self._h, len(views), cams, opts, ptrs, _stream_ptr(stream)),
                    "ar_render_batch")
        return outs

    def capture(self, views, **shared) -> Graph:
        """Capture a render of the views into a HIP graph (ar_graph_capture)."""
        cams, opts = _view_arrays(views, shared)
        outs = self._outputs(views, shared)
        ptrs = (C.c_void_p * len(views))(*[_ptr(o) for o in outs])
        g = C.c_void_p()
        self._check(self._lib.ar_graph_capture(self._h, len(views), cams, opts, ptrs, C.byref(g)), "ar_graph_capture")
        return Graph(self._lib, g, outs)

    # ---------------------------------------------------------------- learned guiding
    def mlp_configure(self, learning_rate=5e-3, batch=1 << 18, iters_per_frame=1, seed=1):
        o = _lib.ar_mlp_opts(float(learning_rate), int(batch), int(iters_per_frame), int(seed))
        self._check(self._lib.ar_mlp_configure(self._h, C.byref(o)), "ar_mlp_configure")

    def mlp_params(self):
        import numpy as np
        p = np.zeros(_lib.MLP_PARAMS, np.float32)
        self._check(self._lib.ar_mlp_params_get(self._h, p.ctypes.data, p.size), "ar_mlp_params_get")
        return p

    def mlp_records(self, cap=1 << 18):
        """The last learned-guided frame's training records (ar_mlp_records_get):
        dict of x (n, 6), dir (n, 3), weight (n,) and total (records before the
        batch was drawn)."""
        import numpy as np
        x = np.zeros((cap, 6), np.float32)
        d = np.zeros((cap, 3), np.float32)
        w = np.zeros(cap, np.float32)
        n, t = C.c_uint32(0), C.c_uint32(0)
        self._check(self._lib.ar_mlp_records_get(self._h, x.ctypes.data, d.ctypes.data, w.ctypes.data, cap,
                                                 C.byref(n), C.byref(t)), "ar_mlp_records_get")
        k = min(n.value, cap)
        return {"x": x[:k], "dir": d[:k], "weight": w[:k], "total": t.value}

    def mlp_set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, np.float32)
        self._check(self._lib.ar_mlp_params_set(self._h, p.ctypes.data, p.size), "ar_mlp_params_set")

    def mlp_forward(self, x5, stream=None):
        """Raw outputs (n, 40) of inputs x5 (n, 5), float32 device tensors."""
        import torch
        x5 = x5.contiguous()
        out = torch.empty((x5.shape[0], 40), dtype=torch.float32, device=x5.device)
        self._check(self._lib.ar_mlp_forward(self._h, _ptr(x5), x5.shape[0], _ptr(out), _stream_ptr(stream)),
                    "ar_mlp_forward")
        return out

    def mlp_gradient(self, x5, dirs, weight, stream=None):
        """(gradient (12720,), loss (1,)) of mean(-weight * log q(dir | x))."""
        import torch
        g = torch.empty(_lib.MLP_PARAMS, dtype=torch.float32, device=x5.device)
        loss = torch.empty(1, dtype=torch.float32, device=x5.device)
        self._check(self._lib.ar_mlp_gradient(self._h, _ptr(x5.contiguous()), _ptr(dirs.contiguous()),
                                              _ptr(weight.contiguous()), x5.shape[0], _ptr(g), _ptr(loss),
                                              _stream_ptr(stream)), "ar_mlp_gradient")
        return g, loss

    def mlp_train(self, x5, dirs, weight, steps, stream=None):
        import torch
        loss = torch.empty(1, dtype=torch.float32, device=x5.device)
        self._check(self._lib.ar_mlp_train(self._h, _ptr(x5.contiguous()), _ptr(dirs.contiguous()),
                                           _ptr(weight.contiguous()), x5.shape[0], int(steps), _ptr(loss),
                                           _stream_ptr(stream)), "ar_mlp_train")
        return loss

    # ---------------------------------------------------------------- guiding
    def guiding_configure(self, max_probes: int):
        self._check(self._lib.ar_guiding_configure(self._h, int(max_probes)), "ar_guiding_configure")

    def guiding_reset(self):
        self._check(self._lib.ar_guiding_reset(self._h), "ar_guiding_reset")

    def guiding_info(self) -> dict:
        info = _lib.ar_guiding_info()
        self._check(self._lib.ar_guiding_info_get(self._h, C.byref(info)), "ar_guiding_info_get")
        return {k: getattr(info, k) for k, _ in info._fields_}

    def guiding_snapshot(self):
        """(keys u64 (n,), texels f32 (n, 85, 4), counts u64 (n,)) on the host."""
        import numpy as np
        n = C.c_uint64()
        info = self.guiding_info()
        cap = int(info["probes"])
        keys = np.zeros(cap, np.uint64)
        tex = np.zeros((cap, 85, 4), np.float32)
        cnt = np.zeros(cap, np.uint64)
        self._check(self._lib.ar_guiding_snapshot(self._h, keys.ctypes.data, tex.ctypes.data, cnt.ctypes.data, cap,
                                                  C.byref(n)), "ar_guiding_snapshot")
        m = int(n.value)
        return keys[:m], tex[:m], cnt[:m]

    def guiding_restore(self, keys, texels, counts):
        import numpy as np
        keys = np.ascontiguousarray(keys, np.uint64)
        texels = np.ascontiguousarray(texels, np.float32)
        counts = np.ascontiguousarray(counts, np.uint64)
        self._check(self._lib.ar_guiding_restore(self._h, keys.ctypes.data, texels.ctypes.data, counts.ctypes.data,
                                                 keys.size), "ar_guiding_restore")

    def bvh_info(self) -> dict:
        info = ar_bvh_info()
        self._check(self._lib.ar_bvh_info_get(self._h, C.byref(info)), "ar_bvh_info_get")
        return {k: getattr(info, k) for k, _ in info._fields_}

    # -- signed-distance volumes ----------------------------------------------
    def sdf_volume_bake(self, shape, dims, params=(0.0, 0.0, 0.0, 0.0)) -> int:
        """ar_sdf_volume_bake: shape 'torus' (params R, r) or 'diamond'; dims (nx, ny, nz)."""
        shape = {"torus": _lib.AR_SDF_SHAPE_TORUS, "diamond": _lib.AR_SDF_SHAPE_DIAMOND}.get(shape, shape)
        prm = (C.c_float * 4)(*[float(x) for x in (list(params) + [0.0] * 4)[:4]])
        dm = (C.c_uint32 * 3)(*[int(x) for x in dims])
        idx = C.c_uint32()
        self._check(self._lib.ar_sdf_volume_bake(self._h, int(shape), prm, dm, C.byref(idx)), "ar_sdf_volume_bake")
        return idx.value

    def sdf_volume_add(self, voxels) -> int:
        """ar_sdf_volume_add of a (nz, ny, nx) float32 array."""
        import numpy as np
        v = np.ascontiguousarray(voxels, np.float32)
        dm = (C.c_uint32 * 3)(v.shape[2], v.shape[1], v.shape[0])
        idx = C.c_uint32()
        self._check(self._lib.ar_sdf_volume_add(self._h, v.ctypes.data, dm, C.byref(idx)), "ar_sdf_volume_add")
        return idx.value

    def sdf_volume_read(self, index):
        """ar_sdf_volume_read: the volume as a (nz, ny, nx) float32 array."""
        import numpy as np
        dm = (C.c_uint32 * 3)()
        self._check(self._lib.ar_sdf_volume_read(self._h, int(index), None, dm), "ar_sdf_volume_read")
        out = np.zeros((dm[2], dm[1], dm[0]), np.float32)
        self._check(self._lib.ar_sdf_volume_read(self._h, int(index), out.ctypes.data, None), "ar_sdf_volume_read")
        return out

    def sdf_volume_clear(self):
        self._check(self._lib.ar_sdf_volume_clear(self._h), "ar_sdf_volume_clear")

    # -- traversal -----------------------------------------------------------
    @staticmethod
    def _rays_c(rays) -> ar_rays:
        """rays: float32 CUDA tensor of shape (8, n) with rows RAY_FIELDS, or
        (9, n) with the per-ray SDF march tolerance in row 8."""
        assert rays.dim() == 2 and rays.shape[0] in (8, 9) and rays.is_contiguous() and rays.is_cuda
        import torch
        assert rays.dtype == torch.float32
        r = ar_rays()
        for k, name in enumerate(RAY_FIELDS):
            setattr(r, name, _ptr(rays[k]))
        r.meps = _ptr(rays[8]) if rays.shape[0] == 9 else None
        return r

    @staticmethod
    def alloc_hits(n: int, device="cuda"):
        import torch
        return {"prim": torch.empty(n, dtype=torch.int32, device=device),
                "tuv": torch.empty((3, n), dtype=torch.float32, device=device),
                "flags": torch.empty(n, dtype=torch.int32, device=device)}

    @staticmethod
    def _hits_c(h) -> ar_hits:
        c = ar_hits()
        c.prim = _ptr(h["prim"])
        c.t = _ptr(h["tuv"][0])
        c.u = _ptr(h["tuv"][1])
        c.v = _ptr(h["tuv"][2])
        c.flags = _ptr(h["flags"])
        return c

    def trace_closest(self, rays, hits=None, stream=None, counters=None):
        n = rays.shape[1]
        if hits is None:
            hits = self.alloc_hits(n, rays.device)
        rc, hc = self._rays_c(rays), self._hits_c(hits)
        if counters is not None:
            self._check(self._lib.ar_trace_closest_counted(self._h, C.byref(rc), n, C.byref(hc),
                                                           _ptr(counters[0]), _ptr(counters[1]),
                                                           _stream_ptr(stream)), "ar_trace_closest_counted")
        else:
            self._check(self._lib.ar_trace_closest(self._h, C.byref(rc), n, C.byref(hc), _stream_ptr(stream)),
                        "ar_trace_closest")
        return hits

    def trace_closest_noalloc(self, rays, hits=None, stream=None):
        """ar_trace_closest_noalloc: closest hit with the library's fixed
        per-device scratch array (no device allocation in the call)."""
        n = rays.shape[1]
        if hits is None:
            hits = self.alloc_hits(n, rays.device)
        rc, hc = self._rays_c(rays), self._hits_c(hits)
        self._check(self._lib.ar_trace_closest_noalloc(self._h, C.byref(rc), n, C.byref(hc), _stream_ptr(stream)),
                    "ar_trace_closest_noalloc")
        return hits

    def trace_any(self, rays, occluded=None, stream=None):
        import torch
        n = rays.shape[1]
        if occluded is None:
            occluded = torch.empty((n + 31) // 32, dtype=torch.int32, device=rays.device)
        rc = self._rays_c(rays)
        self._check(self._lib.ar_trace_any(self._h, C.byref(rc), n, _ptr(occluded), _stream_ptr(stream)),
                    "ar_trace_any")
        return occluded

    def make_shadow_rays(self, rays, hits, target, tmin=1e-4, tmax_scale=1e-4, out=None, stream=None):
        import torch
        n = rays.shape[1]
        if out is None:
            out = torch.empty_like(rays)
        rc, hc, oc = self._rays_c(rays), self._hits_c(hits), self._rays_c(out)
        tgt = (C.c_float * 3)(*[float(x) for x in target])
        self._check(self._lib.ar_make_shadow_rays(self._h, C.byref(rc), C.byref(hc), n, C.byref(tgt), tmin,
                                                  tmax_scale, C.byref(oc), _stream_ptr(stream)),
                    "ar_make_shadow_rays")
        return out

    # -- integrator ----------------------------------------------------------
    def render_progressive(self, frame, accum, odd=None, even=None, spp=1, cam_pos=(0.0, 0.0, 1.1),
                           cam_dir=(0.0, 0.0, -1.0), vfov=60.0, seed=1, max_depth=5, nee=True, sky=AR_SKY_NONE,
                           env=(0.0, 0.0, 0.0), rr_min_depth=3, spectral=0, output=0, gbuffer=None, stats=None,
                           guiding=0, frame_spp=0, frame_growth=0, stream=None):
        """ar_render_progressive: fold spp samples (indices frame ..) into the
        running means accum (and odd / even), (h, w, 3) float32 device tensors;
        gbuffer (h, w, 16) / stats (h, w, 12) keep the denoiser's records;
        guiding updates the context's guide after every frame (frame_spp
        samples, times frame_growth from one frame to the next)."""
        h, w = accum.shape[0], accum.shape[1]
        cam, o = _camera_opts(w, h, spp, cam_pos, cam_dir, vfov, seed, max_depth, nee, sky, env, rr_min_depth, 0, 0,
                              spectral, output, guiding, frame_spp, frame_growth)
        opt = lambda t: _ptr(t) if t is not None else None
        self._check(self._lib.ar_render_progressive(self._h, C.byref(cam), C.byref(o), int(frame), _ptr(accum),
                                                    opt(odd), opt(even), opt(gbuffer), opt(stats),
                                                    _stream_ptr(stream)),
                    "ar_render_progressive")
        return accum

    def denoise(self, hdr, gbuffer, stats, out=None, stream=None):
        """ar_denoise: (h, w, 3) float32 device image filtered with its
        (h, w, 16) G-buffer and (h, w, 12) statistics records."""
        import torch
        h, w = hdr.shape[0], hdr.shape[1]
        if out is None:
            out = torch.empty_like(hdr)
        self._check(self._lib.ar_denoise(self._h, _ptr(hdr), _ptr(gbuffer), _ptr(stats), _ptr(out), w, h,
                                         _stream_ptr(stream)), "ar_denoise")
        return out

    # ---- spectral basis network (the reference's RGB / wavelength encoders)

    def spectral_mlp_set(self, params):
        """ar_spectral_mlp_set: AR_SMLP_PARAMS float32 parameters (host array)."""
        import numpy as np
        p = np.ascontiguousarray(params, np.float32).reshape(-1)
        if p.size != _lib.SMLP_PARAMS:
            raise ValueError(f"expected {_lib.SMLP_PARAMS} parameters, got {p.size}")
        self._check(self._lib.ar_spectral_mlp_set(self._h, p.ctypes.data), "ar_spectral_mlp_set")

    def spectral_mlp_init(self, seed=0, stream=None):
        """Sets spectral_mlp_init(seed) parameters, then makes sure training
        can start: the reference's clamp passes no gradient where the dot
        product is outside (0, 1), and a draw whose outputs all sit at 0 (or
        at 1) would never move.  On 4096 probe rows the wavelength output
        layer is negated while most outputs are 0 and halved while most are
        1.  Returns the parameters set."""
        import numpy as np
        import torch
        w = spectral_mlp_init(seed)
        rng = np.random.default_rng(seed + 1)
        probe = np.empty((4096, 4), np.float32)
        probe[:, :3] = rng.random((4096, 3))
        probe[:, 3] = rng.uniform(360.0, 780.0, 4096)
        xp = torch.from_numpy(probe).cuda()
        for _ in range(12):
            self.spectral_mlp_set(w)
            out = self.spectral_mlp_eval(xp, stream=stream).cpu().numpy()
            if np.mean(out <= 0.0) > 0.5:
                w[15968:16096] *= -1.0
            elif np.mean(out >= 1.0) > 0.5:
                w[15968:16096] *= 0.5
            else:
                break
        self.spectral_mlp_set(w)
        return w

    def spectral_set_basis(self, network):
        """ar_spectral_set_basis: the basis network's tables (True) or the
        sigmoid uplift (False) for spectral renders."""
        self._check(self._lib.ar_spectral_set_basis(self._h, _lib.AR_BASIS_NETWORK if network else
                                                    _lib.AR_BASIS_SIGMOID), "ar_spectral_set_basis")

    def spectral_mlp_get(self, ema=False):
        """ar_spectral_mlp_get: the parameters (or their EMA) as a numpy array."""
        import numpy as np
        out = np.empty(_lib.SMLP_PARAMS, np.float32)
        self._check(self._lib.ar_spectral_mlp_get(self._h, out.ctypes.data, 1 if ema else 0), "ar_spectral_mlp_get")
        return out

    def spectral_mlp_eval(self, x, out=None, stream=None):
        """ar_spectral_mlp_eval of (n, 4) float32 device rows (r, g, b, nm)."""
        import torch
        x = x.contiguous()
        if out is None:
            out = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        self._check(self._lib.ar_spectral_mlp_eval(self._h, _ptr(x), _ptr(out), x.shape[0], _stream_ptr(stream)),
                    "ar_spectral_mlp_eval")
        return out

    def spectral_mlp_build_tables(self, rgb_res=512, lambda_res=840, stream=None):
        """ar_spectral_mlp_build_tables (the reference's resolutions by default)."""
        self._check(self._lib.ar_spectral_mlp_build_tables(self._h, int(rgb_res), int(lambda_res),
                                                           _stream_ptr(stream)), "ar_spectral_mlp_build_tables")

    def spectral_mlp_read_tables(self, rgb_res, lambda_res):
        """ar_spectral_mlp_read_tables: ((res, res, res, 4) [b][g][r], (lambda_res, 4)) numpy arrays."""
        import numpy as np
        rgb = np.empty((rgb_res, rgb_res, rgb_res, 4), np.float32)
        lam = np.empty((lambda_res, 4), np.float32)
        self._check(self._lib.ar_spectral_mlp_read_tables(self._h, rgb.ctypes.data, lam.ctypes.data),
                    "ar_spectral_mlp_read_tables")
        return rgb, lam

    def spectral_mlp_query(self, x, out=None, stream=None):
        """ar_spectral_mlp_query (Interact through the tables) of (n, 4) device rows."""
        import torch
        x = x.contiguous()
        if out is None:
            out = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        self._check(self._lib.ar_spectral_mlp_query(self._h, _ptr(x), _ptr(out), x.shape[0], _stream_ptr(stream)),
                    "ar_spectral_mlp_query")
        return out

    def spectral_mlp_gradient(self, x, y, batch, step=0, seed=0, noise=False, stream=None):
        """ar_spectral_mlp_gradient: (gradient, loss) device tensors of one batch."""
        import torch
        grad = torch.empty(_lib.SMLP_PARAMS, dtype=torch.float32, device=x.device)
        loss = torch.empty(1, dtype=torch.float32, device=x.device)
        self._check(self._lib.ar_spectral_mlp_gradient(self._h, _ptr(x), _ptr(y), x.shape[0], int(batch), int(step),
                                                       int(seed) & 0xFFFFFFFF, 1 if noise else 0, _ptr(grad),
                                                       _ptr(loss), _stream_ptr(stream)), "ar_spectral_mlp_gradient")
        return grad, loss

    def spectral_mlp_train(self, x, y, steps, batch=4096, first_step=0, total_steps=0, seed=0, noise=True,
                           lr=5e-4, weight_decay=1e-3, beta1=0.9, beta2=0.95, eps=1e-8, clip_norm=1.0,
                           warmup_ratio=0.001, min_lr_ratio=5e-6, ema_decay=0.999, stream=None):
        """ar_spectral_mlp_train; defaults are the reference's (config.ini,
        train_spectrum.py).  Returns the per-step losses (device tensor)."""
        import torch
        o = _lib.ar_smlp_train_opts(int(steps), int(batch), int(first_step), int(total_steps),
                                    int(seed) & 0xFFFFFFFF, 1 if noise else 0, lr, weight_decay, beta1, beta2, eps,
                                    clip_norm, warmup_ratio, min_lr_ratio, ema_decay)
        hist = torch.empty(max(1, int(steps)), dtype=torch.float32, device=x.device)
        self._check(self._lib.ar_spectral_mlp_train(self._h, _ptr(x), _ptr(y), x.shape[0], C.byref(o), _ptr(hist),
                                                    _stream_ptr(stream)), "ar_spectral_mlp_train")
        return hist[:int(steps)]

    def spectral_uplift_dataset(self, n, seed=0, device="cuda", stream=None):
        """ar_spectral_uplift_dataset: (x (n, 4), y (n,)) device tensors."""
        import torch
        x = torch.empty((int(n), 4), dtype=torch.float32, device=device)
        y = torch.empty(int(n), dtype=torch.float32, device=device)
        self._check(self._lib.ar_spectral_uplift_dataset(self._h, int(n), int(seed) & 0xFFFFFFFF, _ptr(x), _ptr(y),
                                                         _stream_ptr(stream)), "ar_spectral_uplift_dataset")
        return x, y

    def tonemap(self, hdr, exposure=1.0, tone=0, xyz_input=False, out=None, stream=None):
        """ar_tonemap: (h, w, 3) float32 device image -> (h, w) int32 tensor of
        0xAARRGGBB pixels."""
        import torch
        h, w = hdr.shape[0], hdr.shape[1]
        if out is None:
            out = torch.empty((h, w), dtype=torch.int32, device=hdr.device)
        self._check(self._lib.ar_tonemap(self._h, _ptr(hdr), _ptr(out), w, h, float(exposure), int(tone),
                                         1 if xyz_input else 0, _stream_ptr(stream)), "ar_tonemap")
        return out

    def luminance_mse(self, a, b, stream=None) -> float:
        """ar_luminance_mse of two (h, w, 3) float32 device images."""
        res = C.c_double()
        npix = a.shape[0] * a.shape[1]
        self._check(self._lib.ar_luminance_mse(self._h, _ptr(a), _ptr(b), npix, C.byref(res), _stream_ptr(stream)),
                    "ar_luminance_mse")
        return res.value

    def render(self, width, height, spp, cam_pos=(0.0, 0.0, 1.1), cam_dir=(0.0, 0.0, -1.0), vfov=60.0, seed=1,
               max_depth=5, nee=True, sky=AR_SKY_NONE, env=(0.0, 0.0, 0.0), rr_min_depth=3, sample_offset=0,
               max_paths_per_pass=0, spectral=0, output=0, guiding=0, frame_spp=0, frame_growth=0, out=None,
               stream=None):
        """Mean radiance image (height, width, 3) float32 on the GPU (ar_render)."""
        import torch
        cam, o = _camera_opts(width, height, spp, cam_pos, cam_dir, vfov, seed, max_depth, nee, sky, env,
                              rr_min_depth, sample_offset, max_paths_per_pass, spectral, output, guiding, frame_spp,
                              frame_growth)
        if out is None:
            out = torch.empty((height, width, 3), dtype=torch.float32, device=f"cuda:{self.device}")
        self._check(self._lib.ar_render(self._h, C.byref(cam), C.byref(o), _ptr(out), _stream_ptr(stream)),
                    "ar_render")
        return out


    def present(self, hdr=None, gbuffer=None, stats=None, view=0, exposure=1.0, tone=AR_TONE_ACES, xyz_input=False,
                width=None, height=None, out=None, stream=None):
        """ar_present: (h, w) int32 device image of 0xAARRGGBB pixels from
        hdr (h, w, 3) through the tone curve (view 0), denoised first (view
        1), or a debug view of the gbuffer (h, w, 16) / stats (h, w, 12)
        records (views 2..10)."""
        import torch
        ref = next(t for t in (hdr, gbuffer, stats) if t is not None)
        h = height if height is not None else ref.shape[0]
        w = width if width is not None else ref.shape[1]
        if out is None:
            out = torch.empty((h, w), dtype=torch.int32, device=ref.device)
        opt = lambda t: _ptr(t) if t is not None else None
        self._check(self._lib.ar_present(self._h, opt(hdr), opt(gbuffer), opt(stats), _ptr(out), w, h,
                                         float(exposure), int(tone), int(bool(xyz_input)), int(view),
                                         _stream_ptr(stream)), "ar_present")
        return out

    def render_counters(self, enable=True):
        """ar_render_counters: count box / triangle tests and shaded vertices
        of this context's renders (measurement only; slower)."""
        self._check(self._lib.ar_render_counters(self._h, 1 if enable else 0), "ar_render_counters")

    def render_counters_read(self) -> dict:
        """ar_render_counters_read: totals since the last read (waits)."""
        import numpy as np
        out = np.zeros(5, np.uint64)
        self._check(self._lib.ar_render_counters_read(self._h, out.ctypes.data), "ar_render_counters_read")
        keys = ("closest_box_tests", "closest_tri_tests", "vertices", "shadow_box_tests", "shadow_tri_tests")
        return {k: int(v) for k, v in zip(keys, out)}

    def render_host(self, width, height, spp, exposure=1.0, tone=AR_TONE_ACES, hdr=True, ldr=True, **kw):
        """ar_render_host: renders into the context's own device image and
        returns (hdr, ldr) numpy arrays ((height, width, 3) float32 and
        (height, width) uint32 0xAARRGGBB; either may be skipped -> None)."""
        import numpy as np
        cam, o = _camera_opts(width, height, spp, **kw)
        h = np.empty((height, width, 3), np.float32) if hdr else None
        l = np.empty((height, width), np.uint32) if ldr else None
        self._check(self._lib.ar_render_host(self._h, C.byref(cam), C.byref(o), h.ctypes.data if hdr else None,
                                             l.ctypes.data if ldr else None, float(exposure), int(tone)),
                    "ar_render_host")
        return h, l


def write_bmp(path, ldr):
    """ar_write_bmp: (h, w) 0xAARRGGBB pixels (numpy uint32 or int32 torch
    tensor) to a 32-bit top-down BMP."""
    import numpy as np
    a = ldr.cpu().numpy() if hasattr(ldr, "cpu") else np.asarray(ldr)
    a = np.ascontiguousarray(a).view(np.uint32)
    rc = _lib.load().ar_write_bmp(str(path).encode(), a.ctypes.data, a.shape[1], a.shape[0])
    if rc != AR_OK:
        raise ArError(rc, f"ar_write_bmp({path})")


def make_batch(scenes) -> Context:
    """Merge scene contexts (scene set, BVH built, same width) into a batch
    context (ar_batch_create)."""
    lib = _lib.load()
    hs = (C.c_void_p * len(scenes))(*[c._h.value if hasattr(c._h, "value") else c._h for c in scenes])
    out = C.c_void_p()
    rc = lib.ar_batch_create(hs, len(scenes), C.byref(out))
    if rc != AR_OK:
        msg = lib.ar_last_error(scenes[0]._h)
        raise ArError(rc, f"ar_batch_create: {msg.decode() if msg else ''}")
    b = Context._adopt(out, scenes[0].device)
    b._keep.append(list(scenes))
    return b


def autotune(scenes, views, probe_spp=1, **shared) -> dict:
    """ar_autotune_batch: fastest (BVH width, tile) for rendering the views of
    the scenes as one batch; leaves the scenes built at that width."""
    lib = _lib.load()
    cams, opts = _view_arrays(views, shared)
    hs = (C.c_void_p * len(scenes))(*[c._h.value if hasattr(c._h, "value") else c._h for c in scenes])
    res = _lib.ar_tune_result()
    rc = lib.ar_autotune_batch(hs, len(scenes), cams, opts, int(probe_spp), C.byref(res))
    if rc != AR_OK:
        msg = lib.ar_last_error(scenes[0]._h)
        raise ArError(rc, f"ar_autotune_batch: {msg.decode() if msg else ''}")
    return {"width": res.width, "tile": res.tile,
            "ms": {f"w{w}_t{t}": res.ms[i][j] for i, w in enumerate((2, 4, 8)) for j, t in enumerate((64, 128, 256, 512))}}
