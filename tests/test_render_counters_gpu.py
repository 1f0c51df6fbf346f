"""The integrator's ops-model counters (ar_render_counters): a counted render
gives the same image bit for bit (the counting traversal variant finds the
same hits); the shaded-vertex count of a one-bounce render is exactly the
number of camera paths, and of a deeper render at least that; the closest
traversals' tests per camera ray equal those of ar_trace_closest_counted on
the same camera rays; counters reset on read and stay off when disabled."""
import numpy as np
import pytest

from aperturerendererv3_amd.workloads import CORNELL_CAMERA, cornell_box

pytestmark = pytest.mark.gpu

CAM = dict(cam_pos=CORNELL_CAMERA["pos"], cam_dir=CORNELL_CAMERA["dir"], vfov=CORNELL_CAMERA["vfov"])


@pytest.fixture(scope="module")
def ctx():
    from aperturerendererv3_amd import Context
    c = Context(0)
    V, F, M, mats = cornell_box()
    c.set_meshes([{"vertices": V, "indices": F, "tri_material": M, "materials": mats}])
    c.build_bvh()
    yield c
    c.close()


@pytest.mark.parametrize("width", [2, 4, 8])
def test_counted_render_is_identical(ctx, width):
    import torch
    ctx.build_bvh(width=width)
    kw = dict(seed=3, max_depth=5, **CAM)
    ref = ctx.render(64, 48, 8, **kw)
    ctx.render_counters(True)
    got = ctx.render(64, 48, 8, **kw)
    c = ctx.render_counters_read()
    ctx.render_counters(False)
    torch.cuda.synchronize()
    assert torch.equal(ref, got)
    paths = 64 * 48 * 8
    print(f"width {width}: {c}", flush=True)
    assert c["vertices"] >= paths and c["closest_box_tests"] > c["vertices"] and c["closest_tri_tests"] > 0
    assert c["shadow_box_tests"] > 0 and c["shadow_tri_tests"] > 0
    ctx.render(64, 48, 8, **kw)          # counting is off: nothing more is added
    assert all(v == 0 for v in ctx.render_counters_read().values())
    ctx.build_bvh(width=8)


def test_one_bounce_counts_camera_paths(ctx):
    ctx.render_counters(True)
    ctx.render(40, 30, 3, seed=1, max_depth=1, nee=False, **CAM)
    c = ctx.render_counters_read()
    ctx.render_counters(False)
    assert c["vertices"] == 40 * 30 * 3
    assert c["shadow_box_tests"] == 0 and c["shadow_tri_tests"] == 0


def test_closest_counts_match_trace_counted(ctx):
    """Camera rays of a one-bounce render: the same box / triangle test totals
    as tracing those rays with ar_trace_closest_counted (same traversal)."""
    import torch
    from aperturerendererv3_amd.workloads import camera_dirs
    w, h = 32, 24
    ctx.render_counters(True)
    ctx.render(w, h, 1, seed=1, max_depth=1, nee=False, **CAM)
    c = ctx.render_counters_read()
    ctx.render_counters(False)
    assert c["vertices"] == w * h and c["closest_box_tests"] > 0
    # the camera rays are jittered per pixel by the render's random numbers, so
    # compare the per-ray average with pixel-centre rays of the same view
    d = camera_dirs(w, h, np.asarray(CAM["cam_dir"], np.float64), CAM["vfov"]).reshape(-1, 3)
    n = d.shape[0]
    rays = np.zeros((8, n), np.float32)
    rays[0:3] = np.asarray(CAM["cam_pos"], np.float32)[:, None]
    rays[4:7] = d.T
    rays[7] = np.inf
    r = torch.from_numpy(rays).cuda()
    cb = torch.empty(n, dtype=torch.int32, device="cuda")
    ct = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.trace_closest(r, counters=(cb, ct))
    torch.cuda.synchronize()
    per_ray = cb.double().mean().item()
    got = c["closest_box_tests"] / (w * h)
    assert abs(got - per_ray) / per_ray < 0.1, (got, per_ray)
