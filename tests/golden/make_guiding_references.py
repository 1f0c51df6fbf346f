"""Writes the FP64 oracle reference images of the guiding workloads and of
the Cornell box.

  guiding_room_fp64_320x180.npz           the occluded-light room
                                          (workloads.occluded_room, 200k
                                          triangles, OCCLUDED_CAMERA) at
                                          320 x 180 pixels
  guiding_spectral_crop_fp64_320x180.npz  the 500k-triangle spectral scene
                                          (workloads.spectral_scene,
                                          SPECTRAL_CAMERA): the 320 x 180
                                          centre window (pixels 800..1119 x
                                          450..629) of the 1920 x 1080 image,
                                          each pixel traced as in the full
                                          frame

Both: 4 hero wavelengths, linear sRGB output, max depth 8, NEE on, no sky,
16,384 samples per pixel (32 chunks of 512, sample offsets 512 k).

  cornell_fp64_256x256.npz                the 32-triangle Cornell box
                                          (workloads.cornell_box,
                                          CORNELL_CAMERA) at 256 x 256: RGB,
                                          max depth 5, NEE on, 65,536 samples
                                          per pixel (32 chunks of 2,048)

All: seed 12345 (independent of the seeds the tests render with), the
oracle's FP64 integrator.  Each file holds `mean` (float32, height x width x
3), `var_mean` (float32, the variance of `mean` estimated from the 32 chunk
means: the reference's own noise) and `spp`.

Usage: python make_guiding_references.py room|spectral|cornell [--threads N]
       [--cache DIR]   (DIR keeps finished chunks so a run can resume)
"""
import argparse
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

import oracle_lib as orc  # noqa: E402
from aperturerendererv3_amd import workloads as W  # noqa: E402

SEED, CHUNKS = 12345, 32
WIDTH, HEIGHT = 320, 180
GUIDED = dict(spp=16384, max_depth=8, spectral=4, output=1)
SCENES = {
    "room": dict(out="guiding_room_fp64_320x180.npz", full=(WIDTH, HEIGHT), crop=None, **GUIDED),
    "spectral": dict(out="guiding_spectral_crop_fp64_320x180.npz", full=(1920, 1080),
                     crop=((1920 - WIDTH) // 2, (1080 - HEIGHT) // 2, WIDTH, HEIGHT), **GUIDED),
    "cornell": dict(out="cornell_fp64_256x256.npz", full=(256, 256), crop=None, spp=65536, max_depth=5, spectral=0,
                    output=0),
}


def scene(name):
    if name == "room":
        V, F, M, mats = W.occluded_room()
        return V, F, M, mats, W.OCCLUDED_CAMERA
    if name == "cornell":
        V, F, M, mats = W.cornell_box()
        return V, F, M, mats, W.CORNELL_CAMERA
    V, F, M, mats = W.spectral_scene()
    return V, F, M, mats, W.SPECTRAL_CAMERA


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("scene", choices=sorted(SCENES))
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--cache", type=Path, default=None)
    a = ap.parse_args()
    cfg = SCENES[a.scene]
    V, F, M, mats, cam = scene(a.scene)
    if a.cache:
        a.cache.mkdir(parents=True, exist_ok=True)
    chunks = []
    t0 = time.time()
    chunk = cfg["spp"] // CHUNKS
    for k in range(CHUNKS):
        path = a.cache / f"{a.scene}_{k:02d}.npy" if a.cache else None
        if path is not None and path.exists():
            chunks.append(np.load(path))
            continue
        img = orc.render(V, F, M, mats, cfg["full"][0], cfg["full"][1], chunk, max_depth=cfg["max_depth"], nee=1,
                         seed=SEED, sample_offset=k * chunk, precision=64, cam_pos=cam["pos"], cam_dir=cam["dir"],
                         vfov=cam["vfov"], spectral=cfg["spectral"], output=cfg["output"], threads=a.threads,
                         crop=cfg["crop"])
        if path is not None:
            np.save(path, img)
        chunks.append(img)
        print(f"{a.scene}: {(k + 1) * chunk} spp, {time.time() - t0:.0f} s", flush=True)
    c = np.stack(chunks)
    mean = c.mean(axis=0)
    var_mean = c.var(axis=0, ddof=1) / c.shape[0]
    np.savez(HERE / cfg["out"], mean=mean.astype(np.float32), var_mean=var_mean.astype(np.float32),
             spp=np.int64(cfg["spp"]))
    print(f"wrote {cfg['out']}: mean {mean.mean():.4g}, reference noise "
          f"{np.mean(var_mean / (mean ** 2 + 0.01)):.3g} (relMSE units)")


if __name__ == "__main__":
    main()
