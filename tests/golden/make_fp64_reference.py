"""Writes the FP64 reference radiance fixtures (4096 samples per pixel, the
oracle in double precision, next-event estimation) on pixel lattices of the
two benchmark workloads:

  fp64_cornell_512.npz        Cornell box, 512x512, depth 4, lattice stride 16
                              (32 x 32 = 1024 pixels)
  fp64_sphere_field_1080p.npz sphere field (seed 0), 1920x1080, depth 8,
                              lattice stride 32 (60 x 34 = 2040 pixels)
  fp64_instanced_2160p.npz    instanced field (seed 0), 3840x2160, depth 12,
                              lattice stride 64 (60 x 34 = 2040 pixels)

Each file holds `pixels` (linear indices), `radiance` (float64, pixels x 3),
`stderr` (per-pixel standard error of the mean, float64) and the settings.
The samples use PCG seed 0x5eed, distinct from every seed the tests and the
benchmark render with.

  python tests/golden/make_fp64_reference.py [--threads N] [--spp 4096]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import binding  # noqa: E402
from tinytrace_amd import Config, LightSampling, _abi, build_constants, scenes  # noqa: E402

SEED = 0x5EED
BATCH = 16  # samples per oracle call (sample_count)

WORKLOADS = {
    "fp64_cornell_512": dict(scene=lambda: scenes.cornell_box(), size=(512, 512), depth=4, stride=16),
    "fp64_sphere_field_1080p": dict(scene=lambda: scenes.sphere_field(seed=0), size=(1920, 1080), depth=8, stride=32),
    "fp64_instanced_2160p": dict(scene=lambda: scenes.instanced_field(seed=0), size=(3840, 2160), depth=12, stride=64),
}


def lattice(width, height, stride):
    xs, ys = np.meshgrid(np.arange(stride // 2, width, stride), np.arange(stride // 2, height, stride))
    return (ys * width + xs).ravel().astype(np.uint32)


def reference(name, spp, threads):
    w = WORKLOADS[name]
    scene, camera = w["scene"]()
    width, height = w["size"]
    oracle = binding.OracleScene(scene, 64)
    pixels = lattice(width, height, w["stride"])
    total = np.zeros((pixels.size, 3))
    total_sq = np.zeros((pixels.size, 3))
    for frame in range(spp // BATCH):
        k = build_constants(camera, Config(bounce_count=w["depth"], sample_count=BATCH,
                                           light_sampling=LightSampling.NEXT_EVENT_ESTIMATION),
                            width, height, frame, 0)
        batch, _ = oracle.render_pixels(k, pixels, rng=_abi.TT_RNG_PCG, seed=SEED, threads=threads)
        batch = batch.astype(np.float64)  # mean of BATCH samples
        total += batch
        total_sq += batch * batch
    n = spp // BATCH
    mean = total / n
    var = np.maximum(total_sq / n - mean * mean, 0.0) / max(n - 1, 1)
    return dict(pixels=pixels, radiance=mean, stderr=np.sqrt(var), width=width, height=height,
                depth=w["depth"], stride=w["stride"], spp=spp, seed=SEED)


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--threads", type=int, default=0)
    p.add_argument("--spp", type=int, default=4096)
    p.add_argument("--only", choices=sorted(WORKLOADS), default=None)
    args = p.parse_args()
    for name in WORKLOADS:
        if args.only and name != args.only:
            continue
        t0 = time.time()
        data = reference(name, args.spp, args.threads)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **data)
        print(f"{name}: {data['pixels'].size} pixels x {args.spp} spp in {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
