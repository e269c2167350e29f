"""Helpers for comparisons against the FP64 reference fixtures written by
tests/golden/make_fp64_reference.py (4096-spp double-precision oracle
radiance on pixel lattices)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def z_scores(frames, ref):
    """Per-pixel-channel z of the mean of independent frames (frames x
    pixels x 3) against the reference mean, using both standard errors."""
    frames = np.asarray(frames, dtype=np.float64)
    mean = frames.mean(0)
    se = frames.std(0, ddof=1) / np.sqrt(frames.shape[0])
    return (mean - ref["radiance"]) / np.sqrt(se ** 2 + ref["stderr"] ** 2 + 1e-30)


def assert_unbiased(frames, ref):
    """The frames' mean agrees with the reference in expectation.  Per pixel
    the z-scores spread ~1 with few outliers; over the lattice the mean
    radiance agrees within 4 standard errors.  Radiance near small bright
    lights is heavy-tailed, so sample means of few frames run low and a
    handful of large |z| are expected: callers use enough frames (about 100
    for the sphere field) that the tail is sampled."""
    frames = np.asarray(frames, dtype=np.float64)
    z = z_scores(frames, ref)
    assert 0.75 < z.std() < 1.3, z.std()
    assert (np.abs(z) > 4).mean() < 0.015, (np.abs(z) > 4).mean()
    mean = frames.mean(0)
    se = frames.std(0, ddof=1) / np.sqrt(frames.shape[0])
    global_se = np.sqrt((se ** 2).sum() + (ref["stderr"] ** 2).sum()) / mean.size
    global_z = (mean.mean() - ref["radiance"].mean()) / global_se
    assert abs(global_z) < 4.0, global_z


def rmse(image_rgb, ref):
    return float(np.sqrt(((np.asarray(image_rgb, dtype=np.float64) - ref["radiance"]) ** 2).mean()))
