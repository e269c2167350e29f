"""The FP32 oracle against the FP64 reference fixtures at independent seeds:
the two precisions estimate the same radiance (the float restatement adds no
bias), and the fixtures match the scene generators and lattices they name."""
import numpy as np
import pytest

from fp64_reference import assert_unbiased, load, rmse
from oracle import binding
from tinytrace_amd import Config, _abi, build_constants, scenes


def lattice(width, height, stride):
    xs, ys = np.meshgrid(np.arange(stride // 2, width, stride), np.arange(stride // 2, height, stride))
    return (ys * width + xs).ravel()


@pytest.mark.parametrize("name,size,stride", [("fp64_cornell_512", (512, 512), 16),
                                              ("fp64_sphere_field_1080p", (1920, 1080), 32)])
def test_fixture_layout(name, size, stride):
    ref = load(name)
    assert (int(ref["width"]), int(ref["height"])) == size
    assert int(ref["stride"]) == stride and int(ref["spp"]) == 4096
    np.testing.assert_array_equal(ref["pixels"], lattice(*size, stride))
    assert ref["radiance"].dtype == np.float64 and ref["radiance"].shape == (ref["pixels"].size, 3)
    assert np.all(np.isfinite(ref["radiance"])) and np.all(ref["radiance"] >= 0)


def _frames(scene, camera, ref, frames, seed):
    oracle = binding.OracleScene(scene, 32)
    out = []
    for f in range(frames):
        k = build_constants(camera, Config(bounce_count=int(ref["depth"])), int(ref["width"]), int(ref["height"]),
                            1000 + f, 0)
        rgb, _ = oracle.render_pixels(k, ref["pixels"], rng=_abi.TT_RNG_PCG, seed=seed, threads=0)
        out.append(rgb)
    return np.array(out)


def test_float_oracle_matches_fp64_cornell():
    ref = load("fp64_cornell_512")
    scene, camera = scenes.cornell_box()
    frames = _frames(scene, camera, ref, 64, seed=9)
    assert_unbiased(frames, ref)
    # Monte Carlo error shrinks as 1/sqrt(frames).
    assert rmse(frames.mean(0), ref) < 0.25 * rmse(frames[0], ref)


def test_float_oracle_matches_fp64_sphere_field():
    ref = load("fp64_sphere_field_1080p")
    scene, camera = scenes.sphere_field(seed=0)
    frames = _frames(scene, camera, ref, 128, seed=4)
    assert_unbiased(frames, ref)
