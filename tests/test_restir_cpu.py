"""ReSTIR DI in the oracle: FP16 reservoir storage and agreement of the
reservoir estimator with uniform next-event estimation (both estimate the
same direct light; the reservoir estimator is unbiased with respect to its
unshadowed target, so the converged images agree to within the visibility
term's small bias)."""
import numpy as np
import pytest

from oracle import binding
from tinytrace_amd import Config, LightSampling, build_constants, scenes


def test_half_conversion_is_round_to_nearest_even():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(200000) * 10.0 ** rng.uniform(-9, 5, 200000),
                        np.array([0.0, -0.0, 65504.0, 65520.0, 1e6, 6e-8, 3e-8, 1e-9, np.inf, -np.inf])]).astype(np.float32)
    got = binding.math_f32("half", x)
    with np.errstate(over="ignore"):
        expected = x.astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), expected.view(np.uint32))


def _converged(scene, camera, light_sampling, frames, size, restir_state=None, oracle=None):
    w, h = size
    img = None
    for f in range(frames):
        k = build_constants(camera, Config(bounce_count=4, light_sampling=light_sampling), w, h, f + 1, f + 1)
        if light_sampling == LightSampling.RESTIR:
            img, _ = oracle.render_restir(k, restir_state, rng=1, seed=0, previous=img)
        else:
            img, _ = oracle.render(k, rng=1, seed=0, previous=img)
    return img[..., :3]


@pytest.mark.parametrize("scene_name", ["cornell", "spheres"])
def test_restir_agrees_with_nee(scene_name):
    if scene_name == "cornell":
        scene, camera = scenes.cornell_box()
    else:
        scene, camera = scenes.sphere_field(seed=2, spheres=40, lights=24, subdivisions=2)
    oracle = binding.OracleScene(scene, 32)
    size = (64, 48)
    restir = _converged(scene, camera, LightSampling.RESTIR, 24, size, oracle.restir_state(*size), oracle)
    nee = _converged(scene, camera, LightSampling.NEXT_EVENT_ESTIMATION, 24, size, oracle=oracle)
    assert abs(restir.mean() - nee.mean()) <= 0.02 * nee.mean()


def _direct_frames(oracle, camera, w, h, frames, restir=None, first_frame=1):
    """Per-frame images of the direct light only (bounce_count 1)."""
    out = []
    state = oracle.restir_state(w, h)
    for f in range(frames):
        if restir is None:
            k = build_constants(camera, Config(bounce_count=1), w, h, first_frame + f, 0)
            img, _ = oracle.render(k, rng=1, seed=0)
        else:
            k = build_constants(camera, Config(bounce_count=1, light_sampling=LightSampling.RESTIR), w, h,
                                first_frame + f, 0)
            img, _ = oracle.render_restir(k, state, rng=1, seed=0, candidates=restir[0], temporal=restir[1],
                                          spatial=restir[2])
        out.append(img[..., :3].copy())
    return np.array(out)


@pytest.fixture(scope="module")
def many_lights():
    scene, camera = scenes.sphere_field(seed=0, spheres=60, lights=256, subdivisions=2)
    camera.fov = 0.6
    return binding.OracleScene(scene, 32), camera


def test_single_candidate_reservoir_equals_nee_in_expectation(many_lights):
    """With one candidate and no reuse the reservoir weight is the inverse
    source density, so the estimator is uniform NEE with other random
    numbers: per-pixel means agree within their standard errors."""
    oracle, camera = many_lights
    n = 96
    nee = _direct_frames(oracle, camera, 32, 24, n)
    one = _direct_frames(oracle, camera, 32, 24, n, restir=(1, False, False), first_frame=5000)
    se = np.sqrt(nee.var(0) / n + one.var(0) / n) + 1e-9
    z = (one.mean(0) - nee.mean(0)) / se
    assert abs(z.mean()) < 0.1
    assert (np.abs(z) > 4).mean() < 0.01


def test_restir_reduces_direct_light_noise_with_many_lights(many_lights):
    """With many small lights (256 emissive triangles), 32 candidates cut the
    per-pixel variance of the direct light several times over, and one frame
    with temporal + spatial reuse is closer to the converged image than one
    NEE frame.  Errors are mean absolute errors: NEE's fireflies near lights
    make squared errors a measure of a handful of pixels."""
    oracle, camera = many_lights
    w, h, n = 32, 24, 64
    nee = _direct_frames(oracle, camera, w, h, n)
    ris = _direct_frames(oracle, camera, w, h, n, restir=(32, False, False), first_frame=7000)
    assert np.median(ris.var(0) / (nee.var(0) + 1e-12)) < 0.25
    reference = 0.5 * (nee.mean(0) + ris.mean(0))
    reused = _direct_frames(oracle, camera, w, h, 6, restir=(32, True, True), first_frame=9000)[-1]
    err_r = np.abs(reused - reference).mean()
    err_n = np.abs(nee[0] - reference).mean()
    assert err_r < 0.6 * err_n, (err_r, err_n)
