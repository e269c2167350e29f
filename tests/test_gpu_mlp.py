"""Radiance-cache MLP on the GPU against the float64 oracle
(oracle/mlp_ref.py), which rounds to bfloat16 where the kernels do:

  * encoding: within 1e-6 of the float64 formula of the same float32
    coordinate (the kernels reduce 2^i u exactly before multiplying by pi);
  * inference: within 1% (a bf16 activation can round the other way when
    fp32 and float64 accumulation straddle a rounding boundary);
  * loss gradient: within 1% in norm per layer of the float64 gradient,
    which finite differences pin on the CPU;
  * Adam: the fp32 update of the oracle from the device gradient;
  * training fits a smooth target, and the cached render stays within a few
    percent of the FP64 Cornell reference.
"""
import numpy as np
import pytest

from fp64_reference import load, rmse
from oracle import mlp_ref
from tinytrace_amd import Config, Context, LightSampling, Renderer, RenderSettings, _abi, scenes
from tinytrace_amd.cache import RadianceCache, render_cached

pytestmark = pytest.mark.gpu

LO, HI = (-2.0, -2.0, -2.0), (2.0, 2.0, 2.0)


def _records(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 16), dtype=np.float32)
    r[:, 0:3] = rng.uniform(-2.5, 2.5, (n, 3))
    d = rng.normal(size=(n, 3))
    r[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    r[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    r[:, 9:12] = rng.uniform(0, 1, (n, 3))
    r[:, 12] = rng.uniform(0.3, 1, n)
    r[:, 13] = rng.integers(0, 2, n)
    r[:, 14] = rng.integers(1, 4, n)
    return r


@pytest.fixture(scope="module")
def cache():
    c = RadianceCache(Context(), seed=3)
    c.set_bounds(LO, HI)
    return c


def test_encoding_matches_formula(cache):
    r = _records(500, 0)
    np.testing.assert_allclose(cache.encode(r), mlp_ref.encode(r, LO, HI), atol=1e-6)


def test_inference_matches_oracle(cache):
    r = _records(1000, 1)
    feats = cache.encode(r)
    got = cache.infer(r)
    want = mlp_ref.infer(cache.params, feats, bf16_points=True)
    np.testing.assert_allclose(got, want, rtol=1e-2, atol=1e-4)


def test_gradient_matches_oracle(cache):
    r = _records(3000, 2)
    t = np.random.default_rng(3).uniform(0, 3, (3000, 3))
    g, loss = cache.gradient(r, t)
    want, want_loss = mlp_ref.gradient(cache.params, cache.encode(r), t, bf16_points=True)
    assert abs(loss - want_loss) <= 2e-3 * want_loss
    for k, (a, b) in enumerate(mlp_ref.layer_slices()):
        err = np.linalg.norm(g[a:b] - want[a:b]) / np.linalg.norm(want[a:b])
        assert err < 1e-2, (k, err)
    exact, _ = mlp_ref.gradient(cache.params, mlp_ref.encode(r, LO, HI), t)
    assert np.linalg.norm(g - exact) / np.linalg.norm(exact) < 0.05


def test_adam_step_matches_oracle():
    c = RadianceCache(Context(), seed=5, learning_rate=1e-3)
    c.set_bounds(LO, HI)
    r = _records(700, 4)
    t = np.random.default_rng(5).uniform(0, 3, (700, 3))
    p0 = c.params
    g, _ = c.gradient(r, t)
    c.train_step(r, t)
    want, _, _ = mlp_ref.adam(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1)
    np.testing.assert_allclose(c.params, want, rtol=1e-6, atol=1e-9)


def test_weights_after_adam_steps_equal_weights_set_from_host():
    """An Adam step writes the bf16 copies the kernels read along with the
    fp32 parameters: after several steps, inference (which reads only the
    copies) and a further gradient equal, bit for bit, those of a cache
    whose parameters are set from the host to the same fp32 values."""
    c = RadianceCache(Context(), seed=9, learning_rate=2e-3)
    c.set_bounds(LO, HI)
    r = _records(2000, 8)
    t = np.random.default_rng(9).uniform(0, 3, (2000, 3))
    for _ in range(5):
        c.train_step(r, t)
    fresh = RadianceCache(Context(), seed=1)
    fresh.set_bounds(LO, HI)
    fresh.params = c.params
    assert not np.array_equal(c.params, RadianceCache(Context(), seed=9).params)
    np.testing.assert_array_equal(c.infer(r), fresh.infer(r))
    g, loss = c.gradient(r, t)
    g_fresh, loss_fresh = fresh.gradient(r, t)
    np.testing.assert_array_equal(g, g_fresh)
    assert loss == loss_fresh


def test_training_fits_a_smooth_target():
    c = RadianceCache(Context(), seed=7, learning_rate=3e-3)
    c.set_bounds(LO, HI)
    r = _records(1 << 14, 6)
    t = np.stack([1.0 + 0.5 * np.sin(2 * r[:, 0]), 0.5 + 0.25 * r[:, 1] ** 2, 0.2 + 0.1 * r[:, 6]], axis=1)
    first = c.train_step(r, t)
    for _ in range(300):
        last = c.train_step(r, t)
    assert last < 0.1 * first, (first, last)
    pred = c.infer(r[:512])
    assert np.abs(pred - t[:512]).mean() < 0.1


def test_inference_equals_encoding_then_network():
    """The inference kernel encodes in registers (each half-wave its own
    features); the encode entry point encodes in one pass.  Inference from
    the kernel must equal the oracle network on the entry point's features
    for every record of a full tile, so the two encoders agree."""
    c = RadianceCache(Context(), seed=4)
    c.set_bounds(LO, HI)
    r = _records(64, 11)
    r[:, 3:6] = np.clip(r[:, 3:6] * 1.7, -1, 1)  # directions off the unit sphere too
    want = mlp_ref.infer(c.params, mlp_ref.bf16(c.encode(r).astype(np.float32)), bf16_points=True)
    np.testing.assert_allclose(c.infer(r), want, rtol=1e-2, atol=1e-4)


def _cornell(seed):
    scene, camera = scenes.cornell_box()
    renderer = Renderer((512, 512), settings=RenderSettings(rng=_abi.TT_RNG_PCG, seed=seed))
    renderer.set_scene(scene)
    renderer.camera = camera
    renderer.set_config(Config(bounce_count=4, sample_count=1, light_sampling=LightSampling.NEXT_EVENT_ESTIMATION))
    return renderer


def test_judge_reports_the_cache_error_and_training_lowers_it():
    """Judge paths (1% of the paths at a cache vertex) are traced plainly and
    compared with the cache's prediction at their vertex.  Untrained, the
    mean prediction is far off the traced mean; after training on 24 frames
    it is within 6% (the judged loss itself stays near the noise floor of
    one-sample traced targets)."""
    renderer = _cornell(4)
    cache = RadianceCache(renderer.context, seed=1)
    first = render_cached(renderer, cache)
    assert 1500 < first["judge_samples"] < 4000, first
    assert np.isfinite(first["judge_loss"]) and first["judge_traced"] > 0
    assert abs(first["judge_cached"] / first["judge_traced"] - 1) > 0.5, first
    for _ in range(24):
        render_cached(renderer, cache, stats=False)
    last = render_cached(renderer, cache)
    assert abs(last["judge_cached"] / last["judge_traced"] - 1) < 0.06, last
    assert last["judge_loss"] < 1.5 * first["judge_loss"], (first, last)
    assert 40000 < last["samples"] <= 65536 * 5 // 4 + 1024, last
    assert last["infer_ms"] > 0 and last["train_ms"] > 0
    none = render_cached(renderer, cache, judge_fraction=-1.0)
    assert none["judge_samples"] == 0


def test_trained_cache_beats_untrained_cache():
    """The same cached frames with the trained network against an untrained
    one (learning rate ~0, so it keeps its initial weights): the trained
    cache's image is closer to the FP64 Cornell reference."""
    ref = load("fp64_cornell_512")

    def image(lr, warm):
        renderer = _cornell(6)
        cache = RadianceCache(renderer.context, seed=2, learning_rate=lr)
        for _ in range(warm):
            render_cached(renderer, cache, stats=False)
        renderer.reset_accumulation()
        for _ in range(8):
            render_cached(renderer, cache, stats=False)
        return renderer.target.download().reshape(-1, 4)[ref["pixels"], :3]

    trained = image(1e-3, 24)
    untrained = image(1e-30, 24)
    assert np.all(np.isfinite(trained)) and np.all(np.isfinite(untrained))
    assert rmse(trained, ref) < 0.7 * rmse(untrained, ref), (rmse(trained, ref), rmse(untrained, ref))


def _cached_run(ordered):
    renderer = _cornell(6)
    cache = RadianceCache(renderer.context, seed=2)
    for _ in range(16):
        render_cached(renderer, cache, stats=False, ordered_training=ordered)
    renderer.reset_accumulation()
    for _ in range(8):
        render_cached(renderer, cache, stats=False, ordered_training=ordered)
    return renderer.target.download().reshape(-1, 4)[:, :3]


def test_cached_frames_repeat_within_a_drift_bound():
    """Two runs of the same cached frames with the same seeds.  Paths are
    deterministic, but training records are appended in the order shading
    waves finish, so the records' chunks, the gradient partials and with
    them the weights can differ in their last bits from run to run.  The
    two images must differ by far less than either differs from the FP64
    reference."""
    ref = load("fp64_cornell_512")
    a, b = _cached_run(False), _cached_run(False)
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
    drift = float(np.sqrt(np.mean((a - b) ** 2)))
    error = rmse(a[ref["pixels"]], ref)
    print(f"run-to-run RMSE {drift:.3g}, identical pixels {np.mean(a == b):.4f}, RMSE vs FP64 {error:.4g}")
    assert drift < 0.05 * error, (drift, error)


def test_ordered_training_repeats_bit_for_bit():
    """With ordered_training the records are scattered by (path, record
    index), counted by a scan and gathered in that order before the Adam
    steps, so two runs give the same image bit for bit, and it stays as
    close to the FP64 reference as the default order's."""
    ref = load("fp64_cornell_512")
    a, b = _cached_run(True), _cached_run(True)
    assert np.all(np.isfinite(a))
    drift = float(np.sqrt(np.mean((a - b) ** 2)))
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (drift, float(np.mean(a == b)))
    error, default_error = rmse(a[ref["pixels"]], ref), rmse(_cached_run(False)[ref["pixels"]], ref)
    assert error < 1.1 * default_error, (error, default_error)


def test_cached_render_close_to_reference():
    ref = load("fp64_cornell_512")
    scene, camera = scenes.cornell_box()
    renderer = Renderer((512, 512), settings=RenderSettings(rng=_abi.TT_RNG_PCG, seed=2))
    renderer.set_scene(scene)
    renderer.camera = camera
    renderer.set_config(Config(bounce_count=4, sample_count=1, light_sampling=LightSampling.NEXT_EVENT_ESTIMATION))
    cache = RadianceCache(renderer.context, seed=1)
    for _ in range(24):  # warm up: train on 24 frames
        render_cached(renderer, cache, stats=False)
    stats = render_cached(renderer, cache)
    assert stats["samples"] > 30000 and stats["queries"] > 100000
    assert np.isfinite(stats["loss"])
    renderer.reset_accumulation()
    for _ in range(16):
        render_cached(renderer, cache)
    image = renderer.target.download().reshape(-1, 4)[ref["pixels"], :3]
    assert np.all(np.isfinite(image))
    assert abs(image.mean() / ref["radiance"].mean() - 1) < 0.05
    assert rmse(image, ref) < 0.1


def test_cached_restir_render_on_sphere_field():
    """ReSTIR at the primary vertex with the cache ending paths afterwards:
    the image stays finite and near the FP64 reference."""
    ref = load("fp64_sphere_field_1080p")
    scene, camera = scenes.sphere_field(seed=0)
    renderer = Renderer((1920, 1080), settings=RenderSettings(rng=_abi.TT_RNG_PCG, seed=3, restir_candidates=32))
    renderer.set_scene(scene)
    renderer.camera = camera
    renderer.set_config(Config(bounce_count=8, sample_count=1, light_sampling=LightSampling.RESTIR))
    cache = RadianceCache(renderer.context, seed=2)
    for _ in range(16):
        render_cached(renderer, cache, stats=False)
    renderer.reset_accumulation()
    for _ in range(8):
        stats = render_cached(renderer, cache)
    image = renderer.target.download().reshape(-1, 4)[ref["pixels"], :3]
    assert np.all(np.isfinite(image)) and stats["queries"] > 0
    assert abs(image.mean() / ref["radiance"].mean() - 1) < 0.15
