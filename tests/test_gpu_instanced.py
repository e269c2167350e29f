"""Parity on the instanced scene (64 meshes of 5k-200k triangles, 12 000
instances, 4096 emissive triangles) and on the Cornell box at its full
benchmark configuration (512x512, 64 passes, depth 4)."""
import numpy as np
import pytest

from test_gpu_parity import assert_hits_equal, assert_radiance_close, axis_rays, random_rays, render_pair, restir_pair
from tinytrace_amd import Config, DeviceScene, LightSampling, RenderSettings, _abi, scenes

pytestmark = pytest.mark.gpu


def _oracle():
    from oracle import binding

    return binding


@pytest.fixture(scope="module")
def instanced():
    scene, camera = scenes.instanced_field(seed=0)
    return scene, camera, _oracle().OracleScene(scene, 32, use_bvh=True)


@pytest.mark.parametrize("cull", [True, False])
def test_trace_instanced_bit_exact(context, instanced, cull):
    scene, _, oracle = instanced
    dev = DeviceScene(context, scene)
    info = dev.info()
    assert info["unique_triangle_count"] == 3215512 and info["triangle_count"] == 596951286
    rng = np.random.default_rng(23)
    rays = np.concatenate([random_rays(rng, 200000, (100, 100, 100), 110.0),
                           axis_rays(rng, 100000, (100, 60, 100), 110.0)])
    assert_hits_equal(dev.trace(rays, cull), oracle.trace(rays, cull))


def test_render_instanced_matches_oracle(instanced):
    scene, camera, _ = instanced
    config = Config(bounce_count=12, sample_count=1, light_sampling=LightSampling.NEXT_EVENT_ESTIMATION)
    image, reference = render_pair(scene, camera, config, 160, 90, 2, RenderSettings(rng=_abi.TT_RNG_PCG, seed=5))
    assert_radiance_close(image, reference)


def test_render_instanced_restir_matches_oracle(instanced):
    scene, camera, _ = instanced
    config = Config(bounce_count=12, sample_count=1, light_sampling=LightSampling.RESTIR)
    settings = RenderSettings(rng=_abi.TT_RNG_PCG, seed=1, restir_candidates=32)
    for image, reference in restir_pair(scene, camera, config, 96, 54, 3, settings):
        assert_radiance_close(image, reference)


def test_render_cornell_full_configuration_matches_oracle():
    scene, camera = scenes.cornell_box()
    config = Config(bounce_count=4, sample_count=1, light_sampling=LightSampling.NEXT_EVENT_ESTIMATION)
    image, reference = render_pair(scene, camera, config, 512, 512, 64, RenderSettings(rng=_abi.TT_RNG_PCG, seed=0))
    assert_radiance_close(image, reference)
