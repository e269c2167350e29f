"""The texture path on the GPU against the oracle: the textured Cornell box
(colour, roughness, IOR, metalness and normal-map channels, trilinear
sampling with wrap, first-hit level of detail from ray differentials),
per pixel, for next-event estimation and hit-light-only, one and two samples
per pixel; emission textures on lights (NEE at level 4 at the sample's
coordinate) and on directly seen emitters; and ReSTIR DI on textured scenes
(textured primary vertices, G-buffer records with the texture coordinate and
level of detail, textured light emission), frame by frame."""
import math

import numpy as np
import pytest

from test_gpu_parity import assert_radiance_close, render_pair, restir_pair
from test_texture_cpu import _wall_and_split_light
from tinytrace_amd import (AssetScene, Camera, Config, DeviceScene, LightSampling, Material, Renderer, RenderSettings,
                           TinytraceError, _abi, scenes)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sampling,spp,rng", [(LightSampling.NEXT_EVENT_ESTIMATION, 1, _abi.TT_RNG_PCG),
                                              (LightSampling.ON_HIT, 1, _abi.TT_RNG_REFERENCE),
                                              (LightSampling.NEXT_EVENT_ESTIMATION, 2, _abi.TT_RNG_PCG)])
def test_render_textured_cornell_matches_oracle(sampling, spp, rng):
    scene, camera = scenes.textured_cornell_box()
    config = Config(bounce_count=4, sample_count=spp, light_sampling=sampling)
    image, reference = render_pair(scene, camera, config, 128, 112, 2, RenderSettings(rng=rng, seed=4))
    assert_radiance_close(image, reference)


@pytest.mark.parametrize("reuse", [(False, False), (True, True)])
def test_render_restir_textured_cornell_matches_oracle(reuse):
    scene, camera = scenes.textured_cornell_box()
    config = Config(bounce_count=3, sample_count=1, light_sampling=LightSampling.RESTIR)
    settings = RenderSettings(rng=_abi.TT_RNG_PCG, seed=5, restir_candidates=8, restir_temporal=reuse[0],
                              restir_spatial=reuse[1])
    for image, reference in restir_pair(scene, camera, config, 64, 48, 3, settings):
        assert_radiance_close(image, reference)


def test_render_restir_textured_lights_match_oracle():
    """ReSTIR with textured emitters: the candidates' targets and the final
    shading take each light sample's emission from its textures."""
    rng = np.random.default_rng(9)
    tex = rng.random((16, 16, 3)).astype(np.float32)
    light = Material(base_color=(0, 0, 0), specular_roughness=0.5, emission_luminance=4.0)
    scene = _wall_and_split_light([(light, {"emission_color": tex})] * 2, [(0.3, 0.6), (0.8, 0.2)])
    camera = Camera(position=(0.0, 0.0, 0.0), yaw=math.pi / 2, pitch=0.0, fov=math.pi / 2)
    config = Config(bounce_count=3, sample_count=1, light_sampling=LightSampling.RESTIR)
    settings = RenderSettings(rng=_abi.TT_RNG_PCG, seed=2, restir_candidates=8, restir_temporal=True,
                              restir_spatial=True)
    for image, reference in restir_pair(scene, camera, config, 64, 48, 3, settings):
        assert reference[..., :3].max() > 0.0
        assert_radiance_close(image, reference)


def test_textured_scene_loads(context):
    # Emission textures load.
    s = AssetScene()
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    s.insert_mesh(p, p[:, :2], [0, 1, 2], None, None)
    m = s.insert_material(Material(emission_luminance=1.0), {"emission_color": np.ones((2, 2, 3), np.float32)})
    s.add_model(s.add_node(), 0, m)
    DeviceScene(context, s)


@pytest.mark.parametrize("sampling", [LightSampling.NEXT_EVENT_ESTIMATION, LightSampling.ON_HIT])
def test_render_emission_textures_match_oracle(sampling):
    """A textured two-triangle light (checkerboard | constant colour, a
    luminance ramp) lighting a wall, and a camera-facing textured emitter in
    view, against the oracle per pixel."""
    rng = np.random.default_rng(8)
    tex = np.zeros((64, 64, 3), np.float32)
    yy, xx = np.mgrid[0:64, 0:32]
    tex[:, :32] = ((xx + yy) % 2)[..., None]
    tex[:, 32:] = rng.random(3)
    ramp = np.tile(np.linspace(2.0, 9.0, 64, dtype=np.float32), (64, 1))
    light = Material(base_color=(0, 0, 0), specular_roughness=0.5, emission_luminance=1.0)
    scene = _wall_and_split_light([(light, {"emission_color": tex, "emission_luminance": ramp})] * 2,
                                  [(0.3, 0.6), (0.8, 0.2)])
    # A small emitter in view with its texture spread over it.
    p = np.array([[-1.0, -1.0, -2.5], [-0.2, -1.0, -2.5], [-0.2, -0.2, -2.5], [-1.0, -0.2, -2.5]], np.float32)
    scene.insert_mesh(p, (p[:, :2] + 1.0) * 1.25, [0, 1, 2, 0, 2, 3], np.tile([0, 0, 1.0], (4, 1)), None)
    m = scene.insert_material(Material(base_color=(0.2, 0.2, 0.2), specular_roughness=0.4, emission_luminance=2.0),
                              {"emission_color": tex})
    scene.add_model(scene.add_node(), 3, m)
    camera = Camera(position=(0.0, 0.0, 0.0), yaw=math.pi / 2, pitch=0.0, fov=math.pi / 2)
    config = Config(bounce_count=3, sample_count=1, light_sampling=sampling)
    image, reference = render_pair(scene, camera, config, 96, 72, 2, RenderSettings(rng=_abi.TT_RNG_PCG, seed=5))
    assert reference[..., :3].max() > 1.0
    assert_radiance_close(image, reference)
