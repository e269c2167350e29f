"""The texture path on the GPU against the oracle: the textured Cornell box
(colour, roughness, IOR, metalness and normal-map channels, trilinear
sampling with wrap, first-hit level of detail from ray differentials),
per pixel, for next-event estimation and hit-light-only, one and two samples
per pixel; and the refusals (ReSTIR and emission textures)."""
import numpy as np
import pytest

from test_gpu_parity import assert_radiance_close, render_pair
from tinytrace_amd import (AssetScene, Config, DeviceScene, LightSampling, Material, Renderer, RenderSettings,
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


def test_textured_scene_refusals(context):
    scene, camera = scenes.textured_cornell_box()
    r = Renderer((32, 32))
    r.set_scene(scene)
    r.camera = camera
    r.set_config(Config(light_sampling=LightSampling.RESTIR))
    with pytest.raises(TinytraceError) as e:
        r.render_to_target()
    assert e.value.status == 8  # TT_UNSUPPORTED
    s = AssetScene()
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    s.insert_mesh(p, p[:, :2], [0, 1, 2], None, None)
    m = s.insert_material(Material(emission_luminance=1.0), {"emission_color": np.ones((2, 2, 3), np.float32)})
    s.add_model(s.add_node(), 0, m)
    with pytest.raises(TinytraceError) as e:
        DeviceScene(context, s)
    assert e.value.status == 8  # TT_UNSUPPORTED
