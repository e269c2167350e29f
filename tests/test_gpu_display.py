"""The display pass on the GPU (tt_target_display) against the oracle's,
byte for byte, with and without the tonemap; its "post_process" timing."""
import numpy as np
import pytest

from tinytrace_amd import Config, LightSampling, Renderer, RenderSettings, _abi, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tonemap", [True, False])
def test_display_matches_oracle(tonemap):
    from oracle import binding

    scene, camera = scenes.cornell_box()
    r = Renderer((96, 80), settings=RenderSettings(rng=_abi.TT_RNG_PCG, seed=1))
    r.set_scene(scene)
    r.camera = camera
    r.set_config(Config(bounce_count=4, light_sampling=LightSampling.NEXT_EVENT_ESTIMATION, tonemap=tonemap))
    for _ in range(3):
        r.render_to_target()
    image = r.display()
    np.testing.assert_array_equal(image, binding.display(r.target.download(), tonemap))
    assert image[..., :3].max() > 0 and np.all(image[..., 3] == 255)
    assert r.timings()["post_process"] > 0.0
