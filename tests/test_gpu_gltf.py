"""The reference's own Cornell box scene (tests/golden/cornell_box.gltf),
imported by the library's glTF importer, rendered on the GPU and by the
oracle from the same flattened arrays: per-pixel parity for next-event
estimation, hit-light-only and ReSTIR."""
import os

import pytest

from test_gpu_parity import assert_radiance_close, render_pair, restir_pair
from tinytrace_amd import Config, GltfImporter, LightSampling, RenderSettings, _abi, scenes

pytestmark = pytest.mark.gpu

CORNELL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornell_box.gltf")


@pytest.mark.parametrize("sampling", [LightSampling.NEXT_EVENT_ESTIMATION, LightSampling.ON_HIT])
def test_render_gltf_cornell_matches_oracle(sampling):
    scene = GltfImporter(CORNELL).new_scene()
    _, camera = scenes.cornell_box()
    config = Config(bounce_count=4, sample_count=1, light_sampling=sampling)
    image, reference = render_pair(scene, camera, config, 128, 128, 3, RenderSettings(rng=_abi.TT_RNG_PCG, seed=3))
    assert image[..., :3].mean() > 0.01  # the camera sees the lit interior
    assert_radiance_close(image, reference)


def test_render_gltf_cornell_restir_matches_oracle():
    scene = GltfImporter(CORNELL).new_scene()
    _, camera = scenes.cornell_box()
    config = Config(bounce_count=4, sample_count=1, light_sampling=LightSampling.RESTIR)
    settings = RenderSettings(rng=_abi.TT_RNG_PCG, seed=2, restir_candidates=32)
    for image, reference in restir_pair(scene, camera, config, 96, 96, 3, settings):
        assert_radiance_close(image, reference)
