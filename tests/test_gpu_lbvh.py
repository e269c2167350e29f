"""The device LBVH builder: closest hits bit-exact against the oracle (the
traversal's result does not depend on the tree, only on conservative boxes),
renders equal to the host-SAH scene's, and it builds far faster."""
import numpy as np
import pytest

from test_gpu_parity import assert_hits_equal, assert_radiance_close, edge_rays, random_rays
from tinytrace_amd import Config, DeviceScene, LightSampling, Renderer, RenderSettings, _abi, scenes

pytestmark = pytest.mark.gpu


def _oracle():
    from oracle import binding

    return binding


@pytest.mark.parametrize("cull", [True, False])
def test_lbvh_trace_cornell_bit_exact(context, cull):
    scene, _ = scenes.cornell_box()
    dev = DeviceScene(context, scene, builder="lbvh")
    oracle = _oracle().OracleScene(scene, 32, use_bvh=False)
    rng = np.random.default_rng(3)
    rays = np.concatenate([random_rays(rng, 50000, (0, 1, 0), 1.2), edge_rays(scene.arrays(), rng)])
    assert_hits_equal(dev.trace(rays, cull), oracle.trace(rays, cull))


@pytest.mark.parametrize("cull", [True, False])
def test_lbvh_trace_sphere_field_bit_exact(context, cull):
    scene, _ = scenes.sphere_field(seed=3, spheres=60, lights=16, subdivisions=3)
    dev = DeviceScene(context, scene, builder="lbvh")
    oracle = _oracle().OracleScene(scene, 32, use_bvh=True)
    rng = np.random.default_rng(13)
    rays = np.concatenate([random_rays(rng, 100000, (50, 50, 50), 60.0), edge_rays(scene.arrays(), rng, 1)])
    assert_hits_equal(dev.trace(rays, cull), oracle.trace(rays, cull))


def test_lbvh_trace_instanced_bit_exact(context):
    scene, _ = scenes.instanced_field(seed=0)
    dev = DeviceScene(context, scene, builder="lbvh")
    oracle = _oracle().OracleScene(scene, 32, use_bvh=True)
    rng = np.random.default_rng(29)
    rays = random_rays(rng, 100000, (100, 100, 100), 110.0)
    assert_hits_equal(dev.trace(rays, True), oracle.trace(rays, True))


def test_lbvh_render_equals_sah_render():
    scene, camera = scenes.sphere_field(seed=1, spheres=120, lights=32, subdivisions=3)
    config = Config(bounce_count=6, sample_count=1, light_sampling=LightSampling.RESTIR)
    images = []
    for builder in ("sah", "lbvh"):
        r = Renderer((160, 90), settings=RenderSettings(rng=_abi.TT_RNG_PCG, seed=4))
        r.set_scene(scene, builder=builder)
        r.camera = camera
        r.set_config(config)
        for _ in range(3):
            image = r.render_to_texture()
        images.append(image)
    np.testing.assert_array_equal(images[0], images[1])


def test_lbvh_builds_faster_than_host_sah(context):
    scene, _ = scenes.sphere_field(seed=0)
    sah = DeviceScene(context, scene, builder="sah").info()
    lbvh = DeviceScene(context, scene, builder="lbvh").info()
    assert lbvh["unique_triangle_count"] == sah["unique_triangle_count"] == 4096258
    assert lbvh["build_ms"] < sah["build_ms"]
    # LBVH trees carry one triangle per leaf: more nodes, a higher SAH cost.
    assert lbvh["node_count"] > sah["node_count"] and lbvh["sah_cost"] > 0
    print("build ms: sah", sah["build_ms"], "lbvh", lbvh["build_ms"], "sah cost", sah["sah_cost"], lbvh["sah_cost"])
