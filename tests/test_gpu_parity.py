"""GPU parity: the gfx950 kernels against the CPU oracle, through the C ABI.

Bars (integer/index work bit-exact; radiance within 1e-3 relative):
  * device scalar math (div, sqrt, sin, cos, exp, log, acos, pow): bit-identical
    to the host evaluation the oracle uses;
  * closest-hit instance, primitive, t and barycentrics: bit-exact against the
    oracle's brute force / BVH traversal;
  * BSDF prepare + evaluate: bit-exact against the oracle;
  * rendered radiance: every channel within 1e-3 relative (floor 1e-3) of the
    oracle at the same seeds.
"""
import math

import numpy as np
import pytest

from tinytrace_amd import Config, DeviceScene, LightSampling, Material, Renderer, RenderSettings, _abi, scenes
from tinytrace_amd._abi import lib, ptr

pytestmark = pytest.mark.gpu

REL_TOL = 1e-3


def _oracle():
    from oracle import binding

    return binding


def device_math(context, op, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=np.float32)
    out = np.zeros_like(x)
    _abi.check(lib().tt_debug_math(context.handle, op, ptr(x), ptr(y), ptr(out), x.size), context.handle, "math")
    return out


def test_device_arithmetic_is_ieee(context):
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(200000) * 10.0 ** rng.uniform(-20, 20, 200000)).astype(np.float32)
    y = (rng.standard_normal(200000) * 10.0 ** rng.uniform(-20, 20, 200000)).astype(np.float32)
    with np.errstate(all="ignore"):
        assert np.array_equal(device_math(context, 8, x, y).view(np.uint32), (x / y).view(np.uint32))
        assert np.array_equal(device_math(context, 9, x).view(np.uint32), (np.float32(1) / x).view(np.uint32))
        ax = np.abs(x)
        assert np.array_equal(device_math(context, 7, ax).view(np.uint32), np.sqrt(ax).view(np.uint32))


@pytest.mark.parametrize("name,op,lo,hi", [("sin", 0, -8.0, 8.0), ("cos", 1, -8.0, 8.0), ("exp", 2, -80.0, 80.0),
                                           ("log", 3, 1e-30, 1e6), ("acos", 4, -1.0, 1.0), ("asin", 5, -1.0, 1.0)])
def test_device_transcendentals_match_host(context, name, op, lo, hi):
    rng = np.random.default_rng(op)
    x = rng.uniform(lo, hi, 100000).astype(np.float32)
    if name == "log":
        x = np.exp(rng.uniform(math.log(lo), math.log(hi), 100000)).astype(np.float32)
    host = _oracle().math_f32(name, x)
    dev = device_math(context, op, x)
    assert np.array_equal(host.view(np.uint32), dev.view(np.uint32))


@pytest.mark.parametrize("name,op", [("sinpi", 11), ("cospi", 12)])
def test_device_sincospi_matches_host(context, name, op):
    rng = np.random.default_rng(op)
    x = np.concatenate([rng.uniform(-2, 2, 50000), rng.uniform(0, 4096, 50000)]).astype(np.float32)
    host = _oracle().math_f32(name, x)
    dev = device_math(context, op, x)
    assert np.array_equal(host.view(np.uint32), dev.view(np.uint32))


def random_rays(rng, n, center, spread, tmin=1e-3, tmax=1e4):
    rays = np.zeros(n, dtype=_abi.RAY_DTYPE)
    rays["origin"] = np.asarray(center) + rng.uniform(-spread, spread, (n, 3))
    d = rng.standard_normal((n, 3))
    rays["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["tmin"] = tmin
    rays["tmax"] = tmax
    return rays


def axis_rays(rng, n, center, spread, tmin=1e-3, tmax=1e4):
    """Rays whose directions have exact zero components: along the axes and
    in the coordinate planes (the box test must cull on such an axis by the
    ray's constant coordinate, and stay conservative)."""
    rays = random_rays(rng, n, center, spread, tmin, tmax)
    d = rng.standard_normal((n, 3))
    zero = rng.integers(0, 3, n)
    d[np.arange(n), zero] = 0.0
    axis = rng.random(n) < 0.3  # a third along one axis only
    keep = rng.integers(0, 3, n)
    d[axis] = 0.0
    d[axis, keep[axis]] = rng.choice([-1.0, 1.0], axis.sum())
    rays["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    # Some origins on integer and zero coordinates (planes of the scenes' boxes).
    snap = rng.random(n) < 0.2
    rays["origin"][snap] = np.round(rays["origin"][snap])
    return rays


def edge_rays(arrays, rng, n_per_tri=4):
    """Rays from random points aimed exactly at triangle vertices and edge
    midpoints (world space), where a non-watertight test would leak."""
    inst = arrays["instances"]
    pos = arrays["positions"].astype(np.float64) / 32767.0
    idx = arrays["indices"]
    meshes = arrays["meshes"]
    targets = []
    for i in range(inst.shape[0]):
        m = meshes[inst["mesh_index"][i]]
        t = inst["transform"][i].reshape(4, 4).T.astype(np.float64)
        tri = idx[m["index_offset"]:m["index_offset"] + m["index_count"]].reshape(-1, 3) + m["vertex_offset"]
        for a, b, c in tri[:64]:
            p = pos[[a, b, c]]
            world = (t[:3, :3] @ p.T).T + t[:3, 3]
            targets += [world[0], world[1], (world[0] + world[1]) / 2, (world[1] + world[2]) / 2]
    targets = np.array(targets)
    n = targets.shape[0] * n_per_tri
    rays = np.zeros(n, dtype=_abi.RAY_DTYPE)
    tg = np.repeat(targets, n_per_tri, axis=0)
    org = tg + rng.normal(0, 1.5, tg.shape)
    d = tg - org
    rays["origin"] = org
    rays["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["tmin"] = 1e-3
    rays["tmax"] = 1e4
    return rays


def assert_hits_equal(a, b):
    for field in ("instance", "primitive"):
        mismatch = np.nonzero(a[field] != b[field])[0]
        assert mismatch.size == 0, f"{field} differs for {mismatch.size} rays, first {mismatch[:5]}"
    hit = a["instance"] != 0xFFFFFFFF
    assert np.array_equal(a["t"][hit].view(np.uint32), b["t"][hit].view(np.uint32))
    assert np.array_equal(a["barycentrics"][hit].view(np.uint32), b["barycentrics"][hit].view(np.uint32))


@pytest.mark.parametrize("cull", [True, False])
def test_trace_cornell_bit_exact(context, cull):
    scene, camera = scenes.cornell_box()
    dev = DeviceScene(context, scene)
    oracle = _oracle().OracleScene(scene, 32, use_bvh=False)
    rng = np.random.default_rng(7)
    rays = np.concatenate([random_rays(rng, 50000, (0, 1, 0), 1.2), edge_rays(scene.arrays(), rng)])
    assert_hits_equal(dev.trace(rays, cull), oracle.trace(rays, cull))


@pytest.mark.parametrize("cull", [True, False])
def test_trace_sphere_field_bit_exact(context, cull):
    scene, camera = scenes.sphere_field(seed=3, spheres=60, lights=16, subdivisions=3)
    dev = DeviceScene(context, scene)
    oracle = _oracle().OracleScene(scene, 32, use_bvh=True)
    rng = np.random.default_rng(11)
    rays = np.concatenate([random_rays(rng, 100000, (50, 50, 50), 60.0), edge_rays(scene.arrays(), rng, 1)])
    assert_hits_equal(dev.trace(rays, cull), oracle.trace(rays, cull))


@pytest.mark.parametrize("cull", [True, False])
def test_trace_axis_aligned_rays_bit_exact(context, cull):
    scene, camera = scenes.sphere_field(seed=3, spheres=60, lights=16, subdivisions=3)
    dev = DeviceScene(context, scene)
    oracle = _oracle().OracleScene(scene, 32, use_bvh=True)
    rays = axis_rays(np.random.default_rng(12 + int(cull)), 100000, (50, 50, 50), 60.0)
    hits = dev.trace(rays, cull)
    assert (hits["instance"] != 0xFFFFFFFF).mean() > 0.05
    assert_hits_equal(hits, oracle.trace(rays, cull))


MATERIALS = {
    "diffuse": Material(base_color=(0.7, 0.5, 0.3), specular_weight=0.0, specular_roughness=1.0),
    "plastic": Material(base_color=(0.2, 0.6, 0.3), specular_roughness=0.4),
    "metal": Material(base_color=(0.9, 0.6, 0.3), base_metalness=1.0, specular_roughness=0.3),
    "coated": Material(base_color=(0.5, 0.1, 0.1), coat_weight=1.0, specular_roughness=0.6),
    "rough_diffuse": Material(base_color=(0.6, 0.6, 0.6), base_diffuse_roughness=0.7),
    "glass": Material(transmission_weight=1.0, specular_roughness=0.2),
    "fuzz": Material(fuzz_weight=0.8, fuzz_roughness=0.4),
    "anisotropic": Material(base_metalness=0.5, specular_roughness=0.5, specular_roughness_anisotropy=0.6,
                            specular_rotation=0.3),
}


@pytest.mark.parametrize("name", sorted(MATERIALS))
def test_bsdf_evaluate_bit_exact(context, name):
    mc = MATERIALS[name].constants()
    rng = np.random.default_rng(5)
    n = 20000
    wi = rng.standard_normal((n, 3))
    wi[:, 2] = np.abs(wi[:, 2]) * np.where(rng.random(n) < 0.9, 1, -1)
    wo = rng.standard_normal((n, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    q = np.zeros(n, dtype=_abi.BSDF_QUERY_DTYPE)
    q["wi"], q["wo"] = wi, wo
    out = np.zeros(n, dtype=_abi.BSDF_RESULT_DTYPE)
    import ctypes

    _abi.check(lib().tt_bsdf_evaluate(context.handle, ctypes.byref(mc), ptr(q), ptr(out), n, 0x12345678),
               context.handle, "bsdf")
    ref = _oracle().bsdf_evaluate(mc, q["wi"], q["wo"], rng_state=0x12345678)
    assert np.array_equal(np.nan_to_num(out["bsdf"]).view(np.uint32), np.nan_to_num(ref["bsdf"]).view(np.uint32))
    assert np.array_equal(np.nan_to_num(out["density"]).view(np.uint32), np.nan_to_num(ref["density"]).view(np.uint32))


def render_pair(scene, camera, config, width, height, passes, settings):
    """GPU passes through the Renderer and the oracle's on the same constants."""
    renderer = Renderer((width, height), settings=settings)
    renderer.set_scene(scene)
    renderer.camera = camera
    renderer.set_config(config)
    oracle = _oracle().OracleScene(scene, 32)
    reference = None
    image = None
    for _ in range(passes):
        k = renderer.constants()
        image = renderer.render_to_texture()
        reference, _ = oracle.render(k, rng=settings.rng, seed=settings.seed, previous=reference)
    return image, reference


def assert_radiance_close(image, reference):
    a, b = image[..., :3], reference[..., :3]
    assert np.all(np.isfinite(a)) == np.all(np.isfinite(b))
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
    worst = np.unravel_index(np.argmax(rel), rel.shape)
    assert rel.max() <= REL_TOL, f"max relative error {rel.max():.3g} at {worst}: {a[worst]} vs {b[worst]}"
    assert np.all(image[..., 3] == 1.0)


@pytest.mark.parametrize("light_sampling", [LightSampling.NEXT_EVENT_ESTIMATION, LightSampling.ON_HIT])
@pytest.mark.parametrize("rng_kind", [_abi.TT_RNG_REFERENCE, _abi.TT_RNG_PCG])
def test_render_cornell_matches_oracle(light_sampling, rng_kind):
    scene, camera = scenes.cornell_box()
    config = Config(bounce_count=4, sample_count=2, light_sampling=light_sampling)
    image, reference = render_pair(scene, camera, config, 96, 64, 3, RenderSettings(rng=rng_kind, seed=9))
    assert_radiance_close(image, reference)


def test_render_sphere_field_matches_oracle():
    scene, camera = scenes.sphere_field(seed=0, spheres=120, lights=32, subdivisions=3)
    config = Config(bounce_count=8, sample_count=1)
    image, reference = render_pair(scene, camera, config, 160, 90, 2, RenderSettings(rng=_abi.TT_RNG_PCG))
    assert_radiance_close(image, reference)


def restir_pair(scene, camera, config, width, height, frames, settings):
    """ReSTIR frames through the Renderer and the oracle; every frame's image
    is compared, so temporal and spatial reuse see the device's own history."""
    renderer = Renderer((width, height), settings=settings)
    renderer.set_scene(scene)
    renderer.camera = camera
    renderer.set_config(config)
    oracle = _oracle().OracleScene(scene, 32)
    state = oracle.restir_state(width, height)
    pairs = []
    for _ in range(frames):
        k = renderer.constants()
        renderer.reset_accumulation()
        image = renderer.render_to_texture()
        reference, _ = oracle.render_restir(k, state, rng=settings.rng, seed=settings.seed,
                                            candidates=settings.restir_candidates,
                                            temporal=bool(settings.restir_temporal),
                                            spatial=bool(settings.restir_spatial))
        pairs.append((image, reference))
    return pairs


@pytest.mark.parametrize("reuse", [(False, False), (True, False), (True, True)])
def test_render_restir_cornell_matches_oracle(reuse):
    scene, camera = scenes.cornell_box()
    config = Config(bounce_count=3, sample_count=1, light_sampling=LightSampling.RESTIR)
    settings = RenderSettings(rng=_abi.TT_RNG_PCG, seed=3, restir_candidates=8, restir_temporal=reuse[0],
                              restir_spatial=reuse[1])
    for image, reference in restir_pair(scene, camera, config, 64, 48, 3, settings):
        assert_radiance_close(image, reference)


def test_render_restir_sphere_field_matches_oracle():
    scene, camera = scenes.sphere_field(seed=1, spheres=80, lights=96, subdivisions=2)
    config = Config(bounce_count=4, sample_count=1, light_sampling=LightSampling.RESTIR)
    settings = RenderSettings(rng=_abi.TT_RNG_PCG, seed=0, restir_candidates=32, restir_temporal=True,
                              restir_spatial=True)
    for image, reference in restir_pair(scene, camera, config, 96, 64, 3, settings):
        assert_radiance_close(image, reference)


def test_render_is_deterministic():
    scene, camera = scenes.cornell_box()
    renderer = Renderer((64, 64))
    renderer.set_scene(scene)
    renderer.camera = camera
    k = renderer.constants()
    a = renderer.render_to_texture()
    renderer.frame_index = 0
    renderer.reset_accumulation()
    b = renderer.render_to_texture()
    assert np.array_equal(a, b)
    assert k.frame_index == 0


@pytest.mark.parametrize("light_sampling", [LightSampling.NEXT_EVENT_ESTIMATION, LightSampling.RESTIR])
def test_render_odd_size_matches_oracle(light_sampling):
    """37 x 29 pixels: the primary rays' 8-row bands end in a 5-row band and
    the pixel count is no multiple of a wave, so the tile order's last band
    and the passes' partial workgroups are both exercised (3 samples a pixel
    with next-event estimation; ReSTIR with both reuses, frame by frame)."""
    scene, camera = scenes.cornell_box()
    if light_sampling == LightSampling.RESTIR:
        config = Config(bounce_count=3, sample_count=1, light_sampling=light_sampling)
        settings = RenderSettings(rng=_abi.TT_RNG_PCG, seed=5, restir_candidates=8, restir_temporal=True,
                                  restir_spatial=True)
        for image, reference in restir_pair(scene, camera, config, 37, 29, 3, settings):
            assert_radiance_close(image, reference)
    else:
        config = Config(bounce_count=4, sample_count=3, light_sampling=light_sampling)
        image, reference = render_pair(scene, camera, config, 37, 29, 2, RenderSettings(rng=_abi.TT_RNG_PCG, seed=5))
        assert_radiance_close(image, reference)
