"""The benchmark configurations at their full sizes on the GPU:

* the 4M-triangle sphere field (1920x1080 workload's scene, seed 0): closest
  hits of 200 000 rays bit-exact against the oracle, including rays aimed at
  triangle vertices and edge midpoints;
* the serving mix at its real resolutions (256^2, 512^2, 1920x1080,
  3840x2160) and pass counts (1 to 64): every request's image equals the
  progressive render of the same passes bit for bit;
* the instanced scene at 3840x2160, depth 12, ReSTIR DI with the radiance
  cache: finite, its mean within 5 % of the FP64 fixture, and at equal
  frame time closer to the fixture than plain rendering.
"""
import time
from types import SimpleNamespace

import numpy as np
import pytest

from fp64_reference import load, rmse
from test_gpu_parity import assert_hits_equal, edge_rays, random_rays
from test_gpu_serving import NEE, progressive
from tinytrace_amd import Config, DeviceScene, LightSampling, Renderer, RenderSettings, Target, _abi, scenes
from tinytrace_amd.serving import Server, orbit_camera

pytestmark = pytest.mark.gpu


def _oracle():
    from oracle import binding

    return binding


@pytest.mark.parametrize("cull", [True, False])
def test_sphere_field_full_hits_bit_exact(context, cull):
    scene, _ = scenes.sphere_field(seed=0)
    dev = DeviceScene(context, scene)
    assert dev.info()["triangle_count"] == 4096258
    oracle = _oracle().OracleScene(scene, 32, use_bvh=True)
    rng = np.random.default_rng(31 + int(cull))
    edges = edge_rays(scene.arrays(), rng, 1)
    edges = edges[rng.choice(edges.shape[0], 50000, replace=False)]
    rays = np.concatenate([random_rays(rng, 150000, (50, 50, 50), 70.0), edges])
    print(f"sphere field: {rays.shape[0]} rays, cull={cull}", flush=True)
    hits = dev.trace(rays, cull)
    assert (hits["instance"] != 0xFFFFFFFF).mean() > 0.3
    assert_hits_equal(hits, oracle.trace(rays, cull))


def test_serving_at_the_mix_resolutions_equals_progressive_renders():
    """Requests at each resolution of the serving mix, with 1 to 64 passes,
    over two resident Cornell-box scenes and orbit cameras, served together:
    each equals its own progressive render bit for bit."""
    ctx_renderer = Renderer((8, 8))
    ctx = ctx_renderer.context
    resident = [DeviceScene(ctx, scenes.cornell_box(v)[0]) for v in range(2)]
    rng = np.random.default_rng(17)
    requests = [(0, 256, 256, 64), (1, 512, 512, 16), (0, 1920, 1080, 4), (1, 3840, 2160, 1), (0, 1920, 1080, 16)]
    server = Server(ctx, pool_paths=4 << 20)
    slots = [server.add_scene(s) for s in resident]
    cams = [orbit_camera(rng) for _ in requests]
    ids = [server.submit(slots[s], cams[i], NEE, w, h, p, seed=100 + i) for i, (s, w, h, p) in enumerate(requests)]
    server.run_until_done(ids)
    for i, (s, w, h, p) in enumerate(requests):
        print(f"request {i}: {w}x{h}, {p} passes", flush=True)
        got = server.take_result(ids[i])
        target = Target(ctx, w, h)
        want = progressive(SimpleNamespace(context=ctx, scene=resident[s], target=target), cams[i], NEE, w, h, p,
                           seed=100 + i)
        target.close()
        assert np.isfinite(got).all() and got[..., :3].max() > 0
        np.testing.assert_array_equal(got, want)


def test_instanced_restir_with_cache_at_2160p():
    from tinytrace_amd.cache import RadianceCache, render_cached

    ref = load("fp64_instanced_2160p")
    width, height, depth = int(ref["width"]), int(ref["height"]), int(ref["depth"])
    assert (width, height, depth) == (3840, 2160, 12)
    scene, camera = scenes.instanced_field(seed=0)
    renderer = Renderer((width, height), settings=RenderSettings(rng=_abi.TT_RNG_PCG, seed=3, restir_candidates=32))
    renderer.set_scene(scene)
    renderer.camera = camera
    renderer.set_config(Config(bounce_count=depth, sample_count=1, light_sampling=LightSampling.RESTIR))
    px = ref["pixels"]

    def frames(render, n):
        out, t = [], 0.0
        renderer.frame_index = 5000
        for _ in range(n):
            renderer.reset_accumulation()
            renderer.target.download()
            t0 = time.perf_counter()
            render()
            img = renderer.target.download()
            t += time.perf_counter() - t0
            out.append(img.reshape(-1, 4)[px, :3].astype(np.float64))
        return np.array(out), t / n

    cache = RadianceCache(renderer.context, seed=0)
    for _ in range(8):
        render_cached(renderer, cache, stats=False)
    cached, cached_s = frames(lambda: render_cached(renderer, cache, stats=False), 4)
    plain, plain_s = frames(renderer.render_to_target, 2)
    cache.close()
    print(f"2160p: cached {cached_s * 1e3:.1f} ms/frame, plain {plain_s * 1e3:.1f} ms/frame", flush=True)
    assert np.isfinite(cached).all()
    assert abs(cached.mean() / ref["radiance"].mean() - 1.0) < 0.05
    assert cached_s < plain_s
    # Equal time: two cached frames per plain frame.
    assert rmse(cached.mean(0), ref) < rmse(plain.mean(0), ref)
    assert rmse(cached[:2].mean(0), ref) < rmse(plain[0], ref)
