"""Serving loop on the GPU: a request's image equals the progressive render
of the same passes bit for bit, is the same alone and in any mix of other
requests (scenes, sizes, pass counts, configurations) and batch shapes, and
requests wait for the HBM budget instead of failing."""
import ctypes as C

import numpy as np
import pytest

from tinytrace_amd import Config, DeviceScene, LightSampling, Renderer, RenderSettings, _abi, scenes
from tinytrace_amd.api import build_constants
from tinytrace_amd.serving import Server, orbit_camera, request_mix, serve_poisson

pytestmark = pytest.mark.gpu

NEE = Config(bounce_count=3, sample_count=1, light_sampling=LightSampling.NEXT_EVENT_ESTIMATION)


def progressive(renderer, camera, config, width, height, passes, seed):
    """tt_render_ex with frame_index = k and accumulated_frame_count = k + 1."""
    lib = _abi.lib()
    settings = RenderSettings(rng=_abi.TT_RNG_PCG, seed=seed)
    for k in range(passes):
        c = build_constants(camera, config, width, height, k, k + 1)
        _abi.check(lib.tt_render_ex(renderer.context.handle, renderer.scene.handle, C.byref(c),
                                    C.byref(settings._desc()), renderer.target.handle), renderer.context.handle, "render")
    return renderer.target.download()


def test_request_equals_progressive_render():
    scene, camera = scenes.cornell_box()
    renderer = Renderer((64, 48))
    renderer.set_scene(scene)
    expected = progressive(renderer, camera, NEE, 64, 48, 4, seed=7)
    server = Server(renderer.context, pool_paths=5000)  # several batches per pass
    slot = server.add_scene(renderer.scene)
    rid = server.submit(slot, camera, NEE, 64, 48, 4, seed=7)
    server.run_until_done([rid])
    np.testing.assert_array_equal(server.take_result(rid), expected)


def test_alone_equals_in_mix():
    renderer = Renderer((8, 8))
    ctx = renderer.context
    cornell = [DeviceScene(ctx, scenes.cornell_box(v)[0]) for v in range(2)]
    spheres = DeviceScene(ctx, scenes.sphere_field(seed=3, spheres=30, lights=8, subdivisions=1)[0])
    _, camera = scenes.cornell_box()
    rng = np.random.default_rng(5)

    alone = Server(ctx, pool_paths=1 << 16)
    slots = [alone.add_scene(s) for s in cornell]
    rid = alone.submit(slots[1], camera, NEE, 80, 60, 6, seed=11)
    alone.run_until_done([rid])
    reference = alone.take_result(rid)

    mixed = Server(ctx, pool_paths=7000)
    slots = [mixed.add_scene(s) for s in cornell] + [mixed.add_scene(spheres)]
    others = []
    for i in range(4):
        others.append(mixed.submit(slots[i % 3], orbit_camera(rng), NEE, 40 + 8 * i, 30 + 4 * i, 1 + i, seed=i))
    on_hit = Config(bounce_count=2, sample_count=1, light_sampling=LightSampling.ON_HIT)
    others.append(mixed.submit(slots[0], camera, on_hit, 64, 64, 3, seed=1))
    target = mixed.submit(slots[1], camera, NEE, 80, 60, 6, seed=11)
    mixed.step()
    others.append(mixed.submit(slots[1], orbit_camera(rng), NEE, 100, 50, 2, seed=4))
    mixed.run_until_done(others + [target])
    np.testing.assert_array_equal(mixed.take_result(target), reference)
    assert mixed.stats()["batches"] > 10


def test_admission_waits_for_budget():
    renderer = Renderer((8, 8))
    scene, camera = scenes.cornell_box()
    pool = 4096
    target_bytes = 128 * 128 * 16
    path_bytes = 11 * 16 + 8
    server = Server(renderer.context, hbm_budget_bytes=pool * path_bytes + target_bytes + target_bytes // 2,
                    pool_paths=pool)
    slot = server.add_scene(DeviceScene(renderer.context, scene))
    ids = [server.submit(slot, camera, NEE, 128, 128, 2, seed=s) for s in range(3)]
    assert [server.status(i)["state"] for i in ids] == [_abi.TT_REQUEST_QUEUED] * 3
    server.run_until_done(ids)
    st = [server.status(i) for i in ids]
    for a, b in zip(st, st[1:]):
        assert b["admitted_ns"] >= a["done_ns"]
    with pytest.raises(_abi.TinytraceError) as e:
        server.submit(slot, camera, NEE, 256, 256, 1)
    assert e.value.status == 7  # TT_OVER_BUDGET


def test_poisson_mix_completes():
    renderer = Renderer((8, 8))
    ctx = renderer.context
    server = Server(ctx, pool_paths=1 << 18)
    slots = [server.add_scene(DeviceScene(ctx, scenes.cornell_box(v)[0])) for v in range(4)]
    specs = request_mix(24, rate_per_s=200.0, scenes=4, seed=1, resolutions=((64, 64), (128, 128), (192, 108),
                                                                              (384, 216)))
    out = serve_poisson(server, slots, specs, Config(bounce_count=4, sample_count=1,
                                                     light_sampling=LightSampling.NEXT_EVENT_ESTIMATION))
    assert out["requests"] == 24 and 0 < out["p50_ms"] <= out["p99_ms"]
    assert out["paths"] == sum(s.width * s.height * s.passes for s in specs)


RESTIR = Config(bounce_count=3, sample_count=1, light_sampling=LightSampling.RESTIR)


def _restir_scene():
    return scenes.sphere_field(seed=2, spheres=40, lights=24, subdivisions=2)


@pytest.mark.parametrize("pool,depth", [(20000, 0), (6144, 3), (6144, 1)],
                         ids=["passes-share-a-batch", "one-pass-per-batch-three-in-flight", "one-in-flight"])
def test_restir_request_equals_progressive_restir_render(pool, depth):
    """A ReSTIR request keeps its reservoirs across its passes: its image
    equals tt_render_ex's progressive ReSTIR render of the same passes on a
    fresh target, whether its passes share batches or follow each other on
    different lanes."""
    scene, camera = _restir_scene()
    renderer = Renderer((96, 64))
    renderer.set_scene(scene)
    expected = progressive(renderer, camera, RESTIR, 96, 64, 5, seed=3)
    server = Server(renderer.context, pool_paths=pool, batches_in_flight=depth)
    slot = server.add_scene(renderer.scene)
    rid = server.submit(slot, camera, RESTIR, 96, 64, 5, seed=3)
    server.run_until_done([rid])
    got = server.take_result(rid)
    assert np.isfinite(got).all() and got[..., :3].max() > 0
    np.testing.assert_array_equal(got, expected)


def test_restir_request_alone_equals_in_mix():
    scene, camera = _restir_scene()
    renderer = Renderer((8, 8))
    ctx = renderer.context
    restir_scene = DeviceScene(ctx, scene)
    cornell = DeviceScene(ctx, scenes.cornell_box()[0])
    rng = np.random.default_rng(9)

    alone = Server(ctx, pool_paths=1 << 15)
    s = alone.add_scene(restir_scene)
    rid = alone.submit(s, camera, RESTIR, 80, 60, 4, seed=21)
    alone.run_until_done([rid])
    reference = alone.take_result(rid)

    mixed = Server(ctx, pool_paths=12000)
    slots = [mixed.add_scene(restir_scene), mixed.add_scene(cornell)]
    others = [mixed.submit(slots[0], orbit_camera(rng, (50, 50, 50), 60.0), RESTIR, 64, 48, 3, seed=1),
              mixed.submit(slots[1], orbit_camera(rng), NEE, 100, 50, 2, seed=2)]
    target = mixed.submit(slots[0], camera, RESTIR, 80, 60, 4, seed=21)
    mixed.step()
    others.append(mixed.submit(slots[0], orbit_camera(rng, (50, 50, 50), 60.0), RESTIR, 40, 30, 6, seed=5,
                               restir_temporal=False))
    others.append(mixed.submit(slots[1], orbit_camera(rng), NEE, 64, 64, 3, seed=6))
    mixed.run_until_done(others + [target])
    np.testing.assert_array_equal(mixed.take_result(target), reference)


def test_restir_request_refusals():
    scene, camera = _restir_scene()
    renderer = Renderer((8, 8))
    server = Server(renderer.context, pool_paths=4096)
    slot = server.add_scene(DeviceScene(renderer.context, scene))
    with pytest.raises(_abi.TinytraceError) as e:
        server.submit(slot, camera, RESTIR, 128, 64, 1)  # a whole pass does not fit the pool
    assert e.value.status == 7  # TT_OVER_BUDGET
    textured, tcam = scenes.textured_cornell_box()
    tslot = server.add_scene(DeviceScene(renderer.context, textured))
    with pytest.raises(_abi.TinytraceError) as e:
        server.submit(tslot, tcam, NEE, 32, 32, 1)
    assert e.value.status == 8  # TT_UNSUPPORTED


# ------------------------------------------------------ radiance cache --
def _trained_cornell_cache(renderer, camera, frames=12):
    """A radiance cache trained by cached renders of the Cornell box."""
    from tinytrace_amd.cache import RadianceCache, render_cached

    renderer.camera = camera
    renderer.set_config(Config(bounce_count=4, sample_count=1, light_sampling=LightSampling.NEXT_EVENT_ESTIMATION))
    cache = RadianceCache(renderer.context, seed=3)
    for _ in range(frames):
        render_cached(renderer, cache, stats=False)
    renderer.target.download()
    return cache


@pytest.mark.parametrize("ordered", [False, True], ids=["slot_order", "ordered_training"])
def test_served_with_cache_queries_trains_and_stays_close_to_plain(ordered):
    """Requests on a scene slot with a radiance cache end their paths at cache
    queries, train the cache batch by batch (records in slot order, or
    gathered in path order first), report judge statistics, and keep the
    image's mean radiance near the plainly served one's."""
    scene, camera = scenes.cornell_box()
    renderer = Renderer((256, 256), settings=RenderSettings(rng=_abi.TT_RNG_PCG, seed=1))
    renderer.set_scene(scene)
    cache = _trained_cornell_cache(renderer, camera)
    before = cache.params
    config = Config(bounce_count=4, sample_count=1, light_sampling=LightSampling.NEXT_EVENT_ESTIMATION)
    rng = np.random.default_rng(2)
    cams = [camera, orbit_camera(rng), orbit_camera(rng)]
    sizes = [(256, 256, 6), (160, 120, 4), (96, 96, 8)]

    def serve(with_cache):
        server = Server(renderer.context, pool_paths=1 << 17)
        slot = server.add_scene(renderer.scene)
        if with_cache:
            server.set_cache(slot, cache, train_samples=8192, train_steps=2, ordered_training=ordered)
        ids = [server.submit(slot, cams[i], config, w, h, p, seed=5 + i) for i, (w, h, p) in enumerate(sizes)]
        server.run_until_done(ids)
        images = [server.take_result(i) for i in ids]
        return images, server.stats()

    plain, plain_stats = serve(False)
    cached, stats = serve(True)
    assert plain_stats["cache_queries"] == 0 and plain_stats["judge_samples"] == 0
    assert stats["cache_queries"] > 0 and stats["cache_samples"] > 0 and stats["judge_samples"] > 0
    assert abs(stats["judge_cached"] - stats["judge_traced"]) < 0.2 * stats["judge_traced"]
    assert not np.array_equal(cache.params, before)  # the batches trained it
    for (w, h, p), a, b in zip(sizes, cached, plain):
        assert np.isfinite(a).all() and a.shape == (h, w, 4)
        ma, mb = a[..., :3].mean(), b[..., :3].mean()
        assert abs(ma - mb) < 0.1 * mb, (w, h, p, ma, mb)
    cache.close()


def test_served_restir_with_cache_on_instanced_scene():
    """ReSTIR requests on the instanced field served with the cache (the
    instanced workload as specified: reservoirs resident, paths ended at the
    cache): finite images near the plainly served mean, queries made."""
    from tinytrace_amd.cache import RadianceCache

    scene, camera = scenes.instanced_field(seed=0)
    renderer = Renderer((8, 8), settings=RenderSettings(rng=_abi.TT_RNG_PCG, seed=1, restir_candidates=32))
    renderer.set_scene(scene)
    config = Config(bounce_count=12, sample_count=1, light_sampling=LightSampling.RESTIR)
    cache = RadianceCache(renderer.context, seed=0)

    def serve(with_cache, passes):
        server = Server(renderer.context, pool_paths=1 << 18)
        slot = server.add_scene(renderer.scene)
        if with_cache:
            server.set_cache(slot, cache)
        ids = [server.submit(slot, camera, config, 320, 180, passes, seed=9, restir_candidates=32),
               server.submit(slot, camera, config, 256, 256, passes, seed=4, restir_candidates=32)]
        server.run_until_done(ids)
        return [server.take_result(i) for i in ids], server.stats()

    serve(True, 24)  # trains the cache
    cached, stats = serve(True, 8)
    plain, _ = serve(False, 8)
    assert stats["cache_queries"] > 0 and stats["judge_samples"] > 0
    for a, b in zip(cached, plain):
        assert np.isfinite(a).all()
        # Robust statistics: a plainly traced pixel can carry a path whose
        # BSDF value over its floored pdf is 1e10 times the pixel's radiance
        # (near-specular lobes; the cache ends such paths).
        la, lb = a[..., :3].sum(-1).ravel(), b[..., :3].sum(-1).ravel()
        assert abs(np.median(la) - np.median(lb)) < 0.1 * np.median(lb)
        ta = np.sort(la)[: int(0.999 * la.size)].mean()
        tb = np.sort(lb)[: int(0.999 * lb.size)].mean()
        assert abs(ta - tb) < 0.15 * tb, (ta, tb)
    cache.close()


def test_set_cache_refusals():
    from tinytrace_amd import Context, TinytraceError
    from tinytrace_amd.cache import RadianceCache

    scene, _ = scenes.cornell_box()
    renderer = Renderer((8, 8))
    renderer.set_scene(scene)
    server = Server(renderer.context)
    slot = server.add_scene(renderer.scene)
    cache = RadianceCache(renderer.context, seed=0)
    with pytest.raises(TinytraceError):
        server.set_cache(slot + 1, cache)  # unknown slot
    with pytest.raises(TinytraceError):
        server.set_cache(slot, cache, train_steps=65)
    other = Context(0)
    foreign = RadianceCache(other, seed=0)
    with pytest.raises(TinytraceError):
        server.set_cache(slot, foreign)
    server.set_cache(slot, cache)
    server.set_cache(slot, None)  # plain again
    foreign.close()
    other.close()
    cache.close()
