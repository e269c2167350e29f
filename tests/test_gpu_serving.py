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
