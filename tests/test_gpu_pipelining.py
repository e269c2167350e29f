"""Renders enqueued back to back overlap on the device (consecutive renders
take two or three lanes in turn; with enough hardware queues each bounce's
shadow rays also overlap the next bounce's extension).  Their results must equal, bit for bit, renders that
are waited for one by one: progressive accumulation order, ReSTIR temporal
history across frames, two targets interleaved on one context, and API calls
(download, clear) between renders."""
import ctypes as C

import numpy as np
import pytest

from tinytrace_amd import Config, LightSampling, Renderer, RenderSettings, Target, _abi, scenes

pytestmark = pytest.mark.gpu

MODES = [LightSampling.NEXT_EVENT_ESTIMATION, LightSampling.RESTIR, LightSampling.ON_HIT]


@pytest.fixture(autouse=True, params=["4", "5", "8"],
                ids=["two-lanes", "three-lanes", "three-lanes-side-streams"])
def hardware_queues(request, monkeypatch):
    """Contexts created in the test take two or three lanes, with or without
    shadow side streams: the library decides from GPU_MAX_HW_QUEUES at
    context creation."""
    monkeypatch.setenv("GPU_MAX_HW_QUEUES", request.param)
    return request.param


def make_renderer(mode, size=(320, 200)):
    scene, camera = scenes.sphere_field(seed=1, spheres=60, lights=16, subdivisions=2)
    r = Renderer(size, settings=RenderSettings(rng=_abi.TT_RNG_PCG, seed=7))
    r.set_scene(scene)
    r.camera = camera
    r.set_config(Config(bounce_count=6, sample_count=1, light_sampling=mode))
    return r


def render_into(r, target):
    """One render of r's scene, camera and frame counters into `target`."""
    k = r.constants()
    _abi.check(_abi.lib().tt_render_ex(r.context.handle, r.scene.handle, C.byref(k), C.byref(r.settings._desc()),
                                       target.handle), r.context.handle, "tt_render_ex")


@pytest.mark.parametrize("mode", MODES)
def test_back_to_back_frames_equal_waited_frames(mode):
    frames = 7
    a = make_renderer(mode)
    for _ in range(frames):
        a.render_to_target()  # no wait in between
    pipelined = a.target.download()
    b = make_renderer(mode)
    for _ in range(frames):
        b.render_to_target()
        b.target.download()  # waits for the render
    waited = b.target.download()
    assert np.isfinite(pipelined).all() and pipelined[..., :3].max() > 0
    np.testing.assert_array_equal(pipelined, waited)


@pytest.mark.parametrize("mode", [LightSampling.RESTIR, LightSampling.NEXT_EVENT_ESTIMATION])
def test_two_targets_interleaved_on_one_context(mode):
    """Frames alternate between two targets of one context; each target's
    accumulation and ReSTIR history equal those of a renderer that renders it
    alone."""
    frames = 5
    r = make_renderer(mode)
    second = Target(r.context, *r.extent)
    cam2 = r.camera
    for _ in range(frames):
        render_into(r, r.target)
        render_into(r, second)  # same constants: same frame on both targets
        r.accumulated_frame_count += 1
        r.frame_index += 1
    got_a, got_b = r.target.download(), second.download()
    alone = make_renderer(mode)
    alone.camera = cam2
    for _ in range(frames):
        alone.render_to_target()
        alone.target.download()
    want = alone.target.download()
    np.testing.assert_array_equal(got_a, want)
    np.testing.assert_array_equal(got_b, want)


def render_mode_into(r, target, mode):
    """render_into with r's constants but another light-sampling mode."""
    k = r.constants()
    k.light_sampling = int(mode)
    _abi.check(_abi.lib().tt_render_ex(r.context.handle, r.scene.handle, C.byref(k), C.byref(r.settings._desc()),
                                       target.handle), r.context.handle, "tt_render_ex")


def test_restir_frames_interleaved_with_other_renders():
    """ReSTIR frames on one target alternate with NEE frames on a second
    target of the same context, so consecutive ReSTIR frames run on lanes
    with another render's lane between them.  Each ReSTIR frame's passes
    must still run after the previous ReSTIR frame's last read of the
    target's reservoirs: both targets equal frames waited for one by one."""
    frames = 6

    def run(wait):
        r = make_renderer(LightSampling.RESTIR)
        second = Target(r.context, *r.extent)
        for _ in range(frames):
            render_mode_into(r, r.target, LightSampling.RESTIR)
            if wait:
                r.target.download()
            render_mode_into(r, second, LightSampling.NEXT_EVENT_ESTIMATION)
            if wait:
                second.download()
            r.accumulated_frame_count += 1
            r.frame_index += 1
        return r.target.download(), second.download()

    got_restir, got_nee = run(wait=False)
    want_restir, want_nee = run(wait=True)
    assert np.isfinite(got_restir).all() and got_restir[..., :3].max() > 0
    np.testing.assert_array_equal(got_restir, want_restir)
    np.testing.assert_array_equal(got_nee, want_nee)


def test_api_calls_between_renders():
    """A download and an accumulation reset between back-to-back renders see
    the renders before them and not those after."""
    r = make_renderer(LightSampling.NEXT_EVENT_ESTIMATION)
    r.render_to_target()
    r.render_to_target()
    mid = r.target.download()
    r.reset_accumulation()  # the next render overwrites the target
    r.render_to_target()
    r.render_to_target()
    end = r.target.download()
    w = make_renderer(LightSampling.NEXT_EVENT_ESTIMATION)
    w.render_to_target()
    w.target.download()
    w.render_to_target()
    np.testing.assert_array_equal(mid, w.target.download())
    w.reset_accumulation()
    w.render_to_target()
    w.target.download()
    w.render_to_target()
    np.testing.assert_array_equal(end, w.target.download())


def test_statistics_sum_over_back_to_back_renders():
    """tt_render_stats_get sums the renders with collect_stats since the last
    read: three renders enqueued back to back count the same rays and
    launches as the same three renders read one at a time."""
    import dataclasses

    from tinytrace_amd import TinytraceError

    frames, size = 3, (320, 200)
    a = make_renderer(LightSampling.NEXT_EVENT_ESTIMATION, size)
    a.settings = dataclasses.replace(a.settings, collect_stats=True)
    for _ in range(frames):
        a.render_to_target()
    trace_ms = a.timings()["trace"]
    s = a.stats()
    b = make_renderer(LightSampling.NEXT_EVENT_ESTIMATION, size)
    b.settings = dataclasses.replace(b.settings, collect_stats=True)
    each = []
    for _ in range(frames):
        b.render_to_target()
        each.append(b.stats())
    paths = size[0] * size[1]
    assert s["paths"] == frames * paths
    for key in ("extension_rays", "shadow_rays", "extend_launches", "shadow_launches"):
        assert s[key] == sum(e[key] for e in each), key
    assert s["extension_rays"] >= frames * paths  # every path's primary ray
    assert s["extend_launches"] == frames * 6 and s["shadow_launches"] == frames * 6
    assert s["extend_ms"] > 0 and s["shadow_ms"] > 0 and s["integrate_ms"] > 0 and trace_ms > 0
    with pytest.raises(TinytraceError):
        a.stats()  # nothing rendered since the last read


def test_serving_batches_equal_progressive_render():
    """Serving batches (with shadow side streams when the context has them)
    give the progressive render's image bit for bit."""
    from test_gpu_serving import NEE, progressive
    from tinytrace_amd.serving import Server

    scene, camera = scenes.cornell_box()
    renderer = Renderer((64, 48))
    renderer.set_scene(scene)
    expected = progressive(renderer, camera, NEE, 64, 48, 4, seed=7)
    server = Server(renderer.context, pool_paths=5000)  # several batches per pass
    slot = server.add_scene(renderer.scene)
    rid = server.submit(slot, camera, NEE, 64, 48, 4, seed=7)
    server.run_until_done([rid])
    np.testing.assert_array_equal(server.take_result(rid), expected)
