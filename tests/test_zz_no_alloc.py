"""tt_trace_closest_static, the variant of tt_trace_closest that allocates
nothing: its hits equal tt_trace_closest's bit for bit, on two calls in a row
where the second is smaller (the scratch keeps the first call's contents),
and counts above TT_TRACE_STATIC_MAX_RAYS are refused."""
import numpy as np
import pytest

from test_gpu_parity import edge_rays, random_rays
from tinytrace_amd import DeviceScene, _abi, scenes
from tinytrace_amd._abi import lib, ptr

pytestmark = pytest.mark.gpu

MAX_RAYS = _abi.TT_TRACE_STATIC_MAX_RAYS


def trace_static(context, dev, rays, cull):
    rays = np.ascontiguousarray(rays, dtype=_abi.RAY_DTYPE)
    hits = np.zeros(rays.shape[0], dtype=_abi.HIT_DTYPE)
    _abi.check(lib().tt_trace_closest_static(context.handle, dev.handle, ptr(rays), ptr(hits), rays.shape[0],
                                             _abi.TT_TRACE_CULL_BACK_FACES if cull else 0),
               context.handle, "tt_trace_closest_static")
    return hits


def assert_same_bytes(a, b):
    assert a.shape == b.shape
    ab = a.view(np.uint8).reshape(a.shape[0], -1)
    bb = b.view(np.uint8).reshape(b.shape[0], -1)
    diff = np.nonzero((ab != bb).any(axis=1))[0]
    assert diff.size == 0, f"hits of {diff.size} rays differ, first {diff[:5]}"
    assert (a["instance"] != 0xFFFFFFFF).any()


@pytest.mark.parametrize("cull", [True, False])
def test_static_trace_equals_trace_sphere_field(context, cull):
    scene, _ = scenes.sphere_field(seed=5, spheres=60, lights=16, subdivisions=3)
    dev = DeviceScene(context, scene)
    rng = np.random.default_rng(40 + int(cull))
    first = np.concatenate([random_rays(rng, 60000, (50, 50, 50), 60.0), edge_rays(scene.arrays(), rng, 1)])
    first = first[:MAX_RAYS]
    second = random_rays(rng, 3001, (50, 50, 50), 60.0)
    for rays in (first, second):
        print(f"static trace: sphere field, cull={cull}, {rays.shape[0]} rays", flush=True)
        assert_same_bytes(trace_static(context, dev, rays, cull), dev.trace(rays, cull))


def test_static_trace_equals_trace_instanced(context):
    # The instanced field's two-level paths are the deepest, so its stacks
    # reach the spill area.
    scene, _ = scenes.instanced_field(seed=0)
    dev = DeviceScene(context, scene)
    rng = np.random.default_rng(42)
    for n in (MAX_RAYS, 777):
        rays = random_rays(rng, n, (100, 100, 100), 100.0)
        print(f"static trace: instanced field, {n} rays", flush=True)
        assert_same_bytes(trace_static(context, dev, rays, True), dev.trace(rays, True))


def test_static_trace_refuses_more_than_its_maximum(context):
    scene, _ = scenes.cornell_box()
    dev = DeviceScene(context, scene)
    rays = random_rays(np.random.default_rng(1), MAX_RAYS + 1, (0, 1, 0), 1.0)
    hits = np.zeros(rays.shape[0], dtype=_abi.HIT_DTYPE)
    status = lib().tt_trace_closest_static(context.handle, dev.handle, ptr(rays), ptr(hits), rays.shape[0], 0)
    assert status == 4  # TT_INVALID_ARGUMENT
    assert "at most 65536" in lib().tt_last_error(context.handle).decode()
    assert not hits.view(np.uint8).any()
    # Nothing to trace is not an error, and the entry point still works after
    # a refusal.
    assert lib().tt_trace_closest_static(context.handle, dev.handle, ptr(rays), ptr(hits), 0, 0) == 0
    small = rays[:1000]
    assert_same_bytes(trace_static(context, dev, small, False), dev.trace(small, False))
