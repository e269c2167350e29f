"""Serving loop (tt_server) and the request-mix workload it is measured on.

`Server` wraps the C ABI's serving entry points: resident scenes, requests
admitted under an HBM budget and rendered by continuous batching over one
shared pool of path slots.  `serve_poisson` replays a Poisson arrival stream
of requests drawn from a resolution/sample-count mix against a server and
reports request latency percentiles.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import time
from typing import Dict, List, Sequence

import numpy as np

from . import _abi
from ._abi import check, lib
from .api import Camera, Config, Context, DeviceScene, build_constants


class Server:
    """tt_server: resident scenes + request queue + batch scheduler."""

    def __init__(self, context: Context, hbm_budget_bytes: int = 0, pool_paths: int = 0, batches_in_flight: int = 0):
        self.context = context
        h = C.c_void_p()
        desc = _abi.ServerDesc(hbm_budget_bytes, pool_paths, batches_in_flight)
        check(lib().tt_server_create(context.handle, C.byref(desc), C.byref(h)), context.handle, "tt_server_create")
        self.handle = h
        _abi.track(self, 0)
        self._scenes: List[DeviceScene] = []
        self._sizes: Dict[int, tuple] = {}
        self._caches: Dict[int, object] = {}  # caches serving scene slots, kept alive with the server

    def close(self):
        if getattr(self, "handle", None):
            lib().tt_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def add_scene(self, scene: DeviceScene) -> int:
        slot = C.c_uint32()
        check(lib().tt_server_add_scene(self.handle, scene.handle, C.byref(slot)), self.context.handle,
              "tt_server_add_scene")
        self._scenes.append(scene)  # keeps the scene alive while the server is
        return slot.value

    def set_cache(self, scene: int, cache, train_samples: int = 0, train_steps: int = 0, judge_fraction: float = 0.0,
                  records_per_path: int = 0, ordered_training: bool = False) -> None:
        """tt_server_set_cache: requests on scene slot `scene` are served
        with the radiance cache `cache` (a cache.RadianceCache; None: plainly
        again), which every batch of them trains (settings per batch as
        render_cached's per frame; 0: the defaults)."""
        settings = _abi.CacheSettings(train_samples, train_steps, judge_fraction, records_per_path,
                                      int(ordered_training))
        check(lib().tt_server_set_cache(self.handle, scene, None if cache is None else cache.handle,
                                        C.byref(settings)), self.context.handle, "tt_server_set_cache")
        if cache is None:
            self._caches.pop(scene, None)
        else:
            self._caches[scene] = cache

    def submit(self, scene: int, camera: Camera, config: Config, width: int, height: int, passes: int,
               seed: int = 0, rng: int = _abi.TT_RNG_PCG, restir_candidates: int = 32, restir_temporal: bool = True,
               restir_spatial: bool = True) -> int:
        """Queues a request; with config.light_sampling RESTIR the request
        keeps its reservoirs resident across its passes (the reuse settings
        as RenderSettings)."""
        k = build_constants(camera, dataclasses.replace(config, sample_count=1), width, height, 0, 0)
        q = _abi.ServerRequest(scene, rng, seed, passes, k, restir_candidates, int(restir_temporal),
                               int(restir_spatial), 0)
        rid = C.c_uint64()
        check(lib().tt_server_submit(self.handle, C.byref(q), C.byref(rid)), self.context.handle,
              "tt_server_submit")
        self._sizes[rid.value] = (width, height)
        return rid.value

    def step(self) -> int:
        done = C.c_uint32()
        check(lib().tt_server_step(self.handle, C.byref(done)), self.context.handle, "tt_server_step")
        return done.value

    def status(self, rid: int) -> dict:
        s = _abi.RequestStatus()
        check(lib().tt_server_status(self.handle, rid, C.byref(s)), self.context.handle, "tt_server_status")
        return {f: getattr(s, f) for f, _ in _abi.RequestStatus._fields_}

    def take_result(self, rid: int) -> np.ndarray:
        w, h = self._sizes.pop(rid)
        out = np.empty((h, w, 4), dtype=np.float32)
        check(lib().tt_server_take_result(self.handle, rid, out.ctypes.data, out.nbytes), self.context.handle,
              "tt_server_take_result")
        return out

    def stats(self) -> dict:
        s = _abi.ServerStats()
        check(lib().tt_server_stats_get(self.handle, C.byref(s)), self.context.handle, "tt_server_stats_get")
        return {f: getattr(s, f) for f, _ in _abi.ServerStats._fields_}

    def run_until_done(self, ids: Sequence[int]) -> None:
        pending = set(ids)
        while pending:
            self.step()
            pending = {i for i in pending if self.status(i)["state"] != _abi.TT_REQUEST_DONE}


# ------------------------------------------------------------------ workload --
RESOLUTIONS = ((256, 256), (512, 512), (1920, 1080), (3840, 2160))
RESOLUTION_P = (0.4, 0.3, 0.2, 0.1)
PASSES = (1, 4, 16, 64)


@dataclasses.dataclass
class RequestSpec:
    arrival_s: float
    scene: int
    width: int
    height: int
    passes: int
    camera: Camera
    seed: int


def orbit_camera(rng: np.random.Generator, center=(0.0, 1.0, 0.0), radius=3.5, fov=0.7) -> Camera:
    """A camera uniform on the sphere of `radius` around `center`, looking at it."""
    while True:
        v = rng.normal(size=3)
        n = np.linalg.norm(v)
        if n > 1e-6:
            break
    v = v / n
    # Keep clear of the poles, where yaw is undefined.
    v[1] = float(np.clip(v[1], -0.98, 0.98))
    v = v / np.linalg.norm(v)
    position = np.asarray(center) + radius * v
    return Camera.looking_at(tuple(position), center, fov=fov)


def request_mix(count: int, rate_per_s: float, scenes: int, seed: int = 0,
                resolutions=RESOLUTIONS, resolution_p=RESOLUTION_P, passes=PASSES) -> List[RequestSpec]:
    """Poisson arrivals; resolution and samples per pixel drawn from the mix;
    scene uniform over the resident scenes, which render with seed = scene."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += rng.exponential(1.0 / rate_per_s)
        w, h = resolutions[rng.choice(len(resolutions), p=resolution_p)]
        s = int(rng.integers(scenes))
        out.append(RequestSpec(t, s, w, h, int(passes[rng.integers(len(passes))]), orbit_camera(rng), seed=s))
    return out


def serve_poisson(server: Server, scene_slots: Sequence[int], specs: Sequence[RequestSpec], config: Config,
                  keep_images: bool = False, keep: Sequence[int] = ()) -> dict:
    """Replays `specs` against `server` in real time: a request is submitted
    once its arrival time has passed, the server steps while work is pending,
    and latency runs from the arrival time to the end of the download.
    Images are kept for every request (keep_images) or for the spec indices
    in `keep`; per-request latencies come back in spec order."""
    t0 = time.perf_counter()
    ids: Dict[int, int] = {}
    arrival_ns: Dict[int, int] = {}
    next_spec = 0
    done: Dict[int, int] = {}
    images = {}
    offset_ns = None
    while len(done) < len(specs):
        now = time.perf_counter() - t0
        while next_spec < len(specs) and specs[next_spec].arrival_s <= now:
            sp = specs[next_spec]
            before = time.perf_counter() - t0
            rid = server.submit(scene_slots[sp.scene], sp.camera, config, sp.width, sp.height, sp.passes, seed=sp.seed)
            if offset_ns is None:
                # Server clock (ns since the server was created) at this
                # script's t0; submit itself takes microseconds.
                offset_ns = server.status(rid)["submitted_ns"] - int(before * 1e9)
            ids[rid] = next_spec
            arrival_ns[rid] = offset_ns + int(sp.arrival_s * 1e9)
            next_spec += 1
        finished = server.step()
        if finished:
            for rid in list(ids):
                if rid in done:
                    continue
                st = server.status(rid)
                if st["state"] == _abi.TT_REQUEST_DONE:
                    done[rid] = st["done_ns"]
                    image = server.take_result(rid)
                    if keep_images or ids[rid] in keep:
                        images[ids[rid]] = image
        elif next_spec < len(specs) and not server.stats()["active"] and not server.stats()["queued"]:
            # Idle until the next arrival.
            wait = specs[next_spec].arrival_s - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.05))
    latency_ms = np.array([(done[r] - arrival_ns[r]) / 1e6 for r in ids])
    by_spec = [0.0] * len(specs)
    for r, i in ids.items():
        by_spec[i] = (done[r] - arrival_ns[r]) / 1e6
    wall_s = time.perf_counter() - t0
    stats = server.stats()
    result = {
        "requests": len(specs),
        "p50_ms": float(np.percentile(latency_ms, 50)),
        "p99_ms": float(np.percentile(latency_ms, 99)),
        "mean_ms": float(latency_ms.mean()),
        "max_ms": float(latency_ms.max()),
        "wall_s": wall_s,
        "paths": int(stats["paths"]),
        "mpaths_per_s": stats["paths"] / wall_s / 1e6,
        "busy_fraction": stats["busy_ms"] / 1e3 / wall_s,
        "batches": int(stats["batches"]),
    }
    result["latency_ms"] = by_spec
    if keep_images or keep:
        result["images"] = images
    return result
