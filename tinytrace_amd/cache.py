"""Radiance cache (tt_mlp): the online-trained MLP and cached rendering.

`RadianceCache` wraps the network (parameters, encoding, inference,
gradients, Adam steps on host arrays); `render_cached` renders one frame
through a Renderer with the cache ending paths at their first rough vertex
after the primary one, judging it against plainly traced paths and
training it on the frame's training paths.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _abi
from ._abi import check, lib
from .api import Context, Renderer, RenderSettings

RECORD_FLOATS = 16


class RadianceCache:
    def __init__(self, context: Context, seed: int = 0, learning_rate: float = 1e-3):
        self.context = context
        h = C.c_void_p()
        desc = _abi.MlpDesc(seed, learning_rate, 0)
        check(lib().tt_mlp_create(context.handle, C.byref(desc), C.byref(h)), context.handle, "tt_mlp_create")
        self.handle = h
        _abi.track(self, 0)

    def close(self):
        if getattr(self, "handle", None):
            lib().tt_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def set_bounds(self, lo, hi) -> None:
        lo = (C.c_float * 3)(*lo)
        hi = (C.c_float * 3)(*hi)
        check(lib().tt_mlp_set_bounds(self.handle, lo, hi), self.context.handle, "tt_mlp_set_bounds")

    @property
    def params(self) -> np.ndarray:
        out = np.empty(_abi.MLP_PARAMS, dtype=np.float32)
        check(lib().tt_mlp_params(self.handle, out.ctypes.data, None, out.size), self.context.handle, "params")
        return out

    @params.setter
    def params(self, values) -> None:
        v = np.ascontiguousarray(values, dtype=np.float32)
        check(lib().tt_mlp_params(self.handle, None, v.ctypes.data, v.size), self.context.handle, "params")

    def _run(self, op, records, targets=None, out_shape=None):
        r = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, RECORD_FLOATS)
        t = None if targets is None else np.ascontiguousarray(
            np.pad(np.asarray(targets, dtype=np.float32).reshape(r.shape[0], -1)[:, :3], ((0, 0), (0, 1))))
        out = None if out_shape is None else np.empty(out_shape, dtype=np.float32)
        loss = C.c_double()
        check(lib().tt_mlp_run(self.handle, op, r.ctypes.data, None if t is None else t.ctypes.data, r.shape[0],
                               None if out is None else out.ctypes.data, C.byref(loss)), self.context.handle,
              "tt_mlp_run")
        return out, loss.value

    def encode(self, records) -> np.ndarray:
        n = np.asarray(records).reshape(-1, RECORD_FLOATS).shape[0]
        return self._run(0, records, out_shape=(n, _abi.MLP_IN))[0]

    def infer(self, records) -> np.ndarray:
        n = np.asarray(records).reshape(-1, RECORD_FLOATS).shape[0]
        return self._run(1, records, out_shape=(n, 3))[0]

    def gradient(self, records, targets):
        """(gradient of the mean relative-L2 loss, loss)."""
        return self._run(2, records, targets, out_shape=(_abi.MLP_PARAMS,))

    def train_step(self, records, targets) -> float:
        """One Adam step; returns the loss before it."""
        return self._run(3, records, targets)[1]


def render_cached(renderer: Renderer, cache: RadianceCache, train_samples: int = 65536, train_steps: int = 4,
                  judge_fraction: float = 0.01, records_per_path: int = 1, settings: Optional[RenderSettings] = None,
                  stats: bool = True, ordered_training: bool = False) -> Optional[dict]:
    """One frame of `renderer` with the radiance cache (accumulates like
    render_to_target).  With `stats` the call waits for the frame and its
    training and returns the frame's cache statistics; without, it returns
    None as soon as the frame is enqueued (frames then overlap like plain
    ones).  judge_fraction < 0: no judge paths.  ordered_training: train on
    the samples in path order, so repeated runs give the same images bit for
    bit."""
    k = renderer.constants()
    st = settings or renderer.settings
    cs = _abi.CacheSettings(train_samples, train_steps, judge_fraction, records_per_path, int(ordered_training))
    out = _abi.CacheStats() if stats else None
    check(lib().tt_render_cached(renderer.context.handle, renderer.scene.handle, C.byref(k), C.byref(st._desc()),
                                 cache.handle, C.byref(cs), renderer.target.handle,
                                 C.byref(out) if stats else None),
          renderer.context.handle, "tt_render_cached")
    renderer.accumulated_frame_count += 1
    renderer.frame_index += 1
    if not stats:
        return None
    return {f: getattr(out, f) for f, _ in _abi.CacheStats._fields_ if f != "reserved"}
