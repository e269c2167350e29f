"""Python mirror of tinytrace's library surface over the C ABI.

Names and argument meanings follow the reference's Rust API:
  Renderer       src/lib.rs:31-211   (new, set_scene, config, set_config, resize,
                                      render_to_target, render_to_texture,
                                      reset_accumulation, timings)
  Camera         src/camera.rs:6-72  (forward, right, move_by, different_from)
  Config         src/config.rs:23-43
  LightSampling  src/config.rs:4-12
  asset.Scene    src/asset/mod.rs:76-205 (insert_mesh, insert_material, root node)
  Material       src/asset/material.rs:8-175
Errors raise TinytraceError carrying the C status (mirrors tinytrace::Error).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import enum
import math
import os
from typing import List, Optional, Sequence

import numpy as np

from . import _abi
from ._abi import TinytraceError, check, lib, ptr


class LightSampling(enum.IntEnum):
    ON_HIT = 1
    NEXT_EVENT_ESTIMATION = 2
    RESTIR = 3


@dataclasses.dataclass
class Config:
    bounce_count: int = 4
    sample_count: int = 1
    tonemap: bool = True
    light_sampling: LightSampling = LightSampling.NEXT_EVENT_ESTIMATION


@dataclasses.dataclass
class CameraMove:
    translation: Sequence[float] = (0.0, 0.0, 0.0)
    yaw: float = 0.0
    pitch: float = 0.0


@dataclasses.dataclass
class Camera:
    position: Sequence[float] = (0.0, 0.0, 0.0)
    yaw: float = 0.0
    pitch: float = 0.0
    z_near: float = 0.1
    z_far: float = 400.0
    fov: float = math.pi / 2

    UP = (0.0, 1.0, 0.0)

    def forward(self) -> np.ndarray:
        f = np.array([math.cos(self.yaw) * math.cos(self.pitch), math.sin(self.pitch),
                      math.sin(self.yaw) * math.cos(self.pitch)], dtype=np.float32)
        return f / np.linalg.norm(f)

    def right(self) -> np.ndarray:
        r = -np.cross(self.forward(), np.array(self.UP, dtype=np.float32))
        return r / np.linalg.norm(r)

    def move_by(self, delta: CameraMove) -> None:
        self.position = tuple(float(p + d) for p, d in zip(self.position, delta.translation))
        self.yaw = math.fmod(self.yaw - delta.yaw, 2 * math.pi)
        self.pitch = min(max(self.pitch - delta.pitch, -1.553), 1.553)

    def different_from(self, other: "Camera") -> bool:
        t = 1e-4
        return (max(abs(a - b) for a, b in zip(self.position, other.position)) > t or abs(self.yaw - other.yaw) > t
                or abs(self.pitch - other.pitch) > t or abs(self.fov - other.fov) > t)

    @staticmethod
    def looking_at(position, target, fov=math.pi / 2) -> "Camera":
        """A camera at `position` whose primary rays go toward `target`.

        The reference's camera_ray_direction (integrate.glsl:17-20) sends the
        centre ray along -forward(), so forward points away from the target.
        """
        d = np.asarray(position, dtype=np.float64) - np.asarray(target, dtype=np.float64)
        d /= np.linalg.norm(d)
        return Camera(position=tuple(float(x) for x in position), yaw=math.atan2(d[2], d[0]),
                      pitch=math.asin(max(-1.0, min(1.0, d[1]))), fov=fov)

    def _desc(self) -> _abi.CameraDesc:
        d = _abi.CameraDesc()
        d.position[:] = [float(x) for x in self.position]
        d.yaw, d.pitch, d.z_near, d.z_far, d.fov = self.yaw, self.pitch, self.z_near, self.z_far, self.fov
        return d


def build_constants(camera: Camera, config: Config, width: int, height: int, frame_index: int,
                    accumulated_frame_count: int) -> _abi.Constants:
    """Constants as Renderer::render_to_target builds them (src/lib.rs:125-142)."""
    cfg = _abi.ConfigDesc(config.bounce_count, config.sample_count, int(bool(config.tonemap)),
                          int(config.light_sampling))
    k = _abi.Constants()
    check(lib().tt_constants_build(C.byref(camera._desc()), C.byref(cfg), width, height, frame_index,
                                   accumulated_frame_count, C.byref(k)), None, "tt_constants_build")
    return k


@dataclasses.dataclass
class Material:
    """asset::Material with constant parameters and the reference defaults."""
    base_weight: float = 1.0
    base_color: Sequence[float] = (0.8, 0.8, 0.8)
    base_metalness: float = 0.0
    base_diffuse_roughness: float = 0.0
    specular_weight: float = 1.0
    specular_color: Sequence[float] = (1.0, 1.0, 1.0)
    specular_roughness: float = 0.3
    specular_roughness_anisotropy: float = 0.0
    specular_ior: float = 1.5
    specular_rotation: float = 0.0
    transmission_weight: float = 0.0
    transmission_color: Sequence[float] = (1.0, 1.0, 1.0)
    transmission_depth: float = 0.0
    coat_weight: float = 0.0
    coat_color: Sequence[float] = (1.0, 1.0, 1.0)
    coat_roughness: float = 0.0
    coat_roughness_anisotropy: float = 0.0
    coat_ior: float = 1.6
    coat_darkening: float = 1.0
    coat_rotation: float = 0.0
    fuzz_weight: float = 0.0
    fuzz_color: Sequence[float] = (1.0, 1.0, 1.0)
    fuzz_roughness: float = 0.5
    emission_luminance: float = 0.0
    emission_color: Sequence[float] = (1.0, 1.0, 1.0)
    geometry_opacity: float = 1.0
    geometry_normal: Sequence[float] = (0.0, 0.0, 1.0)

    def _desc(self) -> _abi.MaterialDesc:
        d = _abi.MaterialDesc()
        for name, k in _abi.MATERIAL_DESC_FIELDS:
            v = getattr(self, name)
            if k == 1:
                setattr(d, name, float(v))
            else:
                getattr(d, name)[:] = [float(x) for x in v]
        return d

    def constants(self) -> _abi.MaterialConstants:
        """MaterialConstants as Material::constants builds them (material.rs:233-266)."""
        m = _abi.MaterialConstants()
        for name in ("base_color", "specular_color", "transmission_color", "coat_color", "fuzz_color",
                     "emission_color", "geometry_normal"):
            getattr(m, name)[:] = [float(x) for x in getattr(self, name)] + [0.0]
        for name, k in _abi.MATERIAL_DESC_FIELDS:
            if k == 1:
                setattr(m, name, float(getattr(self, name)))
        return m


class GltfImporter:
    """gltf::GltfImporter (src/asset/gltf.rs:44-72, 419-452): a .gltf file
    (external or base64 data-URI buffers) imported into an AssetScene."""

    def __init__(self, path):
        self.path = os.fspath(path)
        if not os.path.isfile(self.path):
            raise FileNotFoundError(self.path)

    def insert_in_scene(self, scene: "AssetScene") -> None:
        scene.import_gltf(self.path)

    def new_scene(self, block_compression: bool = True) -> "AssetScene":
        scene = AssetScene(block_compression)
        self.insert_in_scene(scene)
        return scene


class AssetScene:
    """asset::Scene: processed meshes, materials and a node tree (node 0 = root).
    Colour and normal textures are stored block-compressed (BC1 / BC5), as
    the reference stores them; block_compression=False keeps their exact
    quantized bytes (RGBA8 / RG8)."""

    def __init__(self, block_compression: bool = True):
        h = C.c_void_p()
        check(lib().tt_asset_scene_create(C.byref(h)), None, "tt_asset_scene_create")
        self._handle = h
        self._desc: Optional[_abi.SceneDesc] = None
        if not block_compression:
            check(lib().tt_asset_set_block_compression(h, 0), None, "tt_asset_set_block_compression")

    def __del__(self):
        if getattr(self, "_handle", None):
            lib().tt_asset_scene_destroy(self._handle)
            self._handle = None

    def insert_mesh(self, positions, texture_coordinates, indices=None, normals=None, tangents=None) -> int:
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        uv = np.ascontiguousarray(texture_coordinates, dtype=np.float32).reshape(-1, 2)
        d = _abi.MeshDesc()
        d.positions, d.texture_coordinates, d.vertex_count = ptr(pos), ptr(uv), pos.shape[0]
        keep = [pos, uv]
        if indices is not None:
            idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
            d.indices, d.index_count = ptr(idx), idx.size
            keep.append(idx)
        if normals is not None:
            n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
            d.normals = ptr(n)
            keep.append(n)
        if tangents is not None:
            t = np.ascontiguousarray(tangents, dtype=np.float32).reshape(-1, 4)
            d.tangents = ptr(t)
            keep.append(t)
        index = C.c_uint32()
        check(lib().tt_asset_insert_mesh(self._handle, C.byref(d), C.byref(index)), None, "insert_mesh")
        self._desc = None
        return index.value

    def insert_material(self, material: Material, textures: Optional[dict] = None) -> int:
        """Scene::insert_material; `textures` maps channel names
        (_abi.TEXTURE_CHANNELS) to (H, W, 3) colour / normal or (H, W)
        scalar float arrays (Param::Texture)."""
        index = C.c_uint32()
        if not textures:
            check(lib().tt_asset_insert_material(self._handle, C.byref(material._desc()), C.byref(index)), None,
                  "insert_material")
        else:
            keep, descs = [], []
            for name, texels in textures.items():
                ch = _abi.TEXTURE_CHANNELS.index(name)
                t = np.ascontiguousarray(texels, dtype=np.float32)
                if t.ndim != (3 if ch <= 6 else 2) or (ch <= 6 and t.shape[2] != 3):
                    raise ValueError(f"texture {name}: expected shape (H, W{', 3' if ch <= 6 else ''})")
                keep.append(t)
                descs.append(_abi.TextureDesc(ch, t.shape[1], t.shape[0], t.ctypes.data))
            arr = (_abi.TextureDesc * len(descs))(*descs)
            check(lib().tt_asset_insert_material_textured(self._handle, C.byref(material._desc()), arr, len(descs),
                                                          C.byref(index)), None, "insert_material")
        self._desc = None
        return index.value

    def add_node(self, transform=None, parent: int = 0) -> int:
        m = np.eye(4, dtype=np.float32) if transform is None else np.asarray(transform, dtype=np.float32)
        cols = np.ascontiguousarray(m.T.reshape(-1))  # column-major
        index = C.c_uint32()
        check(lib().tt_asset_add_node(self._handle, parent, ptr(cols), C.byref(index)), None, "add_node")
        self._desc = None
        return index.value

    def add_model(self, node: int, mesh: int, material: int) -> None:
        check(lib().tt_asset_add_model(self._handle, node, mesh, material), None, "add_model")
        self._desc = None

    def import_gltf(self, path) -> None:
        """Append a .gltf file's default scene under the root (GltfImporter::insert_in_scene)."""
        check(lib().tt_asset_import_gltf(self._handle, os.fsencode(path)), None, "tt_asset_import_gltf")
        self._desc = None

    def flatten(self) -> _abi.SceneDesc:
        """The tt_scene_desc that scene::Scene::new would upload (valid while this object lives)."""
        if self._desc is None:
            d = _abi.SceneDesc()
            check(lib().tt_asset_flatten(self._handle, C.byref(d)), None, "flatten")
            self._desc = d
        return self._desc

    def arrays(self) -> dict:
        """numpy views (copies) of the flattened reference-layout arrays."""
        d = self.flatten()

        def view(p, count, dtype):
            if count == 0:
                return np.zeros(0, dtype=dtype)
            buf = (C.c_char * (count * np.dtype(dtype).itemsize)).from_address(p)
            return np.frombuffer(buf, dtype=dtype).copy()

        return {
            "positions": view(d.positions, d.position_count * 3, np.int16).reshape(-1, 3),
            "vertices": view(d.vertices, d.vertex_count, _abi.VERTEX_DTYPE),
            "indices": view(d.indices, d.index_count, np.uint32),
            "meshes": view(d.meshes, d.mesh_count, _abi.MESH_DTYPE),
            "materials": view(d.materials, d.material_count, _abi.MATERIAL_DTYPE),
            "instances": view(d.instances, d.instance_count, _abi.INSTANCE_DTYPE),
            "emissive_triangles": view(d.emissive_triangles, d.emissive_triangle_count, _abi.EMISSIVE_DTYPE),
            "textures": view(d.textures, d.texture_count, _abi.TEXTURE_RECORD_DTYPE),
            "texture_data": view(d.texture_data, d.texture_data_bytes, np.uint8),
        }


class Context:
    """A tt_context: one HIP device and stream."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        desc = _abi.ContextDesc(device, 0)
        check(lib().tt_context_create(C.byref(desc), C.byref(h)), None, "tt_context_create")
        self.handle = h
        _abi.track(self, 2)

    def close(self):
        if getattr(self, "handle", None):
            lib().tt_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class DeviceScene:
    """A resident scene (tt_scene): arrays uploaded and BVH built, with the
    host binned-SAH builder ("sah") or the device LBVH builder ("lbvh")."""

    BUILDERS = {"sah": 0, "lbvh": 1}

    def __init__(self, context: Context, asset: AssetScene, builder: str = "sah"):
        self.context = context
        h = C.c_void_p()
        opts = _abi.SceneOptions(self.BUILDERS[builder], 0)
        check(lib().tt_scene_create_ex(context.handle, C.byref(asset.flatten()), C.byref(opts), C.byref(h)),
              context.handle, "tt_scene_create_ex")
        self.handle = h
        _abi.track(self, 1)

    def info(self) -> dict:
        i = _abi.SceneInfo()
        check(lib().tt_scene_info_get(self.handle, C.byref(i)), None, "tt_scene_info_get")
        return {f: getattr(i, f) for f, _ in _abi.SceneInfo._fields_}

    def close(self):
        if getattr(self, "handle", None):
            lib().tt_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def trace(self, rays: np.ndarray, cull: bool) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=_abi.RAY_DTYPE)
        hits = np.zeros(rays.shape[0], dtype=_abi.HIT_DTYPE)
        check(lib().tt_trace_closest(self.context.handle, self.handle, ptr(rays), ptr(hits), rays.shape[0],
                                     _abi.TT_TRACE_CULL_BACK_FACES if cull else 0), self.context.handle, "trace")
        return hits


class Target:
    def __init__(self, context: Context, width: int, height: int):
        self.context = context
        self.width, self.height = width, height
        h = C.c_void_p()
        check(lib().tt_target_create(context.handle, width, height, C.byref(h)), context.handle, "tt_target_create")
        self.handle = h
        _abi.track(self, 1)

    def download(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), dtype=np.float32)
        check(lib().tt_target_download(self.context.handle, self.handle, ptr(out), out.nbytes), self.context.handle,
              "tt_target_download")
        return out

    def close(self):
        if getattr(self, "handle", None):
            lib().tt_target_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


@dataclasses.dataclass
class RenderSettings:
    rng: int = _abi.TT_RNG_REFERENCE
    seed: int = 0
    restir_candidates: int = 32
    restir_temporal: int = 1
    restir_spatial: int = 1
    collect_stats: bool = False

    def _desc(self) -> _abi.RenderSettings:
        return _abi.RenderSettings(self.rng, self.seed, self.restir_candidates, self.restir_temporal,
                                   self.restir_spatial, _abi.TT_RENDER_COLLECT_STATS if self.collect_stats else 0)


class Renderer:
    """tinytrace::Renderer (src/lib.rs:31-211) on the gfx950 path.

    One deliberate difference: headless rendering here advances frame_index
    on every render_to_target (the reference advances it only in present(),
    crates/tinytrace_backend/src/lib.rs:393-410), so successive passes use
    fresh random numbers.  The progressive-accumulation formula is the
    reference's (integrate.glsl:341-346).
    """

    def __init__(self, extent=(1024, 1024), device: int = 0, settings: Optional[RenderSettings] = None):
        self.context = Context(device)
        self.extent = tuple(extent)
        self.target = Target(self.context, *self.extent)
        self.camera = Camera()
        self.config = Config()
        self.settings = settings or RenderSettings()
        self.accumulated_frame_count = 0
        self.frame_index = 0
        self._asset = AssetScene()
        self.scene = DeviceScene(self.context, self._asset)

    def set_scene(self, scene: AssetScene, builder: str = "sah") -> None:
        self.scene.close()
        self._asset = scene
        self.scene = DeviceScene(self.context, scene, builder)

    def set_config(self, config: Config) -> None:
        if config != self.config:
            self.reset_accumulation()
        self.config = dataclasses.replace(config)

    def resize(self, extent) -> None:
        self.target.close()
        self.extent = tuple(extent)
        self.target = Target(self.context, *self.extent)
        self.reset_accumulation()

    def reset_accumulation(self) -> None:
        self.accumulated_frame_count = 0

    def constants(self) -> _abi.Constants:
        return build_constants(self.camera, self.config, self.extent[0], self.extent[1], self.frame_index,
                               self.accumulated_frame_count)

    def render_to_target(self) -> None:
        k = self.constants()
        check(lib().tt_render_ex(self.context.handle, self.scene.handle, C.byref(k), C.byref(self.settings._desc()),
                                 self.target.handle), self.context.handle, "tt_render_ex")
        self.accumulated_frame_count += 1
        self.frame_index += 1

    def render_to_texture(self) -> np.ndarray:
        self.render_to_target()
        return self.target.download()

    def display(self) -> np.ndarray:
        """The post-process pass on the target (display.glsl, prepare_to_present):
        neutral_tonemap when config.tonemap, sRGB-encoded RGBA8 (H, W, 4)."""
        k = self.constants()
        out = np.empty((self.target.height, self.target.width, 4), dtype=np.uint8)
        check(lib().tt_target_display(self.context.handle, self.target.handle, C.byref(k), out.ctypes.data,
                                      out.nbytes), self.context.handle, "tt_target_display")
        return out

    def timings(self) -> dict:
        """Renderer::timings (src/lib.rs:167-178) in milliseconds: integrate,
        post_process (after display()) and, with collect_stats, trace."""
        out = {}
        for name in ("integrate", "post_process", "trace"):
            v = C.c_double()
            if lib().tt_timing_ms(self.context.handle, name.encode(), C.byref(v)) == _abi.TT_OK:
                out[name] = v.value
        return out

    def stats(self) -> dict:
        """tt_render_stats_get: ray counts and launch timings summed over the
        renders with collect_stats since the last call."""
        s = _abi.RenderStats()
        check(lib().tt_render_stats_get(self.context.handle, C.byref(s)), self.context.handle, "stats")
        return {f: getattr(s, f) for f, _ in _abi.RenderStats._fields_}
