"""tinytrace_amd: tinytrace's path-integration hot path on MI355X (gfx950).

The product is the native library libtinytrace_amd.so (HIP kernels + C ABI,
include/tinytrace.h); this package is its Python mirror of the reference's
Rust API and the procedural scene builders used by tests and the benchmark.
"""
from ._abi import LIBRARY_PATH, TinytraceError, lib
from .api import (AssetScene, Camera, CameraMove, Config, Context, DeviceScene, GltfImporter, LightSampling, Material,
                  Renderer, RenderSettings, Target, build_constants)
from .serving import Server

__all__ = [
    "LIBRARY_PATH", "TinytraceError", "lib", "AssetScene", "Camera", "CameraMove", "Config", "Context",
    "DeviceScene", "GltfImporter", "LightSampling", "Material", "Renderer", "RenderSettings", "Server", "Target", "build_constants",
]
