"""glTF import (src/asset/gltf.rs) against an independent reading of the
same files: the reference's own Cornell box scene (tests/golden/
cornell_box.{gltf,bin}, copied from the reference's scenes/ folder as a
fixture) read here with json + numpy, and small synthetic documents for TRS
nodes, data-URI buffers and the refusals (textures: test_gltf_texture_cpu).

Tangents: the reference generates missing tangents with MikkTSpace
(mikktspace 0.3.0), which is not restated here, so for the UV-less Cornell
meshes the tangent direction is parity unpinned; the tests check what the
library promises instead (a unit tangent orthogonal to the normal, equal to
orthonormal(n) where the texture coordinates span no area)."""
import base64
import json
import os
import shutil

import numpy as np
import pytest

from oracle import binding
from tinytrace_amd import AssetScene, GltfImporter, Material, _abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CORNELL = os.path.join(GOLDEN, "cornell_box.gltf")

COMPONENT = {5126: np.float32, 5125: np.uint32, 5123: np.uint16}
WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}


def read_accessor(doc, buffers, index):
    """Accessor contents by the glTF 2.0 rules (view offset + accessor
    offset, byteStride or tight packing)."""
    a = doc["accessors"][index]
    view = doc["bufferViews"][a["bufferView"]]
    dtype = np.dtype(COMPONENT[a["componentType"]])
    width = WIDTH[a["type"]]
    size = dtype.itemsize * width
    stride = view.get("byteStride", size)
    start = view.get("byteOffset", 0) + a.get("byteOffset", 0)
    data = buffers[view["buffer"]]
    rows = [np.frombuffer(data, dtype=dtype, count=width, offset=start + i * stride) for i in range(a["count"])]
    return np.array(rows).reshape(a["count"], width)


def load(path):
    with open(path) as f:
        doc = json.load(f)
    base = os.path.dirname(path)
    buffers = []
    for b in doc["buffers"]:
        with open(os.path.join(base, b["uri"]), "rb") as f:
            buffers.append(f.read())
    return doc, buffers


def node_matrices(doc):
    """World matrix of every mesh node, depth first from the default scene."""
    out = []

    def visit(i, parent):
        n = doc["nodes"][i]
        m = np.array(n["matrix"], dtype=np.float64).reshape(4, 4).T if "matrix" in n else np.eye(4)
        world = parent @ m
        if "mesh" in n:
            out.append((n["mesh"], world))
        for c in n.get("children", []):
            visit(c, world)

    for root in doc["scenes"][doc.get("scene", 0)]["nodes"]:
        visit(root, np.eye(4))
    return out


def world_positions(arrays, instance):
    """Flattened snorm16 positions of an instance mapped to world space by its
    position transform."""
    inst = arrays["instances"][instance]
    mesh = arrays["meshes"][inst["mesh_index"]]
    q = arrays["positions"][mesh["vertex_offset"]:mesh["vertex_offset"] + mesh["vertex_count"]].astype(np.float64)
    q = np.maximum(q / 32767.0, -1.0)
    m = inst["transform"].reshape(4, 4).T.astype(np.float64)
    return (m @ np.concatenate([q, np.ones((len(q), 1))], axis=1).T).T[:, :3]


def decode_frames(encoded):
    out = np.zeros(7, dtype=np.float32)
    frames = []
    for e in encoded:
        binding.lib().oracle_decode_tangent_frame(int(e), out.ctypes.data)
        frames.append(out.copy())
    return np.array(frames)


@pytest.fixture(scope="module")
def cornell():
    return GltfImporter(CORNELL).new_scene().arrays()


def test_cornell_structure(cornell):
    doc, _ = load(CORNELL)
    assert len(cornell["meshes"]) == len(doc["meshes"]) == 8
    assert len(cornell["materials"]) == len(doc["materials"]) == 8
    assert len(cornell["instances"]) == 8
    # Depth first: node 2..9 hold meshes 0..7 with materials 0..7.
    np.testing.assert_array_equal(cornell["instances"]["mesh_index"], np.arange(8))
    np.testing.assert_array_equal(cornell["instances"]["material_index"], np.arange(8))
    # 6 quads and two 10-triangle boxes.
    assert len(cornell["indices"]) == 6 * 6 + 2 * 30


def test_cornell_indices_and_world_positions(cornell):
    doc, buffers = load(CORNELL)
    for inst, (mesh, world) in enumerate(node_matrices(doc)):
        prim = doc["meshes"][mesh]["primitives"][0]
        rec = cornell["meshes"][mesh]
        idx = cornell["indices"][rec["index_offset"]:rec["index_offset"] + rec["index_count"]]
        np.testing.assert_array_equal(idx, read_accessor(doc, buffers, prim["indices"]).reshape(-1))
        p = read_accessor(doc, buffers, prim["attributes"]["POSITION"]).astype(np.float64)
        expected = (world @ np.concatenate([p, np.ones((len(p), 1))], axis=1).T).T[:, :3]
        # snorm16 in the mesh's bounding sphere: |error| <= radius / 32767.
        np.testing.assert_allclose(world_positions(cornell, inst), expected, atol=rec["radius"] / 32767.0 * 1.01 + 1e-6)


def test_cornell_normals_and_generated_tangents(cornell):
    doc, buffers = load(CORNELL)
    for mesh in range(8):
        prim = doc["meshes"][mesh]["primitives"][0]
        rec = cornell["meshes"][mesh]
        n = read_accessor(doc, buffers, prim["attributes"]["NORMAL"]).astype(np.float64)
        frames = decode_frames(cornell["vertices"]["tangent_frame"][rec["vertex_offset"]:
                                                                     rec["vertex_offset"] + rec["vertex_count"]])
        np.testing.assert_allclose(frames[:, :3], n, atol=5e-3)  # 10-bit octahedral normals
        # No texture coordinates: the tangent is orthonormal(n), so its
        # 11-bit angle around the (quantized) normal's orthonormal vector is
        # 0 up to one step either way round, and the sign is the negated +1.
        t = frames[:, 3:6] / np.linalg.norm(frames[:, 3:6], axis=1, keepdims=True)
        assert np.abs(np.sum(t * frames[:, :3], axis=1)).max() < 1e-2
        angle = (cornell["vertices"]["tangent_frame"][rec["vertex_offset"]:rec["vertex_offset"] +
                                                       rec["vertex_count"]] >> 20) & 0x7ff
        assert np.all(np.minimum(angle, 2047 - angle) <= 1), angle
        np.testing.assert_array_equal(frames[:, 6], -1.0)


def test_cornell_materials_follow_the_gltf_mapping(cornell):
    """gltf.rs:242-381: base colour and metalness/roughness factors (glTF
    defaults 1), emissive factor (default 0) and KHR_materials_emissive_strength
    as the luminance; everything else keeps asset::Material's defaults."""
    doc, _ = load(CORNELL)
    for i, g in enumerate(doc["materials"]):
        pbr = g.get("pbrMetallicRoughness", {})
        expected = Material(
            base_color=tuple(pbr.get("baseColorFactor", [1, 1, 1, 1])[:3]),
            base_metalness=pbr.get("metallicFactor", 1.0),
            specular_roughness=pbr.get("roughnessFactor", 1.0),
            emission_color=tuple(g.get("emissiveFactor", [0, 0, 0])),
            emission_luminance=g.get("extensions", {}).get("KHR_materials_emissive_strength", {})
            .get("emissiveStrength", 0.0),
        )
        want = np.frombuffer(bytes(expected.constants()), dtype=np.float32)
        np.testing.assert_array_equal(cornell["materials"]["constants"][i], want)
    # Only the light (material 4, luminance 40) is emissive: its quad's two
    # triangles are the scene's emissive triangles.
    assert len(cornell["emissive_triangles"]) == 2
    assert set(cornell["emissive_triangles"]["instance_index"]) == {4}


def _write(tmp_path, doc, name="scene.gltf"):
    path = os.path.join(tmp_path, name)
    with open(path, "w") as f:
        json.dump(doc, f)
    return path


def _triangle_doc(node, uri):
    data = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float32).tobytes()
    return {
        "asset": {"version": "2.0"},
        "scene": 0,
        "scenes": [{"nodes": [0]}],
        "nodes": [dict(node, mesh=0)],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0}}]}],
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"}],
        "bufferViews": [{"buffer": 0, "byteLength": len(data)}],
        "buffers": [{"byteLength": len(data), "uri": uri(data)}],
    }


def test_trs_node_and_data_uri_buffer(tmp_path):
    """A data-URI buffer, an unindexed primitive with generated normals and
    the default material, and a node given as translation / rotation / scale:
    world = T R S p with R from the unit quaternion (x, y, z, w)."""
    angle = 0.7
    axis = np.array([1.0, 2.0, 2.0]) / 3.0
    quat = list(axis * np.sin(angle / 2)) + [np.cos(angle / 2)]
    node = {"translation": [1.0, -2.0, 0.5], "rotation": quat, "scale": [2.0, 0.5, 3.0]}
    uri = lambda b: "data:application/octet-stream;base64," + base64.b64encode(b).decode()
    path = _write(str(tmp_path), _triangle_doc(node, uri))
    a = GltfImporter(path).new_scene().arrays()
    assert len(a["instances"]) == 1 and len(a["indices"]) == 3
    x, y, z, w = quat
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r @ np.diag(node["scale"])
    m[:3, 3] = node["translation"]
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float64)
    expected = (m @ np.concatenate([p, np.ones((3, 1))], axis=1).T).T[:, :3]
    np.testing.assert_allclose(world_positions(a, 0), expected, atol=3e-4)
    default = np.frombuffer(bytes(Material().constants()), dtype=np.float32)
    np.testing.assert_array_equal(a["materials"]["constants"][0], default)
    # generate_normals: the face normal (0, 0, 1).
    frames = decode_frames(a["vertices"]["tangent_frame"])
    np.testing.assert_allclose(frames[:, :3], np.tile([0, 0, 1.0], (3, 1)), atol=5e-3)


def test_tangents_follow_texture_coordinates():
    """A planar quad with uv = xy: MikkTSpace's tangent is +x with sign +1
    (cross(n, t) = +y = the bitangent); the reference negates the sign."""
    s = AssetScene()
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], dtype=np.float32)
    s.insert_mesh(pos, pos[:, :2], [0, 1, 2, 0, 2, 3], np.tile([0, 0, 1.0], (4, 1)), None)
    m = s.insert_material(Material())
    s.add_model(s.add_node(), 0, m)
    frames = decode_frames(s.arrays()["vertices"]["tangent_frame"])
    t = frames[:, 3:6] / np.linalg.norm(frames[:, 3:6], axis=1, keepdims=True)
    np.testing.assert_allclose(t, np.tile([1.0, 0, 0], (4, 1)), atol=1e-2)
    np.testing.assert_array_equal(frames[:, 6], -1.0)
    # Mirrored mapping (u = -x): the tangent flips to -x and the sign with it.
    s2 = AssetScene()
    uv = np.stack([-pos[:, 0], pos[:, 1]], axis=1)
    s2.insert_mesh(pos, uv, [0, 1, 2, 0, 2, 3], np.tile([0, 0, 1.0], (4, 1)), None)
    frames = decode_frames(s2.arrays()["vertices"]["tangent_frame"])
    t = frames[:, 3:6] / np.linalg.norm(frames[:, 3:6], axis=1, keepdims=True)
    np.testing.assert_allclose(t, np.tile([-1.0, 0, 0], (4, 1)), atol=1e-2)
    np.testing.assert_array_equal(frames[:, 6], 1.0)


def _status(fn):
    try:
        fn()
    except _abi.TinytraceError as e:
        return e.status
    return 0


def test_refusals(tmp_path):
    tmp = str(tmp_path)
    unsupported, invalid = 8, 4
    # Binary glTF (the reference leaves it unimplemented).
    glb = os.path.join(tmp, "scene.gltf")
    with open(glb, "wb") as f:
        f.write(b"glTF\x02\x00\x00\x00" + b"\x00" * 16)
    assert _status(lambda: AssetScene().import_gltf(glb)) == unsupported
    # A texture reference with no textures array.
    shutil.copy(os.path.join(GOLDEN, "cornell_box.bin"), tmp)
    doc, _ = load(CORNELL)
    doc["materials"][0]["pbrMetallicRoughness"]["baseColorTexture"] = {"index": 0}
    assert _status(lambda: AssetScene().import_gltf(_write(tmp, doc, "textured.gltf"))) == invalid
    # Points instead of triangles.
    doc, _ = load(CORNELL)
    doc["meshes"][0]["primitives"][0]["mode"] = 0
    assert _status(lambda: AssetScene().import_gltf(_write(tmp, doc, "points.gltf"))) == unsupported
    # Wrong attribute format (verify_accessor).
    doc, _ = load(CORNELL)
    doc["accessors"][0]["type"] = "VEC4"
    assert _status(lambda: AssetScene().import_gltf(_write(tmp, doc, "format.gltf"))) == unsupported
    # Accessor past the end of its view, missing buffer file, malformed JSON.
    doc, _ = load(CORNELL)
    doc["accessors"][0]["count"] = 10000
    assert _status(lambda: AssetScene().import_gltf(_write(tmp, doc, "overrun.gltf"))) == invalid
    doc, _ = load(CORNELL)
    doc["buffers"][0]["uri"] = "missing.bin"
    assert _status(lambda: AssetScene().import_gltf(_write(tmp, doc, "nobuffer.gltf"))) == invalid
    bad = os.path.join(tmp, "bad.gltf")
    with open(bad, "w") as f:
        f.write('{"asset": {"version": "2.0"}, "nodes": [')
    assert _status(lambda: AssetScene().import_gltf(bad)) == invalid
    with pytest.raises(FileNotFoundError):
        GltfImporter(os.path.join(tmp, "absent.gltf"))
