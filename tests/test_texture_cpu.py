"""Texture path on the CPU: processing into the reference's formats and mip
chains (src/asset/texture.rs, material.rs:305-357), the oracle's sampler
(textureLod with linear filtering and repeat addressing) and its first-hit
level of detail (integrate.glsl:89-107, 134-139), pinned by closed forms.

Colour and normal textures are block-compressed (BC1 / BC5) by this
library's own encoder; tests of exact texel bytes build their scenes
uncompressed.  Parity unpinned: the bytes of the reference's encoder
(texpresso 2.0.1 ClusterFit, a dependency the reference does not vendor).
The BC decode is pinned by known-answer blocks in test_bc_cpu.py."""
import math

import numpy as np
import pytest

from oracle import binding
from tinytrace_amd import AssetScene, Camera, Config, Material, _abi, build_constants

CH = {name: i for i, name in enumerate(_abi.TEXTURE_CHANNELS)}


def quantize_unorm8(v):
    return np.floor(np.clip(np.asarray(v, np.float32), 0.0, 1.0) * np.float32(255.0) + np.float32(0.5)).astype(np.uint8)


def quad_scene(textures, material=None, depth=2.0, half=50.0, uv_scale=0.5, compress=False):
    """A camera-facing quad (normal +z) at z = -depth with uv = uv_scale * (x, y);
    colour and normal textures uncompressed unless `compress`."""
    s = AssetScene(block_compression=compress)
    p = np.array([[-half, -half, -depth], [half, -half, -depth], [half, half, -depth], [-half, half, -depth]],
                 np.float32)
    uv = p[:, :2] * uv_scale
    n = np.tile([0.0, 0.0, 1.0], (4, 1))
    t = np.tile([1.0, 0.0, 0.0, 1.0], (4, 1))
    s.insert_mesh(p, uv, [0, 1, 2, 0, 2, 3], n, t)
    m = s.insert_material(material or Material(), textures)
    s.add_model(s.add_node(), 0, m)
    return s


def texture_bytes(arrays, index, level):
    rec = arrays["textures"][index]
    texel = {0: 4, 1: 2, 2: 1, 3: 2}[int(rec["format"])]
    w, h = int(rec["width"]) >> level, int(rec["height"]) >> level
    o = int(rec["offsets"][level])
    return arrays["texture_data"][o:o + w * h * texel]


def test_colour_texture_levels_and_format():
    rng = np.random.default_rng(1)
    tex = rng.random((2, 4, 3)).astype(np.float32)  # H = 2, W = 4
    a = quad_scene({"base_color": tex}).arrays()
    rec = a["textures"][0]
    assert (rec["format"], rec["width"], rec["height"], rec["levels"]) == (_abi.TT_TEXTURE_FORMAT_RGBA8_SRGB, 4, 2, 2)
    assert np.all(rec["offsets"][:2] % 16 == 0)
    lvl0 = texture_bytes(a, 0, 0).reshape(2, 4, 4)
    np.testing.assert_array_equal(lvl0[..., :3], quantize_unorm8(tex))
    assert np.all(lvl0[..., 3] == 255)
    # Level 1 (2 x 1): ((((0 + t(2x, 0)) + t(2x, 1)) + t(2x + 1, 0)) + t(2x + 1, 1)) / 4 in float.
    avg = np.zeros((1, 2, 3), np.float32)
    for x in range(2):
        acc = np.zeros(3, np.float32)
        for dx, dy in ((0, 0), (0, 1), (1, 0), (1, 1)):
            acc = acc + tex[dy, 2 * x + dx]
        avg[0, x] = acc / np.float32(4.0)
    np.testing.assert_array_equal(texture_bytes(a, 0, 1).reshape(1, 2, 4)[..., :3], quantize_unorm8(avg))
    # The textured channel's constant is its type's default, zero.
    consts = a["materials"]["constants"][0]
    np.testing.assert_array_equal(consts[0:3], 0.0)
    assert a["materials"]["textures"][0][CH["base_color"]] == 0


def test_channel_formats_and_insertion_order():
    rng = np.random.default_rng(2)
    rough = rng.random((8, 8)).astype(np.float32)
    ior = rng.uniform(1.0, 2.5, (8, 8)).astype(np.float32)
    normals = rng.standard_normal((8, 8, 3)).astype(np.float32)
    normals /= np.linalg.norm(normals, axis=2, keepdims=True)
    a = quad_scene({"specular_ior": ior, "specular_roughness": rough, "geometry_normal": normals}).arrays()
    t = a["materials"]["textures"][0]
    # insert_material visits geometry_normal, then specular_roughness, then
    # (after the unorm scalars) specular_ior.
    assert (t[CH["geometry_normal"]], t[CH["specular_roughness"]], t[CH["specular_ior"]]) == (0, 1, 2)
    formats = a["textures"]["format"]
    assert list(formats) == [_abi.TT_TEXTURE_FORMAT_RG8_UNORM, _abi.TT_TEXTURE_FORMAT_R8_UNORM,
                             _abi.TT_TEXTURE_FORMAT_R16_FLOAT]
    assert list(a["textures"]["levels"]) == [4, 4, 4]  # 8, 4, 2, 1
    np.testing.assert_array_equal(texture_bytes(a, 1, 0).reshape(8, 8), quantize_unorm8(rough))
    np.testing.assert_array_equal(texture_bytes(a, 2, 0).view(np.float16).reshape(8, 8), ior.astype(np.float16))
    # Octahedral normals: decoding the bytes gives the normal back to ~1/255.
    rg = texture_bytes(a, 0, 0).reshape(8, 8, 2).astype(np.float64) / 255.0
    s = rg * 2 - 1
    n = np.stack([s[..., 0], s[..., 1], 1 - np.abs(s[..., 0]) - np.abs(s[..., 1])], axis=-1)
    tfold = np.clip(-n[..., 2], 0, 1)
    n[..., 0] += np.where(n[..., 0] >= 0, -tfold, tfold)
    n[..., 1] += np.where(n[..., 1] >= 0, -tfold, tfold)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    assert np.abs(n - normals).max() < 0.02


@pytest.fixture(scope="module")
def ramp_oracle():
    """R8 texture 16 x 4 whose texel (x, y) holds (x + 1) / 17 (scalar
    specular_roughness), so bilinear weights are readable from samples."""
    w, h = 16, 4
    tex = np.tile((np.arange(w, dtype=np.float32) + 1) / 17.0, (h, 1))
    scene = quad_scene({"specular_roughness": tex})
    return scene, binding.OracleScene(scene, 32), tex


def test_sampler_bilinear_wrap_and_levels(ramp_oracle):
    scene, o, tex = ramp_oracle
    a = scene.arrays()
    lvl0 = texture_bytes(a, 0, 0).reshape(4, 16).astype(np.float32) / np.float32(255.0)
    # Texel centres give the texel; between centres the linear blend.
    for x in range(15):
        for f in (0.0, 0.25, 0.5):
            u = (x + 0.5 + f) / 16.0
            got = o.sample_texture(0, u, 0.3, 0.0)[0]
            want = lvl0[0, x] * (1 - f) + lvl0[0, x + 1] * f
            assert abs(got - want) < 2e-6, (x, f)
    # Repeat addressing: between the last and the first texel, and u + k.
    edge = o.sample_texture(0, 15.75 / 16.0, 0.5, 0.0)[0]
    assert abs(edge - (lvl0[0, 15] * 0.75 + lvl0[0, 0] * 0.25)) < 2e-6
    for u in (0.13, 0.77):
        ref = o.sample_texture(0, u, 0.4, 0.0)
        for k in (-3.0, -1.0, 1.0, 7.0):
            np.testing.assert_allclose(o.sample_texture(0, u + k, 0.4 + k, 0.0), ref, atol=2e-6)
    # Level of detail: 0.5 blends levels 0 and 1 equally; below 0 (or NaN)
    # is level 0; beyond the last level clamps to it.
    u = 0.4
    l0, l1 = o.sample_texture(0, u, 0.5, 0.0)[0], o.sample_texture(0, u, 0.5, 1.0)[0]
    assert abs(o.sample_texture(0, u, 0.5, 0.5)[0] - 0.5 * (l0 + l1)) < 2e-6
    assert o.sample_texture(0, u, 0.5, -3.0)[0] == l0
    assert o.sample_texture(0, u, 0.5, float("nan"))[0] == l0
    top = int(a["textures"]["levels"][0]) - 1
    assert o.sample_texture(0, u, 0.5, 40.0)[0] == o.sample_texture(0, u, 0.5, float(top))[0]
    # R8 returns (value, 0, 0, 1).
    np.testing.assert_array_equal(o.sample_texture(0, u, 0.5, 0.0)[1:], [0.0, 0.0, 1.0])


def test_constant_texture_samples_exactly():
    tex = np.full((6, 10, 3), 0.5, np.float32)
    o = binding.OracleScene(quad_scene({"base_color": tex}), 32)
    code = int(quantize_unorm8(0.5))
    srgb = code / 255.0
    lin = srgb / 12.92 if srgb <= 0.04045 else ((srgb + 0.055) / 1.055) ** 2.4
    for u, v, lod in ((0.1, 0.9, 0.0), (3.7, -2.2, 1.3), (0.5, 0.5, 9.0)):
        s = o.sample_texture(0, u, v, lod)
        np.testing.assert_allclose(s[:3], np.float32(lin), rtol=0, atol=0)
        assert s[3] == 1.0


def test_first_hit_level_of_detail_closed_form():
    """A camera-facing quad at distance d with uv = k (x, y): one pixel over
    moves uv by k 2 d tan(fov/2) aspect / W in x (k 2 d tan(fov/2) / H in y);
    the reference scales those by the screen size, so at the screen centre
    lod = log2(k 2 d tan(fov/2) max(aspect, 1))."""
    w, h = 200, 100
    for d, k, fov in ((2.0, 0.5, math.pi / 2), (5.0, 0.3, 0.8), (1.0, 4.0, 1.2)):
        scene = quad_scene({"base_color": np.full((4, 4, 3), 0.5, np.float32)}, depth=d, uv_scale=k)
        o = binding.OracleScene(scene, 32)
        cam = Camera(position=(0.0, 0.0, 0.0), yaw=math.pi / 2, pitch=0.0, fov=fov)
        kc = build_constants(cam, Config(), w, h, 0, 0)
        lod = o.primary_lod(kc, 0.0, 0.0)
        want = math.log2(k * 2 * d * math.tan(fov / 2) * (w / h))
        assert abs(lod - want) < 2e-3, (d, k, fov, lod, want)


def srgb_decode(code):
    s = code / 255.0
    return s / 12.92 if s <= 0.04045 else ((s + 0.055) / 1.055) ** 2.4


def test_emission_texture_seen_directly():
    """material_emissive at a first hit (integrate.glsl:305): a camera-filling
    emitter whose colour texture is A on the left half and B on the right,
    with a luminance texture L: one bounce shows A L and B L away from the
    seam."""
    w, h = 96, 64
    a_col, b_col, lum = np.float32([0.2, 0.5, 0.9]), np.float32([0.7, 0.1, 0.4]), 3.0
    tex = np.zeros((32, 64, 3), np.float32)
    tex[:, :32] = a_col
    tex[:, 32:] = b_col
    scene = quad_scene({"emission_color": tex, "emission_luminance": np.full((32, 64), lum, np.float32)},
                       Material(base_color=(0, 0, 0)), depth=2.0, half=2.0, uv_scale=0.25)
    # uv = (x, y) / 4 + wrap: the quad x in [-2, 2] covers u in [-0.5, 0.5],
    # i.e. the texture's right half (u < 0 wraps) on x < 0 and its left
    # half on x > 0.
    o = binding.OracleScene(scene, 32)
    cam = Camera(position=(0.0, 0.0, 0.0), yaw=math.pi / 2, pitch=0.0, fov=math.pi / 2)
    k = build_constants(cam, Config(bounce_count=1, sample_count=1), w, h, 0, 0)
    img, _ = o.render(k, rng=_abi.TT_RNG_PCG, seed=1)
    want_a = np.float32([srgb_decode(c) for c in quantize_unorm8(a_col)]) * lum
    want_b = np.float32([srgb_decode(c) for c in quantize_unorm8(b_col)]) * lum
    # The quad spans the full height and the middle 2/3 of the width (the
    # field of view is vertical); screen x may grow either way, so find which
    # side shows which half.
    left, right = img[4:-4, 24:44, :3], img[4:-4, 52:72, :3]
    if abs(left[0, 0, 0] - want_a[0]) > abs(left[0, 0, 0] - want_b[0]):
        left, right = right, left
    np.testing.assert_allclose(left.reshape(-1, 3), np.tile(want_a, (left.size // 3, 1)), rtol=1e-5)
    np.testing.assert_allclose(right.reshape(-1, 3), np.tile(want_b, (right.size // 3, 1)), rtol=1e-5)


def _wall_and_split_light(light_materials, light_uvs):
    """A diffuse wall at z = -3 facing +z and a 1 x 1 emitter at z = -1.5
    facing it, made of two one-triangle meshes (materials and constant
    per-triangle texture coordinates given), outside a camera at the origin
    looking down -z with a 90 degree field of view (textures uncompressed,
    so the closed form holds to rounding)."""
    s = AssetScene(block_compression=False)
    wall = np.array([[-3, -3, -3], [3, -3, -3], [3, 3, -3], [-3, 3, -3]], np.float32)
    s.insert_mesh(wall, wall[:, :2] / 6.0, [0, 1, 2, 0, 2, 3], np.tile([0, 0, 1.0], (4, 1)), None)
    quad = np.array([[1.6, 0.2, -1.5], [2.6, 0.2, -1.5], [2.6, 1.2, -1.5], [1.6, 1.2, -1.5]], np.float32)
    for k, tri in enumerate(([0, 2, 1], [0, 3, 2])):  # wound to face -z
        p = quad[tri]
        uv = np.tile(np.float32(light_uvs[k]), (3, 1))
        s.insert_mesh(p, uv, [0, 1, 2], np.tile([0, 0, -1.0], (3, 1)), None)
    wall_m = s.insert_material(Material(base_color=(0.8, 0.8, 0.8), specular_roughness=0.5))
    s.add_model(s.add_node(), 0, wall_m)
    for k, (mat, tex) in enumerate(light_materials):
        s.add_model(s.add_node(), 1 + k, s.insert_material(mat, tex))
    return s


def test_textured_light_samples_level_4_at_the_sample_coordinate():
    """direct_light_contribution's material_emissive at the light sample's
    interpolated texture coordinate and level 4 (integrate.glsl:189,
    scene.glsl:130-133).  The emitter's texture is a 0/1 checkerboard on its
    left half and 0.25 on its right; level 4 (4 x 4) is 0.5 | 0.25, while
    level 0 at the chosen coordinates is a checker texel (0).  Each light
    triangle has one constant coordinate, so NEE must match, to rounding, a
    scene whose two light triangles have the constant emission
    decode(quantize(0.5)) and decode(quantize(0.25)) x L."""
    tex = np.zeros((64, 64, 3), np.float32)
    yy, xx = np.mgrid[0:64, 0:32]
    tex[:, :32] = ((xx + yy) % 2)[..., None]
    tex[:, 32:] = 0.25
    lum = 6.0
    uv_a, uv_b = (16.5 / 64, 0.5 + 0.5 / 64), (48.5 / 64, 0.5 + 0.5 / 64)
    textured = Material(base_color=(0, 0, 0), specular_roughness=0.5, emission_luminance=lum)
    a = _wall_and_split_light([(textured, {"emission_color": tex})] * 2, [uv_a, uv_b])
    ca = np.float32(srgb_decode(quantize_unorm8(0.5)))
    cb = np.float32(srgb_decode(quantize_unorm8(0.25)))
    const = lambda c: (Material(base_color=(0, 0, 0), specular_roughness=0.5, emission_luminance=lum,
                                emission_color=(c, c, c)), None)
    b = _wall_and_split_light([const(ca), const(cb)], [uv_a, uv_b])
    assert len(a.arrays()["emissive_triangles"]) == 2
    cam = Camera(position=(0.0, 0.0, 0.0), yaw=math.pi / 2, pitch=0.0, fov=math.pi / 2)
    k = build_constants(cam, Config(bounce_count=3, sample_count=1), 64, 48, 0, 0)
    img_a, _ = binding.OracleScene(a, 32).render(k, rng=_abi.TT_RNG_PCG, seed=2)
    img_b, _ = binding.OracleScene(b, 32).render(k, rng=_abi.TT_RNG_PCG, seed=2)
    assert img_b[..., :3].mean() > 0.01
    np.testing.assert_allclose(img_a, img_b, rtol=1e-5, atol=1e-7)
    # Emission from level 0 would be black on the first triangle: the two
    # triangles' contributions differ, so a swapped coordinate shows too.
    c = _wall_and_split_light([const(cb), const(ca)], [uv_a, uv_b])
    img_c, _ = binding.OracleScene(c, 32).render(k, rng=_abi.TT_RNG_PCG, seed=2)
    assert np.abs(img_c - img_b).max() > 1e-3
