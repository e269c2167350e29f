"""Block-compressed textures (BC1 colour, BC5 normal maps) on the CPU.

The reference stores colour textures as BC1 sRGB and normal maps as BC5
(src/asset/texture.rs:86-117) through texpresso's ClusterFit encoder, a
dependency it does not vendor: the bytes of that encoder are parity
unpinned.  This library encodes with its own encoder and decodes, in the
device sampler and in the oracle, by the public block formats as
include/tinytrace.h states them.  Pinned here:

* the oracle's decode against known-answer blocks worked out by hand
  (endpoint expansion by bit replication, the palette entries, texel
  addressing, BC1's three-colour mode, BC4's six-value mode);
* every decoded palette entry within half a step of the format's real-valued
  mix (2/3 c0 + 1/3 c1, (8 - k) / 7 e0 + (k - 1) / 7 e1, ...), so a wrong
  weight or index fails;
* the encoder's error on smooth and noisy textures, its determinism and the
  level layout, decoded through the oracle.
The device decode is pinned against the oracle by the textured render
parity tests (test_gpu_texture.py, test_gpu_gltf_texture.py).
"""
import numpy as np
import pytest

from oracle import binding
from tinytrace_amd import AssetScene, Material, _abi

BC1, BC5 = _abi.TT_TEXTURE_FORMAT_BC1_SRGB, _abi.TT_TEXTURE_FORMAT_BC5_UNORM


def bc1_block(c0, c1, indices):
    bits = 0
    for t, k in enumerate(indices):
        bits |= k << (2 * t)
    return bytes([c0 & 0xFF, c0 >> 8, c1 & 0xFF, c1 >> 8]) + int(bits).to_bytes(4, "little")


def bc4_block(e0, e1, indices):
    bits = 0
    for t, k in enumerate(indices):
        bits |= k << (3 * t)
    return bytes([e0, e1]) + int(bits).to_bytes(6, "little")


def decode_level(fmt, data, w, h):
    return np.array([[binding.block_texel(fmt, data, w, x, y) for x in range(w)] for y in range(h)])


def test_bc1_known_answers():
    red, blue = 0xF800, 0x001F
    # Four-colour mode (c0 > c1): indices 0, 1, 2, 3 repeating.
    blk = bc1_block(red, blue, [t % 4 for t in range(16)])
    want = {0: (255, 0, 0, 255), 1: (0, 0, 255, 255), 2: (170, 0, 85, 255), 3: (85, 0, 170, 255)}
    for t in range(16):
        np.testing.assert_array_equal(binding.block_texel(BC1, blk, 4, t % 4, t // 4), want[t % 4])
    # Three-colour mode (c0 <= c1): the midpoint and transparent black.
    blk = bc1_block(blue, red, [2, 3] + [0] * 14)
    np.testing.assert_array_equal(binding.block_texel(BC1, blk, 4, 0, 0), (128, 0, 128, 255))
    np.testing.assert_array_equal(binding.block_texel(BC1, blk, 4, 1, 0), (0, 0, 0, 0))
    np.testing.assert_array_equal(binding.block_texel(BC1, blk, 4, 2, 0), (0, 0, 255, 255))
    # Bit replication: r5 = 16 -> 132, g6 = 32 -> 130, b5 = 16 -> 132; g6 = 63 -> 255, r5 = 1 -> 8.
    blk = bc1_block(0x8410, 0x07E0 | (1 << 11), [0, 1] + [0] * 14)
    np.testing.assert_array_equal(binding.block_texel(BC1, blk, 4, 0, 0), (132, 130, 132, 255))
    np.testing.assert_array_equal(binding.block_texel(BC1, blk, 4, 1, 0), (8, 255, 0, 255))


def test_bc1_texel_addressing_across_blocks():
    """An 8 x 8 level is 2 x 2 blocks, row-major; texel (x, y) is texel
    4 (y % 4) + x % 4 of block (x / 4, y / 4)."""
    colours = [0xF800, 0x07E0, 0x001F, 0xFFFF]  # one endpoint colour per block
    data = b""
    for b in range(4):
        # Texel t takes index 0 (the block's colour) except t = 9, index 1 (black).
        data += bc1_block(colours[b], 0x0000, [1 if t == 9 else 0 for t in range(16)])
    img = decode_level(BC1, data, 8, 8)
    expect = {0: (255, 0, 0), 1: (0, 255, 0), 2: (0, 0, 255), 3: (255, 255, 255)}
    for y in range(8):
        for x in range(8):
            b = (y // 4) * 2 + x // 4
            t = (y % 4) * 4 + x % 4
            want = (0, 0, 0) if t == 9 else expect[b]
            np.testing.assert_array_equal(img[y, x, :3], want, err_msg=f"{x},{y}")


def test_bc5_known_answers():
    # Eight-value red channel (e0 > e1), six-value green channel (e0 <= e1).
    red = bc4_block(255, 0, [t % 8 for t in range(16)])
    green = bc4_block(0, 255, [t % 8 for t in range(16)])
    blk = red + green
    want_r = [255, 0, 219, 182, 146, 109, 73, 36]
    want_g = [0, 255, 51, 102, 153, 204, 0, 255]
    for t in range(16):
        c = binding.block_texel(BC5, blk, 4, t % 4, t // 4)
        assert (c[0], c[1], c[2], c[3]) == (want_r[t % 8], want_g[t % 8], 0, 255), t


def test_palettes_within_half_a_step_of_the_formats_mix():
    """Every palette entry, for random endpoints, within 0.5 (+ rounding
    ties) of the real-valued mix the formats define; a wrong weight or a
    shifted index misses by far more."""
    rng = np.random.default_rng(0)
    for _ in range(300):
        c0, c1 = (int(v) for v in rng.integers(0, 65536, 2))
        blk = bc1_block(c0, c1, list(range(4)) * 4)
        ends = []
        for c in (c0, c1):
            r5, g6, b5 = c >> 11, (c >> 5) & 63, c & 31
            exact = np.array([r5 * 255 / 31, g6 * 255 / 63, b5 * 255 / 31])
            # Bit replication is the endpoint's exact 8-bit value to within
            # 0.68 (5 bits) and 0.72 (6 bits).
            rep = np.array([(r5 << 3) | (r5 >> 2), (g6 << 2) | (g6 >> 4), (b5 << 3) | (b5 >> 2)], np.float64)
            assert np.all(np.abs(rep - exact) <= 0.72)
            ends.append(rep)
        a, b = ends
        if c0 > c1:
            mixes = [a, b, (2 * a + b) / 3, (a + 2 * b) / 3]
        else:
            mixes = [a, b, (a + b) / 2, np.zeros(3)]
        for k in range(4):
            got = binding.block_texel(BC1, blk, 4, k, 0)[:3].astype(np.float64)
            assert np.all(np.abs(got - mixes[k]) <= 0.5 + 1e-9), (c0, c1, k, got, mixes[k])
    for _ in range(300):
        e0, e1 = (int(v) for v in rng.integers(0, 256, 2))
        blk = bc4_block(e0, e1, list(range(8)) * 2) + bc4_block(0, 0, [0] * 16)
        if e0 > e1:
            mixes = [e0, e1] + [((8 - k) * e0 + (k - 1) * e1) / 7 for k in range(2, 8)]
        else:
            mixes = [e0, e1] + [((6 - k) * e0 + (k - 1) * e1) / 5 for k in range(2, 6)] + [0, 255]
        for k in range(8):
            got = binding.block_texel(BC5, blk, 4, k % 4, k // 4)[0]
            assert abs(got - mixes[k]) <= 0.5 + 1e-9, (e0, e1, k, got, mixes[k])


def test_known_answers_fail_a_slipped_decoder():
    """The known-answer blocks tell apart plausible slips: 3-bit indices read
    from the high end, BC1's interpolants swapped, or the palette mixes
    truncated rather than rounded."""
    red = bc4_block(255, 0, [t % 8 for t in range(16)])
    bits = int.from_bytes(red[2:8], "little")
    high_first = [(bits >> (45 - 3 * t)) & 7 for t in range(16)]
    assert high_first != [t % 8 for t in range(16)]
    assert (2 * 255 + 0) // 3 != 170 or (0 + 2 * 255) // 3 != 85  # truncation
    np.testing.assert_array_equal(
        binding.block_texel(BC1, bc1_block(0xF800, 0x001F, [2] * 16), 4, 0, 0)[:3], (170, 0, 85))


def _compressed(textures, compress=True):
    s = AssetScene(block_compression=compress)
    p = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0]], np.float32)
    s.insert_mesh(p, p[:, :2], [0, 1, 2], None, None)
    s.add_model(s.add_node(), 0, s.insert_material(Material(), textures))
    return s.arrays()


def _level(a, i, level):
    rec = a["textures"][i]
    w, h = max(1, int(rec["width"]) >> level), max(1, int(rec["height"]) >> level)
    o = int(rec["offsets"][level])
    per = 8 if rec["format"] == BC1 else 16
    n = ((w + 3) // 4) * ((h + 3) // 4) * per
    return a["texture_data"][o:o + n].tobytes(), w, h


def quantize(v):
    return np.floor(np.clip(v.astype(np.float32), 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.int64)


def test_layout_and_determinism():
    rng = np.random.default_rng(1)
    tex = rng.random((13, 22, 3)).astype(np.float32)
    normals = rng.standard_normal((9, 6, 3)).astype(np.float32)
    normals /= np.linalg.norm(normals, axis=2, keepdims=True)
    a = _compressed({"base_color": tex, "geometry_normal": normals})
    b = _compressed({"base_color": tex, "geometry_normal": normals})
    np.testing.assert_array_equal(a["texture_data"], b["texture_data"])
    recs = a["textures"]
    assert list(recs["format"]) == [BC1, BC5]
    assert np.all(recs["offsets"][0][: recs["levels"][0]] % 16 == 0)
    # Levels: 22 x 13, 11 x 6, 5 x 3, 2 x 1 -> blocks 6x4, 3x2, 2x1, 1x1.
    assert recs["levels"][0] == 4
    sizes = np.diff(np.append(recs["offsets"][0][:4], recs["offsets"][1][0]))
    assert list(sizes) == [6 * 4 * 8, 3 * 2 * 8, 2 * 1 * 8, 16]  # the last one padded to 16 B


def test_bc1_encoder_error():
    """Decoded against the quantized bytes: a gradient along one colour line
    within 6 (the 565 grid's steps are 8 and 4); a gradient spanning a plane
    of colours (not representable by one line a block) within 24, mean 6;
    independent random colours (BC1's worst case) within a mean of 50; a flat
    colour within the 565 rounding."""
    y, x = np.mgrid[0:32, 0:32].astype(np.float32)
    t = (x + y) / 62
    line = np.stack([0.1 + 0.8 * t, 0.2 + 0.5 * t, 0.9 - 0.6 * t], axis=-1)
    plane = np.stack([x / 31, y / 31, (x + y) / 62], axis=-1)
    rng = np.random.default_rng(3)
    noise = rng.random((32, 32, 3)).astype(np.float32)
    flat = np.full((8, 8, 3), 0.5, np.float32)
    for tex, max_err, mean_err in ((line, 6, 2.5), (plane, 24, 6), (noise, 255, 50), (flat, 5, 5)):
        a = _compressed({"base_color": tex})
        data, w, h = _level(a, 0, 0)
        got = decode_level(BC1, data, w, h)[..., :3].astype(np.int64)
        err = np.abs(got - quantize(tex))
        assert err.max() <= max_err and err.mean() <= mean_err, (err.max(), err.mean())
        assert np.all(decode_level(BC1, data, w, h)[..., 3] == 255)  # always four-colour or flat


def test_bc5_normal_map_error():
    """A smooth normal map through BC5 (octahedral R, G): every decoded
    normal within 2 degrees of the input."""
    y, x = np.mgrid[0:16, 0:16].astype(np.float32)
    n = np.stack([np.sin(x / 10) * 0.5, np.cos(y / 7) * 0.5, np.ones_like(x)], axis=-1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    a = _compressed({"geometry_normal": n})
    data, w, h = _level(a, 0, 0)
    rg = decode_level(BC5, data, w, h)[..., :2].astype(np.float64) / 255.0
    s = rg * 2 - 1
    d = np.stack([s[..., 0], s[..., 1], 1 - np.abs(s[..., 0]) - np.abs(s[..., 1])], axis=-1)
    tfold = np.clip(-d[..., 2], 0, 1)
    d[..., 0] += np.where(d[..., 0] >= 0, -tfold, tfold)
    d[..., 1] += np.where(d[..., 1] >= 0, -tfold, tfold)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    angle = np.degrees(np.arccos(np.clip((d * n).sum(-1), -1, 1)))
    assert angle.max() < 2.0, angle.max()


def test_sampler_reads_the_decoded_blocks():
    """The oracle's sampler at texel centres of level 0 returns the sRGB
    decode of the block texels."""
    rng = np.random.default_rng(4)
    tex = rng.random((8, 8, 3)).astype(np.float32)
    s = AssetScene()
    p = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0]], np.float32)
    s.insert_mesh(p, p[:, :2], [0, 1, 2], None, None)
    s.add_model(s.add_node(), 0, s.insert_material(Material(), {"base_color": tex}))
    a = s.arrays()
    data, w, h = _level(a, 0, 0)
    o = binding.OracleScene(s, 32)
    c = np.arange(256) / 255.0
    srgb = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    for y in range(0, 8, 3):
        for x in range(0, 8, 3):
            got = o.sample_texture(0, (x + 0.5) / w, (y + 0.5) / h, 0.0)
            want = srgb[binding.block_texel(BC1, data, w, x, y)[:3]]
            np.testing.assert_allclose(got[:3], want, atol=1e-6)
