"""glTF textures (src/asset/gltf.rs:86-137, 242-381): PNG decoding as the
reference configures the png crate (ALPHA | normalize_to_color8) and the
mapping of glTF texture slots onto material channels.

PNG files come from two independent encoders: Pillow, and a small encoder
written here from the PNG specification that can choose every row filter,
Adam7 interlacing, low bit depths, 16-bit samples and tRNS keys, and that
compresses with zlib as stored, fixed-Huffman or dynamic-Huffman blocks.
The expected RGBA is derived from the encoded samples by the specification's
rules (low depths scaled by 255 / (2^d - 1), 16-bit samples cut to their high
byte, tRNS keys transparent), not by decoding.

Material mapping is pinned by equivalence: importing a glTF with textures
gives bit for bit the material and texture arrays that insert_material gives
for float textures built here in numpy as the reference's transform closures
(factor x byte / 255, byte / 255 x 2 - 1, 1 - base colour)."""
import base64
import json
import os
import struct
import zlib

import numpy as np
import pytest

from tinytrace_amd import AssetScene, GltfImporter, Material, _abi

CH = {name: i for i, name in enumerate(_abi.TEXTURE_CHANNELS)}
ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


# ------------------------------------------------------------ PNG writer --
def _chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xffffffff)


def _pack_row(row, depth):
    """One scanline of samples (flattened) at `depth` bits, big-endian."""
    if depth == 8:
        return bytes(int(v) for v in row)
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in row)
    out, acc, bits = bytearray(), 0, 0
    for v in row:
        acc = (acc << depth) | int(v)
        bits += depth
        if bits == 8:
            out.append(acc)
            acc, bits = 0, 0
    if bits:
        out.append(acc << (8 - bits))
    return bytes(out)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _filter(kind, row, prev, bpp):
    out = bytearray(len(row))
    for i in range(len(row)):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        pred = [0, a, b, (a + b) // 2, _paeth(a, b, c)][kind]
        out[i] = (row[i] - pred) & 0xff
    return bytes(out)


def write_png(samples, ctype, depth, palette=None, trns=None, interlace=False, filters=(0, 1, 2, 3, 4),
              mode="dynamic", idat_pieces=3, corrupt_adler=False):
    """samples: (H, W, C) raw sample values.  Filters cycle per row."""
    h, w, c = samples.shape
    assert c == CHANNELS[ctype]
    bpp = max(1, c * depth // 8)
    raw = bytearray()
    k = 0
    for x0, y0, dx, dy in (ADAM7 if interlace else [(0, 0, 1, 1)]):
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        prev = bytes(len(_pack_row(sub[0].reshape(-1), depth)))
        for r in range(sub.shape[0]):
            row = _pack_row(sub[r].reshape(-1), depth)
            kind = filters[k % len(filters)]
            k += 1
            raw += bytes([kind]) + _filter(kind, row, prev, bpp)
            prev = row
    if mode == "stored":
        z = zlib.compressobj(0)
    elif mode == "fixed":
        z = zlib.compressobj(9, zlib.DEFLATED, 15, 9, zlib.Z_FIXED)
    else:
        z = zlib.compressobj(9)
    data = z.compress(bytes(raw)) + z.flush()
    if corrupt_adler:
        data = data[:-1] + bytes([data[-1] ^ 0xff])
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        out += _chunk(b"tRNS", trns)
    step = max(1, len(data) // idat_pieces + 1)
    for i in range(0, len(data), step):
        out += _chunk(b"IDAT", data[i:i + step])
    return out + _chunk(b"IEND", b"")


def expected_rgba(samples, ctype, depth, palette=None, trns=None):
    """RGBA8 by the PNG rules the reference's decoder setup applies; grey is
    replicated into R, G, B."""
    s = samples.astype(np.int64)
    h, w, _ = s.shape
    out = np.zeros((h, w, 4), np.int64)
    to8 = (lambda v: v >> 8) if depth == 16 else (lambda v: v * 255 // ((1 << depth) - 1))
    if ctype == 3:
        pal = np.asarray(palette, np.int64).reshape(-1, 3)
        out[..., :3] = pal[s[..., 0]]
        alpha = np.full(len(pal), 255, np.int64)
        if trns is not None:
            alpha[:len(trns)] = list(trns)
        out[..., 3] = alpha[s[..., 0]]
    elif ctype in (0, 4):
        out[..., :3] = to8(s[..., :1])
        if ctype == 4:
            out[..., 3] = to8(s[..., 1])
        else:
            key = struct.unpack(">H", trns)[0] if trns is not None else None
            out[..., 3] = np.where(s[..., 0] == key, 0, 255) if key is not None else 255
    else:
        out[..., :3] = to8(s[..., :3])
        if ctype == 6:
            out[..., 3] = to8(s[..., 3])
        elif trns is not None:
            key = np.array(struct.unpack(">HHH", trns))
            out[..., 3] = np.where(np.all(s[..., :3] == key, axis=-1), 0, 255)
        else:
            out[..., 3] = 255
    return out.astype(np.uint8)


# ------------------------------------------------------------ glTF docs --
def _uri(data, mime="application/octet-stream"):
    return f"data:{mime};base64," + base64.b64encode(data).decode()


def quad_doc(materials, images, textures=None):
    """One quad (two triangles, uv over [0, 1]^2) per material, in a
    data-URI buffer."""
    p = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    uv = p[:, :2].copy()
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    data = p.tobytes() + uv.tobytes() + idx.tobytes()
    n = len(materials)
    return {
        "asset": {"version": "2.0"},
        "scene": 0,
        "scenes": [{"nodes": list(range(n))}],
        "nodes": [{"mesh": i} for i in range(n)],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "indices": 2, "material": i}]}
                   for i in range(n)],
        "materials": materials,
        "textures": textures if textures is not None else [{"source": i} for i in range(len(images))],
        "images": images,
        "accessors": [
            {"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
            {"bufferView": 1, "componentType": 5126, "count": 4, "type": "VEC2"},
            {"bufferView": 2, "componentType": 5125, "count": 6, "type": "SCALAR"},
        ],
        "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 48},
                        {"buffer": 0, "byteOffset": 48, "byteLength": 32},
                        {"buffer": 0, "byteOffset": 80, "byteLength": 24}],
        "buffers": [{"byteLength": len(data), "uri": _uri(data)}],
    }


def write_doc(tmp, doc, name="scene.gltf"):
    path = os.path.join(tmp, name)
    with open(path, "w") as f:
        json.dump(doc, f)
    return path


def texture_level(arrays, index, level=0):
    rec = arrays["textures"][index]
    texel = {0: 4, 1: 2, 2: 1, 3: 2}[int(rec["format"])]
    w, h = int(rec["width"]) >> level, int(rec["height"]) >> level
    o = int(rec["offsets"][level])
    return arrays["texture_data"][o:o + w * h * texel].reshape(h, w, texel)


def decoded_through_import(tmp, png):
    """RGBA8 of `png` as the importer decodes it: R, G, B from a base colour
    texture (RGBA8 level 0, factor 1, imported uncompressed) and A from a
    specular weight texture (R8 level 0, factor 1) sharing the image."""
    with open(os.path.join(tmp, "image.png"), "wb") as f:
        f.write(png)
    doc = quad_doc([{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}},
                     "extensions": {"KHR_materials_specular": {"specularTexture": {"index": 0}}}}],
                   [{"uri": "image.png"}])
    a = GltfImporter(write_doc(tmp, doc)).new_scene(block_compression=False).arrays()
    assert a["textures"]["format"][0] == _abi.TT_TEXTURE_FORMAT_RGBA8_SRGB
    assert a["textures"]["format"][1] == _abi.TT_TEXTURE_FORMAT_R8_UNORM
    rgb = texture_level(a, 0)[..., :3]
    alpha = texture_level(a, 1)[..., 0]
    return np.concatenate([rgb, alpha[..., None]], axis=-1)


def test_byte_round_trip_is_exact():
    """quantize_unorm(v / 255) = v for every byte, so level 0 of a factor-1
    texture holds the decoded bytes themselves."""
    v = np.arange(256, dtype=np.float32)
    q = np.floor(np.clip(v / np.float32(255.0), 0, 1) * np.float32(255.0) + np.float32(0.5))
    np.testing.assert_array_equal(q, v)


CASES = [
    # (ctype, depth, interlace, compression)
    (6, 8, False, "dynamic"), (6, 8, True, "fixed"), (6, 16, False, "stored"), (6, 16, True, "dynamic"),
    (2, 8, False, "fixed"), (2, 8, True, "stored"), (2, 16, False, "dynamic"),
    (0, 1, False, "dynamic"), (0, 2, True, "dynamic"), (0, 4, False, "fixed"), (0, 8, True, "dynamic"),
    (0, 16, False, "dynamic"),
    (4, 8, False, "dynamic"), (4, 16, True, "fixed"),
    (3, 1, False, "dynamic"), (3, 2, True, "fixed"), (3, 4, False, "stored"), (3, 8, True, "dynamic"),
]


@pytest.mark.parametrize("ctype,depth,interlace,mode", CASES)
@pytest.mark.parametrize("size", [(13, 11), (1, 1), (3, 2)])
def test_png_formats_decode_by_the_spec(tmp_path, ctype, depth, interlace, mode, size):
    w, h = size
    rng = np.random.default_rng(ctype * 100 + depth + 7 * interlace + w)
    c = CHANNELS[ctype]
    palette = trns = None
    if ctype == 3:
        entries = min(1 << depth, 37)
        palette = rng.integers(0, 256, (entries, 3))
        trns = bytes(rng.integers(0, 256, entries // 2).astype(np.uint8))
        samples = rng.integers(0, entries, (h, w, 1))
    else:
        samples = rng.integers(0, 1 << depth, (h, w, c))
        if ctype == 0:
            trns = struct.pack(">H", int(samples[0, 0, 0]))
        elif ctype == 2:
            trns = struct.pack(">HHH", *[int(v) for v in samples[h // 2, w // 2]])
    png = write_png(samples, ctype, depth, palette, trns, interlace, mode=mode)
    got = decoded_through_import(str(tmp_path), png)
    np.testing.assert_array_equal(got, expected_rgba(samples, ctype, depth, palette, trns))


def test_png_from_pillow(tmp_path):
    """Images Pillow encodes in its modes RGB, RGBA, L, LA, P (+ tRNS), 1
    and I;16, with its own filter choice and compression levels."""
    Image = pytest.importorskip("PIL.Image")
    import io
    rng = np.random.default_rng(5)
    h, w = 9, 14
    rgba = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    cases = []
    cases.append((Image.fromarray(np.ascontiguousarray(rgba[..., :3])), {}, expected_rgba(rgba[..., :3], 2, 8)))
    cases.append((Image.fromarray(rgba), {"compress_level": 0}, rgba))
    g = np.ascontiguousarray(rgba[..., 0])
    cases.append((Image.fromarray(g), {"compress_level": 1}, expected_rgba(g[..., None], 0, 8)))
    la = np.ascontiguousarray(rgba[..., :2])
    cases.append((Image.frombytes("LA", (w, h), la.tobytes()), {}, expected_rgba(la, 4, 8)))
    bits = (g > 127)
    cases.append((Image.fromarray(bits), {}, expected_rgba(bits[..., None].astype(np.int64), 0, 1)))
    g16 = rng.integers(0, 65536, (h, w)).astype(np.uint16)
    img16 = Image.frombytes("I;16", (w, h), g16.astype("<u2").tobytes())
    cases.append((img16, {}, expected_rgba(g16[..., None], 0, 16)))
    pal_idx = rng.integers(0, 20, (h, w)).astype(np.uint8)
    palette = rng.integers(0, 256, (20, 3)).astype(np.uint8)
    p = Image.frombytes("P", (w, h), pal_idx.tobytes())
    p.putpalette(palette.reshape(-1).tolist())
    alpha = bytes(rng.integers(0, 256, 20).astype(np.uint8))
    cases.append((p, {"transparency": alpha}, expected_rgba(pal_idx[..., None], 3, 8, palette, alpha)))
    for k, (img, opts, want) in enumerate(cases):
        buf = io.BytesIO()
        img.save(buf, "PNG", **opts)
        got = decoded_through_import(str(tmp_path), buf.getvalue())
        np.testing.assert_array_equal(got, want, err_msg=f"case {k} mode {img.mode}")


def test_image_sources_buffer_view_and_data_uri(tmp_path):
    rng = np.random.default_rng(3)
    rgba = rng.integers(0, 256, (5, 6, 4))
    png = write_png(rgba, 6, 8)
    want = texture_level(_import_one(tmp_path, png, "file"), 0)
    for source in ("data", "view"):
        np.testing.assert_array_equal(texture_level(_import_one(tmp_path, png, source), 0), want)
    np.testing.assert_array_equal(want[..., :3], rgba[..., :3])


def _import_one(tmp_path, png, source):
    tmp = str(tmp_path)
    doc = quad_doc([{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}], [{}])
    if source == "file":
        with open(os.path.join(tmp, "t.png"), "wb") as f:
            f.write(png)
        doc["images"][0] = {"uri": "t.png"}
    elif source == "data":
        doc["images"][0] = {"uri": _uri(png, "image/png")}
    else:
        geometry = base64.b64decode(doc["buffers"][0]["uri"].split(",", 1)[1])
        pad = (-len(geometry)) % 4
        data = geometry + b"\0" * pad + png
        doc["buffers"][0] = {"byteLength": len(data), "uri": _uri(data)}
        doc["bufferViews"].append({"buffer": 0, "byteOffset": len(geometry) + pad, "byteLength": len(png)})
        doc["images"][0] = {"bufferView": 3, "mimeType": "image/png"}
    return GltfImporter(write_doc(tmp, doc, f"{source}.gltf")).new_scene(block_compression=False).arrays()


def _unorm(b):
    return np.asarray(b, np.float32) / np.float32(255.0)


def test_material_mapping_matches_insert_material(tmp_path):
    """Every texture slot the reference reads, with non-trivial factors,
    against insert_material on float textures built by the reference's
    transform closures.  Two materials: the first textures everything but
    transmission; the second has a transmission texture (which the reference
    assigns to specular weight) and a textured base colour (so the
    transmission colour is 1 - base colour, texel by texel)."""
    tmp = str(tmp_path)
    rng = np.random.default_rng(11)
    imgs = [rng.integers(0, 256, (4 + k, 8 - k, 4)) for k in range(6)]
    for k, im in enumerate(imgs):
        with open(os.path.join(tmp, f"i{k}.png"), "wb") as f:
            f.write(write_png(im, 6, 8))
    base_f = [0.9, 0.5, 0.25, 1.0]
    metal_f, rough_f = 0.75, 0.6
    emis_f = [0.5, 1.0, 0.125]
    spec_f, spec_c = 0.8, [0.3, 0.6, 0.9]
    trans_f = 0.7
    m0 = {
        "pbrMetallicRoughness": {"baseColorFactor": base_f, "baseColorTexture": {"index": 0},
                                 "metallicFactor": metal_f, "roughnessFactor": rough_f,
                                 "metallicRoughnessTexture": {"index": 1}},
        "normalTexture": {"index": 2},
        "emissiveFactor": emis_f, "emissiveTexture": {"index": 3},
        "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 4.0},
                       "KHR_materials_specular": {"specularFactor": spec_f, "specularTexture": {"index": 4},
                                                  "specularColorFactor": spec_c,
                                                  "specularColorTexture": {"index": 5}},
                       "KHR_materials_ior": {"ior": 1.33}},
    }
    m1 = {
        "pbrMetallicRoughness": {"baseColorFactor": base_f, "baseColorTexture": {"index": 5}},
        "extensions": {"KHR_materials_transmission": {"transmissionFactor": trans_f,
                                                      "transmissionTexture": {"index": 4}}},
    }
    doc = quad_doc([m0, m1], [{"uri": f"i{k}.png"} for k in range(6)])
    got = GltfImporter(write_doc(tmp, doc)).new_scene().arrays()

    f32 = np.float32
    s = AssetScene()
    p = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    s.insert_mesh(p, p[:, :2], [0, 1, 2, 0, 2, 3], None, None)
    base0 = f32(base_f[:3]) * _unorm(imgs[0][..., :3])
    t0 = {
        "base_color": base0,
        "base_metalness": _unorm(imgs[1][..., 2]) * f32(metal_f),
        "specular_roughness": _unorm(imgs[1][..., 1]) * f32(rough_f),
        "geometry_normal": _unorm(imgs[2][..., :3]) * f32(2.0) - f32(1.0),
        "emission_color": f32(emis_f) * _unorm(imgs[3][..., :3]),
        "specular_weight": _unorm(imgs[4][..., 3]) * f32(spec_f),
        "specular_color": f32(spec_c) * _unorm(imgs[5][..., :3]),
    }
    mat0 = Material(base_color=base_f[:3], base_metalness=metal_f, specular_roughness=rough_f,
                    emission_color=emis_f, emission_luminance=4.0, specular_weight=spec_f,
                    specular_color=spec_c, specular_ior=1.33)
    base1 = f32(base_f[:3]) * _unorm(imgs[5][..., :3])
    t1 = {
        "base_color": base1,
        "specular_weight": f32(trans_f) * _unorm(imgs[4][..., 0]),
        "transmission_color": f32(1.0) - base1,
    }
    mat1 = Material(base_color=base_f[:3], base_metalness=1.0, specular_roughness=1.0, emission_color=(0, 0, 0),
                    transmission_color=[1.0 - f32(c) for c in base_f[:3]])
    s.add_model(s.add_node(), 0, s.insert_material(mat0, t0))
    s.add_model(s.add_node(), 0, s.insert_material(mat1, t1))
    want = s.arrays()
    for key in ("materials", "textures", "texture_data"):
        np.testing.assert_array_equal(got[key], want[key], err_msg=key)
    assert len(got["textures"]) == 10  # 7 + 3: the comparison covers every slot


def test_shared_material_imports_once(tmp_path):
    tmp = str(tmp_path)
    with open(os.path.join(tmp, "t.png"), "wb") as f:
        f.write(write_png(np.full((2, 2, 4), 200), 6, 8))
    m = {"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}
    doc = quad_doc([m, m], [{"uri": "t.png"}])
    for prim in doc["meshes"]:
        prim["primitives"][0]["material"] = 0
    a = GltfImporter(write_doc(tmp, doc)).new_scene().arrays()
    assert len(a["materials"]) == 1 and len(a["textures"]) == 1
    assert list(a["instances"]["material_index"]) == [0, 0]


def _status(fn):
    try:
        fn()
    except _abi.TinytraceError as e:
        return e.status
    return 0


def test_texture_refusals(tmp_path):
    tmp = str(tmp_path)
    unsupported, invalid = 8, 4
    good = write_png(np.full((4, 4, 4), 9), 6, 8)

    def status(images, textures=None, files=None, name="t.gltf"):
        for fname, data in (files or {}).items():
            with open(os.path.join(tmp, fname), "wb") as f:
                f.write(data)
        doc = quad_doc([{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}], images, textures)
        return _status(lambda: AssetScene().import_gltf(write_doc(tmp, doc, name)))

    assert status([{"uri": "ok.png"}], files={"ok.png": good}) == 0
    # Not PNG: a .jpg file, a JPEG buffer view, a JPEG data URI.
    assert status([{"uri": "t.jpg"}], files={"t.jpg": good}) == unsupported
    assert status([{"uri": _uri(good, "image/jpeg")}]) == unsupported
    # A texture without a source image.
    assert status([{"uri": "ok.png"}], textures=[{}]) == unsupported
    # Missing file, out-of-range texture, corrupt bytes, a bad CRC, bad Adler-32.
    assert status([{"uri": "absent.png"}]) == invalid
    assert status([{"uri": "ok.png"}], textures=[{"source": 3}]) == invalid
    assert status([{"uri": "junk.png"}], files={"junk.png": b"not a png at all"}) == invalid
    crc = bytearray(good)
    crc[20] ^= 1  # inside IHDR
    assert status([{"uri": "crc.png"}], files={"crc.png": bytes(crc)}) == invalid
    assert status([{"uri": "cut.png"}], files={"cut.png": good[:len(good) - 20]}) == invalid
    tampered = write_png(np.full((4, 4, 4), 9), 6, 8, corrupt_adler=True)
    assert status([{"uri": "adler.png"}], files={"adler.png": tampered}) == invalid
