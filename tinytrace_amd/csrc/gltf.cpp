// glTF 2.0 import into an asset scene: the counterpart of the reference's
// GltfImporter (src/asset/gltf.rs).
//
//   GltfImporter::new        gltf.rs:50-72    .gltf JSON + external or data-URI
//                                             buffers (binary .glb is refused,
//                                             as the reference leaves it todo)
//   buffer_data / accessors  gltf.rs:74-119   strided views, f32 attributes,
//                                             u16/u32 indices (verify_accessor,
//                                             gltf.rs:455-463)
//   load_mesh                gltf.rs:197-240  positions, texture coordinates
//                                             (zeros when absent), indices,
//                                             optional normals and tangents,
//                                             cached per (mesh, primitive)
//   load_material            gltf.rs:242-381  pbrMetallicRoughness factors
//                                             and textures (base colour,
//                                             metalness from B, roughness
//                                             from G), normal and emissive
//                                             textures, emissive factor,
//                                             KHR_materials_emissive_strength,
//                                             KHR_materials_specular (weight
//                                             from A, colour), _ior,
//                                             _transmission (its texture
//                                             lands on specular weight, as in
//                                             the reference); cached per index
//                                             (the default material once)
//   image_data / load_png    gltf.rs:86-137   PNG images from buffer views,
//                                             files or data URIs (png.cpp)
//   load_nodes               gltf.rs:398-417  depth first: a node's models,
//                                             then its children
//   insert_in_scene          gltf.rs:419-452  default scene, else the first
//
// Textures become float texel arrays (factor x dequantized byte, as the
// reference's Texture::transform closures) and go through
// tt_asset_insert_material_textured.  The reference's placeholder 32 x 32
// texture for scenes without any (gltf.rs:442-449) only fills an empty
// descriptor array and is not added.  Mesh processing (bounding sphere,
// snorm16, normals and tangents when absent) is tt_asset_insert_mesh's.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/tinytrace.h"

namespace tt {
// png.cpp: PNG bytes -> RGBA8; an empty string on success.
std::string decode_png(const uint8_t* data, size_t n, uint32_t& width, uint32_t& height, std::vector<uint8_t>& rgba);
}  // namespace tt

namespace {

// ----------------------------------------------------------------- JSON --
struct Json {
    enum Kind { NUL, BOOL, NUMBER, STRING, ARRAY, OBJECT } kind = NUL;
    bool boolean = false;
    double number = 0.0;
    std::string text;
    std::vector<Json> items;
    std::vector<std::pair<std::string, Json>> members;

    const Json* get(const char* key) const {
        if (kind != OBJECT) return nullptr;
        for (const auto& m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
    const Json* at(size_t i) const { return kind == ARRAY && i < items.size() ? &items[i] : nullptr; }
    size_t size() const { return kind == ARRAY ? items.size() : 0; }
    bool is_number() const { return kind == NUMBER; }
};

double number_or(const Json* j, double fallback) { return j && j->is_number() ? j->number : fallback; }

// Non-negative integer member (an index or a count); -1 when absent.
long long index_of(const Json* j) {
    if (!j || !j->is_number() || j->number < 0.0 || j->number != std::floor(j->number) || j->number > 4294967295.0)
        return -1;
    return (long long)j->number;
}

class JsonParser {
  public:
    JsonParser(const char* p, const char* end) : p_(p), end_(end) {}
    bool parse(Json& out) {
        skip();
        if (!value(out, 0)) return false;
        skip();
        return p_ == end_;
    }

  private:
    const char* p_;
    const char* end_;

    void skip() {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) p_++;
    }
    bool literal(const char* word) {
        size_t n = std::strlen(word);
        if (size_t(end_ - p_) < n || std::memcmp(p_, word, n) != 0) return false;
        p_ += n;
        return true;
    }
    static void utf8(std::string& s, uint32_t c) {
        if (c < 0x80) {
            s += char(c);
        } else if (c < 0x800) {
            s += char(0xc0 | (c >> 6));
            s += char(0x80 | (c & 0x3f));
        } else if (c < 0x10000) {
            s += char(0xe0 | (c >> 12));
            s += char(0x80 | ((c >> 6) & 0x3f));
            s += char(0x80 | (c & 0x3f));
        } else {
            s += char(0xf0 | (c >> 18));
            s += char(0x80 | ((c >> 12) & 0x3f));
            s += char(0x80 | ((c >> 6) & 0x3f));
            s += char(0x80 | (c & 0x3f));
        }
    }
    bool hex4(uint32_t& v) {
        if (end_ - p_ < 4) return false;
        v = 0;
        for (int i = 0; i < 4; i++) {
            char c = *p_++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= uint32_t(c - '0');
            else if (c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
            else return false;
        }
        return true;
    }
    bool string(std::string& s) {
        if (p_ >= end_ || *p_ != '"') return false;
        p_++;
        while (p_ < end_) {
            char c = *p_++;
            if (c == '"') return true;
            if (static_cast<unsigned char>(c) < 0x20) return false;
            if (c != '\\') {
                s += c;
                continue;
            }
            if (p_ >= end_) return false;
            char e = *p_++;
            switch (e) {
                case '"': s += '"'; break;
                case '\\': s += '\\'; break;
                case '/': s += '/'; break;
                case 'b': s += '\b'; break;
                case 'f': s += '\f'; break;
                case 'n': s += '\n'; break;
                case 'r': s += '\r'; break;
                case 't': s += '\t'; break;
                case 'u': {
                    uint32_t c0;
                    if (!hex4(c0)) return false;
                    if (c0 >= 0xd800 && c0 < 0xdc00) {
                        uint32_t c1;
                        if (end_ - p_ < 6 || p_[0] != '\\' || p_[1] != 'u') return false;
                        p_ += 2;
                        if (!hex4(c1) || c1 < 0xdc00 || c1 >= 0xe000) return false;
                        c0 = 0x10000 + ((c0 - 0xd800) << 10) + (c1 - 0xdc00);
                    } else if (c0 >= 0xdc00 && c0 < 0xe000) {
                        return false;
                    }
                    utf8(s, c0);
                    break;
                }
                default: return false;
            }
        }
        return false;
    }
    bool number(double& v) {
        const char* start = p_;
        if (p_ < end_ && *p_ == '-') p_++;
        if (p_ >= end_ || !(*p_ >= '0' && *p_ <= '9')) return false;
        if (*p_ == '0') p_++;
        else
            while (p_ < end_ && *p_ >= '0' && *p_ <= '9') p_++;
        if (p_ < end_ && *p_ == '.') {
            p_++;
            if (p_ >= end_ || !(*p_ >= '0' && *p_ <= '9')) return false;
            while (p_ < end_ && *p_ >= '0' && *p_ <= '9') p_++;
        }
        if (p_ < end_ && (*p_ == 'e' || *p_ == 'E')) {
            p_++;
            if (p_ < end_ && (*p_ == '+' || *p_ == '-')) p_++;
            if (p_ >= end_ || !(*p_ >= '0' && *p_ <= '9')) return false;
            while (p_ < end_ && *p_ >= '0' && *p_ <= '9') p_++;
        }
        std::string digits(start, p_);
        char* stop = nullptr;
        v = std::strtod(digits.c_str(), &stop);
        return stop == digits.c_str() + digits.size();
    }
    bool value(Json& out, int depth) {
        if (depth > 256 || p_ >= end_) return false;
        char c = *p_;
        if (c == '{') {
            p_++;
            out.kind = Json::OBJECT;
            skip();
            if (p_ < end_ && *p_ == '}') {
                p_++;
                return true;
            }
            for (;;) {
                skip();
                std::pair<std::string, Json> m;
                if (!string(m.first)) return false;
                skip();
                if (p_ >= end_ || *p_ != ':') return false;
                p_++;
                skip();
                if (!value(m.second, depth + 1)) return false;
                out.members.push_back(std::move(m));
                skip();
                if (p_ < end_ && *p_ == ',') {
                    p_++;
                    continue;
                }
                if (p_ < end_ && *p_ == '}') {
                    p_++;
                    return true;
                }
                return false;
            }
        }
        if (c == '[') {
            p_++;
            out.kind = Json::ARRAY;
            skip();
            if (p_ < end_ && *p_ == ']') {
                p_++;
                return true;
            }
            for (;;) {
                skip();
                out.items.emplace_back();
                if (!value(out.items.back(), depth + 1)) return false;
                skip();
                if (p_ < end_ && *p_ == ',') {
                    p_++;
                    continue;
                }
                if (p_ < end_ && *p_ == ']') {
                    p_++;
                    return true;
                }
                return false;
            }
        }
        if (c == '"') {
            out.kind = Json::STRING;
            return string(out.text);
        }
        if (c == 't' || c == 'f') {
            out.kind = Json::BOOL;
            out.boolean = c == 't';
            return literal(c == 't' ? "true" : "false");
        }
        if (c == 'n') {
            out.kind = Json::NUL;
            return literal("null");
        }
        out.kind = Json::NUMBER;
        return number(out.number);
    }
};

// -------------------------------------------------------------- import --
constexpr int F32 = 5126, U32 = 5125, U16 = 5123;

struct Failure {
    tt_status status;
};

[[noreturn]] void refuse(tt_status s) { throw Failure{s}; }

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.seekg(0, std::ios::end);
    std::streamoff n = f.tellg();
    if (n < 0) return false;
    f.seekg(0, std::ios::beg);
    out.resize(size_t(n));
    if (n > 0) f.read(reinterpret_cast<char*>(out.data()), n);
    return bool(f) || n == 0;
}

bool base64(const std::string& s, size_t from, std::vector<uint8_t>& out) {
    auto sextet = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+') return 62;
        if (c == '/') return 63;
        return -1;
    };
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < s.size(); i++) {
        char c = s[i];
        if (c == '=') break;
        int v = sextet(c);
        if (v < 0) return false;
        acc = (acc << 6) | uint32_t(v);
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back(uint8_t((acc >> bits) & 0xffu));
        }
    }
    return true;
}

class Importer {
  public:
    Importer(const Json& doc, std::vector<std::vector<uint8_t>>& buffers, tt_asset_scene* scene, std::string dir)
        : doc_(doc), buffers_(buffers), scene_(scene), dir_(std::move(dir)) {}

    void run() {
        const Json* scenes = doc_.get("scenes");
        long long which = index_of(doc_.get("scene"));
        const Json* chosen = nullptr;
        if (which >= 0) chosen = scenes ? scenes->at(size_t(which)) : nullptr;
        if (which >= 0 && !chosen) refuse(TT_INVALID_ARGUMENT);
        if (!chosen && scenes) chosen = scenes->at(0);
        if (!chosen) return;  // no scene: nothing to insert
        const Json* roots = chosen->get("nodes");
        for (size_t i = 0; i < (roots ? roots->size() : 0); i++) node(index_of(roots->at(i)), 0, 0);
    }

  private:
    const Json& doc_;
    std::vector<std::vector<uint8_t>>& buffers_;
    tt_asset_scene* scene_;
    std::map<std::pair<long long, long long>, uint32_t> meshes_;
    std::map<long long, uint32_t> materials_;
    std::string dir_;
    struct Image {
        uint32_t width = 0, height = 0;
        std::vector<uint8_t> rgba;
    };
    std::map<long long, Image> images_;

    const Json& element(const char* array, long long i) const {
        const Json* a = doc_.get(array);
        const Json* e = (a && i >= 0) ? a->at(size_t(i)) : nullptr;
        if (!e) refuse(TT_INVALID_ARGUMENT);
        return *e;
    }

    // Accessor contents packed to `components` values of `type` bytes each
    // per element (verify_accessor, gltf.rs:455-463).
    std::vector<uint8_t> accessor(long long index, int component_type, const char* type, size_t& count) const {
        const Json& a = element("accessors", index);
        const Json* t = a.get("type");
        long long ct = index_of(a.get("componentType"));
        if (!t || t->kind != Json::STRING || t->text != type || ct != component_type) refuse(TT_UNSUPPORTED);
        if (a.get("sparse")) refuse(TT_UNSUPPORTED);
        const Json* normalized = a.get("normalized");
        if (normalized && normalized->kind == Json::BOOL && normalized->boolean) refuse(TT_UNSUPPORTED);
        long long n = index_of(a.get("count"));
        if (n < 0) refuse(TT_INVALID_ARGUMENT);
        size_t components = std::strcmp(type, "SCALAR") == 0 ? 1 : size_t(type[3] - '0');
        size_t component_bytes = component_type == U16 ? 2 : 4;
        size_t size = components * component_bytes;
        count = size_t(n);
        long long view_index = index_of(a.get("bufferView"));
        if (view_index < 0) refuse(TT_UNSUPPORTED);  // all-zero accessors are not used by the reference
        const Json& view = element("bufferViews", view_index);
        long long buffer = index_of(view.get("buffer"));
        if (buffer < 0 || size_t(buffer) >= buffers_.size()) refuse(TT_INVALID_ARGUMENT);
        const std::vector<uint8_t>& data = buffers_[size_t(buffer)];
        long long view_offset = index_of(view.get("byteOffset"));
        long long view_length = index_of(view.get("byteLength"));
        long long offset = index_of(a.get("byteOffset"));
        long long stride = index_of(view.get("byteStride"));
        size_t vo = view_offset < 0 ? 0 : size_t(view_offset);
        size_t ao = offset < 0 ? 0 : size_t(offset);
        size_t st = stride <= 0 ? size : size_t(stride);
        if (view_length < 0 || st < size) refuse(TT_INVALID_ARGUMENT);
        if (vo + size_t(view_length) > data.size()) refuse(TT_INVALID_ARGUMENT);
        if (count > 0 && ao + (count - 1) * st + size > size_t(view_length)) refuse(TT_INVALID_ARGUMENT);
        std::vector<uint8_t> out(count * size);
        for (size_t i = 0; i < count; i++) std::memcpy(out.data() + i * size, data.data() + vo + ao + i * st, size);
        return out;
    }

    template <typename T>
    static std::vector<T> as(const std::vector<uint8_t>& bytes) {
        std::vector<T> v(bytes.size() / sizeof(T));
        if (!v.empty()) std::memcpy(v.data(), bytes.data(), v.size() * sizeof(T));
        return v;
    }

    // load_mesh (gltf.rs:197-240).
    uint32_t mesh(long long mesh_index, long long primitive_index, const Json& primitive) {
        auto key = std::make_pair(mesh_index, primitive_index);
        auto hit = meshes_.find(key);
        if (hit != meshes_.end()) return hit->second;
        long long mode = index_of(primitive.get("mode"));
        if (mode >= 0 && mode != 4) refuse(TT_UNSUPPORTED);  // triangles only
        const Json* attributes = primitive.get("attributes");
        if (!attributes) refuse(TT_INVALID_ARGUMENT);
        long long position = index_of(attributes->get("POSITION"));
        if (position < 0) refuse(TT_INVALID_ARGUMENT);
        size_t vertices = 0, n = 0;
        std::vector<float> positions = as<float>(accessor(position, F32, "VEC3", vertices));
        std::vector<float> uv(vertices * 2, 0.0f);  // zeros when absent (gltf.rs:174-193)
        if (long long t = index_of(attributes->get("TEXCOORD_0")); t >= 0) {
            uv = as<float>(accessor(t, F32, "VEC2", n));
            if (n != vertices) refuse(TT_INVALID_ARGUMENT);
        }
        std::vector<float> normals, tangents;
        if (long long a = index_of(attributes->get("NORMAL")); a >= 0) {
            normals = as<float>(accessor(a, F32, "VEC3", n));
            if (n != vertices) refuse(TT_INVALID_ARGUMENT);
        }
        if (long long a = index_of(attributes->get("TANGENT")); a >= 0) {
            tangents = as<float>(accessor(a, F32, "VEC4", n));
            if (n != vertices) refuse(TT_INVALID_ARGUMENT);
        }
        std::vector<uint32_t> indices;
        bool indexed = false;
        if (long long a = index_of(primitive.get("indices")); a >= 0) {
            indexed = true;
            const Json& acc = element("accessors", a);
            long long ct = index_of(acc.get("componentType"));
            if (ct == U32) {
                indices = as<uint32_t>(accessor(a, U32, "SCALAR", n));
            } else if (ct == U16) {
                std::vector<uint16_t> small = as<uint16_t>(accessor(a, U16, "SCALAR", n));
                indices.assign(small.begin(), small.end());
            } else {
                refuse(TT_UNSUPPORTED);
            }
            for (uint32_t i : indices)
                if (i >= vertices) refuse(TT_INVALID_ARGUMENT);
        }
        tt_mesh_desc d{};
        d.positions = positions.data();
        d.texture_coordinates = uv.data();
        d.indices = indexed ? indices.data() : nullptr;
        d.normals = normals.empty() ? nullptr : normals.data();
        d.tangents = tangents.empty() ? nullptr : tangents.data();
        d.vertex_count = uint32_t(vertices);
        d.index_count = indexed ? uint32_t(indices.size()) : 0u;
        uint32_t index = 0;
        tt_status s = tt_asset_insert_mesh(scene_, &d, &index);
        if (s != TT_OK) refuse(s);
        meshes_[key] = index;
        return index;
    }

    // image_data (gltf.rs:86-108) through load_png (gltf.rs:122-137): the
    // source image of texture info `info` as RGBA8.  PNG only: a buffer view
    // with mimeType image/png, a file beside the .gltf, or a base64 data URI.
    // The texture's sampler and texCoord are not read (TEXCOORD_0, repeat
    // addressing), as in the reference.
    const Image& image(const Json& info) {
        const Json& texture = element("textures", index_of(info.get("index")));
        long long source = index_of(texture.get("source"));
        if (source < 0) refuse(TT_UNSUPPORTED);
        auto hit = images_.find(source);
        if (hit != images_.end()) return hit->second;
        const Json& img = element("images", source);
        std::vector<uint8_t> bytes;
        if (long long view_index = index_of(img.get("bufferView")); view_index >= 0) {
            const Json* mime = img.get("mimeType");
            if (!mime || mime->kind != Json::STRING || mime->text != "image/png") refuse(TT_UNSUPPORTED);
            const Json& view = element("bufferViews", view_index);
            long long buffer = index_of(view.get("buffer"));
            if (buffer < 0 || size_t(buffer) >= buffers_.size()) refuse(TT_INVALID_ARGUMENT);
            const std::vector<uint8_t>& data = buffers_[size_t(buffer)];
            long long offset = index_of(view.get("byteOffset")), length = index_of(view.get("byteLength"));
            size_t o = offset < 0 ? 0 : size_t(offset);
            if (length < 0 || o + size_t(length) > data.size()) refuse(TT_INVALID_ARGUMENT);
            bytes.assign(data.begin() + std::ptrdiff_t(o), data.begin() + std::ptrdiff_t(o + size_t(length)));
        } else {
            const Json* uri = img.get("uri");
            if (!uri || uri->kind != Json::STRING) refuse(TT_INVALID_ARGUMENT);
            const std::string& u = uri->text;
            if (u.compare(0, 5, "data:") == 0) {
                size_t comma = u.find(',');
                if (comma == std::string::npos || u.compare(5, 9, "image/png") != 0 ||
                    u.rfind(";base64", comma) == std::string::npos || !base64(u, comma + 1, bytes))
                    refuse(TT_UNSUPPORTED);
            } else {
                size_t dot = u.find_last_of('.'), slash = u.find_last_of('/');
                if (dot != std::string::npos && (slash == std::string::npos || dot > slash + 1) &&
                    u.compare(dot + 1, std::string::npos, "png") != 0)
                    refuse(TT_UNSUPPORTED);
                if (!read_file(dir_ + u, bytes)) refuse(TT_INVALID_ARGUMENT);
            }
        }
        Image im;
        if (!tt::decode_png(bytes.data(), bytes.size(), im.width, im.height, im.rgba).empty()) refuse(TT_INVALID_ARGUMENT);
        return images_.emplace(source, std::move(im)).first->second;
    }

    // A float texture the size of `im`, `comps` values a texel from f(rgba).
    struct FloatTexture {
        uint32_t width = 0, height = 0;
        std::vector<float> texels;
    };
    template <typename F>
    static FloatTexture transform(const Image& im, int comps, F f) {
        FloatTexture t;
        t.width = im.width;
        t.height = im.height;
        t.texels.resize(size_t(im.width) * im.height * size_t(comps));
        for (size_t i = 0; i < size_t(im.width) * im.height; i++) f(&im.rgba[4 * i], &t.texels[i * size_t(comps)]);
        return t;
    }

    // math::dequantize_unorm(v, 8).
    static float unorm8(uint8_t v) { return float(v) / 255.0f; }

    // load_material (gltf.rs:242-381).
    uint32_t material(long long material_index) {
        auto hit = materials_.find(material_index);
        if (hit != materials_.end()) return hit->second;
        tt_material_desc m;
        tt_material_desc_default(&m);
        std::map<uint32_t, FloatTexture> textures;  // by TT_TEXTURE_* channel
        if (material_index >= 0) {
            const Json& g = element("materials", material_index);
            const Json* pbr = g.get("pbrMetallicRoughness");
            // glTF defaults: base colour (1, 1, 1, 1), metallic 1, roughness 1.
            float base[3];
            const Json* factor = pbr ? pbr->get("baseColorFactor") : nullptr;
            for (int c = 0; c < 3; c++) base[c] = float(number_or(factor ? factor->at(size_t(c)) : nullptr, 1.0));
            for (int c = 0; c < 3; c++) m.base_color[c] = base[c];
            if (const Json* info = pbr ? pbr->get("baseColorTexture") : nullptr)
                textures[TT_TEXTURE_BASE_COLOR] = transform(image(*info), 3, [&](const uint8_t* p, float* o) {
                    for (int c = 0; c < 3; c++) o[c] = base[c] * unorm8(p[c]);
                });
            float metalness = float(number_or(pbr ? pbr->get("metallicFactor") : nullptr, 1.0));
            float roughness = float(number_or(pbr ? pbr->get("roughnessFactor") : nullptr, 1.0));
            m.base_metalness = metalness;
            m.specular_roughness = roughness;
            if (const Json* info = pbr ? pbr->get("metallicRoughnessTexture") : nullptr) {
                const Image& im = image(*info);
                textures[TT_TEXTURE_BASE_METALNESS] =
                    transform(im, 1, [&](const uint8_t* p, float* o) { o[0] = unorm8(p[2]) * metalness; });
                textures[TT_TEXTURE_SPECULAR_ROUGHNESS] =
                    transform(im, 1, [&](const uint8_t* p, float* o) { o[0] = unorm8(p[1]) * roughness; });
            }
            if (const Json* info = g.get("normalTexture"))
                textures[TT_TEXTURE_GEOMETRY_NORMAL] = transform(image(*info), 3, [](const uint8_t* p, float* o) {
                    for (int c = 0; c < 3; c++) o[c] = unorm8(p[c]) * 2.0f - 1.0f;
                });
            const Json* emissive = g.get("emissiveFactor");
            float emission[3];
            for (int c = 0; c < 3; c++) emission[c] = float(number_or(emissive ? emissive->at(size_t(c)) : nullptr, 0.0));
            for (int c = 0; c < 3; c++) m.emission_color[c] = emission[c];
            if (const Json* info = g.get("emissiveTexture"))
                textures[TT_TEXTURE_EMISSION_COLOR] = transform(image(*info), 3, [&](const uint8_t* p, float* o) {
                    for (int c = 0; c < 3; c++) o[c] = emission[c] * unorm8(p[c]);
                });
            const Json* ext = g.get("extensions");
            if (const Json* e = ext ? ext->get("KHR_materials_emissive_strength") : nullptr)
                m.emission_luminance = float(number_or(e->get("emissiveStrength"), 1.0));
            if (const Json* e = ext ? ext->get("KHR_materials_specular") : nullptr) {
                float weight = float(number_or(e->get("specularFactor"), 1.0));
                const Json* color = e->get("specularColorFactor");
                float specular[3];
                for (int c = 0; c < 3; c++) specular[c] = float(number_or(color ? color->at(size_t(c)) : nullptr, 1.0));
                m.specular_weight = weight;
                if (const Json* info = e->get("specularTexture"))
                    textures[TT_TEXTURE_SPECULAR_WEIGHT] =
                        transform(image(*info), 1, [&](const uint8_t* p, float* o) { o[0] = unorm8(p[3]) * weight; });
                for (int c = 0; c < 3; c++) m.specular_color[c] = specular[c];
                if (const Json* info = e->get("specularColorTexture"))
                    textures[TT_TEXTURE_SPECULAR_COLOR] = transform(image(*info), 3, [&](const uint8_t* p, float* o) {
                        for (int c = 0; c < 3; c++) o[c] = specular[c] * unorm8(p[c]);
                    });
            }
            if (const Json* e = ext ? ext->get("KHR_materials_ior") : nullptr)
                m.specular_ior = float(number_or(e->get("ior"), 1.5));
            if (const Json* e = ext ? ext->get("KHR_materials_transmission") : nullptr) {
                float weight = float(number_or(e->get("transmissionFactor"), 0.0));
                if (const Json* info = e->get("transmissionTexture")) {
                    // gltf.rs:357-363 assigns the texture to specular.weight
                    // and leaves transmission.weight at its default; kept.
                    textures[TT_TEXTURE_SPECULAR_WEIGHT] =
                        transform(image(*info), 1, [&](const uint8_t* p, float* o) { o[0] = weight * unorm8(p[0]); });
                } else {
                    m.transmission_weight = weight;
                }
                // transmission.color = 1 - base.color, texel by texel when
                // the base colour is a texture.
                for (int c = 0; c < 3; c++) m.transmission_color[c] = 1.0f - m.base_color[c];
                auto b = textures.find(TT_TEXTURE_BASE_COLOR);
                if (b != textures.end()) {
                    FloatTexture t = b->second;
                    for (float& v : t.texels) v = 1.0f - v;
                    textures[TT_TEXTURE_TRANSMISSION_COLOR] = std::move(t);
                }
            }
        }
        std::vector<tt_texture_desc> descs;
        for (const auto& [channel, t] : textures)
            descs.push_back(tt_texture_desc{channel, t.width, t.height, t.texels.data()});
        uint32_t index = 0;
        tt_status s = tt_asset_insert_material_textured(scene_, &m, descs.data(), uint32_t(descs.size()), &index);
        if (s != TT_OK) refuse(s);
        materials_[material_index] = index;
        return index;
    }

    // node.transform().matrix(): the matrix, else T * R * S (column major).
    static void transform(const Json& n, float out[16]) {
        const Json* matrix = n.get("matrix");
        if (matrix && matrix->size() == 16) {
            for (int i = 0; i < 16; i++) out[i] = float(number_or(matrix->at(size_t(i)), 0.0));
            return;
        }
        const Json* t = n.get("translation");
        const Json* r = n.get("rotation");
        const Json* s = n.get("scale");
        float tx = float(number_or(t ? t->at(0) : nullptr, 0.0)), ty = float(number_or(t ? t->at(1) : nullptr, 0.0)),
              tz = float(number_or(t ? t->at(2) : nullptr, 0.0));
        float x = float(number_or(r ? r->at(0) : nullptr, 0.0)), y = float(number_or(r ? r->at(1) : nullptr, 0.0)),
              z = float(number_or(r ? r->at(2) : nullptr, 0.0)), w = float(number_or(r ? r->at(3) : nullptr, 1.0));
        float sx = float(number_or(s ? s->at(0) : nullptr, 1.0)), sy = float(number_or(s ? s->at(1) : nullptr, 1.0)),
              sz = float(number_or(s ? s->at(2) : nullptr, 1.0));
        // Rotation matrix of the unit quaternion (x, y, z, w), columns scaled.
        float x2 = x + x, y2 = y + y, z2 = z + z;
        float xx = x * x2, xy = x * y2, xz = x * z2, yy = y * y2, yz = y * z2, zz = z * z2;
        float wx = w * x2, wy = w * y2, wz = w * z2;
        float cols[16] = {(1.0f - (yy + zz)) * sx, (xy + wz) * sx, (xz - wy) * sx, 0.0f,
                          (xy - wz) * sy, (1.0f - (xx + zz)) * sy, (yz + wx) * sy, 0.0f,
                          (xz + wy) * sz, (yz - wx) * sz, (1.0f - (xx + yy)) * sz, 0.0f,
                          tx, ty, tz, 1.0f};
        std::memcpy(out, cols, sizeof(cols));
    }

    // load_nodes (gltf.rs:398-417): the node's models, then its children.
    void node(long long index, uint32_t parent, int depth) {
        if (depth > 512) refuse(TT_INVALID_ARGUMENT);  // a cycle: glTF node graphs are trees
        const Json& n = element("nodes", index);
        float m[16];
        transform(n, m);
        uint32_t self = 0;
        tt_status s = tt_asset_add_node(scene_, parent, m, &self);
        if (s != TT_OK) refuse(s);
        if (long long mi = index_of(n.get("mesh")); mi >= 0) {
            const Json& g = element("meshes", mi);
            const Json* primitives = g.get("primitives");
            for (size_t p = 0; p < (primitives ? primitives->size() : 0); p++) {
                const Json& prim = *primitives->at(p);
                uint32_t mesh_slot = mesh(mi, (long long)p, prim);
                const Json* mat = prim.get("material");
                uint32_t material_slot = material(mat ? index_of(mat) : -1);
                s = tt_asset_add_model(scene_, self, mesh_slot, material_slot);
                if (s != TT_OK) refuse(s);
            }
        }
        const Json* children = n.get("children");
        for (size_t c = 0; c < (children ? children->size() : 0); c++)
            node(index_of(children->at(c)), self, depth + 1);
    }
};

}  // namespace

extern "C" tt_status tt_asset_import_gltf(tt_asset_scene* scene, const char* path) try {
    if (!scene || !path) return TT_INVALID_ARGUMENT;
    {
        std::vector<uint8_t> text;
        if (!read_file(path, text)) return TT_INVALID_ARGUMENT;
        if (text.size() >= 4 && std::memcmp(text.data(), "glTF", 4) == 0) return TT_UNSUPPORTED;  // .glb
        Json doc;
        JsonParser parser(reinterpret_cast<const char*>(text.data()), reinterpret_cast<const char*>(text.data()) + text.size());
        if (!parser.parse(doc) || doc.kind != Json::OBJECT) return TT_INVALID_ARGUMENT;
        std::string dir(path);
        size_t slash = dir.find_last_of('/');
        dir = slash == std::string::npos ? std::string() : dir.substr(0, slash + 1);
        std::vector<std::vector<uint8_t>> buffers;
        const Json* list = doc.get("buffers");
        for (size_t i = 0; i < (list ? list->size() : 0); i++) {
            const Json* uri = list->at(i)->get("uri");
            if (!uri || uri->kind != Json::STRING) return TT_UNSUPPORTED;  // the .glb binary chunk
            buffers.emplace_back();
            const std::string& u = uri->text;
            if (u.compare(0, 5, "data:") == 0) {
                size_t comma = u.find(',');
                if (comma == std::string::npos || u.rfind(";base64", comma) == std::string::npos ||
                    !base64(u, comma + 1, buffers.back()))
                    return TT_UNSUPPORTED;
            } else if (!read_file(dir + u, buffers.back())) {
                return TT_INVALID_ARGUMENT;
            }
            long long length = index_of(list->at(i)->get("byteLength"));
            if (length < 0 || size_t(length) > buffers.back().size()) return TT_INVALID_ARGUMENT;
        }
        Importer(doc, buffers, scene, dir).run();
        return TT_OK;
    }
} catch (const Failure& f) {
    return f.status;
} catch (const std::bad_alloc&) {
    return TT_OUT_OF_MEMORY;
} catch (...) {
    // Anything else comes from a malformed document.
    return TT_INVALID_ARGUMENT;
}
