// Host side of the library: asset processing, node-tree flattening and the
// camera constants, i.e. what the Rust library does on the CPU before the
// hot path runs.
//
//   Scene::insert_mesh          src/asset/mod.rs:89-161
//   emissive_triangles_in_model src/asset/mod.rs:163-193
//   depth_first_node_tree       src/asset/mod.rs:207-224
//   generate_normals            src/asset/mod.rs:226-242
//   generate_tangents           src/asset/mod.rs:244-262 (see insert_mesh)
//   bounding_sphere             src/asset/mod.rs:308-321
//   TangentFrame::new           src/asset/normal.rs:18-36
//   angle_around_normal         src/asset/normal.rs:38-52
//   orthonormal                 src/asset/normal.rs:54-61
//   octahedron / quantize       src/math/mod.rs:3-41
//   Material defaults/constants src/asset/material.rs:35-175, 233-266
//   Scene::insert_material      src/asset/material.rs:305-357
//   scene::Scene::new flatten   src/scene.rs:32-64
//   Camera / Config             src/camera.rs:6-56, src/config.rs:23-43
//   Renderer constants          src/lib.rs:125-142
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <set>
#include <string>
#include <vector>

#include "abi_guard.h"
#include "../../include/tinytrace.h"
#include "host_math.h"

static_assert(sizeof(tt_constants) == 368, "Constants layout");
static_assert(sizeof(tt_vertex) == 8, "Vertex layout");
static_assert(sizeof(tt_mesh_record) == 32, "ProcessedMesh layout");
static_assert(sizeof(tt_material_constants) == 192, "MaterialConstants layout");
static_assert(sizeof(tt_material) == 304, "ProcessedMaterial layout");
static_assert(sizeof(tt_instance) == 208, "Instance layout");
static_assert(sizeof(tt_emissive_triangle) == 36, "EmissiveTriangle layout");
static_assert(offsetof(tt_emissive_triangle, hash) == 18, "EmissiveTriangle.hash");
static_assert(offsetof(tt_emissive_triangle, instance_index) == 32, "EmissiveTriangle.instance");
static_assert(offsetof(tt_constants, frame_index) == 336, "Constants.frame_index");
static_assert(offsetof(tt_constants, screen_size) == 352, "Constants.screen_size");

namespace tt {

namespace {

const float TAU_F = 6.28318530717958647692f;

// src/math/mod.rs:3-13
void encode_octahedron(Vec3 n, float& u, float& v) {
    float s = std::fabs(n.x) + std::fabs(n.y) + std::fabs(n.z);
    n = n / s;
    auto signum = [](float x) { return std::signbit(x) ? -1.0f : 1.0f; };
    if (n.z < 0.0f) {
        u = (1.0f - std::fabs(n.y)) * signum(n.x) * 0.5f + 0.5f;
        v = (1.0f - std::fabs(n.x)) * signum(n.y) * 0.5f + 0.5f;
    } else {
        u = n.x * 0.5f + 0.5f;
        v = n.y * 0.5f + 0.5f;
    }
}

// src/math/mod.rs:15-22
Vec3 decode_octahedron(float u, float v) {
    float x = u * 2.0f - 1.0f, y = v * 2.0f - 1.0f;
    Vec3 n(x, y, 1.0f - std::fabs(x) - std::fabs(y));
    float t = std::min(std::max(-n.z, 0.0f), 1.0f);
    n.x += n.x >= 0.0f ? -t : t;
    n.y += n.y >= 0.0f ? -t : t;
    return normalize(n);
}

// src/math/mod.rs:24-37
int32_t quantize_unorm(float v, int n) {
    float scale = float((1 << n) - 1);
    v = v >= 0.0f ? v : 0.0f;
    v = v <= 1.0f ? v : 1.0f;
    return int32_t(v * scale + 0.5f);
}
int32_t quantize_snorm(float v, int n) {
    float scale = float((1 << (n - 1)) - 1);
    float round = v >= 0.0f ? 0.5f : -0.5f;
    v = v >= -1.0f ? v : -1.0f;
    v = v <= 1.0f ? v : 1.0f;
    return int32_t(v * scale + round);
}
float dequantize_unorm(uint32_t value, uint32_t n) { return float(value) / float((1 << n) - 1); }

// src/asset/normal.rs:54-61
Vec3 orthonormal(Vec3 n) {
    if (std::fabs(n.x) > std::fabs(n.z)) return normalize(Vec3(-n.y, n.x, 0.0f));
    return normalize(Vec3(0.0f, -n.z, n.y));
}

// src/asset/normal.rs:38-52
float angle_around_normal(Vec3 n, Vec3 a, Vec3 b) {
    float d = dot(a, b);
    float det = dot(n, cross(a, b));
    float angle = std::atan2(det, d);
    return angle < 0.0f ? angle + 2.0f * float(M_PI) : angle;
}

// src/asset/normal.rs:18-36
uint32_t tangent_frame(Vec3 normal, const float tangent[4]) {
    float ou, ov;
    encode_octahedron(normal, ou, ov);
    uint32_t u = uint32_t(quantize_unorm(ou, 10));
    uint32_t v = uint32_t(quantize_unorm(ov, 10));
    Vec3 ortho = orthonormal(decode_octahedron(dequantize_unorm(u, 10), dequantize_unorm(v, 10)));
    float angle = angle_around_normal(normal, ortho, Vec3(tangent[0], tangent[1], tangent[2])) / TAU_F;
    uint32_t a = uint32_t(quantize_unorm(angle, 11));
    return (tangent[3] >= 0.0f ? 1u : 0u) << 31 | a << 20 | v << 10 | u;
}

struct Model {
    uint32_t mesh, material;
};

struct Node {
    Mat4 transform = Mat4::identity();
    std::vector<Model> models;
    std::vector<uint32_t> children;
};

}  // namespace

struct AssetScene {
    std::vector<int16_t> positions;  // 3 per vertex
    std::vector<uint32_t> indices;
    std::vector<tt_vertex> vertices;
    std::vector<tt_material> materials;
    std::vector<tt_mesh_record> meshes;
    std::set<uint32_t> emissive_materials;
    std::vector<Node> nodes;
    // Flattened output.
    std::vector<tt_instance> instances;
    std::vector<tt_emissive_triangle> emissive;
    // Processed textures (scene.textures) and their flattened form.
    struct Texture {
        uint32_t format = 0, width = 0, height = 0;
        std::vector<std::vector<uint8_t>> levels;
    };
    std::vector<Texture> textures;
    bool block_compression = true;  // colour / normal textures as BC1 / BC5
    std::vector<tt_texture_record> texture_records;
    std::vector<uint8_t> texture_bytes;
    std::string error;
    AssetScene() { nodes.emplace_back(); }
};

}  // namespace tt

struct tt_asset_scene : tt::AssetScene {};

using namespace tt;

namespace {

// generate_mip_chain (src/asset/texture.rs:121-143): each level averages the
// 2 x 2 texels (2x, 2y), (2x, 2y + 1), (2x + 1, 2y), (2x + 1, 2y + 1) of
// the one above, summed in that order, until a side is below 2.
std::vector<std::vector<float>> mip_chain(const float* texels, uint32_t width, uint32_t height, int comps,
                                          std::vector<std::pair<uint32_t, uint32_t>>& extents) {
    std::vector<std::vector<float>> chain;
    chain.emplace_back(texels, texels + size_t(width) * height * comps);
    extents.assign(1, {width, height});
    while (extents.back().first >= 2 && extents.back().second >= 2) {
        uint32_t w = extents.back().first, h = extents.back().second;
        uint32_t nw = w >> 1, nh = h >> 1;
        const std::vector<float>& last = chain.back();
        std::vector<float> next(size_t(nw) * nh * comps);
        for (uint32_t y = 0; y < nh; y++)
            for (uint32_t x = 0; x < nw; x++)
                for (int c = 0; c < comps; c++) {
                    auto at = [&](uint32_t xx, uint32_t yy) { return last[(size_t(yy) * w + xx) * comps + c]; };
                    float sum = 0.0f;
                    sum = sum + at(2 * x, 2 * y);
                    sum = sum + at(2 * x, 2 * y + 1);
                    sum = sum + at(2 * x + 1, 2 * y);
                    sum = sum + at(2 * x + 1, 2 * y + 1);
                    next[(size_t(y) * nw + x) * comps + c] = sum / 4.0f;
                }
        chain.push_back(std::move(next));
        extents.push_back({nw, nh});
    }
    return chain;
}

// ---------------------------------------------------- block compression --
// This library's BC1 / BC4 encoder (the reference calls texpresso's
// ClusterFit, a third-party encoder).  Blocks decode as include/tinytrace.h
// states; the encoder picks each texel's index against exactly that decoded
// palette.

// BC1 palette (8-bit RGB) of endpoints c0, c1; entries usable in the mode.
int bc1_palette(uint32_t c0, uint32_t c1, int p[4][3]) {
    const uint32_t e[2] = {c0, c1};
    for (int j = 0; j < 2; j++) {
        int r5 = int(e[j] >> 11), g6 = int((e[j] >> 5) & 63u), b5 = int(e[j] & 31u);
        p[j][0] = (r5 << 3) | (r5 >> 2);
        p[j][1] = (g6 << 2) | (g6 >> 4);
        p[j][2] = (b5 << 3) | (b5 >> 2);
    }
    for (int k = 0; k < 3; k++) {
        if (c0 > c1) {
            p[2][k] = (2 * p[0][k] + p[1][k] + 1) / 3;
            p[3][k] = (p[0][k] + 2 * p[1][k] + 1) / 3;
        } else {
            p[2][k] = (p[0][k] + p[1][k] + 1) / 2;
            p[3][k] = 0;
        }
    }
    return c0 > c1 ? 4 : 3;  // the three-colour mode's fourth entry is transparent black
}

uint32_t rgb565(const double e[3]) {
    auto q = [](double v, int bits) {
        const int top = (1 << bits) - 1;
        long r = std::lround(std::min(255.0, std::max(0.0, v)) * top / 255.0);
        return uint32_t(std::min<long>(top, std::max<long>(0, r)));
    };
    return (q(e[0], 5) << 11) | (q(e[1], 6) << 5) | q(e[2], 5);
}

// Endpoints (c0 >= c1, so four colours unless equal) and nearest-entry
// indices; returns the squared error.
int64_t bc1_fit(const int px[16][3], uint32_t c0, uint32_t c1, uint32_t& bits, uint32_t& e0, uint32_t& e1) {
    if (c0 < c1) std::swap(c0, c1);
    int p[4][3];
    const int n = bc1_palette(c0, c1, p);
    int64_t err = 0;
    bits = 0;
    for (int t = 0; t < 16; t++) {
        int best = 0;
        int64_t best_d = -1;
        for (int k = 0; k < n; k++) {
            int64_t d = 0;
            for (int c = 0; c < 3; c++) d += int64_t(px[t][c] - p[k][c]) * (px[t][c] - p[k][c]);
            if (best_d < 0 || d < best_d) {
                best_d = d;
                best = k;
            }
        }
        bits |= uint32_t(best) << (2 * t);
        err += best_d;
    }
    e0 = c0;
    e1 = c1;
    return err;
}

// One BC1 block of 16 texels (row-major within the block): endpoints from
// the extent of the colours along their principal axis, then one
// least-squares refit of the endpoints to the chosen indices.
void bc1_block(const int px[16][3], uint8_t out[8]) {
    double mean[3] = {0.0, 0.0, 0.0};
    for (int t = 0; t < 16; t++)
        for (int c = 0; c < 3; c++) mean[c] += px[t][c] / 16.0;
    double cov[3][3] = {};
    for (int t = 0; t < 16; t++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) cov[a][b] += (px[t][a] - mean[a]) * (px[t][b] - mean[b]);
    double axis[3] = {1.0, 1.0, 1.0};
    for (int it = 0; it < 8; it++) {
        double v[3];
        for (int a = 0; a < 3; a++) v[a] = cov[a][0] * axis[0] + cov[a][1] * axis[1] + cov[a][2] * axis[2];
        double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (len < 1e-12) break;  // no spread: any axis
        for (int a = 0; a < 3; a++) axis[a] = v[a] / len;
    }
    double lo = 0.0, hi = 0.0;
    for (int t = 0; t < 16; t++) {
        double s = 0.0;
        for (int c = 0; c < 3; c++) s += (px[t][c] - mean[c]) * axis[c];
        lo = std::min(lo, s);
        hi = std::max(hi, s);
    }
    double a[3], b[3];
    for (int c = 0; c < 3; c++) {
        a[c] = mean[c] + axis[c] * hi;
        b[c] = mean[c] + axis[c] * lo;
    }
    uint32_t bits, e0, e1;
    int64_t err = bc1_fit(px, rgb565(a), rgb565(b), bits, e0, e1);
    if (e0 > e1) {
        // Least squares over the four-colour weights of the chosen indices.
        const double wk[4] = {1.0, 0.0, 2.0 / 3.0, 1.0 / 3.0};
        double A = 0.0, B = 0.0, Cc = 0.0, X[3] = {}, Y[3] = {};
        for (int t = 0; t < 16; t++) {
            double w = wk[(bits >> (2 * t)) & 3u];
            A += w * w;
            B += w * (1.0 - w);
            Cc += (1.0 - w) * (1.0 - w);
            for (int c = 0; c < 3; c++) {
                X[c] += w * px[t][c];
                Y[c] += (1.0 - w) * px[t][c];
            }
        }
        const double det = A * Cc - B * B;
        if (std::fabs(det) > 1e-9) {
            for (int c = 0; c < 3; c++) {
                a[c] = (Cc * X[c] - B * Y[c]) / det;
                b[c] = (A * Y[c] - B * X[c]) / det;
            }
            uint32_t bits2, f0, f1;
            int64_t err2 = bc1_fit(px, rgb565(a), rgb565(b), bits2, f0, f1);
            if (err2 < err) {
                bits = bits2;
                e0 = f0;
                e1 = f1;
            }
        }
    }
    out[0] = uint8_t(e0 & 0xffu);
    out[1] = uint8_t(e0 >> 8);
    out[2] = uint8_t(e1 & 0xffu);
    out[3] = uint8_t(e1 >> 8);
    for (int k = 0; k < 4; k++) out[4 + k] = uint8_t(bits >> (8 * k));
}

// One BC4 block: endpoints max > min (eight values), indices to the
// nearest decoded value; a flat block is its value in both endpoints.
void bc4_block(const int v[16], uint8_t out[8]) {
    int e0 = 0, e1 = 255;
    for (int t = 0; t < 16; t++) {
        e0 = std::max(e0, v[t]);
        e1 = std::min(e1, v[t]);
    }
    int p[8];
    p[0] = e0;
    p[1] = e1;
    for (int k = 2; k < 8; k++) p[k] = ((8 - k) * e0 + (k - 1) * e1 + 3) / 7;
    uint64_t bits = 0;
    if (e0 > e1)
        for (int t = 0; t < 16; t++) {
            int best = 0;
            for (int k = 1; k < 8; k++)
                if (std::abs(v[t] - p[k]) < std::abs(v[t] - p[best])) best = k;
            bits |= uint64_t(best) << (3 * t);
        }
    out[0] = uint8_t(e0);
    out[1] = uint8_t(e1);
    for (int k = 0; k < 6; k++) out[2 + k] = uint8_t(bits >> (8 * k));
}

// A level of `comps`-byte texels (w x h) as BC1 (comps 3) or BC5 (comps 2)
// blocks; texels past the edge repeat the edge texel.
std::vector<uint8_t> compress_level(const std::vector<uint8_t>& texels, uint32_t w, uint32_t h, int comps) {
    const uint32_t bx = (w + 3) / 4, by = (h + 3) / 4;
    const size_t block_bytes = comps == 3 ? 8 : 16;
    std::vector<uint8_t> out(size_t(bx) * by * block_bytes);
    for (uint32_t j = 0; j < by; j++)
        for (uint32_t i = 0; i < bx; i++) {
            int px[16][3];
            for (int t = 0; t < 16; t++) {
                uint32_t x = std::min(w - 1, 4 * i + uint32_t(t % 4)), y = std::min(h - 1, 4 * j + uint32_t(t / 4));
                for (int c = 0; c < comps; c++) px[t][c] = texels[(size_t(y) * w + x) * comps + c];
            }
            uint8_t* dst = &out[(size_t(j) * bx + i) * block_bytes];
            if (comps == 3) {
                bc1_block(px, dst);
            } else {
                int r[16], g[16];
                for (int t = 0; t < 16; t++) {
                    r[t] = px[t][0];
                    g[t] = px[t][1];
                }
                bc4_block(r, dst);
                bc4_block(g, dst + 8);
            }
        }
    return out;
}

// ProcessedTexture::from_* (texture.rs:57-119): colour -> BC1 of the
// quantized RGB (the reference appends alpha 1 before compressing),
// normal -> BC5 of the quantized octahedral encoding, scalar -> R8 unorm or
// R16 float (material.rs:312-346 picks which).
AssetScene::Texture process_texture(const tt_texture_desc& t, bool compress) {
    AssetScene::Texture out;
    bool colour = t.channel < TT_TEXTURE_GEOMETRY_NORMAL;
    bool normal = t.channel == TT_TEXTURE_GEOMETRY_NORMAL;
    bool half = t.channel == TT_TEXTURE_SPECULAR_IOR || t.channel == TT_TEXTURE_SPECULAR_ROTATION ||
        t.channel == TT_TEXTURE_TRANSMISSION_DEPTH || t.channel == TT_TEXTURE_COAT_IOR ||
        t.channel == TT_TEXTURE_COAT_ROTATION || t.channel == TT_TEXTURE_EMISSION_LUMINANCE;
    if (colour) out.format = compress ? TT_TEXTURE_FORMAT_BC1_SRGB : TT_TEXTURE_FORMAT_RGBA8_SRGB;
    else if (normal) out.format = compress ? TT_TEXTURE_FORMAT_BC5_UNORM : TT_TEXTURE_FORMAT_RG8_UNORM;
    else out.format = half ? TT_TEXTURE_FORMAT_R16_FLOAT : TT_TEXTURE_FORMAT_R8_UNORM;
    out.width = t.width;
    out.height = t.height;
    int comps = colour || normal ? 3 : 1;
    std::vector<std::pair<uint32_t, uint32_t>> extents;
    std::vector<std::vector<float>> chain = mip_chain(t.texels, t.width, t.height, comps, extents);
    for (size_t li = 0; li < chain.size(); li++) {
        const std::vector<float>& level = chain[li];
        size_t n = level.size() / comps;
        std::vector<uint8_t> bytes;
        for (size_t i = 0; i < n; i++) {
            const float* v = &level[i * comps];
            if (colour) {
                for (int c = 0; c < 3; c++) bytes.push_back(uint8_t(quantize_unorm(v[c], 8)));
                if (!compress) bytes.push_back(uint8_t(quantize_unorm(1.0f, 8)));
            } else if (normal) {
                float u, w;
                encode_octahedron(Vec3(v[0], v[1], v[2]), u, w);
                bytes.push_back(uint8_t(quantize_unorm(u, 8)));
                bytes.push_back(uint8_t(quantize_unorm(w, 8)));
            } else if (half) {
                uint16_t h = float_to_half(v[0]);
                bytes.push_back(uint8_t(h & 0xffu));
                bytes.push_back(uint8_t(h >> 8));
            } else {
                bytes.push_back(uint8_t(quantize_unorm(v[0], 8)));
            }
        }
        if (compress && (colour || normal))
            bytes = compress_level(bytes, extents[li].first, extents[li].second, colour ? 3 : 2);
        out.levels.push_back(std::move(bytes));
    }
    return out;
}

}  // namespace

extern "C" {

void tt_material_desc_default(tt_material_desc* m) {
    *m = tt_material_desc{};
    m->base_weight = 1.0f;
    m->base_color[0] = m->base_color[1] = m->base_color[2] = 0.8f;
    m->base_metalness = 0.0f;
    m->base_diffuse_roughness = 0.0f;
    m->specular_weight = 1.0f;
    m->specular_color[0] = m->specular_color[1] = m->specular_color[2] = 1.0f;
    m->specular_roughness = 0.3f;
    m->specular_roughness_anisotropy = 0.0f;
    m->specular_ior = 1.5f;
    m->specular_rotation = 0.0f;
    m->transmission_weight = 0.0f;
    m->transmission_color[0] = m->transmission_color[1] = m->transmission_color[2] = 1.0f;
    m->transmission_depth = 0.0f;
    m->coat_weight = 0.0f;
    m->coat_color[0] = m->coat_color[1] = m->coat_color[2] = 1.0f;
    m->coat_roughness = 0.0f;
    m->coat_roughness_anisotropy = 0.0f;
    m->coat_ior = 1.6f;
    m->coat_darkening = 1.0f;
    m->coat_rotation = 0.0f;
    m->fuzz_weight = 0.0f;
    m->fuzz_color[0] = m->fuzz_color[1] = m->fuzz_color[2] = 1.0f;
    m->fuzz_roughness = 0.5f;
    m->emission_luminance = 0.0f;
    m->emission_color[0] = m->emission_color[1] = m->emission_color[2] = 1.0f;
    m->geometry_opacity = 1.0f;
    m->geometry_normal[0] = 0.0f;
    m->geometry_normal[1] = 0.0f;
    m->geometry_normal[2] = 1.0f;
}

void tt_camera_default(tt_camera* c) {
    c->position[0] = c->position[1] = c->position[2] = 0.0f;
    c->fov = float(M_PI / 2.0);
    c->z_near = 0.1f;
    c->z_far = 400.0f;
    c->yaw = 0.0f;
    c->pitch = 0.0f;
}

void tt_config_default(tt_config* c) {
    c->light_sampling = TT_LIGHT_SAMPLING_NEXT_EVENT_ESTIMATION;
    c->tonemap = 1;
    c->bounce_count = 4;
    c->sample_count = 1;
}

tt_status tt_constants_build(const tt_camera* camera, const tt_config* config, uint32_t width, uint32_t height,
                             uint32_t frame_index, uint32_t accumulated_frame_count, tt_constants* out) {
    if (!camera || !config || !out || width == 0 || height == 0) return TT_INVALID_ARGUMENT;
    // Camera::forward (src/camera.rs:31-36)
    Vec3 forward = normalize(Vec3(std::cos(camera->yaw) * std::cos(camera->pitch), std::sin(camera->pitch),
                                  std::sin(camera->yaw) * std::cos(camera->pitch)));
    Vec3 eye(camera->position[0], camera->position[1], camera->position[2]);
    Vec3 up(0.0f, 1.0f, 0.0f);
    // Mat4::look_at_rh(eye, eye + forward, up) = look_to_rh(eye, center - eye, up)
    Vec3 center = eye + forward;
    Vec3 f = normalize(center - eye);
    Vec3 s = normalize(cross(f, up));
    Vec3 u = cross(s, f);
    float view_cols[16] = {s.x, u.x, -f.x, 0.0f, s.y, u.y, -f.y, 0.0f, s.z, u.z, -f.z, 0.0f,
                           -dot(eye, s), -dot(eye, u), dot(eye, f), 1.0f};
    Mat4 view = Mat4::from_cols(view_cols);
    // Mat4::perspective_infinite_reverse_rh with z_axis[3] = 1
    float aspect = float(width) / float(height);
    float fl = 1.0f / std::tan(0.5f * camera->fov);
    float proj_cols[16] = {fl / aspect, 0, 0, 0, 0, fl, 0, 0, 0, 0, 0, 1.0f, 0, 0, camera->z_near, 0};
    Mat4 proj = Mat4::from_cols(proj_cols);
    Mat4 proj_view = mul(proj, view);
    Mat4 inverse_view = inverse(view);
    Mat4 inverse_proj = inverse(proj);
    std::memcpy(out->view, view.m, 64);
    std::memcpy(out->proj, proj.m, 64);
    std::memcpy(out->proj_view, proj_view.m, 64);
    std::memcpy(out->inverse_view, inverse_view.m, 64);
    std::memcpy(out->inverse_proj, inverse_proj.m, 64);
    out->camera_position[0] = eye.x;
    out->camera_position[1] = eye.y;
    out->camera_position[2] = eye.z;
    out->camera_position[3] = 0.0f;
    out->frame_index = frame_index;
    out->accumulated_frame_count = accumulated_frame_count;
    out->sample_count = config->sample_count;
    out->bounce_count = config->bounce_count;
    out->screen_size[0] = width;
    out->screen_size[1] = height;
    out->tonemap = config->tonemap;
    out->light_sampling = config->light_sampling;
    return TT_OK;
}

tt_status tt_asset_scene_create(tt_asset_scene** out) {
    if (!out) return TT_INVALID_ARGUMENT;
    *out = new (std::nothrow) tt_asset_scene();
    return *out ? TT_OK : TT_OUT_OF_MEMORY;
}

void tt_asset_scene_destroy(tt_asset_scene* scene) try {
    delete scene;
} catch (...) {
}

tt_status tt_asset_insert_mesh(tt_asset_scene* scene, const tt_mesh_desc* mesh, uint32_t* index) try {
    if (!scene || !mesh || !mesh->positions || !mesh->texture_coordinates) return TT_INVALID_ARGUMENT;
    uint32_t vc = mesh->vertex_count;
    if (vc == 0) return TT_INVALID_ARGUMENT;
    if (mesh->indices) {
        if (mesh->index_count % 3 != 0) return TT_INVALID_ARGUMENT;
        for (uint32_t i = 0; i < mesh->index_count; i++)
            if (mesh->indices[i] >= vc) return TT_INVALID_ARGUMENT;
    } else if (vc % 3 != 0) {
        return TT_INVALID_ARGUMENT;
    }
    auto pos = [&](uint32_t i) { return Vec3(mesh->positions[3 * i], mesh->positions[3 * i + 1], mesh->positions[3 * i + 2]); };

    uint32_t mesh_index = uint32_t(scene->meshes.size());
    uint32_t vertex_offset = uint32_t(scene->positions.size() / 3);

    // bounding_sphere (src/asset/mod.rs:308-321): running mean, max distance.
    Vec3 center;
    for (uint32_t i = 0; i < vc; i++) center = center + (pos(i) - center) / float(i + 1);
    float radius = 0.0f;
    for (uint32_t i = 0; i < vc; i++) radius = std::max(radius, length(pos(i) - center));
    if (!(radius > 0.0f)) return TT_INVALID_ARGUMENT;

    std::vector<int16_t> quantized(3 * size_t(vc));
    for (uint32_t i = 0; i < vc; i++) {
        Vec3 p = (pos(i) - center) / radius;
        if (std::max(std::fabs(p.x), std::max(std::fabs(p.y), std::fabs(p.z))) > 1.0f) return TT_INVALID_ARGUMENT;
        quantized[3 * i + 0] = int16_t(quantize_snorm(p.x, 16));
        quantized[3 * i + 1] = int16_t(quantize_snorm(p.y, 16));
        quantized[3 * i + 2] = int16_t(quantize_snorm(p.z, 16));
    }

    uint32_t index_offset = uint32_t(scene->indices.size());
    uint32_t index_count;
    if (mesh->indices) {
        scene->indices.insert(scene->indices.end(), mesh->indices, mesh->indices + mesh->index_count);
        index_count = mesh->index_count;
    } else {
        for (uint32_t i = 0; i < vc; i++) scene->indices.push_back(i);
        index_count = vc;
    }

    // generate_normals (src/asset/mod.rs:226-242) when absent.
    std::vector<Vec3> normals(vc);
    if (mesh->normals) {
        for (uint32_t i = 0; i < vc; i++)
            normals[i] = Vec3(mesh->normals[3 * i], mesh->normals[3 * i + 1], mesh->normals[3 * i + 2]);
    } else {
        const uint32_t* idx = scene->indices.data() + index_offset;
        for (uint32_t t = 0; t + 2 < index_count; t += 3) {
            uint32_t i0 = idx[t], i1 = idx[t + 1], i2 = idx[t + 2];
            Vec3 n = cross(pos(i1) - pos(i0), pos(i2) - pos(i0));
            normals[i0] = normals[i0] + n;
            normals[i1] = normals[i1] + n;
            normals[i2] = normals[i2] + n;
        }
        for (auto& n : normals) n = normalize(n);
    }

    // generate_tangents (src/asset/mod.rs:244-262) when absent.  The
    // reference runs MikkTSpace (mikktspace 0.3.0) and negates the sign it
    // returns.  This restates MikkTSpace's result for smooth, non-degenerate
    // texture mappings: per vertex, the sum of the triangles' tangents and
    // bitangents from the texture-coordinate derivatives (Lengyel 2001),
    // Gram-Schmidt against the normal, sign = +1 when cross(n, t) points
    // along the bitangent, then negated as the reference does.  Where the
    // texture coordinates span no area (e.g. none were given) the tangent is
    // orthonormal(n) with the negated sign -1.
    std::vector<float> generated;
    const float* tangents = mesh->tangents;
    if (!tangents) {
        std::vector<Vec3> tsum(vc), bsum(vc);
        const uint32_t* idx = scene->indices.data() + index_offset;
        const float* uv = mesh->texture_coordinates;
        for (uint32_t t = 0; t + 2 < index_count; t += 3) {
            uint32_t i0 = idx[t], i1 = idx[t + 1], i2 = idx[t + 2];
            Vec3 e1 = pos(i1) - pos(i0), e2 = pos(i2) - pos(i0);
            float du1 = uv[2 * i1] - uv[2 * i0], dv1 = uv[2 * i1 + 1] - uv[2 * i0 + 1];
            float du2 = uv[2 * i2] - uv[2 * i0], dv2 = uv[2 * i2 + 1] - uv[2 * i0 + 1];
            float r = du1 * dv2 - du2 * dv1;
            if (!(std::fabs(r) > 0.0f) || !std::isfinite(r)) continue;
            float f = 1.0f / r;
            Vec3 tv = (e1 * dv2 - e2 * dv1) * f, bv = (e2 * du1 - e1 * du2) * f;
            for (uint32_t k : {i0, i1, i2}) {
                tsum[k] = tsum[k] + tv;
                bsum[k] = bsum[k] + bv;
            }
        }
        generated.resize(4 * size_t(vc));
        for (uint32_t i = 0; i < vc; i++) {
            Vec3 n = normals[i];
            Vec3 t = tsum[i] - n * dot(n, tsum[i]);
            float len = length(t);
            bool ok = len > 1e-20f && std::isfinite(len);
            t = ok ? t / len : orthonormal(n);
            float sign = ok && dot(cross(n, t), bsum[i]) < 0.0f ? -1.0f : 1.0f;
            generated[4 * size_t(i) + 0] = t.x;
            generated[4 * size_t(i) + 1] = t.y;
            generated[4 * size_t(i) + 2] = t.z;
            generated[4 * size_t(i) + 3] = -sign;
        }
        tangents = generated.data();
    }

    scene->positions.insert(scene->positions.end(), quantized.begin(), quantized.end());
    for (uint32_t i = 0; i < vc; i++) {
        tt_vertex v;
        v.texture_coordinates[0] = float_to_half(mesh->texture_coordinates[2 * i]);
        v.texture_coordinates[1] = float_to_half(mesh->texture_coordinates[2 * i + 1]);
        v.tangent_frame = tangent_frame(normals[i], tangents + 4 * size_t(i));
        scene->vertices.push_back(v);
    }
    tt_mesh_record rec;
    rec.center[0] = center.x;
    rec.center[1] = center.y;
    rec.center[2] = center.z;
    rec.radius = radius;
    rec.vertex_offset = vertex_offset;
    rec.vertex_count = vc;
    rec.index_offset = index_offset;
    rec.index_count = index_count;
    scene->meshes.push_back(rec);
    if (index) *index = mesh_index;
    return TT_OK;
} catch (...) {
    return tt::exception_status(scene ? &scene->error : nullptr);
}

tt_status tt_asset_insert_material(tt_asset_scene* scene, const tt_material_desc* d, uint32_t* index) try {
    return tt_asset_insert_material_textured(scene, d, nullptr, 0, index);
} catch (...) {
    return tt::exception_status(scene ? &scene->error : nullptr);
}

tt_status tt_asset_set_block_compression(tt_asset_scene* scene, uint32_t enable) try {
    if (!scene) return TT_INVALID_ARGUMENT;
    scene->block_compression = enable != 0;
    return TT_OK;
} catch (...) {
    return tt::exception_status(scene ? &scene->error : nullptr);
}

tt_status tt_asset_insert_material_textured(tt_asset_scene* scene, const tt_material_desc* d,
                                            const tt_texture_desc* textures, uint32_t texture_count, uint32_t* index) try {
    if (!scene || !d || (texture_count && !textures)) return TT_INVALID_ARGUMENT;
    const tt_texture_desc* by_channel[TT_TEXTURE_CHANNELS] = {};
    for (uint32_t k = 0; k < texture_count; k++) {
        const tt_texture_desc& t = textures[k];
        if (t.channel >= TT_TEXTURE_CHANNELS || by_channel[t.channel] || !t.texels || t.width == 0 || t.height == 0 ||
            t.width > 32768 || t.height > 32768)
            return TT_INVALID_ARGUMENT;
        by_channel[t.channel] = &t;
    }
    uint32_t i = uint32_t(scene->materials.size());
    if (d->emission_luminance > 0.0f) scene->emissive_materials.insert(i);
    tt_material m;
    std::memset(&m, 0, sizeof(m));
    tt_material_constants& c = m.constants;
    auto v4 = [](float* dst, const float* src) {
        dst[0] = src[0];
        dst[1] = src[1];
        dst[2] = src[2];
        dst[3] = 0.0f;
    };
    v4(c.base_color, d->base_color);
    v4(c.specular_color, d->specular_color);
    v4(c.transmission_color, d->transmission_color);
    v4(c.coat_color, d->coat_color);
    v4(c.fuzz_color, d->fuzz_color);
    v4(c.emission_color, d->emission_color);
    v4(c.geometry_normal, d->geometry_normal);
    c.base_weight = d->base_weight;
    c.base_metalness = d->base_metalness;
    c.base_diffuse_roughness = d->base_diffuse_roughness;
    c.specular_weight = d->specular_weight;
    c.specular_roughness = d->specular_roughness;
    c.specular_roughness_anisotropy = d->specular_roughness_anisotropy;
    c.specular_ior = d->specular_ior;
    c.specular_rotation = d->specular_rotation;
    c.transmission_weight = d->transmission_weight;
    c.transmission_depth = d->transmission_depth;
    c.coat_weight = d->coat_weight;
    c.coat_roughness = d->coat_roughness;
    c.coat_roughness_anisotropy = d->coat_roughness_anisotropy;
    c.coat_ior = d->coat_ior;
    c.coat_darkening = d->coat_darkening;
    c.coat_rotation = d->coat_rotation;
    c.fuzz_weight = d->fuzz_weight;
    c.fuzz_roughness = d->fuzz_roughness;
    c.emission_luminance = d->emission_luminance;
    c.geometry_opacity = d->geometry_opacity;
    for (int t = 0; t < 27; t++) m.textures[t] = TT_INVALID_INDEX;
    m.padding = 0;
    // Textured channels: processed in the order insert_material visits them
    // (material.rs:312-346), each index the next in scene.textures; the
    // channel's constant is its type's default, zero (constant_or_default).
    static const int order[TT_TEXTURE_CHANNELS] = {0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 15,
                                                   17, 18, 19, 21, 23, 24, 26, 25, 20, 13, 16, 14, 22};
    float* constants = reinterpret_cast<float*>(&m.constants);
    for (int k = 0; k < TT_TEXTURE_CHANNELS; k++) {
        int ch = order[k];
        const tt_texture_desc* t = by_channel[ch];
        if (!t) continue;
        m.textures[ch] = uint32_t(scene->textures.size());
        scene->textures.push_back(process_texture(*t, scene->block_compression));
        if (ch < 7) {
            for (int c = 0; c < 3; c++) constants[4 * ch + c] = 0.0f;
        } else {
            constants[28 + (ch - 7)] = 0.0f;
        }
    }
    // is_emissive (material.rs:178-183): a positive constant luminance, or a
    // luminance texture with any positive texel.
    bool emissive = m.constants.emission_luminance > 0.0f;
    if (const tt_texture_desc* t = by_channel[TT_TEXTURE_EMISSION_LUMINANCE]) {
        emissive = false;
        for (size_t k = 0; k < size_t(t->width) * t->height; k++) emissive = emissive || t->texels[k] > 0.0f;
    }
    if (emissive)
        scene->emissive_materials.insert(i);
    else
        scene->emissive_materials.erase(i);
    scene->materials.push_back(m);
    if (index) *index = i;
    return TT_OK;
} catch (...) {
    return tt::exception_status(scene ? &scene->error : nullptr);
}

tt_status tt_asset_add_node(tt_asset_scene* scene, uint32_t parent, const float transform[16], uint32_t* index) try {
    if (!scene || !transform || parent >= scene->nodes.size()) return TT_INVALID_ARGUMENT;
    uint32_t i = uint32_t(scene->nodes.size());
    Node n;
    n.transform = Mat4::from_cols(transform);
    scene->nodes.push_back(n);
    scene->nodes[parent].children.push_back(i);
    if (index) *index = i;
    return TT_OK;
} catch (...) {
    return tt::exception_status(scene ? &scene->error : nullptr);
}

tt_status tt_asset_add_model(tt_asset_scene* scene, uint32_t node, uint32_t mesh, uint32_t material) try {
    if (!scene || node >= scene->nodes.size() || mesh >= scene->meshes.size() || material >= scene->materials.size())
        return TT_INVALID_ARGUMENT;
    scene->nodes[node].models.push_back(Model{mesh, material});
    return TT_OK;
} catch (...) {
    return tt::exception_status(scene ? &scene->error : nullptr);
}

tt_status tt_asset_flatten(tt_asset_scene* scene, tt_scene_desc* out) try {
    if (!scene || !out) return TT_INVALID_ARGUMENT;
    scene->instances.clear();
    scene->emissive.clear();
    // Depth-first pre-order over the node tree (src/asset/mod.rs:207-224).
    struct Item {
        uint32_t node;
        Mat4 parent;
    };
    std::vector<Item> stack{{0, Mat4::identity()}};
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        const Node& node = scene->nodes[it.node];
        Mat4 transform = mul(it.parent, node.transform);
        for (const Model& model : node.models) {
            const tt_mesh_record& mesh = scene->meshes[model.mesh];
            float sphere_cols[16] = {mesh.radius, 0, 0, 0, 0, mesh.radius, 0, 0, 0, 0, mesh.radius, 0,
                                     mesh.center[0], mesh.center[1], mesh.center[2], 1.0f};
            Mat4 position_transform = mul(transform, Mat4::from_cols(sphere_cols));
            uint32_t instance_index = uint32_t(scene->instances.size());
            tt_instance inst;
            Mat4 normal_transform = transpose(inverse(transform));
            Mat4 inverse_transform = inverse(position_transform);
            std::memcpy(inst.transform, position_transform.m, 64);
            std::memcpy(inst.inverse_transform, inverse_transform.m, 64);
            std::memcpy(inst.normal_transform, normal_transform.m, 64);
            inst.vertex_offset = mesh.vertex_offset;
            inst.index_offset = mesh.index_offset;
            inst.material_index = model.material;
            inst.mesh_index = model.mesh;
            scene->instances.push_back(inst);
            // emissive_triangles_in_model (src/asset/mod.rs:163-193)
            if (scene->emissive_materials.count(model.material)) {
                for (uint32_t t = 0; t < mesh.index_count / 3; t++) {
                    tt_emissive_triangle e;
                    std::memset(&e, 0, sizeof(e));
                    for (int c = 0; c < 3; c++) {
                        uint32_t vi = mesh.vertex_offset + scene->indices[mesh.index_offset + 3 * t + c];
                        for (int a = 0; a < 3; a++) e.positions[c][a] = scene->positions[3 * size_t(vi) + a];
                        e.texture_coordinates[c][0] = scene->vertices[vi].texture_coordinates[0];
                        e.texture_coordinates[c][1] = scene->vertices[vi].texture_coordinates[1];
                    }
                    // `%` binds to the offset only, then truncation to u16.
                    e.hash = uint16_t(mesh.index_offset / 3 + t % 65535u);
                    e.instance_index = instance_index;
                    scene->emissive.push_back(e);
                }
            }
        }
        // Children are visited in order: push in reverse.
        for (auto c = node.children.rbegin(); c != node.children.rend(); ++c) stack.push_back({*c, transform});
    }
    out->positions = scene->positions.data();
    out->position_count = uint32_t(scene->positions.size() / 3);
    out->vertices = scene->vertices.data();
    out->vertex_count = uint32_t(scene->vertices.size());
    out->indices = scene->indices.data();
    out->index_count = uint32_t(scene->indices.size());
    out->meshes = scene->meshes.data();
    out->mesh_count = uint32_t(scene->meshes.size());
    out->materials = scene->materials.data();
    out->material_count = uint32_t(scene->materials.size());
    out->instances = scene->instances.data();
    out->instance_count = uint32_t(scene->instances.size());
    out->emissive_triangles = scene->emissive.data();
    out->emissive_triangle_count = uint32_t(scene->emissive.size());
    // Textures: every level at a 16-byte aligned offset.
    scene->texture_records.clear();
    scene->texture_bytes.clear();
    for (const auto& t : scene->textures) {
        tt_texture_record r;
        std::memset(&r, 0, sizeof(r));
        r.format = t.format;
        r.width = t.width;
        r.height = t.height;
        r.levels = uint32_t(t.levels.size());
        for (size_t k = 0; k < t.levels.size(); k++) {
            scene->texture_bytes.resize((scene->texture_bytes.size() + 15) & ~size_t(15), 0);
            r.offsets[k] = uint32_t(scene->texture_bytes.size());
            scene->texture_bytes.insert(scene->texture_bytes.end(), t.levels[k].begin(), t.levels[k].end());
        }
        scene->texture_records.push_back(r);
    }
    scene->texture_bytes.resize((scene->texture_bytes.size() + 15) & ~size_t(15), 0);
    out->textures = scene->texture_records.empty() ? nullptr : scene->texture_records.data();
    out->texture_count = uint32_t(scene->texture_records.size());
    out->texture_data = scene->texture_bytes.empty() ? nullptr : scene->texture_bytes.data();
    out->texture_data_bytes = scene->texture_bytes.size();
    return TT_OK;
} catch (...) {
    return tt::exception_status(scene ? &scene->error : nullptr);
}

}  // extern "C"
