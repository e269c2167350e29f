// Shared device helpers of the wavefront kernels: wave-level queue
// primitives, the surface at a hit (get_ray_hit + create_surface_basis,
// src/integrate/integrate.glsl:22-110, :195-201), emissive-triangle geometry
// (src/includes/scene.glsl:114-136) and the packed ReSTIR reservoir.
#pragma once

#include "device_bsdf.h"
#include "device_scene.h"
#include "kernels.h"

namespace ttd {

TT_DEV uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

TT_DEV uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// Wave-level stream compaction: every active lane calls this; lanes with
// `pred` get consecutive slots from one atomic add per wave.
TT_DEV uint32_t wave_append(uint32_t* counter, bool pred) {
    uint64_t active = __ballot(1);
    uint64_t mask = __ballot(pred);
    uint32_t leader = uint32_t(__ffsll((unsigned long long)active) - 1);
    uint32_t total = uint32_t(__popcll(mask));
    uint32_t base = 0;
    if (lane_id() == leader && total != 0) base = atomicAdd(counter, total);
    base = __shfl(base, int(leader));
    return base + lanes_below(mask);
}

// Persistent-wave work fetch over n items in tt::WORK_PARTS contiguous
// partitions, each with its own cursor on a line of its own.  A wave starts
// on its workgroup's partition (blockIdx % WORK_PARTS: with workgroups
// dispatched round-robin over the eight XCDs, a partition's waves share an
// XCD and its L2), takes 64 items per atomic and, once that partition is
// drained, moves through the others, skipping drained ones after a plain
// load.  Every value here is wave-uniform.
struct WorkFetch {
    uint32_t* set;
    uint32_t n;
    uint32_t part;
    uint32_t visited;
};

TT_DEV WorkFetch work_fetch(uint32_t* set, uint32_t n) {
    return WorkFetch{set, n, blockIdx.x % tt::WORK_PARTS, 0u};
}

// The next batch: items [first, first + count), 1 <= count <= 64; false
// once every partition is drained.
TT_DEV bool next_batch(WorkFetch& w, uint32_t& first, uint32_t& count) {
    while (w.visited < tt::WORK_PARTS) {
        const uint32_t lo = uint32_t(uint64_t(w.n) * w.part / tt::WORK_PARTS);
        const uint32_t len = uint32_t(uint64_t(w.n) * (w.part + 1u) / tt::WORK_PARTS) - lo;
        uint32_t* cursor = w.set + w.part * tt::WORK_STRIDE;
        bool drained = len == 0u;
        if (!drained && w.visited > 0u)
            drained = __builtin_amdgcn_readfirstlane(
                          __hip_atomic_load(cursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >= len;
        if (!drained) {
            // readfirstlane keeps the wave-uniform cursor values in scalar
            // registers (the whole wave runs this loop, lane 0 included).
            uint32_t got = 0;
            if (lane_id() == 0) got = atomicAdd(cursor, 64u);
            got = __builtin_amdgcn_readfirstlane(got);
            if (got < len) {
                first = lo + got;
                count = len - got < 64u ? len - got : 64u;
                return true;
            }
        }
        w.visited++;
        w.part = w.part + 1u == tt::WORK_PARTS ? 0u : w.part + 1u;
    }
    return false;
}

__device__ __host__ inline SceneView view(const tt::DeviceScene& d) {
    SceneView s;
    s.nodes = static_cast<const float4*>(d.nodes);
    s.tris = static_cast<const float4*>(d.tris);
    s.bvh_insts = static_cast<const float4*>(d.bvh_insts);
    s.inst_boxes = static_cast<const float4*>(d.inst_boxes);
    s.tlas_root = d.tlas_root;
    s.emissive_count = d.emissive_count;
    s.instances = static_cast<const float*>(d.instances);
    s.materials = static_cast<const float4*>(d.materials);
    s.positions = static_cast<const int16_t*>(d.positions);
    s.vertices = static_cast<const uint2*>(d.vertices);
    s.indices = static_cast<const uint32_t*>(d.indices);
    s.emissive = static_cast<const uint32_t*>(d.emissive);
    s.lights = static_cast<const float4*>(d.lights);
    s.tex_records = static_cast<const uint4*>(d.tex_records);
    s.tex_data = static_cast<const uint8_t*>(d.tex_data);
    s.srgb_lut = static_cast<const float*>(d.srgb_lut);
    s.texture_count = d.texture_count;
    return s;
}

TT_DEV Stack make_stack(uint32_t* spill) {
    Stack st;
    st.lane = threadIdx.x;
    st.spill = spill + blockIdx.x * TRACE_BLOCK + threadIdx.x;
    st.spill_stride = gridDim.x * TRACE_BLOCK;
    st.sp = 0;
    return st;
}

// ----------------------------------------------------------- tangent frames --
TT_DEV f3 tt_octahedron_decode(float u, float v) {
    float sx = u * 2.0f - 1.0f, sy = v * 2.0f - 1.0f;
    f3 n(sx, sy, 1.0f - abs_(sx) - abs_(sy));
    float t = clamp_(-n.z, 0.0f, 1.0f);
    n.x = n.x + (n.x >= 0.0f ? -t : t);
    n.y = n.y + (n.y >= 0.0f ? -t : t);
    return normalize(n);
}

struct TangentFrame {
    f3 normal, tangent;
    float bitangent_sign;
};

// scene.glsl:29-64
TT_DEV TangentFrame decode_tangent_frame(uint32_t e) {
    float u = float(e & 0x3ffu) / float(1023);
    float v = float((e >> 10) & 0x3ffu) / float(1023);
    TangentFrame f;
    f.normal = tt_octahedron_decode(u, v);
    f3 ortho = abs_(f.normal.x) > abs_(f.normal.z) ? f3(-f.normal.y, f.normal.x, 0.0f) : f3(0.0f, -f.normal.z, f.normal.y);
    float angle = (float((e >> 20) & 0x7ffu) / float(2047)) * TAU;
    f.tangent = ortho * ttm::cos(angle) + cross(f.normal, ortho) * ttm::sin(angle);
    f.bitangent_sign = (e >> 31) == 1u ? 1.0f : -1.0f;
    return f;
}

struct Basis {
    f3 normal, tangent, bitangent;
};

TT_DEV f3 transform_to_basis(const Basis& b, f3 d) { return normalize(d.x * b.tangent + d.y * b.bitangent + d.z * b.normal); }
TT_DEV f3 transform_from_basis(const Basis& b, f3 d) {
    return normalize(f3(dot(d, b.tangent), dot(d, b.bitangent), dot(d, b.normal)));
}
TT_DEV f3 interpolate(f3 b, f3 x, f3 y, f3 z) { return b.x * x + b.y * y + b.z * z; }

TT_DEV f3 dequantize_snorm16(int16_t x, int16_t y, int16_t z) {
    return f3(float(x) / 32767.0f, float(y) / 32767.0f, float(z) / 32767.0f);
}

struct M4 {
    float m[16];
};
TT_DEV M4 load_m4(const float* p) {
    M4 r;
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int c = 0; c < 4; c++) {
        float4 v = q[c];
        r.m[4 * c + 0] = v.x;
        r.m[4 * c + 1] = v.y;
        r.m[4 * c + 2] = v.z;
        r.m[4 * c + 3] = v.w;
    }
    return r;
}

// --------------------------------------------------------------- textures --
// The scene's textures: records (tt_texture_record, 5 x uint4), texel bytes
// and the sRGB decode table.
struct TextureView {
    const uint4* records;
    const uint8_t* data;
    const float* srgb_lut;
};

TT_DEV TextureView texture_view(const SceneView& s) { return TextureView{s.tex_records, s.tex_data, s.srgb_lut}; }

// textureLod with the reference's sampler (scene.rs:121-128: linear
// filtering, repeat addressing), evaluated in software: bilinear on the two
// mip levels around the clamped level of detail, then linear between them.
// The sampler's max_anisotropy of 16 does not enter: textureLod passes an
// explicit level of detail (integrate.glsl:144, the reference's only
// texture read), which leaves no derivative footprint to filter over.

// BC1 texel (x, y) as 8-bit sRGB-encoded RGB and alpha (public block
// format: RGB565 endpoints expanded by bit replication, palette interpolated
// in integers as the header states).
TT_DEV uint4 bc1_texel(const uint8_t* level, uint32_t w, uint32_t x, uint32_t y) {
    const uint32_t blocks_x = (w + 3u) >> 2;
    const uint2 blk = *reinterpret_cast<const uint2*>(level + 8 * (size_t(y >> 2) * blocks_x + (x >> 2)));
    const uint32_t c0 = blk.x & 0xffffu, c1 = blk.x >> 16;
    const uint32_t k = (blk.y >> (2u * (((y & 3u) << 2) | (x & 3u)))) & 3u;
    auto expand = [](uint32_t c, uint32_t& r, uint32_t& g, uint32_t& b) {
        r = (c >> 11) & 31u;
        g = (c >> 5) & 63u;
        b = c & 31u;
        r = (r << 3) | (r >> 2);
        g = (g << 2) | (g >> 4);
        b = (b << 3) | (b >> 2);
    };
    uint32_t r0, g0, b0, r1, g1, b1;
    expand(c0, r0, g0, b0);
    expand(c1, r1, g1, b1);
    if (k == 0u) return make_uint4(r0, g0, b0, 255u);
    if (k == 1u) return make_uint4(r1, g1, b1, 255u);
    if (c0 > c1) {
        if (k == 2u) return make_uint4((2u * r0 + r1 + 1u) / 3u, (2u * g0 + g1 + 1u) / 3u, (2u * b0 + b1 + 1u) / 3u, 255u);
        return make_uint4((r0 + 2u * r1 + 1u) / 3u, (g0 + 2u * g1 + 1u) / 3u, (b0 + 2u * b1 + 1u) / 3u, 255u);
    }
    if (k == 2u) return make_uint4((r0 + r1 + 1u) / 2u, (g0 + g1 + 1u) / 2u, (b0 + b1 + 1u) / 2u, 255u);
    return make_uint4(0u, 0u, 0u, 0u);
}

// BC4 value of texel (x, y) of one 8-byte block.
TT_DEV uint32_t bc4_texel(uint2 blk, uint32_t x, uint32_t y) {
    const uint32_t e0 = blk.x & 0xffu, e1 = (blk.x >> 8) & 0xffu;
    const uint64_t bits = (uint64_t(blk.x) >> 16) | (uint64_t(blk.y) << 16);
    const uint32_t k = uint32_t(bits >> (3u * (((y & 3u) << 2) | (x & 3u)))) & 7u;
    if (k == 0u) return e0;
    if (k == 1u) return e1;
    if (e0 > e1) return ((8u - k) * e0 + (k - 1u) * e1 + 3u) / 7u;
    if (k < 6u) return ((6u - k) * e0 + (k - 1u) * e1 + 2u) / 5u;
    return k == 6u ? 0u : 255u;
}

// Texel (x, y) of a level w texels wide, as the sampler returns it.
TT_DEV float4 texel(const TextureView& s, uint32_t format, const uint8_t* level, uint32_t w, uint32_t x, uint32_t y) {
    const size_t i = size_t(y) * w + x;
    if (format == TT_TEXTURE_FORMAT_BC1_SRGB) {
        uint4 c = bc1_texel(level, w, x, y);
        return make_float4(s.srgb_lut[c.x], s.srgb_lut[c.y], s.srgb_lut[c.z], float(c.w) / 255.0f);
    }
    if (format == TT_TEXTURE_FORMAT_BC5_UNORM) {
        const uint4 blk = *reinterpret_cast<const uint4*>(level + 16 * (size_t(y >> 2) * ((w + 3u) >> 2) + (x >> 2)));
        return make_float4(float(bc4_texel(make_uint2(blk.x, blk.y), x, y)) / 255.0f,
                           float(bc4_texel(make_uint2(blk.z, blk.w), x, y)) / 255.0f, 0.0f, 1.0f);
    }
    if (format == TT_TEXTURE_FORMAT_RGBA8_SRGB) {
        uint32_t v = *reinterpret_cast<const uint32_t*>(level + 4 * i);
        return make_float4(s.srgb_lut[v & 0xffu], s.srgb_lut[(v >> 8) & 0xffu], s.srgb_lut[(v >> 16) & 0xffu],
                           float(v >> 24) / 255.0f);
    }
    if (format == TT_TEXTURE_FORMAT_RG8_UNORM) {
        uint32_t v = *reinterpret_cast<const uint16_t*>(level + 2 * i);
        return make_float4(float(v & 0xffu) / 255.0f, float(v >> 8) / 255.0f, 0.0f, 1.0f);
    }
    if (format == TT_TEXTURE_FORMAT_R8_UNORM) return make_float4(float(level[i]) / 255.0f, 0.0f, 0.0f, 1.0f);
    uint16_t h = *reinterpret_cast<const uint16_t*>(level + 2 * i);
    return make_float4(ttm::half_to_float(h), 0.0f, 0.0f, 1.0f);
}

// Texel index along one axis of extent n: the repeat-wrapped coordinate's
// lower neighbour i0, upper i1 and the weight of i1.
TT_DEV void wrap_axis(float c, uint32_t n, uint32_t& i0, uint32_t& i1, float& a) {
    float f = c - floorf(c);
    if (!(f >= 0.0f && f < 1.0f)) f = 0.0f;
    float x = f * float(n) - 0.5f;
    float fx = floorf(x);
    a = x - fx;
    int i = int(fx);
    i0 = i < 0 ? n - 1u : uint32_t(i);
    i1 = i0 + 1u == n ? 0u : i0 + 1u;
}

TT_DEV float4 lerp4(float4 a, float4 b, float t) {
    return make_float4(a.x * (1.0f - t) + b.x * t, a.y * (1.0f - t) + b.y * t, a.z * (1.0f - t) + b.z * t,
                       a.w * (1.0f - t) + b.w * t);
}

TT_DEV float4 bilinear(const TextureView& s, uint32_t format, const uint8_t* level, uint32_t w, uint32_t h, float u,
                       float v) {
    uint32_t x0, x1, y0, y1;
    float ax, ay;
    wrap_axis(u, w, x0, x1, ax);
    wrap_axis(v, h, y0, y1, ay);
    float4 top = lerp4(texel(s, format, level, w, x0, y0), texel(s, format, level, w, x1, y0), ax);
    float4 bottom = lerp4(texel(s, format, level, w, x0, y1), texel(s, format, level, w, x1, y1), ax);
    return lerp4(top, bottom, ay);
}

TT_DEV float4 sample_texture(const TextureView& s, uint32_t index, float u, float v, float lod) {
    const uint4* rec = s.records + 5 * size_t(index);
    uint4 r0 = rec[0];
    uint32_t format = r0.x, w = r0.y, h = r0.z, levels = r0.w;
    float top = float(levels - 1u);
    float l = lod > 0.0f ? (lod < top ? lod : top) : 0.0f;
    uint32_t l0 = uint32_t(floorf(l));
    uint32_t l1 = l0 + 1u < levels ? l0 + 1u : levels - 1u;
    float t = l - float(l0);
    const uint32_t* offsets = reinterpret_cast<const uint32_t*>(rec + 1);
    float4 a = bilinear(s, format, s.data + offsets[l0], w >> l0, h >> l0, u, v);
    float4 b = bilinear(s, format, s.data + offsets[l1], w >> l1, h >> l1, u, v);
    return lerp4(a, b, t);
}

// load_textures (integrate.glsl:219-267): the channels it overrides, in its
// order (base weight, coat roughness and anisotropy, coat IOR, transmission
// depth and emission are not read there).
TT_DEV void load_textures(const TextureView& s, const float4* material, Mat& m, float u, float v, float lod) {
    const uint32_t* t = reinterpret_cast<const uint32_t*>(material + 12);
    auto fetch = [&](int ch, float4& out) {
        if (t[ch] == 0xffffffffu) return false;
        out = sample_texture(s, t[ch], u, v, lod);
        return true;
    };
    float4 c;
    if (fetch(TT_TEXTURE_BASE_COLOR, c)) m.base_color = f3(c.x, c.y, c.z);
    if (fetch(TT_TEXTURE_SPECULAR_COLOR, c)) m.specular_color = f3(c.x, c.y, c.z);
    if (fetch(TT_TEXTURE_TRANSMISSION_COLOR, c)) m.transmission_color = f3(c.x, c.y, c.z);
    if (fetch(TT_TEXTURE_COAT_COLOR, c)) m.coat_color = f3(c.x, c.y, c.z);
    if (fetch(TT_TEXTURE_FUZZ_COLOR, c)) m.fuzz_color = f3(c.x, c.y, c.z);
    if (fetch(TT_TEXTURE_GEOMETRY_NORMAL, c)) m.geometry_normal = tt_octahedron_decode(c.x, c.y);
    if (fetch(TT_TEXTURE_BASE_METALNESS, c)) m.base_metalness = c.x;
    if (fetch(TT_TEXTURE_BASE_DIFFUSE_ROUGHNESS, c)) m.base_diffuse_roughness = c.x;
    if (fetch(TT_TEXTURE_SPECULAR_WEIGHT, c)) m.specular_weight = c.x;
    if (fetch(TT_TEXTURE_SPECULAR_ROUGHNESS, c)) m.specular_roughness = c.x;
    if (fetch(TT_TEXTURE_SPECULAR_ROUGHNESS_ANISOTROPY, c)) m.specular_roughness_anisotropy = c.x;
    if (fetch(TT_TEXTURE_SPECULAR_IOR, c)) m.specular_ior = c.x;
    if (fetch(TT_TEXTURE_SPECULAR_ROTATION, c)) m.specular_rotation = c.x;
    if (fetch(TT_TEXTURE_TRANSMISSION_WEIGHT, c)) m.transmission_weight = c.x;
    if (fetch(TT_TEXTURE_COAT_WEIGHT, c)) m.coat_weight = c.x;
    if (fetch(TT_TEXTURE_COAT_DARKENING, c)) m.coat_darkening = c.x;
    if (fetch(TT_TEXTURE_COAT_ROTATION, c)) m.coat_rotation = c.x;
    if (fetch(TT_TEXTURE_FUZZ_WEIGHT, c)) m.fuzz_weight = c.x;
    if (fetch(TT_TEXTURE_FUZZ_ROUGHNESS, c)) m.fuzz_roughness = c.x;
    if (fetch(TT_TEXTURE_GEOMETRY_OPACITY, c)) m.geometry_opacity = c.x;
}

// material_emissive's fetches (integrate.glsl:148-154): the emission colour
// (RGB) and luminance (.x) textures at (u, v) and level `lod`.
TT_DEV void load_emission_textures(const TextureView& s, const float4* material, Mat& m, float u, float v,
                                   float lod) {
    const uint32_t* t = reinterpret_cast<const uint32_t*>(material + 12);
    if (t[TT_TEXTURE_EMISSION_COLOR] != 0xffffffffu) {
        float4 c = sample_texture(s, t[TT_TEXTURE_EMISSION_COLOR], u, v, lod);
        m.emission_color = f3(c.x, c.y, c.z);
    }
    if (t[TT_TEXTURE_EMISSION_LUMINANCE] != 0xffffffffu)
        m.emission_luminance = sample_texture(s, t[TT_TEXTURE_EMISSION_LUMINANCE], u, v, lod).x;
}

// Texture code only appears in kernels instantiated for textured scenes:
// a kernel's register count is set by its busiest point, taken or not, and
// the sampler would otherwise cost untextured renders occupancy.

// load_textures, then the hit's emission (material_emissive at the same
// coordinate and level, integrate.glsl:305), on *m.
TT_DEV void texture_hit(TextureView tv, const float4* material, Mat* m, float u, float v, float lod) {
    Mat r = *m;
    load_textures(tv, material, r, u, v, lod);
    load_emission_textures(tv, material, r, u, v, lod);
    *m = r;
}

// material_emissive (integrate.glsl:189) at level 4 at texture coordinate
// (u, v) of a light sample.
TT_DEV f3 textured_emission(TextureView tv, const float4* material, float u, float v) {
    Mat m = load_material(material);
    load_emission_textures(tv, material, m, u, v, 4.0f);
    return m.emission_color * m.emission_luminance;
}

// camera_ray_direction (integrate.glsl:17-20) of normalized device
// coordinates.
TT_DEV f3 camera_ray_direction(const float* inverse_view, const float* inverse_proj, float ndc_x, float ndc_y) {
    const float* ip = inverse_proj;
    f3 v(ip[0] * (-ndc_x) + ip[4] * (-ndc_y) + ip[8] * 1.0f + ip[12] * 1.0f,
         ip[1] * (-ndc_x) + ip[5] * (-ndc_y) + ip[9] * 1.0f + ip[13] * 1.0f,
         ip[2] * (-ndc_x) + ip[6] * (-ndc_y) + ip[10] * 1.0f + ip[14] * 1.0f);
    return normalize(mul_point(inverse_view, v, 0.0f));
}

// triangle_intersection (math.glsl:113-125): barycentrics of a ray's
// crossing of the triangle's plane.
TT_DEV f3 triangle_intersection(const f3 p[3], f3 origin, f3 direction) {
    f3 edge_to_origin = origin - p[0];
    f3 edge_2 = p[2] - p[0];
    f3 edge_1 = p[1] - p[0];
    f3 r = cross(direction, edge_2);
    f3 q = cross(edge_to_origin, edge_1);
    float inverse_det = 1.0f / dot(r, edge_1);
    float v1 = dot(r, edge_to_origin);
    float v2 = dot(q, direction);
    float b = v1 * inverse_det;
    float c = v2 * inverse_det;
    return f3(1.0f - b - c, b, c);
}

// The primary ray of a path, for the mip level of its first hit: camera
// matrices, position, screen size and the sample's device coordinates.
struct PrimaryRay {
    const float* inverse_view;
    const float* inverse_proj;
    f3 origin;
    float width, height;
    float ndc_x, ndc_y;
};

// Texture terms of a hit.  With `primary`, the level of detail of the first
// hit from ray differentials (integrate.glsl:89-107, calculate_mip_level
// :134-139) replaces `lod`; then, if the scene has textures, load_textures
// and the emission textures at the hit's texture coordinate and `lod`.
struct TextureHit {
    bool enabled;
    bool primary;
    PrimaryRay ray;
    float lod;          // in: the path's level of detail; out: the level used
    float u = 0.0f, v = 0.0f;  // out: the hit's texture coordinate
};

// The first hit's level of detail from differentials one pixel over in x
// and in y (integrate.glsl:89-107, 134-139).
TT_DEV float first_hit_lod(PrimaryRay pr, f3 p0, f3 p1, f3 p2, f3 bary, f3 u, f3 v) {
    const f3 pos[3] = {p0, p1, p2};
    float tsx = 2.0f / pr.width, tsy = 2.0f / pr.height;
    f3 hx = triangle_intersection(pos, pr.origin,
                                  camera_ray_direction(pr.inverse_view, pr.inverse_proj, pr.ndc_x + tsx, pr.ndc_y));
    f3 hy = triangle_intersection(pos, pr.origin,
                                  camera_ray_direction(pr.inverse_view, pr.inverse_proj, pr.ndc_x, pr.ndc_y + tsy));
    f3 dx = bary - hx, dy = bary - hy;
    float dxu = (dx.x * u.x + dx.y * u.y + dx.z * u.z) * pr.width;
    float dxv = (dx.x * v.x + dx.y * v.y + dx.z * v.z) * pr.height;
    float dyu = (dy.x * u.x + dy.y * u.y + dy.z * u.z) * pr.width;
    float dyv = (dy.x * v.x + dy.y * v.y + dy.z * v.z) * pr.height;
    float gx = dxu * dxu + dxv * dxv, gy = dyu * dyu + dyv * dyv;
    return 0.5f * ttm::log2(gx < gy ? gy : gx);
}

// The shading point of a closest hit: get_ray_hit (integrate.glsl:22-110)
// and create_surface_basis.
struct Surface {
    f3 position;
    Basis basis;  // surface basis (geometry normal applied)
    f3 wi;        // -ray direction in the surface basis
    uint32_t material;
};

TT_DEV Surface surface_at_hit(const SceneView& s, uint32_t instance, uint32_t prim, float b1, float b2, f3 ray_dir,
                              Mat& m, TextureHit& tex) {
    const float* inst = s.instances + 52 * size_t(instance);
    M4 transform = load_m4(inst);
    M4 normal_transform = load_m4(inst + 32);
    uint4 offs = *reinterpret_cast<const uint4*>(inst + 48);
    uint32_t vertex_offset = offs.x, index_offset = offs.y;
    uint32_t base = 3 * prim + index_offset;
    f3 pos[3];
    TangentFrame fr[3];
    uint32_t uv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        uint32_t vi = s.indices[base + c] + vertex_offset;
        const int16_t* pp = s.positions + 3 * size_t(vi);
        pos[c] = mul_point1(transform.m, dequantize_snorm16(pp[0], pp[1], pp[2]));
        uint2 vx = s.vertices[vi];
        fr[c] = decode_tangent_frame(vx.y);
        uv[c] = vx.x;
    }
    f3 bary(0.0f, b1, b2);
    bary.x = 1.0f - bary.y - bary.z;
    float tu = 0.0f, tv = 0.0f;
    const float4* material = s.materials + 19 * size_t(offs.z);
    bool textured = s.texture_count > 0u && tex.enabled;
    if (textured) {
        f3 u(ttm::half_to_float(uint16_t(uv[0] & 0xffffu)), ttm::half_to_float(uint16_t(uv[1] & 0xffffu)),
             ttm::half_to_float(uint16_t(uv[2] & 0xffffu)));
        f3 v(ttm::half_to_float(uint16_t(uv[0] >> 16)), ttm::half_to_float(uint16_t(uv[1] >> 16)),
             ttm::half_to_float(uint16_t(uv[2] >> 16)));
        tu = bary.x * u.x + bary.y * u.y + bary.z * u.z;
        tv = bary.x * v.x + bary.y * v.y + bary.z * v.z;
        if (tex.primary) tex.lod = first_hit_lod(tex.ray, pos[0], pos[1], pos[2], bary, u, v);
        tex.u = tu;
        tex.v = tv;
    }
    Basis ts;
    ts.normal = normalize(mul_dir(normal_transform.m, interpolate(bary, fr[0].normal, fr[1].normal, fr[2].normal)));
    ts.tangent = normalize(mul_dir(normal_transform.m, interpolate(bary, fr[0].tangent, fr[1].tangent, fr[2].tangent)));
    float sign = sign_(fr[0].bitangent_sign * bary.x + fr[1].bitangent_sign * bary.y + fr[2].bitangent_sign * bary.z);
    ts.bitangent = normalize(sign * cross(ts.normal, ts.tangent));
    Surface sf;
    sf.position = bary.x * pos[0] + bary.y * pos[1] + bary.z * pos[2];
    sf.material = offs.z;
    m = load_material(material);
    if (textured) {
        Mat r = m;
        texture_hit(texture_view(s), material, &r, tu, tv, tex.lod);
        m = r;
    }
    sf.basis.normal = transform_to_basis(ts, m.geometry_normal);
    sf.basis.tangent = normalize(ts.tangent - sf.basis.normal * dot(sf.basis.normal, ts.tangent));
    sf.basis.bitangent = normalize(cross(sf.basis.normal, sf.basis.tangent));
    sf.wi = transform_from_basis(sf.basis, -ray_dir);
    return sf;
}

// The same without textures.
TT_DEV Surface surface_at_hit(const SceneView& s, uint32_t instance, uint32_t prim, float b1, float b2, f3 ray_dir,
                              Mat& m) {
    TextureHit none{};
    return surface_at_hit(s, instance, prim, b1, b2, ray_dir, m, none);
}

// ---------------------------------------------------------------- lights --
struct LightPoint {
    f3 position, normal, emission;
    float area;
    uint32_t instance, hash;
};

// The world-space triangle of emissive record `li` (sample_random_light,
// scene.glsl:114-136; material_emissive): vertices, normal, area, emission.
struct LightTriangle {
    f3 p[3], normal, emission;
    float area;
    uint32_t instance, hash;
    uint32_t textured_material;  // material index if its emission is textured, else 0xffffffff
};

TT_DEV LightTriangle light_triangle(const SceneView& s, uint32_t li) {
    const uint32_t* et = s.emissive + 9 * size_t(li);
    uint32_t w[9];
#pragma unroll
    for (int k = 0; k < 9; k++) w[k] = et[k];
    auto half = [&](int k) { return int16_t((w[k >> 1] >> (16 * (k & 1))) & 0xffffu); };
    LightTriangle t;
    t.hash = w[4] >> 16;
    t.instance = w[8];
    const float* linst = s.instances + 52 * size_t(t.instance);
    M4 ltr = load_m4(linst);
#pragma unroll
    for (int c = 0; c < 3; c++)
        t.p[c] = mul_point1(ltr.m, dequantize_snorm16(half(3 * c), half(3 * c + 1), half(3 * c + 2)));
    t.area = fmax_(0.00001f, 0.5f * length(cross(t.p[1] - t.p[0], t.p[2] - t.p[0])));
    t.normal = normalize(cross(t.p[1] - t.p[0], t.p[2] - t.p[0]));
    uint32_t material = __float_as_uint(linst[50]);
    Mat lm = load_material(s.materials + 19 * size_t(material));
    t.emission = lm.emission_color * lm.emission_luminance;
    const uint32_t* textures = reinterpret_cast<const uint32_t*>(s.materials + 19 * size_t(material) + 12);
    bool textured = s.texture_count > 0u && (textures[TT_TEXTURE_EMISSION_COLOR] != 0xffffffffu ||
                                             textures[TT_TEXTURE_EMISSION_LUMINANCE] != 0xffffffffu);
    t.textured_material = textured ? material : 0xffffffffu;
    return t;
}

// Light record (80 B, built once per scene from light_triangle):
//   (p0, area) (p1, normal.x) (p2, normal.y) (emission, normal.z)
//   (instance, hash, textured material or 0xffffffff, 0)
TT_DEV void store_light_record(float4* r, const LightTriangle& t) {
    r[0] = make_float4(t.p[0].x, t.p[0].y, t.p[0].z, t.area);
    r[1] = make_float4(t.p[1].x, t.p[1].y, t.p[1].z, t.normal.x);
    r[2] = make_float4(t.p[2].x, t.p[2].y, t.p[2].z, t.normal.y);
    r[3] = make_float4(t.emission.x, t.emission.y, t.emission.z, t.normal.z);
    r[4] = make_float4(__uint_as_float(t.instance), __uint_as_float(t.hash), __uint_as_float(t.textured_material), 0.0f);
}

// The point at barycentrics `lb` of emissive triangle `li`, with the light
// record's constant emission (kernels instantiated for untextured scenes,
// and the visibility pass, which needs no emission).
TT_DEV LightPoint light_point(const SceneView& s, uint32_t li, f3 lb) {
    const float4* r = s.lights + 5 * size_t(li);
    float4 a = r[0], b = r[1], c = r[2], d = r[3], e = r[4];
    LightPoint lp;
    lp.area = a.w;
    lp.normal = f3(b.w, c.w, d.w);
    lp.position = interpolate(lb, f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), f3(c.x, c.y, c.z));
    lp.emission = f3(d.x, d.y, d.z);
    lp.instance = __float_as_uint(e.x);
    lp.hash = __float_as_uint(e.y);
    return lp;
}

// light_point for scenes with textures: a light whose material has emission
// textures takes material_emissive at the sample's texture coordinate, level
// 4 (integrate.glsl:189; the coordinate as sample_random_light interpolates
// it, scene.glsl:130-133).
TT_DEV LightPoint light_point_textured(const SceneView& s, uint32_t li, f3 lb) {
    LightPoint lp = light_point(s, li, lb);
    uint32_t textured_material = __float_as_uint(s.lights[5 * size_t(li) + 4].z);
    if (textured_material != 0xffffffffu) {
        const uint32_t* et = s.emissive + 9 * size_t(li);
        uint32_t w0 = et[5], w1 = et[6], w2 = et[7];
        float tu = lb.x * ttm::half_to_float(uint16_t(w0 & 0xffffu)) + lb.y * ttm::half_to_float(uint16_t(w1 & 0xffffu)) +
            lb.z * ttm::half_to_float(uint16_t(w2 & 0xffffu));
        float tv = lb.x * ttm::half_to_float(uint16_t(w0 >> 16)) + lb.y * ttm::half_to_float(uint16_t(w1 >> 16)) +
            lb.z * ttm::half_to_float(uint16_t(w2 >> 16));
        lp.emission = textured_emission(texture_view(s), s.materials + 19 * size_t(textured_material), tu, tv);
    }
    return lp;
}

// ---------------------------------------------------------------- ReSTIR --
// The lobes (bsdf_prepare's stochastic albedo estimates) at a primary vertex
// in ReSTIR mode come from this generator, so the reservoir target at a
// pixel can be rebuilt from its G-buffer record (pixel, frame) anywhere.
TT_DEV Generator restir_lobe_generator(uint32_t pixel, uint32_t frame) {
    return pcg32_seeded(splitmix64(pcg32_key(pixel, frame) ^ 0xa511e9b3c0ffee11ull));
}

// Reservoir (16 B): x = emissive triangle (0xffffffff = empty), y = the
// sample's barycentrics (b0, b1) as unorm16, z = W | M as IEEE half, w = 0.
struct Reservoir {
    uint32_t light;
    uint32_t bary;
    float W, M;
};

TT_DEV f3 reservoir_bary(uint32_t packed) {
    float b0 = float(packed & 0xffffu) / 65535.0f;
    float b1 = float(packed >> 16) / 65535.0f;
    return f3(b0, b1, 1.0f - b0 - b1);
}

TT_DEV uint32_t quantize_bary(f3 b) {
    uint32_t u = uint32_t(b.x * 65535.0f + 0.5f);
    uint32_t v = uint32_t(b.y * 65535.0f + 0.5f);
    if (u > 65535u) u = 65535u;
    if (u + v > 65535u) v = 65535u - u;
    return u | (v << 16);
}

TT_DEV uint4 pack_reservoir(const Reservoir& r) {
    // W saturates at the largest finite half.
    return make_uint4(r.light, r.bary,
                      uint32_t(ttm::float_to_half(fmin_(r.W, 65504.0f))) | (uint32_t(ttm::float_to_half(r.M)) << 16),
                      0u);
}

TT_DEV Reservoir unpack_reservoir(uint4 v) {
    Reservoir r;
    r.light = v.x;
    r.bary = v.y;
    r.W = ttm::half_to_float(uint16_t(v.z & 0xffffu));
    r.M = ttm::half_to_float(uint16_t(v.z >> 16));
    return r;
}

}  // namespace ttd
