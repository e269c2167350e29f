// ORACLE (test infrastructure only; never linked into the product).
//
// CPU restatement of ReSTIR DI at the primary vertex, as this library
// defines it (the reference has no ReSTIR; it replaces the uniform
// next-event estimate of integrate.glsl:156-193 at bounce 0):
//   * Bitterli et al. 2020: RIS over M uniform light candidates (source pdf
//     1 / (count * area)), visibility reuse, one temporal and one spatial
//     reuse;
//   * Lin et al. 2022: generalized balance heuristic weights when combining
//     two reservoirs;
//   * target p_hat_q(y) = luminance(f_q Le) |cos wo| cos_light / d^2 with
//     lobes prepared from a per-pixel generator;
//   * reservoirs store (emissive triangle, unorm16 barycentrics, half W,
//     half M).
// Float only.  Passes run over the whole frame, one pixel at a time.
#pragma once

#include "oracle_scene.hpp"

namespace oracle {

struct RestirSettings {
    uint32_t candidates;
    bool temporal, spatial;
};

struct LightPointF {
    V3<float> position, normal, emission;
    float area;
    uint32_t instance, hash;
};

// scene.glsl:114-136 for emissive record li at barycentrics lb.
inline LightPointF light_point(const Scene<float>& s, uint32_t li, V3<float> lb) {
    const tt_emissive_triangle& tri = s.desc->emissive_triangles[li];
    const Instance<float>& in = s.instances[tri.instance_index];
    V3<float> p[3];
    for (int i = 0; i < 3; i++)
        p[i] = mul_point(in.transform,
                         V3<float>(dequantize_snorm16<float>(tri.positions[i][0]), dequantize_snorm16<float>(tri.positions[i][1]),
                                   dequantize_snorm16<float>(tri.positions[i][2])),
                         1.0f);
    LightPointF lp;
    lp.area = rmax(0.00001f, 0.5f * length(cross(p[1] - p[0], p[2] - p[0])));
    lp.normal = normalize(cross(p[1] - p[0], p[2] - p[0]));
    lp.position = interpolate(lb, p[0], p[1], p[2]);
    // material_emissive at the sample's texture coordinate, level 4
    // (integrate.glsl:189; scene.glsl:130-133); the constants when untextured.
    float tc[2];
    for (int k = 0; k < 2; k++)
        tc[k] = lb.x * ttm::half_to_float(tri.texture_coordinates[0][k]) +
            lb.y * ttm::half_to_float(tri.texture_coordinates[1][k]) +
            lb.z * ttm::half_to_float(tri.texture_coordinates[2][k]);
    lp.emission = s.emissive_at(in.material_index, tc[0], tc[1], 4.0f);
    lp.instance = tri.instance_index;
    lp.hash = tri.hash;
    return lp;
}

inline V3<float> reservoir_bary(uint32_t packed) {
    float b0 = float(packed & 0xffffu) / 65535.0f;
    float b1 = float(packed >> 16) / 65535.0f;
    return V3<float>(b0, b1, 1.0f - b0 - b1);
}

inline uint32_t quantize_bary(V3<float> b) {
    uint32_t u = uint32_t(b.x * 65535.0f + 0.5f);
    uint32_t v = uint32_t(b.y * 65535.0f + 0.5f);
    if (u > 65535u) u = 65535u;
    if (u + v > 65535u) v = 65535u - u;
    return u | (v << 16);
}

struct Res {
    uint32_t light = 0xffffffffu;
    uint32_t bary = 0;
    float W = 0.0f, M = 0.0f;
};

inline void pack(const Res& r, uint32_t* out) {
    out[0] = r.light;
    out[1] = r.bary;
    out[2] = uint32_t(ttm::float_to_half(rmin(r.W, 65504.0f))) | (uint32_t(ttm::float_to_half(r.M)) << 16);
    out[3] = 0;
}

inline Res unpack(const uint32_t* v) {
    Res r;
    r.light = v[0];
    r.bary = v[1];
    r.W = ttm::half_to_float(uint16_t(v[2] & 0xffffu));
    r.M = ttm::half_to_float(uint16_t(v[2] >> 16));
    return r;
}

struct ShadingF {
    V3<float> position;
    Basis<float> basis;
    V3<float> wi;
    uint32_t material;
    Material<float> m;
    Lobes<float> lobes;
    float depth;
    float u = 0.0f, v = 0.0f, lod = 1.0f;  // texture coordinate and level of detail of the hit
    bool valid = false;
};

// Lobes at a primary vertex in ReSTIR mode come from this generator (per
// pixel and frame), so the target can be rebuilt from a G-buffer record.
inline Generator lobe_generator(uint32_t pixel, uint32_t frame) {
    return pcg32_seeded(splitmix64(pcg32_key(pixel, frame) ^ 0xa511e9b3c0ffee11ull));
}

// One generator per ReSTIR pass (salt 1 candidates, 2 temporal, 3 spatial).
inline Generator pass_generator(uint32_t pixel, uint32_t frame, uint32_t seed, uint32_t salt) {
    return init_generator_from_index(pixel, frame, 1, seed ^ (salt * 0x9e3779b9u));
}

// G-buffer record: RECORD floats (position, material bits | normal, valid |
// tangent, depth | -ray direction, frame bits | texture coordinate, level of
// detail, 0).
constexpr size_t RECORD = 20;

inline void write_record(float* g, const ShadingF& q, V3<float> wi_world, uint32_t frame) {
    uint32_t mb = q.material;
    float mf, ff;
    std::memcpy(&mf, &mb, 4);
    std::memcpy(&ff, &frame, 4);
    float vals[RECORD] = {q.position.x, q.position.y, q.position.z, mf, q.basis.normal.x, q.basis.normal.y,
                          q.basis.normal.z, q.valid ? 1.0f : 0.0f, q.basis.tangent.x, q.basis.tangent.y,
                          q.basis.tangent.z, q.depth, wi_world.x, wi_world.y, wi_world.z, ff, q.u, q.v, q.lod, 0.0f};
    std::memcpy(g, vals, sizeof(vals));
}

inline ShadingF read_record(const Scene<float>& s, const float* g, uint32_t pixel) {
    ShadingF q;
    q.valid = g[7] != 0.0f;
    if (!q.valid) return q;
    q.position = V3<float>(g[0], g[1], g[2]);
    std::memcpy(&q.material, &g[3], 4);
    q.basis.normal = V3<float>(g[4], g[5], g[6]);
    q.basis.tangent = V3<float>(g[8], g[9], g[10]);
    q.basis.bitangent = normalize(cross(q.basis.normal, q.basis.tangent));
    q.wi = transform_from_basis(q.basis, V3<float>(g[12], g[13], g[14]));
    q.depth = g[11];
    q.u = g[16];
    q.v = g[17];
    q.lod = g[18];
    q.m = s.material_at(q.material, q.u, q.v, q.lod);
    uint32_t frame;
    std::memcpy(&frame, &g[15], 4);
    Generator lg = lobe_generator(pixel, frame);
    q.lobes = bsdf_prepare(q.m, q.wi, lg);
    return q;
}

inline float restir_target(const Scene<float>& s, const ShadingF& q, uint32_t light, uint32_t bary, float* area = nullptr) {
    LightPointF lp = light_point(s, light, reservoir_bary(bary));
    if (area) *area = lp.area;
    V3<float> to_light = lp.position - q.position;
    V3<float> wo = transform_from_basis(q.basis, to_light);
    float d2 = length_squared(to_light);
    V3<float> dir = to_light / Fn<float>::sqrt(d2);
    float cos_light = saturate(dot(-dir, lp.normal));
    float density;
    V3<float> f = bsdf_evaluate(q.m, q.lobes, q.wi, wo, density);
    V3<float> c = f * lp.emission;
    float lum = dot(c, V3<float>(0.2126f, 0.7152f, 0.0722f));
    float p = ((lum * rabs(cos_theta(wo))) * cos_light) / d2;
    return p > 0.0f && p < INFINITY ? p : 0.0f;
}

inline bool similar(const ShadingF& a, const ShadingF& b) {
    return a.valid && b.valid && dot(a.basis.normal, b.basis.normal) >= 0.9f && rabs(a.depth - b.depth) <= 0.1f * a.depth;
}

inline Res combine(const Scene<float>& s, const ShadingF& q, Res rq, const ShadingF& o, Res ro, Generator& g) {
    float pq_yq = 0.0f, po_yq = 0.0f, pq_yo = 0.0f, po_yo = 0.0f;
    if (rq.light != 0xffffffffu) {
        pq_yq = restir_target(s, q, rq.light, rq.bary);
        po_yq = restir_target(s, o, rq.light, rq.bary);
    }
    if (ro.light != 0xffffffffu) {
        pq_yo = restir_target(s, q, ro.light, ro.bary);
        po_yo = restir_target(s, o, ro.light, ro.bary);
    }
    float dq = rq.M * pq_yq + ro.M * po_yq;
    float d_o = rq.M * pq_yo + ro.M * po_yo;
    float mq = dq > 0.0f ? (rq.M * pq_yq) / dq : 0.0f;
    float mo = d_o > 0.0f ? (ro.M * po_yo) / d_o : 0.0f;
    float wq = mq * pq_yq * rq.W;
    float wo = mo * pq_yo * ro.W;
    float sum = wq + wo;
    float u = random_float<float>(g);
    bool take_other = u * sum < wo;
    Res r;
    r.light = take_other ? ro.light : rq.light;
    r.bary = take_other ? ro.bary : rq.bary;
    float p = take_other ? pq_yo : pq_yq;
    r.W = p > 0.0f ? sum / p : 0.0f;
    r.M = rq.M + ro.M;
    if (r.W == 0.0f && r.light == 0xffffffffu) r = Res();
    return r;
}

// Whether the reference's shadow-ray rule accepts light sample lp from x.
inline bool light_visible(const Scene<float>& s, V3<float> x, V3<float> dir, const LightPointF& lp, TraceCounts& counts) {
    Hit h;
    counts.shadow++;
    if (!s.closest_hit(x, dir, false, 1.0e-3f, 10000.0f, h)) return false;
    if (h.instance != lp.instance) return false;
    return (h.primitive + s.desc->instances[h.instance].index_offset / 3) % 0xffffu == lp.hash;
}

// The path of one pixel with the ReSTIR estimate at bounce 0 and next-event
// estimation afterwards (integrate.glsl:270-321 otherwise unchanged).
inline V3<float> trace_path_restir(const Scene<float>& s, const Camera<float>& c, uint32_t bounce_count, V3<float> dir,
                                   Generator g, bool primary_found, const Hit& primary, const uint32_t* reservoir,
                                   uint32_t pixel, uint32_t frame, TraceCounts& counts, float ndc_x, float ndc_y) {
    V3<float> accumulated;
    V3<float> throughput(1.0f);
    V3<float> ray_dir = dir, ray_org = c.position;
    float lod = 1.0f;  // PathState.mip_level
    for (uint32_t bounce = 0; bounce < bounce_count; bounce++) {
        Hit h;
        bool found;
        if (bounce == 0) {
            found = primary_found;
            h = primary;
        } else {
            counts.extension++;
            found = s.closest_hit(ray_org, ray_dir, true, 1.0e-3f, 1000.0f, h);
        }
        if (!found) {
            accumulated = accumulated + V3<float>(1.0f) * throughput;
            break;
        }
        RayHit<float> hit = get_ray_hit(s, h);
        if (bounce == 0 && s.textured()) lod = first_hit_lod(c, hit, ndc_x, ndc_y);
        const Material<float> m = s.material_at(hit.material_index, hit.tu, hit.tv, lod);
        Basis<float> basis = create_surface_basis(hit.tangent_space, m.geometry_normal);
        V3<float> wi = transform_from_basis(basis, -ray_dir);
        Lobes<float> lobes;
        if (bounce == 0) {
            Generator lg = lobe_generator(pixel, frame);
            lobes = bsdf_prepare(m, wi, lg);
        } else {
            lobes = bsdf_prepare(m, wi, g);
        }
        bool gate = s.desc->emissive_triangle_count > 0 && m.specular_roughness > 0.15f;
        V3<float> emissive = s.emissive_at(hit.material_index, hit.tu, hit.tv, lod);
        bool sampled = false;
        V3<float> direct;
        if (gate && bounce > 0) {
            direct = direct_light_contribution(s, hit.world_position, wi, basis, m, lobes, g, counts);
            sampled = true;
        } else if (gate) {
            Res r = unpack(reservoir);
            if (r.light != 0xffffffffu && r.W > 0.0f) {
                LightPointF lp = light_point(s, r.light, reservoir_bary(r.bary));
                V3<float> to_light = lp.position - hit.world_position;
                V3<float> wo = transform_from_basis(basis, to_light);
                float d2 = length_squared(to_light);
                float distance = Fn<float>::sqrt(d2);
                V3<float> ldir = to_light / distance;
                float cos_light = saturate(dot(-ldir, lp.normal));
                if (light_visible(s, hit.world_position, ldir, lp, counts)) {
                    float density;
                    V3<float> f = bsdf_evaluate(m, lobes, wi, wo, density);
                    float scale = ((rabs(cos_theta(wo)) * cos_light) / d2) * r.W;
                    direct = (f * lp.emission) * scale;
                }
                sampled = true;
            }
        }
        if (sampled) accumulated = accumulated + throughput * direct;
        if (bounce == 0 || !gate) accumulated = accumulated + emissive * throughput;
        if (bounce + 1 == bounce_count) break;
        V3<float> wo, value;
        float density;
        int lobe;
        if (!bsdf_sample(m, lobes, wi, g, wo, density, lobe, value)) break;
        ray_dir = transform_to_basis(basis, wo);
        ray_org = hit.world_position;
        throughput = throughput * (value / rmax(density, K<float>::PDF_EPSILON) * rabs(cos_theta(wo)));
    }
    return accumulated;
}

}  // namespace oracle
