// OpenPBR-style six-lobe BSDF for the gfx950 shading kernel.
//
// Follows the reference's lobe set and combiner:
//   src/material/bsdf.glsl:15-203        lobe weights, probabilities, evaluate, sample
//   src/material/microfacet.glsl:7-54    GGX D, Smith lambda / G1 / G2, VNDF sampling
//   src/material/fresnel.glsl:6-39       Schlick, dielectric Fresnel, OpenPBR Eqs. 102-103
//   src/material/diffuse_brdf.glsl:9-61  energy-compensated Oren-Nayar
//   src/material/specular_brdf.glsl:8-116, specular_btdf.glsl:8-119
//   src/material/metal_brdf.glsl:8-82, coat_brdf.glsl:8-73, fuzz_brdf.glsl:11-73
//   src/includes/sample.glsl:7-30, src/includes/math.glsl:35-60
// Random draws come from a Generator (reference xorshift32 or PCG) in the
// reference's order.  Two things GLSL leaves undefined are defined here:
// lobe densities start at zero, and a lobe sample that produces no direction
// reports failure (the path then ends).
#pragma once

#include "device_math.h"

namespace ttd {

// ---------------------------------------------------------------- random --
// kind 0: the reference's xorshift32 (random.glsl:29-34) in the low 32 bits
// of `state`; kind 1: pcg32 (PCG-XSH-RR 64/32, O'Neill 2014) on the 64-bit
// LCG state with the pcg32 default stream.
struct Generator {
    uint64_t state;
    uint32_t kind;
};

TT_DEV uint32_t jenkins_hash(uint32_t x) {
    x += x << 10;
    x ^= x >> 6;
    x += x << 3;
    x ^= x >> 11;
    x += x << 15;
    return x;
}

constexpr uint64_t PCG32_MULT = 6364136223846793005ull;
constexpr uint64_t PCG32_INC = 0xda3e39cb94b95bdbull;  // pcg32's default stream

// XSH-RR output of the pre-step state.
TT_DEV uint32_t pcg32_output(uint64_t old) {
    uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

// SplitMix64 finaliser (Steele, Lea & Flood 2014): turns (pixel, frame, seed)
// keys into well-spread pcg32 initial states.
TT_DEV uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// pcg32_srandom with the default stream.
TT_DEV Generator pcg32_seeded(uint64_t initstate) {
    uint64_t s = PCG32_INC;  // 0 * MULT + INC
    s = s + initstate;
    s = s * PCG32_MULT + PCG32_INC;
    return Generator{s, 1u};
}

TT_DEV uint64_t pcg32_key(uint32_t index, uint32_t frame) { return (uint64_t(frame) << 32) | index; }

TT_DEV Generator init_generator(uint32_t index, uint32_t frame, uint32_t kind, uint32_t seed) {
    if (kind == 1u) return pcg32_seeded(splitmix64(pcg32_key(index, frame) ^ splitmix64(seed)));
    return Generator{jenkins_hash(index ^ jenkins_hash(frame)), 0u};
}

TT_DEV uint32_t next_u32(Generator& g) {
    if (g.kind == 1u) {
        uint64_t old = g.state;
        g.state = old * PCG32_MULT + PCG32_INC;
        return pcg32_output(old);
    }
    uint32_t s = uint32_t(g.state);
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    g.state = s;
    return s;
}

// random.glsl:25-27
TT_DEV float random_float(Generator& g) {
    return __uint_as_float(0x3f800000u | (next_u32(g) >> 9)) - 1.0f;
}
TT_DEV uint32_t random_uint(Generator& g, uint32_t end) { return next_u32(g) % end; }
TT_DEV f2 random_vec2(Generator& g) {
    float a = random_float(g);
    float b = random_float(g);
    return f2{a, b};
}

// ------------------------------------------------------------- material --
struct Mat {
    f3 base_color, specular_color, transmission_color, coat_color, fuzz_color, emission_color, geometry_normal;
    float base_weight, base_metalness, base_diffuse_roughness, specular_weight, specular_roughness,
        specular_roughness_anisotropy, specular_ior, specular_rotation, transmission_weight, transmission_depth,
        coat_weight, coat_roughness, coat_roughness_anisotropy, coat_ior, coat_darkening, coat_rotation,
        fuzz_weight, fuzz_roughness, emission_luminance, geometry_opacity;
};

// MaterialConstants (material.glsl:10-40), 192 B, 16-B aligned.
TT_DEV Mat load_material(const float4* p) {
    Mat m;
    float4 a[12];
#pragma unroll
    for (int i = 0; i < 12; i++) a[i] = p[i];
    m.base_color = f3(a[0].x, a[0].y, a[0].z);
    m.specular_color = f3(a[1].x, a[1].y, a[1].z);
    m.transmission_color = f3(a[2].x, a[2].y, a[2].z);
    m.coat_color = f3(a[3].x, a[3].y, a[3].z);
    m.fuzz_color = f3(a[4].x, a[4].y, a[4].z);
    m.emission_color = f3(a[5].x, a[5].y, a[5].z);
    m.geometry_normal = f3(a[6].x, a[6].y, a[6].z);
    m.base_weight = a[7].x;
    m.base_metalness = a[7].y;
    m.base_diffuse_roughness = a[7].z;
    m.specular_weight = a[7].w;
    m.specular_roughness = a[8].x;
    m.specular_roughness_anisotropy = a[8].y;
    m.specular_ior = a[8].z;
    m.specular_rotation = a[8].w;
    m.transmission_weight = a[9].x;
    m.transmission_depth = a[9].y;
    m.coat_weight = a[9].z;
    m.coat_roughness = a[9].w;
    m.coat_roughness_anisotropy = a[10].x;
    m.coat_ior = a[10].y;
    m.coat_darkening = a[10].z;
    m.coat_rotation = a[10].w;
    m.fuzz_weight = a[11].x;
    m.fuzz_roughness = a[11].y;
    m.emission_luminance = a[11].z;
    m.geometry_opacity = a[11].w;
    return m;
}

// ---------------------------------------------------------------- sample --
TT_DEV f3 sample_triangle(f2 r) {
    r.x = r.x / 2.0f;
    r.y = r.y / 2.0f;
    float offset = r.y - r.x;
    if (offset > 0.0f) {
        r.y = r.y + offset;
    } else {
        r.x = r.x - offset;
    }
    return f3(r.x, r.y, 1.0f - r.x - r.y);
}

TT_DEV f3 cosine_hemisphere_sample(f2 r) {
    float radius = sqrt_(r.x);
    float theta = 2.0f * PI * r.y;
    float x = radius * ttm::cos(theta), y = radius * ttm::sin(theta);
    float z = sqrt_(fmax_(0.0f, 1.0f - x * x - y * y));
    return f3(x, y, z);
}

TT_DEV float cosine_hemisphere_density(float c) {
    if (c <= 1e-6f) return 1e-6f / PI;
    return c / PI;
}

// ------------------------------------------------------------ local frame --
struct LocalRotation {
    float t00, t01, t10, t11, i00, i01, i10, i11;
};
TT_DEV LocalRotation local_frame_rotation(float angle) {
    if (angle == 0.0f || angle == 2.0f * PI) return LocalRotation{1, 0, 0, 1, 1, 0, 0, 1};
    float c = ttm::cos(angle), s = ttm::sin(angle);
    return LocalRotation{c, s, -s, c, c, -s, s, c};
}
TT_DEV f3 rotate_local(f3 v, const LocalRotation& r) {
    return f3(r.t00 * v.x + r.t10 * v.y, r.t01 * v.x + r.t11 * v.y, v.z);
}
TT_DEV f3 inverse_rotate_local(f3 v, const LocalRotation& r) {
    return f3(r.i00 * v.x + r.i10 * v.y, r.i01 * v.x + r.i11 * v.y, v.z);
}

// ------------------------------------------------------------ microfacet --
TT_DEV float microfacet_distribution(f2 alpha, f3 m) {
    float tan_squared = tan_theta_squared(m);
    if (isinf_(tan_squared)) return 0.0f;
    float e = tan_squared * (pow2(cos_phi(m) / alpha.x) + pow2(sin_phi(m) / alpha.y));
    return 1.0f / (PI * alpha.x * alpha.y * pow2(cos_theta_squared(m)) * pow2(1.0f + e));
}
TT_DEV float microfacet_lambda(f2 alpha, f3 v) {
    float t2 = tan_theta_squared(v);
    if (isinf_(t2)) return 0.0f;
    float a2 = pow2(cos_phi(v) * alpha.x) + pow2(sin_phi(v) * alpha.y);
    return (sqrt_(1.0f + a2 * t2) - 1.0f) / 2.0f;
}
TT_DEV float microfacet_visibility(f2 alpha, f3 i, f3 o) {
    return 1.0f / (1.0f + microfacet_lambda(alpha, o) + microfacet_lambda(alpha, i));
}
TT_DEV float microfacet_masking(f2 alpha, f3 v) { return 1.0f / (1.0f + microfacet_lambda(alpha, v)); }
TT_DEV f3 microfacet_sample(f2 alpha, f3 view, f2 r) {
    view = normalize(f3(view.x * alpha.x, view.y * alpha.y, view.z));
    float phi = 2.0f * PI * r.x;
    float z = (1.0f - r.y) * (1.0f + view.z) - view.z;
    float st = sqrt_(saturate(1.0f - pow2(z)));
    f3 h = f3(st * ttm::cos(phi), st * ttm::sin(phi), z) + view;
    return normalize(f3(h.x * alpha.x, h.y * alpha.y, h.z));
}
TT_DEV f2 microfacet_roughness_to_alpha(float roughness, float anisotropy) {
    float rs = pow2(roughness);
    float ax = rs * sqrt_(2.0f / (1.0f + pow2(1.0f - anisotropy)));
    float ay = (1.0f - anisotropy) * ax;
    return f2{fmax_(1e-4f, ax), fmax_(1e-4f, ay)};
}

// --------------------------------------------------------------- fresnel --
TT_DEV f3 fresnel_schlick(f3 f0, float c) { return f0 + pow5(1.0f - c) * (1.0f - f0); }
TT_DEV float fresnel_at_normal_incident(float ior) { return pow2((ior - 1.0f) / (ior + 1.0f)); }
TT_DEV float dieletric_fresnel_factor(float ior) {
    return ttm::log((10893.0f * ior - 1438.2f) / (-774.4f * pow2(ior) + 10212.0f * ior + 1.0f));
}
TT_DEV float fresnel_dielectric(float cos_i, float ior) {
    float s2i = 1.0f - pow2(cos_i);
    float s2t = s2i / pow2(ior);
    if (s2t >= 1.0f) return 1.0f;
    float cos_t = sqrt_(1.0f - s2t);
    float par = (ior * cos_i - cos_t) / (ior * cos_i + cos_t);
    float per = (cos_i - ior * cos_t) / (cos_i + ior * cos_t);
    return 0.5f * (pow2(par) + pow2(per));
}
TT_DEV float average_dielectric_fresnel(float ior) {
    if (ior > 1.0f) return dieletric_fresnel_factor(ior);
    if (ior < 1.0f) return 1.0f - pow2(ior) * (1.0f - dieletric_fresnel_factor(1.0f / ior));
    return 0.0f;
}

// --------------------------------------------------------------- diffuse --
// diffuse_brdf.glsl:9-16; the GLSL's pow(s, 3.0) is evaluated as s * s * s.
TT_DEV float oren_nayar(float c, float s2) {
    float a = 1.0f - 0.5f * (s2 / (s2 + 0.33f));
    float b = 0.45f * s2 / (s2 + 0.09f);
    float s = sqrt_(1.0f - pow2(c));
    float g = s * (ttm::acos(c) - s * c) + (s / fmax_(1e-7f, c)) * (1.0f - s * s * s) * 2.0f / 3.0f;
    return a + (b * INVERSE_PI) * g;
}
// The factor max(1e-7, 1 - oren_nayar) of one direction in the energy
// compensation; the incoming direction's is computed once per shading point
// (Lobes::diffuse_wi) and reused by every evaluation there.
TT_DEV float oren_nayar_complement(const Mat& m, f3 w) {
    float s2 = pow2(PI / 2.0f * m.base_diffuse_roughness);
    return fmax_(1e-7f, 1.0f - oren_nayar(cos_theta(w), s2));
}
// diffuse_brdf.glsl:18-32, with wi_term = oren_nayar_complement(m, wi).
TT_DEV f3 energy_compensated_oren_nayar(f3 rho, float sigma, f3 wi, f3 wo, float wi_term) {
    float s2 = pow2(sigma);
    float s = dot(wi, wo) - cos_theta(wi) * cos_theta(wo);
    float s_over_t = s > 0.0f ? s / fmax_(cos_theta(wi), cos_theta(wo)) : 0.0f;
    float a = 1.0f - 0.5f * (s2 / (s2 + 0.33f));
    float b = 0.45f * s2 / (s2 + 0.09f);
    float on_o = oren_nayar(cos_theta(wo), s2);
    float average_albedo = a + (2.0f / 3.0f - 64.0f / (45.0f * PI)) * b;
    f3 rho_ms = rho * rho * average_albedo / (1.0f - rho * fmax_(0.0f, 1.0f - average_albedo));
    return (rho * INVERSE_PI) * (a + b * s_over_t) +
        (rho_ms * INVERSE_PI) * fmax_(1e-7f, 1.0f - on_o) * wi_term / fmax_(1e-7f, 1.0f - average_albedo);
}
TT_DEV f3 diffuse_brdf_evaluate(const Mat& m, f3 wi, f3 wo, float wi_term, float& density) {
    if (cos_theta(wi) < DENOM_TOLERANCE || cos_theta(wo) < DENOM_TOLERANCE) return f3();
    density = cosine_hemisphere_density(cos_theta(wo));
    return energy_compensated_oren_nayar(m.base_color * m.base_weight, PI / 2.0f * m.base_diffuse_roughness, wi, wo,
                                         wi_term);
}
TT_DEV f3 diffuse_brdf_sample(const Mat& m, f3 wi, float wi_term, Generator& g, f3& wo, float& density, bool& ok) {
    ok = false;
    if (cos_theta(wi) < DENOM_TOLERANCE) return f3();
    wo = cosine_hemisphere_sample(random_vec2(g));
    ok = true;
    density = cosine_hemisphere_density(cos_theta(wo));
    return energy_compensated_oren_nayar(m.base_color * m.base_weight, PI / 2.0f * m.base_diffuse_roughness, wi, wo,
                                         wi_term);
}
TT_DEV f3 diffuse_brdf_albedo(const Mat& m, f3 wi) {
    if (cos_theta(wi) < DENOM_TOLERANCE) return f3();
    return m.base_color * m.base_weight;
}

// -------------------------------------------------------- specular BRDF --
TT_DEV float specular_ior(const Mat& m) { return m.specular_ior / mixs(1.0f, m.coat_ior, m.coat_weight); }
TT_DEV float specular_ior_ratio(const Mat& m) {
    float ior = specular_ior(m);
    float f0 = fresnel_at_normal_incident(ior);
    float clamped = clamp_(m.specular_weight, 0.0f, 1.0f / fmax_(f0, DENOM_TOLERANCE));
    float eps = sign_(ior - 1.0f) * sqrt_(clamped * f0);
    return (1.0f + eps) / fmax_(1.0f - eps, DENOM_TOLERANCE);
}
TT_DEV f3 specular_fresnel_with_coat(const Mat& m, float c, float ior) {
    float coat_refract_angle = sqrt_(1.0f - (1.0f - pow2(c)) / pow2(m.coat_ior));
    float ior_mix = mixs(m.specular_ior, ior, m.coat_weight);
    return f3(fresnel_dielectric(coat_refract_angle, ior_mix));
}
TT_DEV f3 specular_brdf(const Mat& m, f3 wi, f3 wo, f3 wir, f3 wor, f3 mn, f2 alpha, float ior, float& density) {
    float distribution = microfacet_distribution(alpha, mn);
    float visible_normals = distribution * microfacet_masking(alpha, wir) * fmax_(0.0f, dot(wir, mn)) /
        fmax_(DENOM_TOLERANCE, cos_theta(wir));
    float jacobian = 1.0f / fmax_(abs_(4.0f * dot(wir, mn)), DENOM_TOLERANCE);
    density = visible_normals * jacobian;
    f3 fresnel;
    if (cos_theta(wir) > 0.0f) {
        fresnel = specular_fresnel_with_coat(m, abs_(dot(wir, mn)), ior);
    } else {
        fresnel = f3(fresnel_dielectric(abs_(dot(wir, mn)), ior));
    }
    f3 brdf = fresnel * distribution * microfacet_visibility(alpha, wir, wor) /
        fmax_(4.0f * abs_(cos_theta(wo)) * abs_(cos_theta(wi)), DENOM_TOLERANCE);
    return brdf * m.specular_color;
}
TT_DEV f3 specular_brdf_evaluate(const Mat& m, f3 wi, f3 wo, float& density) {
    if (!is_same_hemisphere(wo, wi)) return f3();
    float ratio = specular_ior_ratio(m);
    float ior = cos_theta(wi) > 0.0f ? ratio : 1.0f / ratio;
    if (abs_(ior - 1.0f) < IOR_EPSILON) return f3();
    f2 alpha = microfacet_roughness_to_alpha(m.specular_roughness, m.specular_roughness_anisotropy);
    LocalRotation rot = local_frame_rotation(2.0f * PI * m.specular_rotation);
    f3 wir = rotate_local(wi, rot), wor = rotate_local(wo, rot);
    f3 mn = normalize(wor + wir);
    if (dot(mn, wir) * cos_theta(wir) < 0.0f || dot(mn, wor) * cos_theta(wor) < 0.0f) return f3();
    return specular_brdf(m, wi, wo, wir, wor, mn, alpha, ior, density);
}
TT_DEV f3 specular_brdf_sample(const Mat& m, f3 wi, Generator& g, f3& wo, float& density, bool& ok) {
    ok = false;
    float ratio = specular_ior_ratio(m);
    float ior = cos_theta(wi) > 0.0f ? ratio : 1.0f / ratio;
    if (abs_(ior - 1.0f) < IOR_EPSILON) return f3();
    f2 alpha = microfacet_roughness_to_alpha(m.specular_roughness, m.specular_roughness_anisotropy);
    LocalRotation rot = local_frame_rotation(2.0f * PI * m.specular_rotation);
    f3 wir = rotate_local(wi, rot);
    f3 mn;
    if (cos_theta(wir) > 0.0f) {
        mn = microfacet_sample(alpha, wir, random_vec2(g));
    } else {
        mn = microfacet_sample(alpha, f3(wir.x, wir.y, -wir.z), random_vec2(g));
        mn.z = mn.z * -1.0f;
    }
    f3 wor = -reflect(wir, mn);
    if (!is_same_hemisphere(wir, wor)) {
        density = 1.0f;
        return f3();
    }
    wo = inverse_rotate_local(wor, rot);
    ok = true;
    return specular_brdf(m, wi, wo, wir, wor, mn, alpha, ior, density);
}
TT_DEV f3 specular_brdf_albedo(const Mat& m, f3 wi, Generator& g) {
    if (abs_(specular_ior_ratio(m) - 1.0f) < IOR_EPSILON) return f3();
    f3 albedo, wo;
    float density = 0.0f;
    bool ok;
    f3 brdf = specular_brdf_sample(m, wi, g, wo, density, ok);
    if (length(brdf) > RADIANCE_EPSILON) albedo = albedo + brdf * abs_(cos_theta(wo)) / fmax_(DENOM_TOLERANCE, density);
    return albedo / 1.0f;
}

// -------------------------------------------------------- specular BTDF --
TT_DEV f3 transmission_tint(const Mat& m) { return m.transmission_depth == 0.0f ? m.transmission_color : f3(1.0f); }
TT_DEV f3 specular_btdf(const Mat& m, f3 mn, f2 alpha, f3 wi, f3 wo, f3 wir, f3 wor, float ior, float& density) {
    float jacobian = pow2(ior) * abs_(cos_theta(wi)) / fmax_(pow2(cos_theta(wo) + ior * cos_theta(wi)), DENOM_TOLERANCE);
    float distribution = microfacet_distribution(alpha, mn);
    float visible_normals = distribution * microfacet_masking(alpha, wir) * fmax_(0.0f, dot(wir, mn)) /
        fmax_(DENOM_TOLERANCE, abs_(cos_theta(wir)));
    density = visible_normals * jacobian;
    float visibility = microfacet_visibility(alpha, wir, wor);
    float transmission = fmax_(0.0f, 1.0f - fresnel_dielectric(abs_(dot(wir, mn)), 1.0f / ior));
    float btdf = transmission * abs_(dot(wir, mn)) * jacobian * visibility * distribution /
        fmax_(abs_(cos_theta(wo)) * abs_(cos_theta(wi)), DENOM_TOLERANCE);
    return btdf * transmission_tint(m);
}
TT_DEV f3 specular_btdf_evaluate(const Mat& m, f3 wi, f3 wo, float& density) {
    if (is_same_hemisphere(wo, wi)) {
        density = 1.0f;
        return f3();
    }
    float ior = cos_theta(wi) > 0.0f ? 1.0f / m.specular_ior : m.specular_ior;
    if (abs_(ior - 1.0f) < IOR_EPSILON) {
        f3 tint = transmission_tint(m);
        density = 1.0f / PDF_EPSILON;
        return tint * density / fmax_(DENOM_TOLERANCE, abs_(cos_theta(wo)));
    }
    LocalRotation rot = local_frame_rotation(2.0f * PI * m.specular_rotation);
    f3 wir = rotate_local(wi, rot), wor = rotate_local(wo, rot);
    f3 mn = -wor - ior * wir;
    if (length_squared(mn) == 0.0f) return f3();
    mn = mn / fmax_(length(mn), 1e-8f);
    if (cos_theta(mn) <= 0.0f) mn = mn * -1.0f;
    f2 alpha = microfacet_roughness_to_alpha(m.specular_roughness, m.specular_roughness_anisotropy);
    return specular_btdf(m, mn, alpha, wi, wo, wir, wor, ior, density);
}
TT_DEV bool refraction_direction(f3 normal, float ior, f3 wi, f3& direction) {
    float cos_in = dot(wi, normal);
    float s2in = fmax_(0.0f, 1.0f - pow2(cos_in));
    float s2tr = pow2(ior) * s2in;
    if (s2tr >= 1.0f) return false;
    float cos_tr = sqrt_(1.0f - s2tr);
    direction = ior * (-wi) + (ior * cos_in - cos_tr) * normal;
    return true;
}
TT_DEV f3 specular_btdf_sample(const Mat& m, f3 wi, Generator& g, f3& wo, float& density, bool& ok) {
    float ior = cos_theta(wi) > 0.0f ? 1.0f / m.specular_ior : m.specular_ior;
    if (abs_(ior - 1.0f) < IOR_EPSILON) {
        wo = -wi;
        ok = true;
        density = 1.0f / PDF_EPSILON;
        f3 tint = transmission_tint(m);
        return tint * density / fmax_(DENOM_TOLERANCE, abs_(cos_theta(wo)));
    }
    ok = false;
    LocalRotation rot = local_frame_rotation(2.0f * PI * m.specular_rotation);
    f3 wir = rotate_local(wi, rot);
    f2 alpha = microfacet_roughness_to_alpha(m.specular_roughness, m.specular_roughness_anisotropy);
    f3 mn;
    if (cos_theta(wir) > 0.0f) {
        mn = microfacet_sample(alpha, wir, random_vec2(g));
    } else {
        mn = microfacet_sample(alpha, f3(wir.x, wir.y, -cos_theta(wir)), random_vec2(g));
        mn.z = mn.z * -1.0f;
    }
    f3 refracted;
    if (!refraction_direction(mn, ior, wir, refracted)) {
        density = PDF_EPSILON;
        return f3();
    }
    f3 wor = -(refracted / fmax_(length(refracted), 1e-8f));
    wo = inverse_rotate_local(wor, rot);
    ok = true;
    return specular_btdf(m, mn, alpha, wi, wo, wir, wor, ior, density);
}
TT_DEV f3 specular_btdf_albedo(const Mat& m, f3 wi, Generator& g) {
    if (abs_(m.specular_ior - 1.0f) < IOR_EPSILON) return transmission_tint(m);
    f3 albedo, wo;
    float density = 0.0f;
    bool ok;
    f3 btdf = specular_btdf_sample(m, wi, g, wo, density, ok);
    if (length(btdf) > RADIANCE_EPSILON) albedo = albedo + btdf * abs_(cos_theta(wo)) / fmax_(DENOM_TOLERANCE, density);
    return albedo / 1.0f;
}

// ----------------------------------------------------------------- metal --
TT_DEV f3 fresnel_metal(float c, f3 f0, f3 tint) {
    const float mu_bar = 1.0f / 7.0f;
    const float denom = mu_bar * pow6(1.0f - mu_bar);
    return fresnel_schlick(f0, c) - c * pow6(1.0f - c) * (1.0f - tint) * fresnel_schlick(f0, mu_bar) / denom;
}
TT_DEV f3 metal_brdf(const Mat& m, f2 alpha, f3 mn, f3 wi, f3 wo, f3 wir, f3 wor, float& density) {
    float distribution = microfacet_distribution(alpha, mn);
    float visible_normals = distribution * microfacet_masking(alpha, wir) * fmax_(0.0f, dot(wir, mn)) /
        fmax_(DENOM_TOLERANCE, cos_theta(wir));
    float jacobian = 1.0f / fmax_(abs_(4.0f * dot(wir, mn)), DENOM_TOLERANCE);
    density = fmax_(PDF_EPSILON, visible_normals * jacobian);
    f3 fresnel = fresnel_metal(abs_(dot(wir, mn)), m.base_color * m.base_weight, m.specular_color * m.specular_weight);
    float visibility = microfacet_visibility(alpha, wir, wor);
    return fresnel * distribution * visibility / fmax_(4.0f * abs_(cos_theta(wo)) * abs_(cos_theta(wi)), DENOM_TOLERANCE);
}
TT_DEV f3 metal_brdf_evaluate(const Mat& m, f3 wi, f3 wo, float& density) {
    if (cos_theta(wi) < DENOM_TOLERANCE || cos_theta(wo) < DENOM_TOLERANCE) {
        density = PDF_EPSILON;
        return f3();
    }
    LocalRotation rot = local_frame_rotation(2.0f * PI * m.specular_rotation);
    f3 wir = rotate_local(wi, rot), wor = rotate_local(wo, rot);
    f2 alpha = microfacet_roughness_to_alpha(m.specular_roughness, m.specular_roughness_anisotropy);
    f3 mn = normalize(wor + wir);
    return metal_brdf(m, alpha, mn, wi, wo, wir, wor, density);
}
TT_DEV f3 metal_brdf_sample(const Mat& m, f3 wi, Generator& g, f3& wo, float& density, bool& ok) {
    ok = false;
    if (cos_theta(wi) < DENOM_TOLERANCE) {
        density = PDF_EPSILON;
        return f3();
    }
    f2 alpha = microfacet_roughness_to_alpha(m.specular_roughness, m.specular_roughness_anisotropy);
    LocalRotation rot = local_frame_rotation(2.0f * PI * m.specular_rotation);
    f3 wir = rotate_local(wi, rot);
    f3 mn = microfacet_sample(alpha, wir, random_vec2(g));
    f3 wor = -reflect(wir, mn);
    if (!is_same_hemisphere(wir, wor)) return f3();
    wo = inverse_rotate_local(wor, rot);
    ok = true;
    return metal_brdf(m, alpha, mn, wi, wo, wir, wor, density);
}
TT_DEV f3 metal_brdf_albedo(const Mat& m, f3 wi, Generator& g) {
    if (cos_theta(wi) < DENOM_TOLERANCE) return f3();
    f3 albedo, wo;
    float density = 0.0f;
    bool ok;
    f3 brdf = metal_brdf_sample(m, wi, g, wo, density, ok);
    if (length(brdf) > RADIANCE_EPSILON) albedo = albedo + brdf * abs_(cos_theta(wo)) / fmax_(PDF_EPSILON, density);
    return albedo / 1.0f;
}

// ------------------------------------------------------------------ coat --
TT_DEV f3 coat_brdf(const Mat& m, f3 wi, f3 wo, f3 wir, f3 wor, float ior, f2 alpha, float& density) {
    f3 mn = normalize(wor + wir);
    float d = dot(wir, mn);
    float distribution = microfacet_distribution(alpha, mn);
    float visible_normals = distribution * microfacet_masking(alpha, wir) * fmax_(0.0f, d) / cos_theta(wir);
    float jacobian = 1.0f / fmax_(abs_(4.0f * d), DENOM_TOLERANCE);
    density = visible_normals * jacobian;
    float fresnel = fresnel_dielectric(abs_(d), ior);
    return f3(fresnel) * distribution * microfacet_visibility(alpha, wir, wor) /
        fmax_(4.0f * abs_(cos_theta(wo)) * abs_(cos_theta(wi)), DENOM_TOLERANCE);
}
TT_DEV f3 coat_brdf_evaluate(const Mat& m, f3 wi, f3 wo, float& density) {
    if (!is_same_hemisphere(wo, wi)) return f3();
    float ior = cos_theta(wi) > 0.0f ? m.coat_ior : 1.0f / m.coat_ior;
    if (abs_(ior - 1.0f) < IOR_EPSILON) return f3();
    f2 alpha = microfacet_roughness_to_alpha(m.coat_roughness, m.coat_roughness_anisotropy);
    LocalRotation rot = local_frame_rotation(2.0f * PI * m.coat_rotation);
    return coat_brdf(m, wi, wo, rotate_local(wi, rot), rotate_local(wo, rot), ior, alpha, density);
}
TT_DEV f3 coat_brdf_sample(const Mat& m, f3 wi, Generator& g, f3& wo, float& density, bool& ok) {
    ok = false;
    float ior = cos_theta(wi) > 0.0f ? m.coat_ior / 1.0f : 1.0f / m.coat_ior;
    if (abs_(ior - 1.0f) < IOR_EPSILON) return f3();
    f2 alpha = microfacet_roughness_to_alpha(m.coat_roughness, m.coat_roughness_anisotropy);
    LocalRotation rot = local_frame_rotation(2.0f * PI * m.coat_rotation);
    f3 wir = rotate_local(wi, rot);
    f3 mn = microfacet_sample(alpha, wir, random_vec2(g));
    if (cos_theta(wir) <= 0.0f) return f3();
    f3 wor = -reflect(wir, mn);
    if (!is_same_hemisphere(wir, wor)) return f3();
    wo = inverse_rotate_local(wor, rot);
    ok = true;
    return coat_brdf(m, wi, wo, wir, wor, ior, alpha, density);
}
TT_DEV f3 coat_brdf_albedo(const Mat& m, f3 wi, Generator& g) {
    float ior = m.coat_ior / 1.0f;
    if (abs_(ior - 1.0f) < IOR_EPSILON) return f3();
    f3 albedo, wo;
    float density = 0.0f;
    bool ok;
    f3 brdf = coat_brdf_sample(m, wi, g, wo, density, ok);
    if (length(brdf) > RADIANCE_EPSILON) albedo = albedo + brdf * abs_(cos_theta(wo)) / fmax_(PDF_EPSILON, density);
    return albedo / 1.0f;
}

// ------------------------------------------------------------------ fuzz --
TT_DEV float fuzz_albedo(float x, float y) {
    float s = y * (0.0206607f + 1.58491f * y) / (0.0379424f + y * (1.32227f + y));
    float mm = y * (-0.193854f + y * (-1.14885f + y * (1.7932f - 0.95943f * y * y))) / (0.046391f + y);
    float o = y * (0.000654023f + (-0.0207818f + 0.119681f * y) * y) / (1.26264f + y * (-1.92021f + y));
    return ttm::exp(-0.5f * pow2((x - mm) / s)) / (s * sqrt_(2.0f * PI)) + o;
}
TT_DEV f2 ltc_inverse_coeffs(float x, float y) {
    float a = (2.58126f * x + 0.813703f * y) * y / (1.0f + 0.310327f * x * x + 2.60994f * x * y);
    float b = sqrt_(1.0f - x) * (y - 1.0f) * y * y * y / (0.0000254053f + 1.71228f * x - 1.71506f * x * y + 1.34174f * y * y);
    return f2{a, b};
}
TT_DEV void orthonormal_basis_ltc(f3 n, f3& t, f3& b) {
    float ls = n.x * n.x + n.y * n.y;
    t = ls > 0.0f ? f3(n.x, n.y, 0.0f) * (1.0f / sqrt_(ls)) : f3(1.0f, 0.0f, 0.0f);
    b = f3(-t.y, t.x, 0.0f);
}
TT_DEV f3 fuzz_density_and_brdf(const Mat& m, f3 direction, f3 wi, float a_inverse, float roughness, float& density) {
    float jacobian = pow2(a_inverse / length_squared(direction));
    density = fmax_(cos_theta(direction), 0.0f) * INVERSE_PI * jacobian;
    float albedo = fuzz_albedo(cos_theta(wi), roughness);
    return m.fuzz_color * albedo * INVERSE_PI * jacobian;
}
TT_DEV f3 fuzz_brdf_evaluate(const Mat& m, f3 wi, f3 wo, float& density) {
    if (cos_theta(wi) < DENOM_TOLERANCE || cos_theta(wo) < DENOM_TOLERANCE) return f3();
    float roughness = clamp_(m.fuzz_roughness, 0.01f, 1.0f);
    f3 t, b;
    orthonormal_basis_ltc(wi, t, b);
    f3 n(0.0f, 0.0f, 1.0f);
    f3 w(dot(t, wo), dot(b, wo), dot(n, wo));
    f2 ic = ltc_inverse_coeffs(cos_theta(wi), roughness);
    f3 direction(ic.x * w.x + ic.y * w.z, ic.x * w.y, w.z);
    return fuzz_density_and_brdf(m, direction, wi, ic.x, roughness, density);
}
TT_DEV f3 fuzz_brdf_sample(const Mat& m, f3 wi, Generator& g, f3& wo, float& density, bool& ok) {
    ok = false;
    if (cos_theta(wi) < DENOM_TOLERANCE) return f3();
    float roughness = clamp_(m.fuzz_roughness, 0.01f, 1.0f);
    f3 direction = cosine_hemisphere_sample(random_vec2(g));
    f2 ic = ltc_inverse_coeffs(cos_theta(wi), roughness);
    f3 w(direction.x / ic.x - direction.z * ic.y / ic.x, direction.y / ic.x, direction.z);
    f3 t, b;
    orthonormal_basis_ltc(wi, t, b);
    f3 wn = normalize(w);
    f3 n(0.0f, 0.0f, 1.0f);
    wo = t * wn.x + b * wn.y + n * wn.z;
    ok = true;
    return fuzz_density_and_brdf(m, direction, wi, ic.x, roughness, density);
}
TT_DEV f3 fuzz_brdf_albedo(const Mat& m, f3 wi) { return f3(fuzz_albedo(cos_theta(wi), m.fuzz_roughness)); }

// ------------------------------------------------------------------ bsdf --
enum { FUZZ = 0, COAT = 1, METAL = 2, SPECULAR_BRDF = 3, SPECULAR_BTDF = 4, DIFFUSE = 5, LOBE_COUNT = 6 };

struct Lobes {
    f3 weights[LOBE_COUNT];
    float probs[LOBE_COUNT];
    float diffuse_wi;  // oren_nayar_complement(m, wi) when the diffuse lobe is active
};

// bsdf.glsl:33-113 (the albedos are only needed while preparing).
TT_DEV Lobes bsdf_prepare(const Mat& m, f3 wi, Generator& g) {
    Lobes l;
    f3 albedos[LOBE_COUNT];
    bool fully_metallic = m.base_metalness == 1.0f;
    bool fully_transmissive = m.transmission_weight == 1.0f;
    albedos[FUZZ] = m.fuzz_weight > 0.0f ? fuzz_brdf_albedo(m, wi) : f3();
    albedos[COAT] = m.coat_weight > 0.0f ? coat_brdf_albedo(m, wi, g) : f3();
    albedos[METAL] = m.base_metalness > 0.0f ? metal_brdf_albedo(m, wi, g) : f3();
    albedos[SPECULAR_BRDF] = !fully_metallic ? specular_brdf_albedo(m, wi, g) : f3();
    albedos[SPECULAR_BTDF] = !fully_metallic && m.transmission_weight > 0.0f ? specular_btdf_albedo(m, wi, g) : f3();
    albedos[DIFFUSE] = !fully_metallic && !fully_transmissive ? diffuse_brdf_albedo(m, wi) : f3();
    l.diffuse_wi = !fully_metallic && !fully_transmissive ? oren_nayar_complement(m, wi) : 1e-7f;

    l.weights[FUZZ] = f3(m.fuzz_weight);
    f3 coated_base_weight = mixv(f3(1.0f), 1.0f - albedos[FUZZ], m.fuzz_weight);
    l.weights[COAT] = coated_base_weight * m.coat_weight;
    f3 base_darkening(1.0f);
    if (m.coat_weight > 0.0f && m.coat_darkening > 0.0f) {
        float fresnel_weight = saturate(m.specular_weight * fresnel_at_normal_incident(specular_ior(m)));
        float dielectric_roughness = mixs(1.0f, m.specular_roughness, fresnel_weight);
        float base_roughness = mixs(dielectric_roughness, m.specular_roughness, m.base_metalness);
        float average_fresnel = 1.0f - (1.0f - average_dielectric_fresnel(m.coat_ior)) / pow2(m.coat_ior);
        float fresnel = fresnel_dielectric(abs_(cos_theta(wi)), m.coat_ior);
        float drc = mixs(fresnel, average_fresnel, base_roughness);
        f3 dielectric_base_albedo = mixv(albedos[DIFFUSE], albedos[SPECULAR_BTDF], m.transmission_weight);
        f3 base_albedo = mixv(dielectric_base_albedo, albedos[METAL], m.base_metalness);
        f3 darkening_factor = (1.0f - drc) / (1.0f - base_albedo * drc);
        base_darkening = mixv(f3(1.0f), darkening_factor, m.coat_weight * m.coat_darkening);
    }
    f3 base_weight = coated_base_weight * mixv(f3(1.0f), base_darkening * m.coat_color * (1.0f - albedos[COAT]), m.coat_weight);
    l.weights[METAL] = base_weight * m.base_metalness;
    f3 dielectric_base_weight = base_weight * f3(1.0f - m.base_metalness);
    l.weights[SPECULAR_BRDF] = dielectric_base_weight;
    l.weights[SPECULAR_BTDF] = dielectric_base_weight * m.transmission_weight;
    f3 opaque = dielectric_base_weight * (1.0f - m.transmission_weight);
    l.weights[DIFFUSE] = opaque * (1.0f - albedos[SPECULAR_BRDF]);

    float total = 0.0f;
#pragma unroll
    for (int i = 0; i < LOBE_COUNT; i++) {
        l.probs[i] = length(l.weights[i] * albedos[i]);
        total = total + l.probs[i];
    }
    total = fmax_(DENOM_TOLERANCE, total);
#pragma unroll
    for (int i = 0; i < LOBE_COUNT; i++) l.probs[i] = l.probs[i] / total;
    return l;
}

// bsdf.glsl:122-147
TT_DEV f3 bsdf_evaluate_lobes(const Mat& m, const Lobes& l, f3 wi, f3 wo, int skip, float d[LOBE_COUNT]) {
    f3 bsdf;
    if (skip != FUZZ && l.probs[FUZZ] > 0.0f) bsdf = bsdf + l.weights[FUZZ] * fuzz_brdf_evaluate(m, wi, wo, d[FUZZ]);
    if (skip != COAT && l.probs[COAT] > 0.0f) bsdf = bsdf + l.weights[COAT] * coat_brdf_evaluate(m, wi, wo, d[COAT]);
    if (skip != METAL && l.probs[METAL] > 0.0f) bsdf = bsdf + l.weights[METAL] * metal_brdf_evaluate(m, wi, wo, d[METAL]);
    if (skip != SPECULAR_BRDF && l.probs[SPECULAR_BRDF] > 0.0f)
        bsdf = bsdf + l.weights[SPECULAR_BRDF] * specular_brdf_evaluate(m, wi, wo, d[SPECULAR_BRDF]);
    if (skip != DIFFUSE && l.probs[DIFFUSE] > 0.0f)
        bsdf = bsdf + l.weights[DIFFUSE] * diffuse_brdf_evaluate(m, wi, wo, l.diffuse_wi, d[DIFFUSE]);
    if (skip != SPECULAR_BTDF && l.probs[SPECULAR_BTDF] > 0.0f)
        bsdf = bsdf + l.weights[SPECULAR_BTDF] * specular_btdf_evaluate(m, wi, wo, d[SPECULAR_BTDF]);
    return bsdf;
}

TT_DEV float bsdf_total_pdf(const Lobes& l, const float d[LOBE_COUNT]) {
    float density = 0.0f;
#pragma unroll
    for (int i = 0; i < LOBE_COUNT; i++) density = density + l.probs[i] * d[i];
    return density;
}

TT_DEV f3 bsdf_evaluate(const Mat& m, const Lobes& l, f3 wi, f3 wo, float& density) {
    float d[LOBE_COUNT] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    f3 bsdf = bsdf_evaluate_lobes(m, l, wi, wo, -1, d);
    density = bsdf_total_pdf(l, d);
    return bsdf;
}

// bsdf.glsl:163-203; false when no direction was produced.  The selected
// lobe is found first and every array is then indexed with compile-time
// indices (dynamic indexing would spill the arrays to scratch memory).
TT_DEV bool bsdf_sample(const Mat& m, const Lobes& l, f3 wi, Generator& g, f3& wo, float& density, f3& value) {
    float random = random_float(g);
    float accumulated = 0.0f;
    int lobe = -1;
#pragma unroll
    for (int i = 0; i < LOBE_COUNT; i++) {
        accumulated = accumulated + l.probs[i];
        if (lobe < 0 && random < accumulated) lobe = i;
    }
    if (lobe < 0) return false;
    float lobe_density = 0.0f;
    f3 lobe_bsdf;
    bool ok = false;
    switch (lobe) {
        case FUZZ: lobe_bsdf = fuzz_brdf_sample(m, wi, g, wo, lobe_density, ok); break;
        case COAT: lobe_bsdf = coat_brdf_sample(m, wi, g, wo, lobe_density, ok); break;
        case METAL: lobe_bsdf = metal_brdf_sample(m, wi, g, wo, lobe_density, ok); break;
        case SPECULAR_BRDF: lobe_bsdf = specular_brdf_sample(m, wi, g, wo, lobe_density, ok); break;
        case SPECULAR_BTDF: lobe_bsdf = specular_btdf_sample(m, wi, g, wo, lobe_density, ok); break;
        default: lobe_bsdf = diffuse_brdf_sample(m, wi, l.diffuse_wi, g, wo, lobe_density, ok); break;
    }
    if (!ok) return false;
    float d[LOBE_COUNT] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    f3 bsdf = bsdf_evaluate_lobes(m, l, wi, wo, lobe, d);
    float total = 0.0f;
    f3 weight;
#pragma unroll
    for (int i = 0; i < LOBE_COUNT; i++) {
        total = total + l.probs[i] * (i == lobe ? lobe_density : d[i]);
        if (i == lobe) weight = l.weights[i];
    }
    density = total;
    value = bsdf + weight * lobe_bsdf;
    return true;
}

}  // namespace ttd
