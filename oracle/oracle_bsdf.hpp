// ORACLE (test infrastructure only; never linked into the product).
//
// Restatement of the reference's OpenPBR-style BSDF:
//   src/material/material.glsl     (constants, MaterialConstants)
//   src/material/microfacet.glsl   (GGX D, Smith lambda/G, VNDF sampling)
//   src/material/fresnel.glsl      (Schlick, dielectric Fresnel, Eqs. 102-103)
//   src/material/diffuse_brdf.glsl (energy-compensated Oren-Nayar)
//   src/material/specular_brdf.glsl, specular_btdf.glsl, metal_brdf.glsl,
//   src/material/coat_brdf.glsl, fuzz_brdf.glsl
//   src/material/bsdf.glsl         (lobe weights/probabilities/evaluate/sample)
//   src/includes/sample.glsl       (cosine hemisphere, triangle sampling)
//   src/includes/random.glsl       (generator)
// Function names follow the reference; each cites its source lines.
//
// Reference behaviour left undefined by GLSL, defined here (and identically in
// the GPU kernels):
//   * LobeDensities starts at zero (bsdf.glsl:157,190 leave it uninitialised).
//   * A lobe sample that returns without writing `wo` reports failure
//     (`ok = false`); bsdf_sample then fails and the path terminates.
//
// Pins (tests/test_bsdf_cpu.py, tests/test_bsdf_pins_cpu.py through
// bsdf_hooks.cpp): closed forms, quadrature, published special cases and
// the reference's fitted constants, each shown to fail on a planted slip.
// specular_btdf_evaluate's Jacobian is the reference's own, with normal-frame
// cosines where Walter 2007 has half-vector dots: it is pinned where the two
// agree (wo = wi refracted through the macro surface, half vector = normal),
// against a finite-difference Jacobian of the half-vector map.
// fuzz_brdf_sample reports density and value at the LTC-space direction; it
// is pinned through its exact relation to fuzz_brdf_evaluate at the sampled
// wo and through the distribution of the sampled wo.
#pragma once

#include "oracle_math.hpp"

namespace oracle {

// ---------------------------------------------------------------- random --
// random.glsl:4-11
inline uint32_t jenkins_hash(uint32_t x) {
    x += x << 10;
    x ^= x >> 6;
    x += x << 3;
    x ^= x >> 11;
    x += x << 15;
    return x;
}

// pcg32 = PCG-XSH-RR 64/32 (O'Neill 2014, "PCG: A Family of Simple Fast
// Space-Efficient Statistically Good Algorithms for Random Number
// Generation", section 6.3.1): 64-bit LCG, multiplier 6364136223846793005,
// output from the pre-step state.  `inc` is the (odd) stream increment;
// pcg32_srandom(initstate, initseq) uses inc = 2 * initseq + 1.
constexpr uint64_t PCG32_MULT = 6364136223846793005ull;
constexpr uint64_t PCG32_INC = 0xda3e39cb94b95bdbull;  // pcg32's default stream

inline uint32_t pcg32_output(uint64_t old) {
    uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}
inline uint64_t pcg32_srandom_state(uint64_t initstate, uint64_t inc) {
    uint64_t s = 0;
    s = s * PCG32_MULT + inc;
    s += initstate;
    return s * PCG32_MULT + inc;
}

// SplitMix64 (Steele, Lea & Flood 2014; the finaliser of Vigna's
// splitmix64.c).
inline uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Generator: `kind` 0 = the reference's xorshift32 (random.glsl:29-34) in the
// low 32 bits of `state`, kind 1 = pcg32 on the default stream.  Both produce
// one u32 per draw; the draw order is the reference's.
struct Generator {
    uint64_t state;
    uint32_t kind;
};

inline Generator pcg32_seeded(uint64_t initstate) { return Generator{pcg32_srandom_state(initstate, PCG32_INC), 1}; }
inline uint64_t pcg32_key(uint32_t index, uint32_t frame) { return (uint64_t(frame) << 32) | index; }

// random.glsl:17-23 (reference), and the pcg32 seeding of this library:
// initstate = splitmix64(((frame << 32) | index) ^ splitmix64(seed)).
inline Generator init_generator_from_index(uint32_t index, uint32_t frame, uint32_t kind, uint32_t seed) {
    if (kind == 1) return pcg32_seeded(splitmix64(pcg32_key(index, frame) ^ splitmix64(seed)));
    return Generator{jenkins_hash(index ^ jenkins_hash(frame)), 0};
}
inline Generator init_generator_from_pixel(uint32_t x, uint32_t y, uint32_t width, uint32_t frame,
                                           uint32_t kind, uint32_t seed) {
    return init_generator_from_index(x + y * width, frame, kind, seed);
}

inline uint32_t next_u32(Generator& g) {
    if (g.kind == 1) {
        uint64_t old = g.state;
        g.state = old * PCG32_MULT + PCG32_INC;
        return pcg32_output(old);
    }
    // random.glsl:29-34
    uint32_t s = uint32_t(g.state);
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    g.state = s;
    return s;
}

// random.glsl:25-27
inline float uint_to_unit_float(uint32_t v) {
    uint32_t bits = 0x3f800000u | (v >> 9);
    float f;
    std::memcpy(&f, &bits, 4);
    return f - 1.0f;
}

// random.glsl:36-42.  The float is produced in single precision (bit cast)
// in both instantiations, then widened.
template <typename R>
inline R random_float(Generator& g) { return R(uint_to_unit_float(next_u32(g))); }
inline uint32_t random_uint(Generator& g, uint32_t end) { return next_u32(g) % end; }
template <typename R>
inline V2<R> random_vec2(Generator& g) {
    R a = random_float<R>(g);  // GLSL evaluates constructor arguments left to right
    R b = random_float<R>(g);
    return {a, b};
}

// ------------------------------------------------------------- constants --
template <typename R>
struct K {
    static constexpr R PI = R(3.1415926535897932);
    static constexpr R TAU = R(6.283185307179586);
    static constexpr R INVERSE_PI = R(1) / R(3.1415926535897932);
    static constexpr R DENOM_TOLERANCE = R(1.0e-10);
    static constexpr R RADIANCE_EPSILON = R(1.0e-12);
    static constexpr R PDF_EPSILON = R(1.0e-6);
    static constexpr R IOR_EPSILON = R(1.0e-5);
};

template <typename R>
struct Material {
    V3<R> base_color, specular_color, transmission_color, coat_color, fuzz_color, emission_color,
        geometry_normal;
    R base_weight, base_metalness, base_diffuse_roughness, specular_weight, specular_roughness,
        specular_roughness_anisotropy, specular_ior, specular_rotation, transmission_weight,
        transmission_depth, coat_weight, coat_roughness, coat_roughness_anisotropy, coat_ior,
        coat_darkening, coat_rotation, fuzz_weight, fuzz_roughness, emission_luminance,
        geometry_opacity;
};

// Reads MaterialConstants (material.glsl:10-40) from its 192-byte layout.
template <typename R>
inline Material<R> load_material(const float* f) {
    Material<R> m;
    auto v = [&](int o) { return V3<R>(R(f[o]), R(f[o + 1]), R(f[o + 2])); };
    m.base_color = v(0);
    m.specular_color = v(4);
    m.transmission_color = v(8);
    m.coat_color = v(12);
    m.fuzz_color = v(16);
    m.emission_color = v(20);
    m.geometry_normal = v(24);
    const float* s = f + 28;
    m.base_weight = s[0];
    m.base_metalness = s[1];
    m.base_diffuse_roughness = s[2];
    m.specular_weight = s[3];
    m.specular_roughness = s[4];
    m.specular_roughness_anisotropy = s[5];
    m.specular_ior = s[6];
    m.specular_rotation = s[7];
    m.transmission_weight = s[8];
    m.transmission_depth = s[9];
    m.coat_weight = s[10];
    m.coat_roughness = s[11];
    m.coat_roughness_anisotropy = s[12];
    m.coat_ior = s[13];
    m.coat_darkening = s[14];
    m.coat_rotation = s[15];
    m.fuzz_weight = s[16];
    m.fuzz_roughness = s[17];
    m.emission_luminance = s[18];
    m.geometry_opacity = s[19];
    return m;
}

// ---------------------------------------------------------------- sample --
// sample.glsl:7-16 (Heitz 2019 low-distortion triangle map)
template <typename R>
inline V3<R> sample_triangle(V2<R> r) {
    r.x = r.x / R(2);
    r.y = r.y / R(2);
    R offset = r.y - r.x;
    if (offset > R(0)) {
        r.y = r.y + offset;
    } else {
        r.x = r.x - offset;
    }
    return V3<R>(r.x, r.y, R(1) - r.x - r.y);
}

// sample.glsl:18-24
template <typename R>
inline V3<R> cosine_hemisphere_sample(V2<R> r) {
    R radius = Fn<R>::sqrt(r.x);
    R theta = R(2) * K<R>::PI * r.y;
    R x = radius * Fn<R>::cos(theta), y = radius * Fn<R>::sin(theta);
    R z = Fn<R>::sqrt(rmax(R(0), R(1) - x * x - y * y));
    return V3<R>(x, y, z);
}

// sample.glsl:26-30
template <typename R>
inline R cosine_hemisphere_density(R c) {
    if (c <= R(1e-6)) return R(1e-6) / K<R>::PI;
    return c / K<R>::PI;
}

// ------------------------------------------------------------ local frame --
// math.glsl:35-60
template <typename R>
struct LocalRotation {
    R t00, t01, t10, t11;  // transform: columns (t00, t01), (t10, t11)
    R i00, i01, i10, i11;  // inverse
};

template <typename R>
inline LocalRotation<R> local_frame_rotation(R angle) {
    LocalRotation<R> r;
    if (angle == R(0) || angle == R(2) * K<R>::PI) {
        r = {R(1), R(0), R(0), R(1), R(1), R(0), R(0), R(1)};
    } else {
        R c = Fn<R>::cos(angle), s = Fn<R>::sin(angle);
        r = {c, s, -s, c, c, -s, s, c};
    }
    return r;
}
template <typename R>
inline V3<R> rotate_local(V3<R> v, const LocalRotation<R>& r) {
    return V3<R>(r.t00 * v.x + r.t10 * v.y, r.t01 * v.x + r.t11 * v.y, v.z);
}
template <typename R>
inline V3<R> inverse_rotate_local(V3<R> v, const LocalRotation<R>& r) {
    return V3<R>(r.i00 * v.x + r.i10 * v.y, r.i01 * v.x + r.i11 * v.y, v.z);
}

// ------------------------------------------------------------ microfacet --
// microfacet.glsl:7-16
template <typename R>
inline R microfacet_distribution(V2<R> alpha, V3<R> m) {
    R tan_squared = tan_theta_squared(m);
    if (Fn<R>::isinf(tan_squared)) return R(0);
    R e = tan_squared * (pow2(cos_phi(m) / alpha.x) + pow2(sin_phi(m) / alpha.y));
    return R(1) / (K<R>::PI * alpha.x * alpha.y * pow2(cos_theta_squared(m)) * pow2(R(1) + e));
}

// microfacet.glsl:18-24
template <typename R>
inline R microfacet_lambda(V2<R> alpha, V3<R> v) {
    R t2 = tan_theta_squared(v);
    if (Fn<R>::isinf(t2)) return R(0);
    R a2 = pow2(cos_phi(v) * alpha.x) + pow2(sin_phi(v) * alpha.y);
    return (Fn<R>::sqrt(R(1) + a2 * t2) - R(1)) / R(2);
}

// microfacet.glsl:27-34
template <typename R>
inline R microfacet_visibility(V2<R> alpha, V3<R> i, V3<R> o) {
    return R(1) / (R(1) + microfacet_lambda(alpha, o) + microfacet_lambda(alpha, i));
}
template <typename R>
inline R microfacet_masking(V2<R> alpha, V3<R> v) {
    return R(1) / (R(1) + microfacet_lambda(alpha, v));
}

// microfacet.glsl:40-47 (VNDF sampling by spherical caps)
template <typename R>
inline V3<R> microfacet_sample(V2<R> alpha, V3<R> view, V2<R> r) {
    view = normalize(V3<R>(view.x * alpha.x, view.y * alpha.y, view.z));
    R phi = R(2) * K<R>::PI * r.x;
    R z = (R(1) - r.y) * (R(1) + view.z) - view.z;
    R st = Fn<R>::sqrt(saturate(R(1) - pow2(z)));
    V3<R> h = V3<R>(st * Fn<R>::cos(phi), st * Fn<R>::sin(phi), z) + view;
    return normalize(V3<R>(h.x * alpha.x, h.y * alpha.y, h.z));
}

// microfacet.glsl:49-54
template <typename R>
inline V2<R> microfacet_roughness_to_alpha(R roughness, R anisotropy) {
    R rs = pow2(roughness);
    R ax = rs * Fn<R>::sqrt(R(2) / (R(1) + pow2(R(1) - anisotropy)));
    R ay = (R(1) - anisotropy) * ax;
    return {rmax(R(1e-4), ax), rmax(R(1e-4), ay)};
}

// --------------------------------------------------------------- fresnel --
// fresnel.glsl:6-39
template <typename R>
inline V3<R> fresnel_schlick(V3<R> f0, R c) {
    return f0 + pow5(R(1) - c) * (R(1) - f0);
}
template <typename R>
inline R fresnel_at_normal_incident(R ior) { return pow2((ior - R(1)) / (ior + R(1))); }
template <typename R>
inline R dieletric_fresnel_factor(R ior) {
    return Fn<R>::log((R(10893.0) * ior - R(1438.2)) /
                      (R(-774.4) * pow2(ior) + R(10212.0) * ior + R(1)));
}
template <typename R>
inline R fresnel_dielectric(R cos_i, R ior) {
    R s2i = R(1) - pow2(cos_i);
    R s2t = s2i / pow2(ior);
    if (s2t >= R(1)) return R(1);
    R cos_t = Fn<R>::sqrt(R(1) - s2t);
    R par = (ior * cos_i - cos_t) / (ior * cos_i + cos_t);
    R per = (cos_i - ior * cos_t) / (cos_i + ior * cos_t);
    return R(0.5) * (pow2(par) + pow2(per));
}
template <typename R>
inline R average_dielectric_fresnel(R ior) {
    if (ior > R(1)) return dieletric_fresnel_factor(ior);
    if (ior < R(1)) return R(1) - pow2(ior) * (R(1) - dieletric_fresnel_factor(R(1) / ior));
    return R(0);
}

// --------------------------------------------------------------- diffuse --
// diffuse_brdf.glsl:9-16; the GLSL's pow(s, 3.0) is evaluated as s * s * s.
template <typename R>
inline R oren_nayar(R c, R s2) {
    R a = R(1) - R(0.5) * (s2 / (s2 + R(0.33)));
    R b = R(0.45) * s2 / (s2 + R(0.09));
    R s = Fn<R>::sqrt(R(1) - pow2(c));
    R g = s * (Fn<R>::acos(c) - s * c) +
        (s / rmax(R(1e-7), c)) * (R(1) - s * s * s) * R(2) / R(3);
    return a + (b * K<R>::INVERSE_PI) * g;
}

// diffuse_brdf.glsl:18-32
template <typename R>
inline V3<R> energy_compensated_oren_nayar(V3<R> rho, R sigma, V3<R> wi, V3<R> wo) {
    R s2 = pow2(sigma);
    R s = dot(wi, wo) - cos_theta(wi) * cos_theta(wo);
    R s_over_t = s > R(0) ? s / rmax(cos_theta(wi), cos_theta(wo)) : R(0);
    R a = R(1) - R(0.5) * (s2 / (s2 + R(0.33)));
    R b = R(0.45) * s2 / (s2 + R(0.09));
    R on_o = oren_nayar(cos_theta(wo), s2);
    R on_i = oren_nayar(cos_theta(wi), s2);
    R average_albedo = a + (R(2) / R(3) - R(64) / (R(45) * K<R>::PI)) * b;
    V3<R> rho_ms = rho * rho * average_albedo / (R(1) - rho * rmax(R(0), R(1) - average_albedo));
    return (rho * K<R>::INVERSE_PI) * (a + b * s_over_t) +
        (rho_ms * K<R>::INVERSE_PI) * rmax(R(1e-7), R(1) - on_o) * rmax(R(1e-7), R(1) - on_i) /
        rmax(R(1e-7), R(1) - average_albedo);
}

// diffuse_brdf.glsl:34-42
template <typename R>
inline V3<R> diffuse_brdf_evaluate(const Material<R>& m, V3<R> wi, V3<R> wo, R& density) {
    if (cos_theta(wi) < K<R>::DENOM_TOLERANCE || cos_theta(wo) < K<R>::DENOM_TOLERANCE) return V3<R>();
    density = cosine_hemisphere_density(cos_theta(wo));
    return energy_compensated_oren_nayar(m.base_color * m.base_weight,
                                         K<R>::PI / R(2) * m.base_diffuse_roughness, wi, wo);
}

// diffuse_brdf.glsl:44-55
template <typename R>
inline V3<R> diffuse_brdf_sample(const Material<R>& m, V3<R> wi, Generator& g, V3<R>& wo, R& density,
                                 bool& ok) {
    if (cos_theta(wi) < K<R>::DENOM_TOLERANCE) {
        ok = false;
        return V3<R>();
    }
    wo = cosine_hemisphere_sample(random_vec2<R>(g));
    ok = true;
    density = cosine_hemisphere_density(cos_theta(wo));
    return energy_compensated_oren_nayar(m.base_color * m.base_weight,
                                         K<R>::PI / R(2) * m.base_diffuse_roughness, wi, wo);
}

// diffuse_brdf.glsl:57-61
template <typename R>
inline V3<R> diffuse_brdf_albedo(const Material<R>& m, V3<R> wi) {
    if (cos_theta(wi) < K<R>::DENOM_TOLERANCE) return V3<R>();
    return m.base_color * m.base_weight;
}

// ------------------------------------------------------ specular (BRDF) --
// specular_brdf.glsl:8-11
template <typename R>
inline R specular_ior(const Material<R>& m) {
    return m.specular_ior / mixs(R(1), m.coat_ior, m.coat_weight);
}

// specular_brdf.glsl:13-21
template <typename R>
inline R specular_ior_ratio(const Material<R>& m) {
    R ior = specular_ior(m);
    R f0 = fresnel_at_normal_incident(ior);
    R clamped = rclamp(m.specular_weight, R(0), R(1) / rmax(f0, K<R>::DENOM_TOLERANCE));
    R eps = rsign(ior - R(1)) * Fn<R>::sqrt(clamped * f0);
    return (R(1) + eps) / rmax(R(1) - eps, K<R>::DENOM_TOLERANCE);
}

// specular_brdf.glsl:23-28
template <typename R>
inline V3<R> specular_fresnel_with_coat(const Material<R>& m, R c, R ior) {
    R coat_refract_angle = Fn<R>::sqrt(R(1) - (R(1) - pow2(c)) / pow2(m.coat_ior));
    R ior_mix = mixs(m.specular_ior, ior, m.coat_weight);
    return V3<R>(fresnel_dielectric(coat_refract_angle, ior_mix));
}

// specular_brdf.glsl:30-49
template <typename R>
inline V3<R> specular_brdf(const Material<R>& m, V3<R> wi, V3<R> wo, V3<R> wir, V3<R> wor, V3<R> mn,
                           V2<R> alpha, R ior, R& density) {
    R distribution = microfacet_distribution(alpha, mn);
    R visible_normals = distribution * microfacet_masking(alpha, wir) * rmax(R(0), dot(wir, mn)) /
        rmax(K<R>::DENOM_TOLERANCE, cos_theta(wir));
    R jacobian = R(1) / rmax(rabs(R(4) * dot(wir, mn)), K<R>::DENOM_TOLERANCE);
    density = visible_normals * jacobian;
    V3<R> fresnel;
    if (cos_theta(wir) > R(0)) {
        fresnel = specular_fresnel_with_coat(m, rabs(dot(wir, mn)), ior);
    } else {
        fresnel = V3<R>(fresnel_dielectric(rabs(dot(wir, mn)), ior));
    }
    V3<R> brdf = fresnel * distribution * microfacet_visibility(alpha, wir, wor) /
        rmax(R(4) * rabs(cos_theta(wo)) * rabs(cos_theta(wi)), K<R>::DENOM_TOLERANCE);
    return brdf * m.specular_color;
}

// specular_brdf.glsl:51-70
template <typename R>
inline V3<R> specular_brdf_evaluate(const Material<R>& m, V3<R> wi, V3<R> wo, R& density) {
    if (!is_same_hemisphere(wo, wi)) return V3<R>();
    R ratio = specular_ior_ratio(m);
    R ior = cos_theta(wi) > R(0) ? ratio : R(1) / ratio;
    if (rabs(ior - R(1)) < K<R>::IOR_EPSILON) return V3<R>();
    V2<R> alpha = microfacet_roughness_to_alpha(m.specular_roughness, m.specular_roughness_anisotropy);
    LocalRotation<R> rot = local_frame_rotation(R(2) * K<R>::PI * m.specular_rotation);
    V3<R> wir = rotate_local(wi, rot), wor = rotate_local(wo, rot);
    V3<R> mn = normalize(wor + wir);
    if (dot(mn, wir) * cos_theta(wir) < R(0) || dot(mn, wor) * cos_theta(wor) < R(0)) return V3<R>();
    return specular_brdf(m, wi, wo, wir, wor, mn, alpha, ior, density);
}

// specular_brdf.glsl:72-101
template <typename R>
inline V3<R> specular_brdf_sample(const Material<R>& m, V3<R> wi, Generator& g, V3<R>& wo, R& density,
                                  bool& ok) {
    ok = false;
    R ratio = specular_ior_ratio(m);
    R ior = cos_theta(wi) > R(0) ? ratio : R(1) / ratio;
    if (rabs(ior - R(1)) < K<R>::IOR_EPSILON) return V3<R>();
    V2<R> alpha = microfacet_roughness_to_alpha(m.specular_roughness, m.specular_roughness_anisotropy);
    LocalRotation<R> rot = local_frame_rotation(R(2) * K<R>::PI * m.specular_rotation);
    V3<R> wir = rotate_local(wi, rot);
    V3<R> mn;
    if (cos_theta(wir) > R(0)) {
        mn = microfacet_sample(alpha, wir, random_vec2<R>(g));
    } else {
        mn = microfacet_sample(alpha, V3<R>(wir.x, wir.y, -wir.z), random_vec2<R>(g));
        mn.z = mn.z * R(-1);
    }
    V3<R> wor = -reflect(wir, mn);
    if (!is_same_hemisphere(wir, wor)) {
        density = R(1);
        return V3<R>();
    }
    wo = inverse_rotate_local(wor, rot);
    ok = true;
    return specular_brdf(m, wi, wo, wir, wor, mn, alpha, ior, density);
}

// specular_brdf.glsl:103-116
template <typename R>
inline V3<R> specular_brdf_albedo(const Material<R>& m, V3<R> wi, Generator& g) {
    if (rabs(specular_ior_ratio(m) - R(1)) < K<R>::IOR_EPSILON) return V3<R>();
    V3<R> albedo;
    V3<R> wo;
    R density = R(0);
    bool ok;
    V3<R> brdf = specular_brdf_sample(m, wi, g, wo, density, ok);
    if (length(brdf) > K<R>::RADIANCE_EPSILON)
        albedo = albedo + brdf * rabs(cos_theta(wo)) / rmax(K<R>::DENOM_TOLERANCE, density);
    return albedo / R(1);
}

// ------------------------------------------------------ specular (BTDF) --
// specular_btdf.glsl:8-26
template <typename R>
inline V3<R> specular_btdf(const Material<R>& m, V3<R> mn, V2<R> alpha, V3<R> wi, V3<R> wo, V3<R> wir,
                           V3<R> wor, R ior, R& density) {
    R jacobian = pow2(ior) * rabs(cos_theta(wi)) /
        rmax(pow2(cos_theta(wo) + ior * cos_theta(wi)), K<R>::DENOM_TOLERANCE);
    R distribution = microfacet_distribution(alpha, mn);
    R visible_normals = distribution * microfacet_masking(alpha, wir) * rmax(R(0), dot(wir, mn)) /
        rmax(K<R>::DENOM_TOLERANCE, rabs(cos_theta(wir)));
    density = visible_normals * jacobian;
    R visibility = microfacet_visibility(alpha, wir, wor);
    R transmission = rmax(R(0), R(1) - fresnel_dielectric(rabs(dot(wir, mn)), R(1) / ior));
    R btdf = transmission * rabs(dot(wir, mn)) * jacobian * visibility * distribution /
        rmax(rabs(cos_theta(wo)) * rabs(cos_theta(wi)), K<R>::DENOM_TOLERANCE);
    return btdf * (m.transmission_depth == R(0) ? m.transmission_color : V3<R>(R(1)));
}

// specular_btdf.glsl:28-54
template <typename R>
inline V3<R> specular_btdf_evaluate(const Material<R>& m, V3<R> wi, V3<R> wo, R& density) {
    if (is_same_hemisphere(wo, wi)) {
        density = R(1);
        return V3<R>();
    }
    R ior = cos_theta(wi) > R(0) ? R(1) / m.specular_ior : m.specular_ior;
    if (rabs(ior - R(1)) < K<R>::IOR_EPSILON) {
        V3<R> tint = m.transmission_depth == R(0) ? m.transmission_color : V3<R>(R(1));
        density = R(1) / K<R>::PDF_EPSILON;
        return tint * density / rmax(K<R>::DENOM_TOLERANCE, rabs(cos_theta(wo)));
    }
    LocalRotation<R> rot = local_frame_rotation(R(2) * K<R>::PI * m.specular_rotation);
    V3<R> wir = rotate_local(wi, rot), wor = rotate_local(wo, rot);
    V3<R> mn = -wor - ior * wir;
    if (length_squared(mn) == R(0)) return V3<R>();
    mn = mn / rmax(length(mn), R(1e-8));  // safe_normalize, math.glsl:67
    if (cos_theta(mn) <= R(0)) mn = mn * R(-1);
    V2<R> alpha = microfacet_roughness_to_alpha(m.specular_roughness, m.specular_roughness_anisotropy);
    return specular_btdf(m, mn, alpha, wi, wo, wir, wor, ior, density);
}

// specular_btdf.glsl:56-67
template <typename R>
inline bool refraction_direction(V3<R> normal, R ior, V3<R> wi, V3<R>& direction) {
    R cos_in = dot(wi, normal);
    R s2in = rmax(R(0), R(1) - pow2(cos_in));
    R s2tr = pow2(ior) * s2in;
    if (s2tr >= R(1)) return false;
    R cos_tr = Fn<R>::sqrt(R(1) - s2tr);
    direction = ior * (-wi) + (ior * cos_in - cos_tr) * normal;
    return true;
}

// specular_btdf.glsl:69-103
template <typename R>
inline V3<R> specular_btdf_sample(const Material<R>& m, V3<R> wi, Generator& g, V3<R>& wo, R& density,
                                  bool& ok) {
    R ior = cos_theta(wi) > R(0) ? R(1) / m.specular_ior : m.specular_ior;
    if (rabs(ior - R(1)) < K<R>::IOR_EPSILON) {
        wo = -wi;
        ok = true;
        density = R(1) / K<R>::PDF_EPSILON;
        V3<R> tint = m.transmission_depth == R(0) ? m.transmission_color : V3<R>(R(1));
        return tint * density / rmax(K<R>::DENOM_TOLERANCE, rabs(cos_theta(wo)));
    }
    ok = false;
    LocalRotation<R> rot = local_frame_rotation(R(2) * K<R>::PI * m.specular_rotation);
    V3<R> wir = rotate_local(wi, rot);
    V2<R> alpha = microfacet_roughness_to_alpha(m.specular_roughness, m.specular_roughness_anisotropy);
    V3<R> mn;
    if (cos_theta(wir) > R(0)) {
        mn = microfacet_sample(alpha, wir, random_vec2<R>(g));
    } else {
        V3<R> reflected = V3<R>(wir.x, wir.y, -cos_theta(wir));
        mn = microfacet_sample(alpha, reflected, random_vec2<R>(g));
        mn.z = mn.z * R(-1);
    }
    V3<R> refracted;
    if (!refraction_direction(mn, ior, wir, refracted)) {
        density = K<R>::PDF_EPSILON;
        return V3<R>();
    }
    V3<R> wor = -(refracted / rmax(length(refracted), R(1e-8)));
    wo = inverse_rotate_local(wor, rot);
    ok = true;
    return specular_btdf(m, mn, alpha, wi, wo, wir, wor, ior, density);
}

// specular_btdf.glsl:105-119 (tests specular_ior, not the ratio)
template <typename R>
inline V3<R> specular_btdf_albedo(const Material<R>& m, V3<R> wi, Generator& g) {
    if (rabs(m.specular_ior - R(1)) < K<R>::IOR_EPSILON)
        return m.transmission_depth == R(0) ? m.transmission_color : V3<R>(R(1));
    V3<R> albedo;
    V3<R> wo;
    R density = R(0);
    bool ok;
    V3<R> btdf = specular_btdf_sample(m, wi, g, wo, density, ok);
    if (length(btdf) > K<R>::RADIANCE_EPSILON)
        albedo = albedo + btdf * rabs(cos_theta(wo)) / rmax(K<R>::DENOM_TOLERANCE, density);
    return albedo / R(1);
}

// ----------------------------------------------------------------- metal --
// metal_brdf.glsl:8-15 (F82-tint, Kutz 2021)
template <typename R>
inline V3<R> fresnel_metal(R c, V3<R> f0, V3<R> tint) {
    const R mu_bar = R(1) / R(7);
    const R denom = mu_bar * pow6(R(1) - mu_bar);
    return fresnel_schlick(f0, c) -
        c * pow6(R(1) - c) * (R(1) - tint) * fresnel_schlick(f0, mu_bar) / denom;
}

// metal_brdf.glsl:17-33
template <typename R>
inline V3<R> metal_brdf(const Material<R>& m, V2<R> alpha, V3<R> mn, V3<R> wi, V3<R> wo, V3<R> wir,
                        V3<R> wor, R& density) {
    R distribution = microfacet_distribution(alpha, mn);
    R visible_normals = distribution * microfacet_masking(alpha, wir) * rmax(R(0), dot(wir, mn)) /
        rmax(K<R>::DENOM_TOLERANCE, cos_theta(wir));
    R jacobian = R(1) / rmax(rabs(R(4) * dot(wir, mn)), K<R>::DENOM_TOLERANCE);
    density = rmax(K<R>::PDF_EPSILON, visible_normals * jacobian);
    V3<R> fresnel = fresnel_metal(rabs(dot(wir, mn)), m.base_color * m.base_weight,
                                  m.specular_color * m.specular_weight);
    R visibility = microfacet_visibility(alpha, wir, wor);
    return fresnel * distribution * visibility /
        rmax(R(4) * rabs(cos_theta(wo)) * rabs(cos_theta(wi)), K<R>::DENOM_TOLERANCE);
}

// metal_brdf.glsl:35-47
template <typename R>
inline V3<R> metal_brdf_evaluate(const Material<R>& m, V3<R> wi, V3<R> wo, R& density) {
    if (cos_theta(wi) < K<R>::DENOM_TOLERANCE || cos_theta(wo) < K<R>::DENOM_TOLERANCE) {
        density = K<R>::PDF_EPSILON;
        return V3<R>();
    }
    LocalRotation<R> rot = local_frame_rotation(R(2) * K<R>::PI * m.specular_rotation);
    V3<R> wir = rotate_local(wi, rot), wor = rotate_local(wo, rot);
    V2<R> alpha = microfacet_roughness_to_alpha(m.specular_roughness, m.specular_roughness_anisotropy);
    V3<R> mn = normalize(wor + wir);
    return metal_brdf(m, alpha, mn, wi, wo, wir, wor, density);
}

// metal_brdf.glsl:49-67
template <typename R>
inline V3<R> metal_brdf_sample(const Material<R>& m, V3<R> wi, Generator& g, V3<R>& wo, R& density,
                               bool& ok) {
    ok = false;
    if (cos_theta(wi) < K<R>::DENOM_TOLERANCE) {
        density = K<R>::PDF_EPSILON;
        return V3<R>();
    }
    V2<R> alpha = microfacet_roughness_to_alpha(m.specular_roughness, m.specular_roughness_anisotropy);
    LocalRotation<R> rot = local_frame_rotation(R(2) * K<R>::PI * m.specular_rotation);
    V3<R> wir = rotate_local(wi, rot);
    V3<R> mn = microfacet_sample(alpha, wir, random_vec2<R>(g));
    V3<R> wor = -reflect(wir, mn);
    if (!is_same_hemisphere(wir, wor)) return V3<R>();
    wo = inverse_rotate_local(wor, rot);
    ok = true;
    return metal_brdf(m, alpha, mn, wi, wo, wir, wor, density);
}

// metal_brdf.glsl:69-82
template <typename R>
inline V3<R> metal_brdf_albedo(const Material<R>& m, V3<R> wi, Generator& g) {
    if (cos_theta(wi) < K<R>::DENOM_TOLERANCE) return V3<R>();
    V3<R> albedo;
    V3<R> wo;
    R density = R(0);
    bool ok;
    V3<R> brdf = metal_brdf_sample(m, wi, g, wo, density, ok);
    if (length(brdf) > K<R>::RADIANCE_EPSILON)
        albedo = albedo + brdf * rabs(cos_theta(wo)) / rmax(K<R>::PDF_EPSILON, density);
    return albedo / R(1);
}

// ------------------------------------------------------------------ coat --
// coat_brdf.glsl:8-22
template <typename R>
inline V3<R> coat_brdf(const Material<R>& m, V3<R> wi, V3<R> wo, V3<R> wir, V3<R> wor, R ior, V2<R> alpha,
                       R& density) {
    V3<R> mn = normalize(wor + wir);
    R d = dot(wir, mn);
    R distribution = microfacet_distribution(alpha, mn);
    R visible_normals = distribution * microfacet_masking(alpha, wir) * rmax(R(0), d) / cos_theta(wir);
    R jacobian = R(1) / rmax(rabs(R(4) * d), K<R>::DENOM_TOLERANCE);
    density = visible_normals * jacobian;
    R fresnel = fresnel_dielectric(rabs(d), ior);
    return V3<R>(fresnel) * distribution * microfacet_visibility(alpha, wir, wor) /
        rmax(R(4) * rabs(cos_theta(wo)) * rabs(cos_theta(wi)), K<R>::DENOM_TOLERANCE);
}

// coat_brdf.glsl:24-37
template <typename R>
inline V3<R> coat_brdf_evaluate(const Material<R>& m, V3<R> wi, V3<R> wo, R& density) {
    if (!is_same_hemisphere(wo, wi)) return V3<R>();
    R ior = cos_theta(wi) > R(0) ? m.coat_ior : R(1) / m.coat_ior;
    if (rabs(ior - R(1)) < K<R>::IOR_EPSILON) return V3<R>();
    V2<R> alpha = microfacet_roughness_to_alpha(m.coat_roughness, m.coat_roughness_anisotropy);
    LocalRotation<R> rot = local_frame_rotation(R(2) * K<R>::PI * m.coat_rotation);
    return coat_brdf(m, wi, wo, rotate_local(wi, rot), rotate_local(wo, rot), ior, alpha, density);
}

// coat_brdf.glsl:39-57
template <typename R>
inline V3<R> coat_brdf_sample(const Material<R>& m, V3<R> wi, Generator& g, V3<R>& wo, R& density,
                              bool& ok) {
    ok = false;
    R ior = cos_theta(wi) > R(0) ? m.coat_ior / R(1) : R(1) / m.coat_ior;
    if (rabs(ior - R(1)) < K<R>::IOR_EPSILON) return V3<R>();
    V2<R> alpha = microfacet_roughness_to_alpha(m.coat_roughness, m.coat_roughness_anisotropy);
    LocalRotation<R> rot = local_frame_rotation(R(2) * K<R>::PI * m.coat_rotation);
    V3<R> wir = rotate_local(wi, rot);
    V3<R> mn = microfacet_sample(alpha, wir, random_vec2<R>(g));
    if (cos_theta(wir) <= R(0)) return V3<R>();
    V3<R> wor = -reflect(wir, mn);
    if (!is_same_hemisphere(wir, wor)) return V3<R>();
    wo = inverse_rotate_local(wor, rot);
    ok = true;
    return coat_brdf(m, wi, wo, wir, wor, ior, alpha, density);
}

// coat_brdf.glsl:59-73
template <typename R>
inline V3<R> coat_brdf_albedo(const Material<R>& m, V3<R> wi, Generator& g) {
    R ior = m.coat_ior / R(1);
    if (rabs(ior - R(1)) < K<R>::IOR_EPSILON) return V3<R>();
    V3<R> albedo;
    V3<R> wo;
    R density = R(0);
    bool ok;
    V3<R> brdf = coat_brdf_sample(m, wi, g, wo, density, ok);
    if (length(brdf) > K<R>::RADIANCE_EPSILON)
        albedo = albedo + brdf * rabs(cos_theta(wo)) / rmax(K<R>::PDF_EPSILON, density);
    return albedo / R(1);
}

// ------------------------------------------------------------------ fuzz --
// fuzz_brdf.glsl:11-16 (Zeltner et al. 2022 fit)
template <typename R>
inline R fuzz_albedo(R x, R y) {
    R s = y * (R(0.0206607) + R(1.58491) * y) / (R(0.0379424) + y * (R(1.32227) + y));
    R m = y * (R(-0.193854) + y * (R(-1.14885) + y * (R(1.7932) - R(0.95943) * y * y))) / (R(0.046391) + y);
    R o = y * (R(0.000654023) + (R(-0.0207818) + R(0.119681) * y) * y) / (R(1.26264) + y * (R(-1.92021) + y));
    return Fn<R>::exp(R(-0.5) * pow2((x - m) / s)) / (s * Fn<R>::sqrt(R(2) * K<R>::PI)) + o;
}

// fuzz_brdf.glsl:18-23
template <typename R>
inline V2<R> ltc_inverse_coeffs(R x, R y) {
    R a = (R(2.58126) * x + R(0.813703) * y) * y / (R(1) + R(0.310327) * x * x + R(2.60994) * x * y);
    R b = Fn<R>::sqrt(R(1) - x) * (y - R(1)) * y * y * y /
        (R(0.0000254053) + R(1.71228) * x - R(1.71506) * x * y + R(1.34174) * y * y);
    return {a, b};
}

// fuzz_brdf.glsl:25-32: columns tangent, bitangent, (0,0,1).
template <typename R>
inline void orthonormal_basis_ltc(V3<R> n, V3<R>& t, V3<R>& b) {
    R ls = n.x * n.x + n.y * n.y;
    t = ls > R(0) ? V3<R>(n.x, n.y, R(0)) * (R(1) / Fn<R>::sqrt(ls)) : V3<R>(R(1), R(0), R(0));
    b = V3<R>(-t.y, t.x, R(0));
}

// fuzz_brdf.glsl:34-42
template <typename R>
inline V3<R> fuzz_density_and_brdf(const Material<R>& m, V3<R> direction, V3<R> wi, R a_inverse, R roughness,
                                   R& density) {
    R jacobian = pow2(a_inverse / length_squared(direction));
    density = rmax(cos_theta(direction), R(0)) * K<R>::INVERSE_PI * jacobian;
    R albedo = fuzz_albedo(cos_theta(wi), roughness);
    return m.fuzz_color * albedo * K<R>::INVERSE_PI * jacobian;
}

// fuzz_brdf.glsl:44-53
template <typename R>
inline V3<R> fuzz_brdf_evaluate(const Material<R>& m, V3<R> wi, V3<R> wo, R& density) {
    if (cos_theta(wi) < K<R>::DENOM_TOLERANCE || cos_theta(wo) < K<R>::DENOM_TOLERANCE) return V3<R>();
    R roughness = rclamp(m.fuzz_roughness, R(0.01), R(1));
    V3<R> t, b;
    orthonormal_basis_ltc(wi, t, b);
    V3<R> n(R(0), R(0), R(1));
    V3<R> w(dot(t, wo), dot(b, wo), dot(n, wo));  // transpose(basis) * wo
    V2<R> ic = ltc_inverse_coeffs(cos_theta(wi), roughness);
    V3<R> direction(ic.x * w.x + ic.y * w.z, ic.x * w.y, w.z);
    return fuzz_density_and_brdf(m, direction, wi, ic.x, roughness, density);
}

// fuzz_brdf.glsl:55-69
template <typename R>
inline V3<R> fuzz_brdf_sample(const Material<R>& m, V3<R> wi, Generator& g, V3<R>& wo, R& density,
                              bool& ok) {
    ok = false;
    if (cos_theta(wi) < K<R>::DENOM_TOLERANCE) return V3<R>();
    R roughness = rclamp(m.fuzz_roughness, R(0.01), R(1));
    V3<R> direction = cosine_hemisphere_sample(random_vec2<R>(g));
    V2<R> ic = ltc_inverse_coeffs(cos_theta(wi), roughness);
    V3<R> w(direction.x / ic.x - direction.z * ic.y / ic.x, direction.y / ic.x, direction.z);
    V3<R> t, b;
    orthonormal_basis_ltc(wi, t, b);
    V3<R> wn = normalize(w);
    V3<R> n(R(0), R(0), R(1));
    wo = t * wn.x + b * wn.y + n * wn.z;
    ok = true;
    return fuzz_density_and_brdf(m, direction, wi, ic.x, roughness, density);
}

// fuzz_brdf.glsl:71-73 (unclamped roughness, no cosine test)
template <typename R>
inline V3<R> fuzz_brdf_albedo(const Material<R>& m, V3<R> wi) {
    return V3<R>(fuzz_albedo(cos_theta(wi), m.fuzz_roughness));
}

// ------------------------------------------------------------------ bsdf --
enum { FUZZ = 0, COAT = 1, METAL = 2, SPECULAR_BRDF = 3, SPECULAR_BTDF = 4, DIFFUSE = 5, LOBE_COUNT = 6 };

template <typename R>
struct Lobes {
    V3<R> weights[LOBE_COUNT];
    V3<R> albedos[LOBE_COUNT];
    R probs[LOBE_COUNT];
};

// bsdf.glsl:33-102
template <typename R>
inline void lobe_weights(const Material<R>& m, V3<R> wi, Generator& g, Lobes<R>& l) {
    bool fully_metallic = m.base_metalness == R(1);
    bool fully_transmissive = m.transmission_weight == R(1);
    l.albedos[FUZZ] = m.fuzz_weight > R(0) ? fuzz_brdf_albedo(m, wi) : V3<R>();
    l.albedos[COAT] = m.coat_weight > R(0) ? coat_brdf_albedo(m, wi, g) : V3<R>();
    l.albedos[METAL] = m.base_metalness > R(0) ? metal_brdf_albedo(m, wi, g) : V3<R>();
    l.albedos[SPECULAR_BRDF] = !fully_metallic ? specular_brdf_albedo(m, wi, g) : V3<R>();
    l.albedos[SPECULAR_BTDF] =
        !fully_metallic && m.transmission_weight > R(0) ? specular_btdf_albedo(m, wi, g) : V3<R>();
    l.albedos[DIFFUSE] = !fully_metallic && !fully_transmissive ? diffuse_brdf_albedo(m, wi) : V3<R>();

    l.weights[FUZZ] = V3<R>(m.fuzz_weight);
    V3<R> coated_base_weight = mixv(V3<R>(R(1)), R(1) - l.albedos[FUZZ], m.fuzz_weight);
    l.weights[COAT] = coated_base_weight * m.coat_weight;

    V3<R> base_darkening(R(1));
    if (m.coat_weight > R(0) && m.coat_darkening > R(0)) {
        R fresnel_weight = saturate(m.specular_weight * fresnel_at_normal_incident(specular_ior(m)));
        R dielectric_roughness = mixs(R(1), m.specular_roughness, fresnel_weight);
        R base_roughness = mixs(dielectric_roughness, m.specular_roughness, m.base_metalness);
        R average_fresnel = R(1) - (R(1) - average_dielectric_fresnel(m.coat_ior)) / pow2(m.coat_ior);
        R fresnel = fresnel_dielectric(rabs(cos_theta(wi)), m.coat_ior);
        R drc = mixs(fresnel, average_fresnel, base_roughness);
        V3<R> dielectric_base_albedo = mixv(l.albedos[DIFFUSE], l.albedos[SPECULAR_BTDF], m.transmission_weight);
        V3<R> base_albedo = mixv(dielectric_base_albedo, l.albedos[METAL], m.base_metalness);
        V3<R> darkening_factor = (R(1) - drc) / (R(1) - base_albedo * drc);
        base_darkening = mixv(V3<R>(R(1)), darkening_factor, m.coat_weight * m.coat_darkening);
    }
    V3<R> base_weight = coated_base_weight *
        mixv(V3<R>(R(1)), base_darkening * m.coat_color * (R(1) - l.albedos[COAT]), m.coat_weight);
    l.weights[METAL] = base_weight * m.base_metalness;
    V3<R> dielectric_base_weight = base_weight * V3<R>(R(1) - m.base_metalness);
    l.weights[SPECULAR_BRDF] = dielectric_base_weight;
    l.weights[SPECULAR_BTDF] = dielectric_base_weight * m.transmission_weight;
    V3<R> opaque = dielectric_base_weight * (R(1) - m.transmission_weight);
    l.weights[DIFFUSE] = opaque * (R(1) - l.albedos[SPECULAR_BRDF]);
}

// bsdf.glsl:104-113
template <typename R>
inline void lobe_probabilities(Lobes<R>& l) {
    R total = R(0);
    for (int i = 0; i < LOBE_COUNT; i++) {
        l.probs[i] = length(l.weights[i] * l.albedos[i]);
        total = total + l.probs[i];
    }
    total = rmax(K<R>::DENOM_TOLERANCE, total);
    for (int i = 0; i < LOBE_COUNT; i++) l.probs[i] = l.probs[i] / total;
}

// bsdf.glsl:115-120
template <typename R>
inline Lobes<R> bsdf_prepare(const Material<R>& m, V3<R> wi, Generator& g) {
    Lobes<R> l;
    lobe_weights(m, wi, g, l);
    lobe_probabilities(l);
    return l;
}

// bsdf.glsl:122-147 (evaluation order fuzz, coat, metal, specular, diffuse, btdf)
template <typename R>
inline V3<R> bsdf_evaluate_lobes(const Material<R>& m, const Lobes<R>& l, V3<R> wi, V3<R> wo, int skip,
                                 R densities[LOBE_COUNT]) {
    V3<R> bsdf;
    if (skip != FUZZ && l.probs[FUZZ] > R(0))
        bsdf = bsdf + l.weights[FUZZ] * fuzz_brdf_evaluate(m, wi, wo, densities[FUZZ]);
    if (skip != COAT && l.probs[COAT] > R(0))
        bsdf = bsdf + l.weights[COAT] * coat_brdf_evaluate(m, wi, wo, densities[COAT]);
    if (skip != METAL && l.probs[METAL] > R(0))
        bsdf = bsdf + l.weights[METAL] * metal_brdf_evaluate(m, wi, wo, densities[METAL]);
    if (skip != SPECULAR_BRDF && l.probs[SPECULAR_BRDF] > R(0))
        bsdf = bsdf + l.weights[SPECULAR_BRDF] * specular_brdf_evaluate(m, wi, wo, densities[SPECULAR_BRDF]);
    if (skip != DIFFUSE && l.probs[DIFFUSE] > R(0))
        bsdf = bsdf + l.weights[DIFFUSE] * diffuse_brdf_evaluate(m, wi, wo, densities[DIFFUSE]);
    if (skip != SPECULAR_BTDF && l.probs[SPECULAR_BTDF] > R(0))
        bsdf = bsdf + l.weights[SPECULAR_BTDF] * specular_btdf_evaluate(m, wi, wo, densities[SPECULAR_BTDF]);
    return bsdf;
}

// bsdf.glsl:149-154
template <typename R>
inline R bsdf_total_pdf(const Lobes<R>& l, const R densities[LOBE_COUNT]) {
    R density = R(0);
    for (int i = 0; i < LOBE_COUNT; i++) density = density + l.probs[i] * densities[i];
    return density;
}

// bsdf.glsl:156-161
template <typename R>
inline V3<R> bsdf_evaluate(const Material<R>& m, const Lobes<R>& l, V3<R> wi, V3<R> wo, R& density) {
    R densities[LOBE_COUNT] = {R(0), R(0), R(0), R(0), R(0), R(0)};
    V3<R> bsdf = bsdf_evaluate_lobes(m, l, wi, wo, -1, densities);
    density = bsdf_total_pdf(l, densities);
    return bsdf;
}

// bsdf.glsl:163-203.  Returns false when no direction was produced.
template <typename R>
inline bool bsdf_sample(const Material<R>& m, const Lobes<R>& l, V3<R> wi, Generator& g, V3<R>& wo,
                        R& density, int& lobe, V3<R>& value) {
    R random = random_float<R>(g);
    R accumulated = R(0);
    for (int i = 0; i < LOBE_COUNT; i++) {
        accumulated = accumulated + l.probs[i];
        if (random < accumulated) {
            R lobe_density = R(0);
            V3<R> lobe_bsdf;
            bool ok = false;
            switch (i) {
                case FUZZ: lobe_bsdf = fuzz_brdf_sample(m, wi, g, wo, lobe_density, ok); break;
                case COAT: lobe_bsdf = coat_brdf_sample(m, wi, g, wo, lobe_density, ok); break;
                case METAL: lobe_bsdf = metal_brdf_sample(m, wi, g, wo, lobe_density, ok); break;
                case SPECULAR_BRDF: lobe_bsdf = specular_brdf_sample(m, wi, g, wo, lobe_density, ok); break;
                case SPECULAR_BTDF: lobe_bsdf = specular_btdf_sample(m, wi, g, wo, lobe_density, ok); break;
                default: lobe_bsdf = diffuse_brdf_sample(m, wi, g, wo, lobe_density, ok); break;
            }
            lobe = i;
            if (!ok) {
                density = R(1);
                value = V3<R>();
                return false;
            }
            R densities[LOBE_COUNT] = {R(0), R(0), R(0), R(0), R(0), R(0)};
            V3<R> bsdf = bsdf_evaluate_lobes(m, l, wi, wo, i, densities);
            densities[i] = lobe_density;
            density = bsdf_total_pdf(l, densities);
            value = bsdf + l.weights[i] * lobe_bsdf;
            return true;
        }
    }
    lobe = -1;
    density = R(1);
    value = V3<R>();
    return false;
}

}  // namespace oracle
