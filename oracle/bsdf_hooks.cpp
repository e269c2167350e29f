// ORACLE (test infrastructure only; never linked into the product).
//
// Entry points into single functions of the oracle's BSDF restatement
// (oracle_bsdf.hpp), so tests can pin each one against closed forms,
// quadrature and published special cases.  Every hook takes `precision`
// (0: the float instantiation the GPU parity tests use, 1: double) and
// returns doubles.  Materials are MaterialConstants in their 192-byte layout
// (src/asset/material.rs:203-231).  Lobe indices follow bsdf.glsl:15-21:
// fuzz 0, coat 1, metal 2, specular BRDF 3, specular BTDF 4, diffuse 5.
//
// This file also compiles on its own (one translation unit, BSDF only), which
// the tests use to build it against deliberately broken copies of
// oracle_bsdf.hpp and show that the pins catch the slip.
#include <cstdint>

#include "oracle_bsdf.hpp"

using namespace oracle;

namespace {

template <typename R>
V3<R> v3(const double* p) {
    return V3<R>(R(p[0]), R(p[1]), R(p[2]));
}

template <typename R>
void out3(V3<R> v, double* p) {
    p[0] = double(v.x);
    p[1] = double(v.y);
    p[2] = double(v.z);
}

template <typename R>
V3<R> lobe_evaluate(const Material<R>& m, int lobe, V3<R> wi, V3<R> wo, R& density) {
    switch (lobe) {
        case FUZZ: return fuzz_brdf_evaluate(m, wi, wo, density);
        case COAT: return coat_brdf_evaluate(m, wi, wo, density);
        case METAL: return metal_brdf_evaluate(m, wi, wo, density);
        case SPECULAR_BRDF: return specular_brdf_evaluate(m, wi, wo, density);
        case SPECULAR_BTDF: return specular_btdf_evaluate(m, wi, wo, density);
        default: return diffuse_brdf_evaluate(m, wi, wo, density);
    }
}

template <typename R>
V3<R> lobe_sample(const Material<R>& m, int lobe, V3<R> wi, Generator& g, V3<R>& wo, R& density, bool& ok) {
    switch (lobe) {
        case FUZZ: return fuzz_brdf_sample(m, wi, g, wo, density, ok);
        case COAT: return coat_brdf_sample(m, wi, g, wo, density, ok);
        case METAL: return metal_brdf_sample(m, wi, g, wo, density, ok);
        case SPECULAR_BRDF: return specular_brdf_sample(m, wi, g, wo, density, ok);
        case SPECULAR_BTDF: return specular_btdf_sample(m, wi, g, wo, density, ok);
        default: return diffuse_brdf_sample(m, wi, g, wo, density, ok);
    }
}

template <typename R>
V3<R> lobe_albedo(const Material<R>& m, int lobe, V3<R> wi, Generator& g) {
    switch (lobe) {
        case FUZZ: return fuzz_brdf_albedo(m, wi);
        case COAT: return coat_brdf_albedo(m, wi, g);
        case METAL: return metal_brdf_albedo(m, wi, g);
        case SPECULAR_BRDF: return specular_brdf_albedo(m, wi, g);
        case SPECULAR_BTDF: return specular_btdf_albedo(m, wi, g);
        default: return diffuse_brdf_albedo(m, wi);
    }
}

template <typename R>
void weights_impl(const float* mc, const double* wi, uint32_t rng, double* weights, double* albedos,
                  double* probs) {
    Material<R> m = load_material<R>(mc);
    Generator g{rng, 0};
    Lobes<R> l = bsdf_prepare(m, v3<R>(wi), g);
    for (int i = 0; i < LOBE_COUNT; i++) {
        out3(l.weights[i], weights + 3 * i);
        out3(l.albedos[i], albedos + 3 * i);
        probs[i] = double(l.probs[i]);
    }
}

}  // namespace

#define BOTH(expr_f, expr_d) (precision ? (expr_d) : (expr_f))

extern "C" {

void oracle_h_fresnel_metal(int precision, double c, const double* f0, const double* tint, double* out) {
    if (precision)
        out3(fresnel_metal(c, v3<double>(f0), v3<double>(tint)), out);
    else
        out3(fresnel_metal(float(c), v3<float>(f0), v3<float>(tint)), out);
}

double oracle_h_fresnel_dielectric(int precision, double c, double ior) {
    return BOTH(double(fresnel_dielectric(float(c), float(ior))), fresnel_dielectric(c, ior));
}

double oracle_h_average_dielectric_fresnel(int precision, double ior) {
    return BOTH(double(average_dielectric_fresnel(float(ior))), average_dielectric_fresnel(ior));
}

double oracle_h_specular_ior(int precision, const float* mc) {
    return BOTH(double(specular_ior(load_material<float>(mc))), specular_ior(load_material<double>(mc)));
}

double oracle_h_specular_ior_ratio(int precision, const float* mc) {
    return BOTH(double(specular_ior_ratio(load_material<float>(mc))), specular_ior_ratio(load_material<double>(mc)));
}

double oracle_h_specular_fresnel_with_coat(int precision, const float* mc, double c, double ior) {
    return BOTH(double(specular_fresnel_with_coat(load_material<float>(mc), float(c), float(ior)).x),
                specular_fresnel_with_coat(load_material<double>(mc), c, ior).x);
}

double oracle_h_oren_nayar(int precision, double c, double s2) {
    return BOTH(double(oren_nayar(float(c), float(s2))), oren_nayar(c, s2));
}

// n (wi, wo) pairs, 3 doubles each; out n x 3.
void oracle_h_energy_compensated_oren_nayar(int precision, const double* rho, double sigma, uint32_t n,
                                            const double* wi, const double* wo, double* out) {
    for (uint32_t i = 0; i < n; i++) {
        if (precision)
            out3(energy_compensated_oren_nayar(v3<double>(rho), sigma, v3<double>(wi + 3 * i), v3<double>(wo + 3 * i)),
                 out + 3 * i);
        else
            out3(energy_compensated_oren_nayar(v3<float>(rho), float(sigma), v3<float>(wi + 3 * i),
                                               v3<float>(wo + 3 * i)),
                 out + 3 * i);
    }
}

double oracle_h_fuzz_albedo(int precision, double x, double y) {
    return BOTH(double(fuzz_albedo(float(x), float(y))), fuzz_albedo(x, y));
}

void oracle_h_ltc_inverse_coeffs(int precision, double x, double y, double* out) {
    if (precision) {
        V2<double> c = ltc_inverse_coeffs(x, y);
        out[0] = c.x;
        out[1] = c.y;
    } else {
        V2<float> c = ltc_inverse_coeffs(float(x), float(y));
        out[0] = c.x;
        out[1] = c.y;
    }
}

// n (wi, wo) pairs; out n x 4 = bsdf rgb, density (0 where the lobe writes
// none).
void oracle_h_lobe_evaluate(int precision, const float* mc, int lobe, uint32_t n, const double* wi, const double* wo,
                            double* out) {
    for (uint32_t i = 0; i < n; i++) {
        double* o = out + 4 * i;
        if (precision) {
            double density = 0.0;
            out3(lobe_evaluate(load_material<double>(mc), lobe, v3<double>(wi + 3 * i), v3<double>(wo + 3 * i), density),
                 o);
            o[3] = density;
        } else {
            float density = 0.0f;
            out3(lobe_evaluate(load_material<float>(mc), lobe, v3<float>(wi + 3 * i), v3<float>(wo + 3 * i), density),
                 o);
            o[3] = density;
        }
    }
}

// Returns whether a direction was produced; wo_out = it, out = bsdf rgb, density.
int oracle_h_lobe_sample(int precision, const float* mc, int lobe, const double* wi, uint32_t rng, double* wo_out,
                         double* out) {
    Generator g{rng, 0};
    bool ok = false;
    if (precision) {
        double density = 0.0;
        V3<double> wo;
        out3(lobe_sample(load_material<double>(mc), lobe, v3<double>(wi), g, wo, density, ok), out);
        out3(wo, wo_out);
        out[3] = density;
    } else {
        float density = 0.0f;
        V3<float> wo;
        out3(lobe_sample(load_material<float>(mc), lobe, v3<float>(wi), g, wo, density, ok), out);
        out3(wo, wo_out);
        out[3] = density;
    }
    return ok ? 1 : 0;
}

// One albedo estimate per generator state in rng[n]; out n x 3.
void oracle_h_lobe_albedo(int precision, const float* mc, int lobe, const double* wi, uint32_t n, const uint32_t* rng,
                          double* out) {
    for (uint32_t i = 0; i < n; i++) {
        Generator g{rng[i], 0};
        if (precision)
            out3(lobe_albedo(load_material<double>(mc), lobe, v3<double>(wi), g), out + 3 * i);
        else
            out3(lobe_albedo(load_material<float>(mc), lobe, v3<float>(wi), g), out + 3 * i);
    }
}

// bsdf_prepare (bsdf.glsl:115-120): lobe weights, albedos (3 per lobe) and
// selection probabilities.
void oracle_h_lobe_weights(int precision, const float* mc, const double* wi, uint32_t rng, double* weights,
                           double* albedos, double* probs) {
    if (precision)
        weights_impl<double>(mc, wi, rng, weights, albedos, probs);
    else
        weights_impl<float>(mc, wi, rng, weights, albedos, probs);
}

}  // extern "C"
