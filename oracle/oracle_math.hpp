// ORACLE (test infrastructure only; never linked into the product).
//
// Scalar vector math of the GLSL restatement, templated on the real type so
// that the same integrator instantiates in float (the parity oracle, the CPU
// baseline) and in double (the FP64 convergence oracle).
//
// Evaluation order is fixed and written out (left to right), because GLSL
// leaves it to the implementation and parity with the GPU needs one order:
//   dot(a, b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
//   mat * vec   = ((c0*v.x + c1*v.y) + c2*v.z) + c3*v.w
//   normalize   = v * (1 / length(v))
//   mix(x,y,a)  = x*(1-a) + y*a          (GLSL spec definition)
//   max(a,b)    = a < b ? b : a          (NaN in a propagates)
// Transcendentals: float uses the host/device-identical functions of
// tt_math.h (Cephes methods, see that header), double uses libm.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#include "../tinytrace_amd/csrc/tt_math.h"

namespace oracle {

template <typename R>
struct Fn;

template <>
struct Fn<float> {
    static float sqrt(float x) { return ttm::sqrt(x); }
    static float sin(float x) { return ttm::sin(x); }
    static float cos(float x) { return ttm::cos(x); }
    static float acos(float x) { return ttm::acos(x); }
    static float log(float x) { return ttm::log(x); }
    static float log2(float x) { return ttm::log2(x); }
    static float exp(float x) { return ttm::exp(x); }
    static float pow(float x, float y) { return ttm::pow(x, y); }
    static bool isinf(float x) { return ttm::isinf(x); }
    static constexpr float epsilon() { return 1.1920929e-7f; }
};

template <>
struct Fn<double> {
    static double sqrt(double x) { return std::sqrt(x); }
    static double sin(double x) { return std::sin(x); }
    static double cos(double x) { return std::cos(x); }
    static double acos(double x) { return std::acos(x); }
    static double log(double x) { return std::log(x); }
    static double log2(double x) { return std::log2(x); }
    static double exp(double x) { return std::exp(x); }
    static double pow(double x, double y) { return std::pow(x, y); }
    static bool isinf(double x) { return std::isinf(x); }
    static constexpr double epsilon() { return 2.220446049250313e-16; }
};

template <typename R>
inline R rmax(R a, R b) { return a < b ? b : a; }
template <typename R>
inline R rmin(R a, R b) { return b < a ? b : a; }
template <typename R>
inline R rclamp(R x, R lo, R hi) { return rmin(rmax(x, lo), hi); }
template <typename R>
inline R saturate(R x) { return rclamp(x, R(0), R(1)); }
template <typename R>
inline R rabs(R x) { return x < R(0) ? -x : (x == R(0) ? R(0) : x); }
template <typename R>
inline R rsign(R x) { return x > R(0) ? R(1) : (x < R(0) ? R(-1) : R(0)); }
template <typename R>
inline R pow2(R x) { return x * x; }
template <typename R>
inline R pow5(R x) { return x * pow2(x) * pow2(x); }
template <typename R>
inline R pow6(R x) { return x * pow5(x); }
template <typename R>
inline R mixs(R x, R y, R a) { return x * (R(1) - a) + y * a; }

template <typename R>
struct V2 {
    R x, y;
};

template <typename R>
struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    explicit V3(R s) : x(s), y(s), z(s) {}
    R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

template <typename R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename R> inline V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <typename R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename R> inline V3<R> operator/(R s, V3<R> a) { return {s / a.x, s / a.y, s / a.z}; }
template <typename R> inline V3<R> operator+(V3<R> a, R s) { return {a.x + s, a.y + s, a.z + s}; }
template <typename R> inline V3<R> operator-(R s, V3<R> a) { return {s - a.x, s - a.y, s - a.z}; }
template <typename R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }

template <typename R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R>
inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R> inline R length_squared(V3<R> a) { return dot(a, a); }
template <typename R> inline R length(V3<R> a) { return Fn<R>::sqrt(dot(a, a)); }
template <typename R> inline V3<R> normalize(V3<R> a) { return a * (R(1) / length(a)); }
template <typename R> inline V3<R> vmax(R s, V3<R> a) { return {rmax(s, a.x), rmax(s, a.y), rmax(s, a.z)}; }
template <typename R> inline V3<R> mixv(V3<R> x, V3<R> y, R a) { return x * (R(1) - a) + y * a; }
template <typename R> inline V3<R> reflect(V3<R> i, V3<R> n) { return i - (R(2) * dot(n, i)) * n; }
template <typename R> inline R cos_theta(V3<R> v) { return v.z; }
template <typename R> inline R cos_theta_squared(V3<R> v) { return pow2(v.z); }
template <typename R> inline R sin_theta_squared(V3<R> v) { return rmax(R(0), R(1) - cos_theta_squared(v)); }
template <typename R> inline R sin_theta(V3<R> v) { return Fn<R>::sqrt(sin_theta_squared(v)); }
template <typename R> inline R tan_theta_squared(V3<R> v) { return sin_theta_squared(v) / cos_theta_squared(v); }
template <typename R>
inline R sin_phi(V3<R> v) {
    R s = sin_theta(v);
    return rabs(s) < R(0.0001) ? R(0) : rclamp(v.y / s, R(-1), R(1));
}
template <typename R>
inline R cos_phi(V3<R> v) {
    R s = sin_theta(v);
    return rabs(s) < R(0.0001) ? R(1) : rclamp(v.x / s, R(-1), R(1));
}
template <typename R> inline bool is_same_hemisphere(V3<R> a, V3<R> b) { return a.z * b.z >= R(0); }
template <typename R> inline R luminance(V3<R> c) { return dot(c, V3<R>(R(0.2126), R(0.7152), R(0.0722))); }

// Column-major 4x4 matrix (glam / GLSL convention); c[col][row].
template <typename R>
struct M4 {
    R c[4][4];
};

template <typename R>
inline M4<R> load_m4(const float* f) {
    M4<R> m;
    for (int col = 0; col < 4; col++)
        for (int row = 0; row < 4; row++) m.c[col][row] = R(f[col * 4 + row]);
    return m;
}

// mat4 * vec4 restricted to the xyz rows.
template <typename R>
inline V3<R> mul_point(const M4<R>& m, V3<R> p, R w) {
    V3<R> r;
    for (int i = 0; i < 3; i++) r[i] = m.c[0][i] * p.x + m.c[1][i] * p.y + m.c[2][i] * p.z + m.c[3][i] * w;
    return r;
}

// mat3(m) * v: upper-left 3x3.
template <typename R>
inline V3<R> mul_dir(const M4<R>& m, V3<R> v) {
    V3<R> r;
    for (int i = 0; i < 3; i++) r[i] = m.c[0][i] * v.x + m.c[1][i] * v.y + m.c[2][i] * v.z;
    return r;
}

// IEEE half -> float (exact).
inline float half_to_float(uint16_t h) {
    uint32_t sign = uint32_t(h & 0x8000u) << 16;
    uint32_t exponent = (h >> 10) & 0x1fu;
    uint32_t mantissa = h & 0x3ffu;
    uint32_t bits;
    if (exponent == 0) {
        if (mantissa == 0) {
            bits = sign;
        } else {
            int e = -1;
            do {
                e++;
                mantissa <<= 1;
            } while ((mantissa & 0x400u) == 0);
            bits = sign | (uint32_t(127 - 15 - e) << 23) | ((mantissa & 0x3ffu) << 13);
        }
    } else if (exponent == 31) {
        bits = sign | 0x7f800000u | (mantissa << 13);
    } else {
        bits = sign | ((exponent + 127 - 15) << 23) | (mantissa << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

}  // namespace oracle
