// Device vector math for the gfx950 kernels.
//
// Evaluation order is fixed and written out because the integrator's results
// are compared against a CPU restatement operation for operation; all HIP
// sources are compiled with -ffp-contract=off so nothing is fused:
//   dot(a, b)  = (a.x*b.x + a.y*b.y) + a.z*b.z
//   mat * vec  = ((c0*v.x + c1*v.y) + c2*v.z) + c3*v.w
//   normalize  = v * (1 / length(v))
//   mix(x,y,a) = x*(1-a) + y*a
//   max(a, b)  = a < b ? b : a
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tt_math.h"

#define TT_DEV __device__ __forceinline__

namespace ttd {

struct f2 {
    float x, y;
};

struct f3 {
    float x, y, z;
    TT_DEV f3() : x(0.0f), y(0.0f), z(0.0f) {}
    TT_DEV f3(float a, float b, float c) : x(a), y(b), z(c) {}
    TT_DEV explicit f3(float s) : x(s), y(s), z(s) {}
};

TT_DEV f3 operator+(f3 a, f3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
TT_DEV f3 operator-(f3 a, f3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
TT_DEV f3 operator*(f3 a, f3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
TT_DEV f3 operator/(f3 a, f3 b) { return f3(a.x / b.x, a.y / b.y, a.z / b.z); }
TT_DEV f3 operator*(f3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
TT_DEV f3 operator*(float s, f3 a) { return f3(s * a.x, s * a.y, s * a.z); }
TT_DEV f3 operator/(f3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
TT_DEV f3 operator/(float s, f3 a) { return f3(s / a.x, s / a.y, s / a.z); }
TT_DEV f3 operator-(float s, f3 a) { return f3(s - a.x, s - a.y, s - a.z); }
TT_DEV f3 operator-(f3 a) { return f3(-a.x, -a.y, -a.z); }

TT_DEV float fmax_(float a, float b) { return a < b ? b : a; }
TT_DEV float fmin_(float a, float b) { return b < a ? b : a; }
TT_DEV float clamp_(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
TT_DEV float saturate(float x) { return clamp_(x, 0.0f, 1.0f); }
TT_DEV float abs_(float x) { return __builtin_fabsf(x); }
TT_DEV float sign_(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
TT_DEV float pow2(float x) { return x * x; }
TT_DEV float pow5(float x) { return x * pow2(x) * pow2(x); }
TT_DEV float pow6(float x) { return x * pow5(x); }
TT_DEV float mixs(float x, float y, float a) { return x * (1.0f - a) + y * a; }
TT_DEV float sqrt_(float x) { return __builtin_sqrtf(x); }

TT_DEV float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
TT_DEV f3 cross(f3 a, f3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
TT_DEV float length_squared(f3 a) { return dot(a, a); }
TT_DEV float length(f3 a) { return sqrt_(dot(a, a)); }
TT_DEV f3 normalize(f3 a) { return a * (1.0f / length(a)); }
TT_DEV f3 mixv(f3 x, f3 y, float a) { return x * (1.0f - a) + y * a; }
TT_DEV f3 reflect(f3 i, f3 n) { return i - (2.0f * dot(n, i)) * n; }
TT_DEV float cos_theta(f3 v) { return v.z; }
TT_DEV float cos_theta_squared(f3 v) { return pow2(v.z); }
TT_DEV float sin_theta_squared(f3 v) { return fmax_(0.0f, 1.0f - cos_theta_squared(v)); }
TT_DEV float sin_theta(f3 v) { return sqrt_(sin_theta_squared(v)); }
TT_DEV float tan_theta_squared(f3 v) { return sin_theta_squared(v) / cos_theta_squared(v); }
TT_DEV float sin_phi(f3 v) {
    float s = sin_theta(v);
    return abs_(s) < 0.0001f ? 0.0f : clamp_(v.y / s, -1.0f, 1.0f);
}
TT_DEV float cos_phi(f3 v) {
    float s = sin_theta(v);
    return abs_(s) < 0.0001f ? 1.0f : clamp_(v.x / s, -1.0f, 1.0f);
}
TT_DEV bool is_same_hemisphere(f3 a, f3 b) { return a.z * b.z >= 0.0f; }
TT_DEV bool isinf_(float x) { return ttm::isinf(x); }

TT_DEV f3 load3(const float* p) { return f3(p[0], p[1], p[2]); }

// Column-major 4x4 (glam/GLSL), c = 16 floats.
TT_DEV f3 mul_point(const float* m, f3 p, float w) {
    return f3(m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12] * w, m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13] * w,
              m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14] * w);
}
TT_DEV f3 mul_point1(const float* m, f3 p) {
    return f3(m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12], m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
              m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]);
}
TT_DEV f3 mul_dir(const float* m, f3 v) {
    return f3(m[0] * v.x + m[4] * v.y + m[8] * v.z, m[1] * v.x + m[5] * v.y + m[9] * v.z,
              m[2] * v.x + m[6] * v.y + m[10] * v.z);
}

constexpr float PI = 3.1415926535897932f;
constexpr float TAU = 6.283185307179586f;
constexpr float INVERSE_PI = 1.0f / 3.1415926535897932f;
constexpr float DENOM_TOLERANCE = 1.0e-10f;
constexpr float RADIANCE_EPSILON = 1.0e-12f;
constexpr float PDF_EPSILON = 1.0e-6f;
constexpr float IOR_EPSILON = 1.0e-5f;

}  // namespace ttd
