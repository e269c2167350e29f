// Host-side vector/matrix helpers with glam 0.27 semantics (column-major
// Mat4, left-to-right accumulation in Mat4 * Vec4 as glam's SSE2 path does).
// Used by the asset pipeline and the camera constants; the device code never
// sees these types.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace tt {

struct Vec3 {
    float x = 0, y = 0, z = 0;
    Vec3() = default;
    Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator/(Vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline Vec3 operator-(Vec3 a) { return {-a.x, -a.y, -a.z}; }
inline float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline float length(Vec3 a) { return std::sqrt(dot(a, a)); }
// glam Vec3::normalize: self * (1 / length)
inline Vec3 normalize(Vec3 a) { return a * (1.0f / length(a)); }

// Column-major 4x4 matrix: m[col*4 + row].
struct Mat4 {
    float m[16];
    static Mat4 identity() {
        Mat4 r;
        for (int i = 0; i < 16; i++) r.m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
        return r;
    }
    static Mat4 from_cols(const float* c) {
        Mat4 r;
        std::memcpy(r.m, c, sizeof(r.m));
        return r;
    }
    float at(int col, int row) const { return m[col * 4 + row]; }
    float& at(int col, int row) { return m[col * 4 + row]; }
};

// glam Mat4 * Mat4: column j = self * other.col(j), each as
// ((x*v.x + y*v.y) + z*v.z) + w*v.w.
inline Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++)
            r.at(j, i) = a.at(0, i) * b.at(j, 0) + a.at(1, i) * b.at(j, 1) + a.at(2, i) * b.at(j, 2) +
                a.at(3, i) * b.at(j, 3);
    return r;
}

inline Mat4 transpose(const Mat4& a) {
    Mat4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) r.at(i, j) = a.at(j, i);
    return r;
}

// Inverse by cofactor expansion (the GLM/glam scalar algorithm: 2x2
// sub-determinants of the lower rows, then the adjugate scaled by
// 1/determinant).
inline Mat4 inverse(const Mat4& a) {
    float m00 = a.at(0, 0), m01 = a.at(0, 1), m02 = a.at(0, 2), m03 = a.at(0, 3);
    float m10 = a.at(1, 0), m11 = a.at(1, 1), m12 = a.at(1, 2), m13 = a.at(1, 3);
    float m20 = a.at(2, 0), m21 = a.at(2, 1), m22 = a.at(2, 2), m23 = a.at(2, 3);
    float m30 = a.at(3, 0), m31 = a.at(3, 1), m32 = a.at(3, 2), m33 = a.at(3, 3);

    float coef00 = m22 * m33 - m32 * m23;
    float coef02 = m12 * m33 - m32 * m13;
    float coef03 = m12 * m23 - m22 * m13;
    float coef04 = m21 * m33 - m31 * m23;
    float coef06 = m11 * m33 - m31 * m13;
    float coef07 = m11 * m23 - m21 * m13;
    float coef08 = m21 * m32 - m31 * m22;
    float coef10 = m11 * m32 - m31 * m12;
    float coef11 = m11 * m22 - m21 * m12;
    float coef12 = m20 * m33 - m30 * m23;
    float coef14 = m10 * m33 - m30 * m13;
    float coef15 = m10 * m23 - m20 * m13;
    float coef16 = m20 * m32 - m30 * m22;
    float coef18 = m10 * m32 - m30 * m12;
    float coef19 = m10 * m22 - m20 * m12;
    float coef20 = m20 * m31 - m30 * m21;
    float coef22 = m10 * m31 - m30 * m11;
    float coef23 = m10 * m21 - m20 * m11;

    float fac0[4] = {coef00, coef00, coef02, coef03};
    float fac1[4] = {coef04, coef04, coef06, coef07};
    float fac2[4] = {coef08, coef08, coef10, coef11};
    float fac3[4] = {coef12, coef12, coef14, coef15};
    float fac4[4] = {coef16, coef16, coef18, coef19};
    float fac5[4] = {coef20, coef20, coef22, coef23};

    float vec0[4] = {m10, m00, m00, m00};
    float vec1[4] = {m11, m01, m01, m01};
    float vec2[4] = {m12, m02, m02, m02};
    float vec3[4] = {m13, m03, m03, m03};

    float inv0[4], inv1[4], inv2[4], inv3[4];
    for (int i = 0; i < 4; i++) {
        inv0[i] = vec1[i] * fac0[i] - vec2[i] * fac1[i] + vec3[i] * fac2[i];
        inv1[i] = vec0[i] * fac0[i] - vec2[i] * fac3[i] + vec3[i] * fac4[i];
        inv2[i] = vec0[i] * fac1[i] - vec1[i] * fac3[i] + vec3[i] * fac5[i];
        inv3[i] = vec0[i] * fac2[i] - vec1[i] * fac4[i] + vec2[i] * fac5[i];
    }
    const float sign_a[4] = {1, -1, 1, -1};
    const float sign_b[4] = {-1, 1, -1, 1};
    Mat4 r;
    for (int i = 0; i < 4; i++) {
        r.at(0, i) = inv0[i] * sign_a[i];
        r.at(1, i) = inv1[i] * sign_b[i];
        r.at(2, i) = inv2[i] * sign_a[i];
        r.at(3, i) = inv3[i] * sign_b[i];
    }
    float dot0 = m00 * r.at(0, 0) + m01 * r.at(1, 0) + m02 * r.at(2, 0) + m03 * r.at(3, 0);
    float rcp = 1.0f / dot0;
    for (int i = 0; i < 16; i++) r.m[i] = r.m[i] * rcp;
    return r;
}

inline Vec3 transform_point(const Mat4& a, Vec3 p) {
    return {a.at(0, 0) * p.x + a.at(1, 0) * p.y + a.at(2, 0) * p.z + a.at(3, 0),
            a.at(0, 1) * p.x + a.at(1, 1) * p.y + a.at(2, 1) * p.z + a.at(3, 1),
            a.at(0, 2) * p.x + a.at(1, 2) * p.y + a.at(2, 2) * p.z + a.at(3, 2)};
}

// IEEE float -> half with round-to-nearest-even (the `half` crate's
// f16::from_f32 behaviour).
inline uint16_t float_to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xff) return uint16_t(sign | 0x7c00u | (man ? (0x200u | (man >> 13)) : 0u));
    int e = int(exp) - 127 + 15;
    if (e >= 0x1f) return uint16_t(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return uint16_t(sign);
        man |= 0x800000u;
        int shift = 14 - e;
        uint32_t half_man = man >> shift;
        uint32_t rem = man & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (half_man & 1u))) half_man++;
        return uint16_t(sign | half_man);
    }
    uint32_t half_man = man >> 13;
    uint32_t rem = man & 0x1fffu;
    uint16_t h = uint16_t(sign | (uint32_t(e) << 10) | half_man);
    if (rem > 0x1000u || (rem == 0x1000u && (half_man & 1u))) h++;
    return h;
}

}  // namespace tt
