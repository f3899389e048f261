// torpedo host math, restated for the MI355X build (header-only, C++17).
//
// Reference: torpedo/math/include/torpedo/math/{common,vec3,vec4,mat4}.h.  Row-major 4x4 matrices,
// FMA-compensated arithmetic where the reference uses it:
//   dop            a*b - c*d with Kahan's error term (common.h:11-17)
//   two_sum        Knuth's TwoSum (common.h:23-29)
//   two_prod       FMA TwoProduct (common.h:35-40)
//   dot            compensated dot product, Graillat et al. "Algorithm 6" (vec3.h:239-252, vec4.h:248-266)
//   cross          components by dop (vec3.h:230-236)
//   normalize      v * (1 / sqrt(dot(v, v))) (vec3.h:254-262; operator/ multiplies by the reciprocal)
//   mul            matrix product with compensated dots (mat4.h:256-264)
// det/inv of mat4 are stubs in the reference (mat4.h:246-254); they are implemented here by
// cofactor expansion.  Translation units using this header must not contract a*b+c into FMAs
// beyond the explicit std::fma calls (build with -ffp-contract=off).
#pragma once

#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>

namespace tpd {

struct vec3 {
    float x{}, y{}, z{};
};
struct vec4 {
    float x{}, y{}, z{}, w{};
};

inline vec3 operator-(const vec3& a, const vec3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator+(const vec3& a, const vec3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator*(const vec3& a, float s) { return {a.x * s, a.y * s, a.z * s}; }

// Row-major 4x4: m[r][c] = data[r * 4 + c].
struct mat4 {
    std::array<float, 16> data{};

    mat4() = default;
    explicit mat4(float diag) { data = {diag, 0, 0, 0, 0, diag, 0, 0, 0, 0, diag, 0, 0, 0, 0, diag}; }
    explicit mat4(const float* d) {
        for (int i = 0; i < 16; ++i) data[i] = d[i];
    }
    mat4(float r0c0, float r0c1, float r0c2, float r0c3, float r1c0, float r1c1, float r1c2, float r1c3, float r2c0,
         float r2c1, float r2c2, float r2c3, float r3c0, float r3c1, float r3c2, float r3c3)
        : data{r0c0, r0c1, r0c2, r0c3, r1c0, r1c1, r1c2, r1c3, r2c0, r2c1, r2c2, r2c3, r3c0, r3c1, r3c2, r3c3} {}

    float& operator()(std::size_t r, std::size_t c) { return data[r * 4 + c]; }
    float operator()(std::size_t r, std::size_t c) const { return data[r * 4 + c]; }
    const float* data_ptr() const { return data.data(); }
    vec4 row(std::size_t r) const { return {data[r * 4], data[r * 4 + 1], data[r * 4 + 2], data[r * 4 + 3]}; }
    vec4 col(std::size_t c) const { return {data[c], data[4 + c], data[8 + c], data[12 + c]}; }
};

namespace math {

inline float dop(float a, float b, float c, float d) {
    const float cd = c * d;
    const float err = std::fma(-c, d, cd);
    const float r = std::fma(a, b, -cd);
    return r + err;
}

inline void two_sum(float a, float b, float& s, float& e) {
    const float sum = a + b;
    const float z = sum - a;
    s = sum;
    e = a - (sum - z) + (b - z);
}

inline void two_prod(float a, float b, float& p, float& e) {
    const float m = a * b;
    p = m;
    e = std::fma(a, b, -m);
}

inline float dot(const vec3& a, const vec3& b) {
    float xx, e0, yy, e1, xy, e3, zz, e2, d, e;
    two_prod(a.x, b.x, xx, e0);
    two_prod(a.y, b.y, yy, e1);
    two_sum(xx, yy, xy, e3);
    const float e4 = e0 + (e3 + e1);
    two_prod(a.z, b.z, zz, e2);
    two_sum(xy, zz, d, e);
    return d + (e4 + (e + e2));
}

inline float dot(const vec4& a, const vec4& b) {
    float xx, e0, yy, e1, xy, e3, zz, e2, xz, e5, ww, e7, d, e;
    two_prod(a.x, b.x, xx, e0);
    two_prod(a.y, b.y, yy, e1);
    two_sum(xx, yy, xy, e3);
    const float e4 = e0 + (e3 + e1);
    two_prod(a.z, b.z, zz, e2);
    two_sum(xy, zz, xz, e5);
    const float e6 = e4 + (e5 + e2);
    two_prod(a.w, b.w, ww, e7);
    two_sum(xz, ww, d, e);
    return d + (e6 + (e + e7));
}

inline vec3 cross(const vec3& a, const vec3& b) {
    return {dop(a.y, b.z, a.z, b.y), dop(a.z, b.x, a.x, b.z), dop(a.x, b.y, a.y, b.x)};
}

inline float norm(const vec3& v) { return std::sqrt(dot(v, v)); }

inline vec3 normalize(const vec3& v) {
    const float inv = 1.0f / norm(v);
    return v * inv;
}

inline mat4 mul(const mat4& a, const mat4& b) {
    mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r(i, j) = dot(a.row(i), b.col(j));
    return r;
}

inline mat4 transpose(const mat4& m) {
    mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r(i, j) = m(j, i);
    return r;
}

// Determinant by expansion over 2x2 minors of the first two rows.
inline float det(const mat4& m) {
    const float s0 = m(0, 0) * m(1, 1) - m(1, 0) * m(0, 1), s1 = m(0, 0) * m(1, 2) - m(1, 0) * m(0, 2);
    const float s2 = m(0, 0) * m(1, 3) - m(1, 0) * m(0, 3), s3 = m(0, 1) * m(1, 2) - m(1, 1) * m(0, 2);
    const float s4 = m(0, 1) * m(1, 3) - m(1, 1) * m(0, 3), s5 = m(0, 2) * m(1, 3) - m(1, 2) * m(0, 3);
    const float c5 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3), c4 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
    const float c3 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2), c2 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
    const float c1 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2), c0 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
    return s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
}

// Inverse by the adjugate; a singular matrix yields non-finite entries (IEEE division by zero).
inline mat4 inv(const mat4& m) {
    const float s0 = m(0, 0) * m(1, 1) - m(1, 0) * m(0, 1), s1 = m(0, 0) * m(1, 2) - m(1, 0) * m(0, 2);
    const float s2 = m(0, 0) * m(1, 3) - m(1, 0) * m(0, 3), s3 = m(0, 1) * m(1, 2) - m(1, 1) * m(0, 2);
    const float s4 = m(0, 1) * m(1, 3) - m(1, 1) * m(0, 3), s5 = m(0, 2) * m(1, 3) - m(1, 2) * m(0, 3);
    const float c5 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3), c4 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
    const float c3 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2), c2 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
    const float c1 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2), c0 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
    const float id = 1.0f / (s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0);
    mat4 r;
    r(0, 0) = (m(1, 1) * c5 - m(1, 2) * c4 + m(1, 3) * c3) * id;
    r(0, 1) = (-m(0, 1) * c5 + m(0, 2) * c4 - m(0, 3) * c3) * id;
    r(0, 2) = (m(3, 1) * s5 - m(3, 2) * s4 + m(3, 3) * s3) * id;
    r(0, 3) = (-m(2, 1) * s5 + m(2, 2) * s4 - m(2, 3) * s3) * id;
    r(1, 0) = (-m(1, 0) * c5 + m(1, 2) * c2 - m(1, 3) * c1) * id;
    r(1, 1) = (m(0, 0) * c5 - m(0, 2) * c2 + m(0, 3) * c1) * id;
    r(1, 2) = (-m(3, 0) * s5 + m(3, 2) * s2 - m(3, 3) * s1) * id;
    r(1, 3) = (m(2, 0) * s5 - m(2, 2) * s2 + m(2, 3) * s1) * id;
    r(2, 0) = (m(1, 0) * c4 - m(1, 1) * c2 + m(1, 3) * c0) * id;
    r(2, 1) = (-m(0, 0) * c4 + m(0, 1) * c2 - m(0, 3) * c0) * id;
    r(2, 2) = (m(3, 0) * s4 - m(3, 1) * s2 + m(3, 3) * s0) * id;
    r(2, 3) = (-m(2, 0) * s4 + m(2, 1) * s2 - m(2, 3) * s0) * id;
    r(3, 0) = (-m(1, 0) * c3 + m(1, 1) * c1 - m(1, 2) * c0) * id;
    r(3, 1) = (m(0, 0) * c3 - m(0, 1) * c1 + m(0, 2) * c0) * id;
    r(3, 2) = (-m(3, 0) * s3 + m(3, 1) * s1 - m(3, 2) * s0) * id;
    r(3, 3) = (m(2, 0) * s3 - m(2, 1) * s1 + m(2, 2) * s0) * id;
    return r;
}

}  // namespace math
}  // namespace tpd
