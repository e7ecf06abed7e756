// sol_math.hpp -- the subset of glam 0.20 the sol ray/scene path uses
// (Vec2/3/4, column-major Mat4, look_at_rh, perspective_rh, inverse,
// transpose, from_scale / from_rotation / from_translation), fp32.
#pragma once

#include <cmath>
#include <cstdint>

namespace sol {

struct Vec2 {
    float x = 0, y = 0;
};
struct Vec3 {
    float x = 0, y = 0, z = 0;
    Vec3 operator+(const Vec3 &o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec3 operator-(const Vec3 &o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vec3 operator*(float s) const { return {x * s, y * s, z * s}; }
    Vec3 operator-() const { return {-x, -y, -z}; }
    static Vec3 splat(float v) { return {v, v, v}; }
};
struct Vec4 {
    float x = 0, y = 0, z = 0, w = 0;
    Vec3 xyz() const { return {x, y, z}; }
};
struct UVec3 {
    uint32_t x = 0, y = 0, z = 0;
};

inline float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline Vec3 normalize(Vec3 a)
{
    const float inv = 1.0f / std::sqrt(dot(a, a));
    return a * inv;
}

// column-major: c[col][row], as glam stores a Mat4
struct Mat4 {
    float c[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};

    static Mat4 identity() { return Mat4{}; }
    static Mat4 zero()
    {
        Mat4 m;
        for (auto &col : m.c)
            for (float &v : col) v = 0.0f;
        return m;
    }
    static Mat4 from_cols_array(const float a[16])
    {
        Mat4 m;
        for (int i = 0; i < 16; ++i) m.c[i / 4][i % 4] = a[i];
        return m;
    }
    static Mat4 from_scale(Vec3 s)
    {
        Mat4 m;
        m.c[0][0] = s.x;
        m.c[1][1] = s.y;
        m.c[2][2] = s.z;
        return m;
    }
    static Mat4 from_translation(Vec3 t)
    {
        Mat4 m;
        m.c[3][0] = t.x;
        m.c[3][1] = t.y;
        m.c[3][2] = t.z;
        return m;
    }
    static Mat4 from_rotation_x(float a)
    {
        Mat4 m;
        const float s = std::sin(a), co = std::cos(a);
        m.c[1][1] = co;
        m.c[1][2] = s;
        m.c[2][1] = -s;
        m.c[2][2] = co;
        return m;
    }
    static Mat4 from_rotation_y(float a)
    {
        Mat4 m;
        const float s = std::sin(a), co = std::cos(a);
        m.c[0][0] = co;
        m.c[0][2] = -s;
        m.c[2][0] = s;
        m.c[2][2] = co;
        return m;
    }
    // unit quaternion (x, y, z, w) rotation
    static Mat4 from_quat(float x, float y, float z, float w)
    {
        Mat4 m;
        const float x2 = x + x, y2 = y + y, z2 = z + z;
        const float xx = x * x2, xy = x * y2, xz = x * z2, yy = y * y2, yz = y * z2, zz = z * z2;
        const float wx = w * x2, wy = w * y2, wz = w * z2;
        m.c[0][0] = 1.0f - (yy + zz);
        m.c[0][1] = xy + wz;
        m.c[0][2] = xz - wy;
        m.c[1][0] = xy - wz;
        m.c[1][1] = 1.0f - (xx + zz);
        m.c[1][2] = yz + wx;
        m.c[2][0] = xz + wy;
        m.c[2][1] = yz - wx;
        m.c[2][2] = 1.0f - (xx + yy);
        return m;
    }

    Mat4 operator*(const Mat4 &b) const
    {
        Mat4 r = zero();
        for (int col = 0; col < 4; ++col)
            for (int row = 0; row < 4; ++row) {
                float s = 0.0f;
                for (int k = 0; k < 4; ++k) s += c[k][row] * b.c[col][k];
                r.c[col][row] = s;
            }
        return r;
    }
    Vec4 operator*(const Vec4 &v) const
    {
        const float in[4] = {v.x, v.y, v.z, v.w};
        float o[4];
        for (int row = 0; row < 4; ++row)
            o[row] = c[0][row] * in[0] + c[1][row] * in[1] + c[2][row] * in[2] + c[3][row] * in[3];
        return {o[0], o[1], o[2], o[3]};
    }
    Mat4 transpose() const
    {
        Mat4 r;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) r.c[i][j] = c[j][i];
        return r;
    }
    // general inverse by cofactors (double accumulation, fp32 result)
    Mat4 inverse() const
    {
        double m[16], inv[16];
        for (int i = 0; i < 16; ++i) m[i] = c[i / 4][i % 4];
        inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
                 m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
        inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
                 m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
        inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
                 m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
        inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
                  m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
        inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
                 m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
        inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
                 m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
        inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
                 m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
        inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
                  m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
        inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
                 m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
        inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
                 m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
        inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
                  m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
        inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
                  m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
        inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
                 m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
        inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
                 m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
        inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
                  m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
        inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
                  m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
        const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
        Mat4 r;
        for (int i = 0; i < 16; ++i) r.c[i / 4][i % 4] = (float)(inv[i] / det);
        return r;
    }
    const float *data() const { return &c[0][0]; }

    // glam Mat4::look_at_rh (src/scene/camera.rs:98-100)
    static Mat4 look_at_rh(Vec3 eye, Vec3 center, Vec3 up)
    {
        const Vec3 f = normalize(center - eye);
        const Vec3 s = normalize(cross(f, up));
        const Vec3 u = cross(s, f);
        Mat4 m;
        m.c[0][0] = s.x;
        m.c[1][0] = s.y;
        m.c[2][0] = s.z;
        m.c[0][1] = u.x;
        m.c[1][1] = u.y;
        m.c[2][1] = u.z;
        m.c[0][2] = -f.x;
        m.c[1][2] = -f.y;
        m.c[2][2] = -f.z;
        m.c[3][0] = -dot(s, eye);
        m.c[3][1] = -dot(u, eye);
        m.c[3][2] = dot(f, eye);
        return m;
    }
    // glam Mat4::perspective_rh, depth in [0, 1] (src/scene/camera.rs:102-106)
    static Mat4 perspective_rh(float fov_y_rad, float aspect, float z_near, float z_far)
    {
        const float h = 1.0f / std::tan(0.5f * fov_y_rad);
        const float w = h / aspect;
        const float r = z_far / (z_near - z_far);
        Mat4 m = zero();
        m.c[0][0] = w;
        m.c[1][1] = h;
        m.c[2][2] = r;
        m.c[2][3] = -1.0f;
        m.c[3][2] = r * z_near;
        return m;
    }
};

}  // namespace sol
