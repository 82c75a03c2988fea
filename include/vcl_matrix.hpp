// vcl_matrix.hpp -- host matrix math of the reference's Rotation / Camera classes
// (236088/vcl nvdiffrast/nvdiffrast/matrix.h:1-96, matrix.cu:3-239), header-only, no GPU
// dependency.  Types keep glm 0.9.9's memory layout so arrays written by glm-based host code
// can be passed through unchanged:
//   vcl::vec3 -- 3 floats (x, y, z);
//   vcl::quat -- 4 floats stored (x, y, z, w) (glm's default storage), constructed as
//                quat(w, x, y, z) like glm::quat, q[i] indexes the storage order;
//   vcl::mat4 -- 16 floats, column-major: m[c][r] is column c, row r (glm::mat4).
// Value-initialised glm::mat4() / glm::quat() are all zeros in the reference build (defaulted
// default constructors), which Rotation::backward relies on (matrix.cu:86); mat4::zero() and
// quat() here are the same.
#pragma once

#include <cmath>
#include <cstring>

namespace vcl {

struct vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
    vec3() = default;
    vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    float& operator[](int i) { return (&x)[i]; }
    float operator[](int i) const { return (&x)[i]; }
};
inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(vec3 a, vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// glm::normalize: v * inversesqrt(dot(v, v))
inline vec3 normalize(vec3 v) { return v * (1.f / std::sqrt(dot(v, v))); }

struct quat {
    float x = 0.f, y = 0.f, z = 0.f, w = 0.f;  // storage order of glm::quat
    quat() = default;
    quat(float w_, float x_, float y_, float z_) : x(x_), y(y_), z(z_), w(w_) {}
    float& operator[](int i) { return (&x)[i]; }
    float operator[](int i) const { return (&x)[i]; }
};
inline float dot(const quat& a, const quat& b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
// Hamilton product p * q (glm::operator*(quat, quat))
inline quat operator*(const quat& p, const quat& q) {
    return quat(p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z, p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y,
                p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z, p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x);
}
// glm::normalize(quat): q / length(q), the identity for a zero-length quaternion
inline quat normalize(const quat& q) {
    float len = std::sqrt(dot(q, q));
    if (len <= 0.f) return quat(1.f, 0.f, 0.f, 0.f);
    float il = 1.f / len;
    return quat(q.w * il, q.x * il, q.y * il, q.z * il);
}
// glm::angleAxis(angle, axis) for a unit axis
inline quat angle_axis(float angle, vec3 axis) {
    float s = std::sin(angle * 0.5f);
    return quat(std::cos(angle * 0.5f), axis.x * s, axis.y * s, axis.z * s);
}

struct mat4 {
    float m[4][4];  // m[column][row]
    mat4() { std::memset(m, 0, sizeof(m)); }
    static mat4 zero() { return mat4(); }
    static mat4 identity() {
        mat4 r;
        for (int i = 0; i < 4; i++) r.m[i][i] = 1.f;
        return r;
    }
    // glm::mat4(x0, y0, z0, w0, x1, ...): 16 values column by column
    static mat4 cols(const float* v16) {
        mat4 r;
        std::memcpy(r.m, v16, sizeof(r.m));
        return r;
    }
    float* data() { return &m[0][0]; }
    const float* data() const { return &m[0][0]; }
    bool operator==(const mat4& o) const { return std::memcmp(m, o.m, sizeof(m)) == 0; }
};
inline mat4 operator*(const mat4& a, const mat4& b) {
    mat4 r;
    for (int c = 0; c < 4; c++)
        for (int i = 0; i < 4; i++) {
            float s = 0.f;
            for (int k = 0; k < 4; k++) s += a.m[k][i] * b.m[c][k];
            r.m[c][i] = s;
        }
    return r;
}
inline void mul_vec4(const mat4& a, const float v[4], float out[4]) {
    for (int i = 0; i < 4; i++) out[i] = a.m[0][i] * v[0] + a.m[1][i] * v[1] + a.m[2][i] * v[2] + a.m[3][i] * v[3];
}
inline mat4 transpose(const mat4& a) {
    mat4 r;
    for (int c = 0; c < 4; c++)
        for (int i = 0; i < 4; i++) r.m[c][i] = a.m[i][c];
    return r;
}

// getRotation (matrix.cu:43-53): the rotation matrix of a unit quaternion, 4x4 homogeneous.
inline mat4 rotation_matrix(const quat& q) {
    float qw2 = q.w * q.w, qx2 = q.x * q.x, qy2 = q.y * q.y, qz2 = q.z * q.z;
    float qwx = q.w * q.x, qwy = q.w * q.y, qwz = q.w * q.z;
    float qxy = q.x * q.y, qyz = q.y * q.z, qzx = q.z * q.x;
    const float v[16] = {qw2 + qx2 - qy2 - qz2, 2.f * (qxy + qwz), 2.f * (qzx - qwy), 0.f,
                         2.f * (qxy - qwz), qw2 - qx2 + qy2 - qz2, 2.f * (qyz + qwx), 0.f,
                         2.f * (qzx + qwy), 2.f * (qyz - qwx), qw2 - qx2 - qy2 + qz2, 0.f,
                         0.f, 0.f, 0.f, 1.f};
    return mat4::cols(v);
}

// getView (matrix.cu:175-180): R(q) with the translation column R * (-eye, 1).
inline mat4 view_matrix(const quat& q, vec3 eye) {
    mat4 v = rotation_matrix(q);
    const float e[4] = {-eye.x, -eye.y, -eye.z, 1.f};
    float t[4];
    mul_vec4(v, e, t);
    for (int i = 0; i < 4; i++) v.m[3][i] = t[i];
    return v;
}

// Camera::setCam (matrix.cu:182-192): the look-at frame f = normalize(eye - center),
// s = normalize(up x f), u = normalize(f x s) as a quaternion (the rows s, u, f of the view
// rotation), with each component's sign taken from the matching off-diagonal pair.
inline quat look_at_quat(vec3 eye, vec3 center, vec3 up) {
    vec3 f = normalize(eye - center);
    vec3 s = normalize(cross(up, f));
    vec3 u = normalize(cross(f, s));
    quat q;
    q.w = .5f * std::sqrt(s.x + u.y + f.z + 1.f);
    q.x = (f.y > u.z ? .5f : -.5f) * std::sqrt(std::fabs(s.x - u.y - f.z + 1.f));
    q.y = (s.z > f.x ? .5f : -.5f) * std::sqrt(std::fabs(-s.x + u.y - f.z + 1.f));
    q.z = (u.x > s.y ? .5f : -.5f) * std::sqrt(std::fabs(-s.x - u.y + f.z + 1.f));
    return q;
}

// Camera::forward's projection (matrix.cu:211-219): sy = znear / size, sx = sy / aspect,
// OpenGL depth mapping of [znear, zfar] to [-1, 1].
inline mat4 projection_matrix(float size, float aspect, float znear, float zfar) {
    float sy = znear / size;
    float sx = sy / aspect;
    float f_n = zfar - znear;
    const float v[16] = {sx, 0.f, 0.f, 0.f, 0.f, sy, 0.f, 0.f, 0.f, 0.f, -(zfar + znear) / f_n, -1.f, 0.f, 0.f, -2.f * zfar * znear / f_n, 0.f};
    return mat4::cols(v);
}

// Camera::backward (matrix.cu:234-239): gradients of mat = P * V * R from the gradient G of mat.
// exact = false keeps the reference's products (G V^T P^T, R^T G P^T, R^T V^T G); exact = true
// gives the derivatives of the product, (P V)^T G, P^T G R^T and G (V R)^T.
inline void camera_backward(const mat4& G, const mat4& R, const mat4& V, const mat4& P, bool exact, mat4& gR, mat4& gV, mat4& gP) {
    if (exact) {
        gR = transpose(P * V) * G;
        gV = transpose(P) * G * transpose(R);
        gP = G * transpose(V * R);
    } else {
        gR = G * transpose(V) * transpose(P);
        gV = transpose(R) * G * transpose(P);
        gP = transpose(R) * transpose(V) * G;
    }
}

// Rotation::backward (matrix.cu:85-99): d loss / d q from the gradient of the rotation
// matrix (its upper 3x3).  The reference sums the 3x3 against dR/dq without the factor 2 of
// dR/dq (matrix.cu:67-83 states it), so its result is half the derivative; exact = true
// gives the derivative.
inline quat rotation_backward(const quat& q, const mat4& G, bool exact) {
    const float* r0 = G.m[0];
    const float* r1 = G.m[1];
    const float* r2 = G.m[2];
    auto d3 = [](const float* a, float x, float y, float z) { return a[0] * x + a[1] * y + a[2] * z; };
    quat g;
    // columns of dR/dq{w,x,y,z} (up to the factor 2): wzy = (w, z, -y), zwx = (-z, w, x),
    // yxw = (y, -x, w), xyz = (x, y, z)
    g.w = d3(r0, q.w, q.z, -q.y) + d3(r1, -q.z, q.w, q.x) + d3(r2, q.y, -q.x, q.w);
    g.x = d3(r0, q.x, q.y, q.z) + d3(r1, q.y, -q.x, q.w) - d3(r2, -q.z, q.w, q.x);
    g.y = -d3(r0, q.y, -q.x, q.w) + d3(r1, q.x, q.y, q.z) + d3(r2, q.w, q.z, -q.y);
    g.z = d3(r0, -q.z, q.w, q.x) - d3(r1, q.w, q.z, -q.y) + d3(r2, q.x, q.y, q.z);
    if (exact) { g.x *= 2.f; g.y *= 2.f; g.z *= 2.f; g.w *= 2.f; }
    return g;
}

// Rotation::step (matrix.cu:111-137): Adam on the four quaternion components, in double,
// with bias correction by rhom^(t+1) (the caller's rhomt/rhovt are advanced first, as in the
// reference), then renormalisation.
struct RotationAdam {
    int it = 0;
    double rhom = 0.9, rhov = 0.999, rhomt = 0.9, rhovt = 0.999, eta = 1e-2, eps = 1e-8;
    quat m, v;
};
inline quat rotation_adam_step(RotationAdam& a, const quat& q, const quat& g) {
    a.it++;
    a.rhomt *= a.rhom;
    a.rhovt *= a.rhov;
    quat r = q;
    for (int i = 0; i < 4; i++) {
        a.m[i] = (float)(a.rhom * a.m[i] + (1 - a.rhom) * g[i]);
        a.v[i] = (float)(a.rhov * a.v[i] + (1 - a.rhov) * g[i] * g[i]);
        double m = a.m[i] / (1 - a.rhomt);
        double v = a.v[i] / (1 - a.rhovt);
        r[i] = (float)(r[i] - m * a.eta / (std::sqrt(v) + a.eps));
    }
    return normalize(r);
}

}  // namespace vcl
