// Host-only C ABI over include/vcl_matrix.hpp: the Rotation / Camera math of the reference
// (matrix.cu:3-239).  No device code; callable without a GPU.
#include "vcl.h"
#include "vcl_matrix.hpp"

using vcl::mat4;
using vcl::quat;
using vcl::vec3;

static quat q_of(const float* q) { return quat(q[3], q[0], q[1], q[2]); }
static void put_q(const quat& q, float* out) { out[0] = q.x; out[1] = q.y; out[2] = q.z; out[3] = q.w; }
static void put_m(const mat4& m, float* out) { std::memcpy(out, m.data(), 16 * sizeof(float)); }

extern "C" int vcl_rotation_matrix(const float* q, float* out16) {
    if (!q || !out16) return VCL_ERR_INVALID;
    put_m(vcl::rotation_matrix(q_of(q)), out16);
    return VCL_OK;
}

extern "C" int vcl_view_matrix(const float* q, const float* eye, float* out16) {
    if (!q || !eye || !out16) return VCL_ERR_INVALID;
    put_m(vcl::view_matrix(q_of(q), vec3(eye[0], eye[1], eye[2])), out16);
    return VCL_OK;
}

extern "C" int vcl_look_at_quat(const float* eye, const float* center, const float* up, float* q_out) {
    if (!eye || !center || !up || !q_out) return VCL_ERR_INVALID;
    put_q(vcl::look_at_quat(vec3(eye[0], eye[1], eye[2]), vec3(center[0], center[1], center[2]), vec3(up[0], up[1], up[2])), q_out);
    return VCL_OK;
}

extern "C" int vcl_projection_matrix(float size, float aspect, float znear, float zfar, float* out16) {
    if (!out16) return VCL_ERR_INVALID;
    put_m(vcl::projection_matrix(size, aspect, znear, zfar), out16);
    return VCL_OK;
}

extern "C" int vcl_camera_matrices(const float* q, const float* eye, float size, float aspect, float znear, float zfar,
                                   const float* rotation16, float* view16, float* projection16, float* mat16) {
    if (!q || !eye || !rotation16 || !view16 || !projection16 || !mat16) return VCL_ERR_INVALID;
    mat4 V = vcl::view_matrix(q_of(q), vec3(eye[0], eye[1], eye[2]));
    mat4 P = vcl::projection_matrix(size, aspect, znear, zfar);
    put_m(V, view16);
    put_m(P, projection16);
    put_m(P * V * mat4::cols(rotation16), mat16);
    return VCL_OK;
}

extern "C" int vcl_camera_backward_host(const float* gmat16, const float* rotation16, const float* view16, const float* projection16,
                                        int exact, float* grot16, float* gview16, float* gproj16) {
    if (!gmat16 || !rotation16 || !view16 || !projection16 || !grot16 || !gview16 || !gproj16) return VCL_ERR_INVALID;
    mat4 gR, gV, gP;
    vcl::camera_backward(mat4::cols(gmat16), mat4::cols(rotation16), mat4::cols(view16), mat4::cols(projection16), exact != 0, gR, gV, gP);
    put_m(gR, grot16);
    put_m(gV, gview16);
    put_m(gP, gproj16);
    return VCL_OK;
}

extern "C" int vcl_rotation_backward_host(const float* q, const float* grot16, int exact, float* gq) {
    if (!q || !grot16 || !gq) return VCL_ERR_INVALID;
    put_q(vcl::rotation_backward(q_of(q), mat4::cols(grot16), exact != 0), gq);
    return VCL_OK;
}

extern "C" int vcl_rotation_adam_step(vcl_rotation_adam* a, float* q, const float* gq) {
    if (!a || !q || !gq) return VCL_ERR_INVALID;
    vcl::RotationAdam s;
    s.it = a->it; s.rhom = a->rhom; s.rhov = a->rhov; s.rhomt = a->rhomt; s.rhovt = a->rhovt; s.eta = a->eta; s.eps = a->eps;
    s.m = q_of(a->m); s.v = q_of(a->v);
    quat r = vcl::rotation_adam_step(s, q_of(q), q_of(gq));
    a->it = s.it; a->rhomt = s.rhomt; a->rhovt = s.rhovt;
    put_q(s.m, a->m); put_q(s.v, a->v);
    put_q(r, q);
    return VCL_OK;
}
