// Closed-form restatements of the reference's camera, light and node matrices.
//
// The reference builds these with glm in FP32 (Camera.cpp:62-98, Level.cpp:464-528,
// Node.cpp:96-108, 225-247).  glm is not vendored in the reference snapshot, so this file computes
// the same closed forms in double precision and rounds once to FP32; agreement with glm's own
// rounding is not claimed.
#include "host_math.h"

#include <cstring>

#include "rosy_amd.h"

namespace rsy {

dmat4 d_identity() {
    dmat4 m{};
    m[0] = m[5] = m[10] = m[15] = 1.0;
    return m;
}

dmat4 d_mul(const dmat4& a, const dmat4& b) {
    dmat4 c{};
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += at(a, k, row) * at(b, col, k);
            at(c, col, row) = s;
        }
    return c;
}

dmat4 d_translate(double x, double y, double z) {
    dmat4 m = d_identity();
    at(m, 3, 0) = x;
    at(m, 3, 1) = y;
    at(m, 3, 2) = z;
    return m;
}

dmat4 d_scale(double s) {
    dmat4 m = d_identity();
    at(m, 0, 0) = s;
    at(m, 1, 1) = s;
    at(m, 2, 2) = s;
    return m;
}

dmat4 d_angle_axis(double angle, double ax, double ay, double az) {
    // Quaternion q = (cos(a/2), axis*sin(a/2)), then the standard quaternion-to-matrix form that
    // glm::mat3_cast uses.
    const double s = std::sin(angle * 0.5), w = std::cos(angle * 0.5);
    const double x = ax * s, y = ay * s, z = az * s;
    const double xx = x * x, yy = y * y, zz = z * z;
    const double xy = x * y, xz = x * z, yz = y * z;
    const double wx = w * x, wy = w * y, wz = w * z;
    dmat4 m = d_identity();
    at(m, 0, 0) = 1.0 - 2.0 * (yy + zz);
    at(m, 0, 1) = 2.0 * (xy + wz);
    at(m, 0, 2) = 2.0 * (xz - wy);
    at(m, 1, 0) = 2.0 * (xy - wz);
    at(m, 1, 1) = 1.0 - 2.0 * (xx + zz);
    at(m, 1, 2) = 2.0 * (yz + wx);
    at(m, 2, 0) = 2.0 * (xz + wy);
    at(m, 2, 1) = 2.0 * (yz - wx);
    at(m, 2, 2) = 1.0 - 2.0 * (xx + yy);
    return m;
}

dmat4 d_transpose(const dmat4& m) {
    dmat4 t{};
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) at(t, c, r) = at(m, r, c);
    return t;
}

dmat4 d_inverse(const dmat4& m) {
    // Cofactor expansion on the row-major view a[i][j] = m[i*4+j] (inverse commutes with
    // transposition, so the storage order does not matter).
    const double* a = m.data();
    double inv[16];
    inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] +
             a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] -
             a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] +
             a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] -
              a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] -
             a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] +
             a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] -
             a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] +
              a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] +
             a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] -
             a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] +
              a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] -
              a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] -
             a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] +
             a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] -
              a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] +
              a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    const double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    dmat4 r{};
    const double id = det != 0.0 ? 1.0 / det : 0.0;
    for (int i = 0; i < 16; ++i) r[i] = inv[i] * id;
    return r;
}

void d_mul_vec(const dmat4& m, const double v[4], double out[4]) {
    for (int r = 0; r < 4; ++r) {
        double s = 0.0;
        for (int k = 0; k < 4; ++k) s += at(m, k, r) * v[k];
        out[r] = s;
    }
}

fmat4 to_float(const dmat4& m) {
    fmat4 f{};
    for (int i = 0; i < 16; ++i) f[i] = static_cast<float>(m[i]);
    return f;
}

dmat4 to_double(const float* m) {
    dmat4 d{};
    for (int i = 0; i < 16; ++i) d[i] = m[i];
    return d;
}

void f_mul(const float* a, const float* b, float* c) {
    float t[16];
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row) {
            float s = a[0 * 4 + row] * b[col * 4 + 0];
            s = std::fmaf(a[1 * 4 + row], b[col * 4 + 1], s);
            s = std::fmaf(a[2 * 4 + row], b[col * 4 + 2], s);
            s = std::fmaf(a[3 * 4 + row], b[col * 4 + 3], s);
            t[col * 4 + row] = s;
        }
    std::memcpy(c, t, sizeof(t));
}

dmat4 light_view_inverse(double sun_distance, double sun_pitch, double sun_yaw) {
    const dmat4 light_translate = d_translate(0.0, 0.0, sun_distance);
    const dmat4 light_line_rot = d_mul(d_angle_axis(sun_yaw, 0.0, -1.0, 0.0), d_angle_axis(-sun_pitch, 1.0, 0.0, 0.0));
    return d_inverse(d_mul(light_line_rot, light_translate));
}

}  // namespace rsy

using namespace rsy;

extern "C" {

void rsy_mat4_mul(const float a[16], const float b[16], float c[16]) { f_mul(a, b, c); }

// Camera.cpp:62-83.  h = 1/tan(fov/2), w = h/aspect, a = -n/(f-n), b = n f/(f-n); the matrix with
// columns (w,0,0,0), (0,h,0,0), (0,0,a,1), (0,0,b,0) times the Vulkan NDC flip diag(1,-1,-1,1).
void rsy_camera_projection(double aspect, double n, double f, double fov, float out[16]) {
    const double h = 1.0 / std::tan(fov * 0.5);
    const double w = h / aspect;
    const double a = -n / (f - n);
    const double b = (n * f) / (f - n);
    dmat4 p{};
    at(p, 0, 0) = w;
    at(p, 1, 1) = h;
    at(p, 2, 2) = a;
    at(p, 2, 3) = 1.0;
    at(p, 3, 2) = b;
    dmat4 ndc = d_identity();
    at(ndc, 1, 1) = -1.0;
    at(ndc, 2, 2) = -1.0;
    const fmat4 r = to_float(d_mul(p, ndc));
    std::memcpy(out, r.data(), sizeof(float) * 16);
}

// Camera.cpp:85-98: view = inverse(T(position) * R_yaw * R_pitch) with
// R_pitch = angleAxis(pitch, -x) and R_yaw = angleAxis(yaw, +y).
void rsy_camera_view_matrix(const float position[3], float pitch, float yaw, float out[16]) {
    const dmat4 rp = d_angle_axis(pitch, -1.0, 0.0, 0.0);
    const dmat4 ry = d_angle_axis(yaw, 0.0, 1.0, 0.0);
    const dmat4 cam = d_mul(d_translate(position[0], position[1], position[2]), d_mul(ry, rp));
    const fmat4 r = to_float(d_inverse(cam));
    std::memcpy(out, r.data(), sizeof(float) * 16);
}

// Level.cpp:464-528.
void rsy_light_matrices(float sun_distance, float sun_pitch, float sun_yaw, float cascade_level,
                        float orthographic_depth, float out_sun_position[4], float out_shadow_projection[16]) {
    const dmat4 light_line_rot =
        d_mul(d_angle_axis(sun_yaw, 0.0, -1.0, 0.0), d_angle_axis(-static_cast<double>(sun_pitch), 1.0, 0.0, 0.0));
    const double sun0[4] = {0.0, 0.0, -static_cast<double>(sun_distance), 1.0};
    double sun[4];
    d_mul_vec(light_line_rot, sun0, sun);
    for (int i = 0; i < 4; ++i) out_sun_position[i] = static_cast<float>(sun[i]);
    dmat4 lp{};
    at(lp, 0, 0) = -2.0 / cascade_level;
    at(lp, 1, 1) = -2.0 / cascade_level;
    at(lp, 2, 2) = -1.0 / orthographic_depth;
    at(lp, 3, 3) = 1.0;
    const fmat4 r = to_float(d_mul(lp, light_view_inverse(sun_distance, sun_pitch, sun_yaw)));
    std::memcpy(out_shadow_projection, r.data(), sizeof(float) * 16);
}

// Node.cpp:96-108 and 225-247 with identity coordinate-system, parent and local transforms:
// W = T * angleAxis(yaw, -y) * S; to-object = inverse(W); normal = transpose(inverse(mat3(W))).
void rsy_object_record_from_trs(const float translate[3], float scale, float yaw, rsy_object_record* out) {
    const dmat4 w = d_mul(d_translate(translate[0], translate[1], translate[2]),
                          d_mul(d_angle_axis(yaw, 0.0, -1.0, 0.0), d_scale(scale)));
    const dmat4 inv = d_inverse(w);
    const fmat4 wf = to_float(w), invf = to_float(inv);
    std::memcpy(out->transform, wf.data(), sizeof(float) * 16);
    std::memcpy(out->to_object_space_transform, invf.data(), sizeof(float) * 16);
    // normal = transpose(inverse(mat3(W))): element (col c, row r) = inverse(mat3 W)(col r, row c).
    // The upper 3x3 of inverse(W) equals inverse(mat3(W)) for an affine W.
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) out->normal_transform[c * 3 + r] = static_cast<float>(at(inv, r, c));
}

}  // extern "C"
