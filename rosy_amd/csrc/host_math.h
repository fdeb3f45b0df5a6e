// Host-side matrix helpers shared by the context and the scene generator.
// Matrices are column-major (m[col*4 + row]), as glm::value_ptr lays them out.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>

namespace rsy {

using dmat4 = std::array<double, 16>;
using fmat4 = std::array<float, 16>;

inline double& at(dmat4& m, int col, int row) { return m[col * 4 + row]; }
inline double at(const dmat4& m, int col, int row) { return m[col * 4 + row]; }

dmat4 d_identity();
dmat4 d_mul(const dmat4& a, const dmat4& b);
dmat4 d_translate(double x, double y, double z);
dmat4 d_scale(double s);
// Rotation matrix of glm::toMat4(glm::angleAxis(angle, axis)) for a unit axis.
dmat4 d_angle_axis(double angle, double ax, double ay, double az);
dmat4 d_transpose(const dmat4& m);
// General 4x4 inverse by cofactors (double precision).
dmat4 d_inverse(const dmat4& m);
void d_mul_vec(const dmat4& m, const double v[4], double out[4]);
fmat4 to_float(const dmat4& m);
dmat4 to_double(const float* m);

// FP32 product c = a * b with the fixed fmaf order used on device and in the oracle:
// c[col][row] = fma(a[3][row], b[col][3], fma(a[2][row], b[col][2], fma(a[1][row], b[col][1], a[0][row]*b[col][0])))
void f_mul(const float* a, const float* b, float* c);

// World-to-light transform inverse(light_line_rot * T(0,0,d)) of Level.cpp:464-477, 525.
dmat4 light_view_inverse(double sun_distance, double sun_pitch, double sun_yaw);

}  // namespace rsy
