// Fixed-point raster rules shared by the oracle's raster and shading halves.  TEST INFRASTRUCTURE.
// (The rules are listed in oracle.h; this file is built with -ffp-contract=off.)
#ifndef ROSY_ORACLE_FIXED_POINT_H
#define ROSY_ORACLE_FIXED_POINT_H

#include <cmath>
#include <cstdint>

namespace orc {

struct C4 {
    float x, y, z, w;
};

// M = A * B by the fmaf chain in k order.
inline void mat_mul_fma(const float* a, const float* b, float* c) {
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row) {
            float s = a[row] * b[col * 4 + 0];
            s = std::fmaf(a[4 + row], b[col * 4 + 1], s);
            s = std::fmaf(a[8 + row], b[col * 4 + 2], s);
            s = std::fmaf(a[12 + row], b[col * 4 + 3], s);
            c[col * 4 + row] = s;
        }
}

// clip row r = fma(M[2][r], z, fma(M[1][r], y, fma(M[0][r], x, M[3][r])))
inline C4 transform(const float* m, const float* p) {
    C4 c;
    float* o = &c.x;
    for (int r = 0; r < 4; ++r)
        o[r] = std::fmaf(m[8 + r], p[2], std::fmaf(m[4 + r], p[1], std::fmaf(m[r], p[0], m[12 + r])));
    return c;
}

inline float plane_dist(const C4& v, int plane) {
    switch (plane) {
        case 0: return v.w - v.z;          // near (reverse-Z: depth <= 1)
        case 1: return v.z;                // far (depth >= 0)
        case 2: return v.x + 4.0f * v.w;   // guard band |x| <= 4w
        case 3: return 4.0f * v.w - v.x;
        case 4: return v.y + 4.0f * v.w;   // guard band |y| <= 4w
        default: return 4.0f * v.w - v.y;
    }
}

// True when the triangle needs no clipping (every vertex inside every plane, w > 0).
inline bool all_inside(const C4 v[3]) {
    for (int k = 0; k < 3; ++k) {
        if (!(v[k].w > 0.f)) return false;
        for (int p = 0; p < 6; ++p)
            if (!(plane_dist(v[k], p) >= 0.f)) return false;
    }
    return true;
}

// X = rint(256 * fma(x/w, W/2, W/2)), Y likewise (8 sub-pixel bits).
inline void snap(const C4& v, uint32_t width, uint32_t height, int32_t& X, int32_t& Y, float& Z) {
    const float hw = 0.5f * float(width), hh = 0.5f * float(height);
    const float xd = v.x / v.w, yd = v.y / v.w;
    Z = v.z / v.w;
    X = int32_t(std::rintf(std::fmaf(xd, hw, hw) * 256.0f));
    Y = int32_t(std::rintf(std::fmaf(yd, hh, hh) * 256.0f));
}

}  // namespace orc

#endif
