// Fixed-point raster rules used by the raster kernels.  Contraction is switched off inside each
// function so that every FP32 operation rounds where the source says, whatever the translation
// unit's -ffp-contract setting.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rsy {

__host__ __device__ inline float plane_dist(const float4& v, int plane) {
#pragma clang fp contract(off)
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
__host__ __device__ inline bool all_inside(const float4& a, const float4& b, const float4& c) {
    bool in = a.w > 0.f && b.w > 0.f && c.w > 0.f;
#pragma unroll
    for (int p = 0; p < 6; ++p) in = in && plane_dist(a, p) >= 0.f && plane_dist(b, p) >= 0.f && plane_dist(c, p) >= 0.f;
    return in;
}

// X = rint(256 * fma(x/w, W/2, W/2)), Y likewise (8 sub-pixel bits); Z = z/w.
__host__ __device__ inline void snap(const float4& v, float hw, float hh, int32_t& X, int32_t& Y, float& Z) {
#pragma clang fp contract(off)
    const float xd = v.x / v.w, yd = v.y / v.w;
    Z = v.z / v.w;
    X = int32_t(rintf(fmaf(xd, hw, hw) * 256.0f));
    Y = int32_t(rintf(fmaf(yd, hh, hh) * 256.0f));
}

}  // namespace rsy
