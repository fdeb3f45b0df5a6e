// Texture sampling shared by the shading and material-decoder kernels: nearest or bilinear,
// CLAMP_TO_EDGE / MIRRORED_REPEAT / REPEAT addressing, sRGB decode before filtering through a
// 256-entry table (Vulkan's filtering order; samplers Graphics.cpp:3421-3432, formats 4000-4015).
// Textures with mip levels are sampled trilinearly at the level of detail of the pixel's texture
// coordinate derivatives (the asset samplers' VK_SAMPLER_MIPMAP_MODE_LINEAR, Graphics.cpp:4002-4012).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "frame_types.h"
#include "rosy_amd.h"

namespace rsy {

__device__ inline int wrap_coord(int i, int n, uint32_t mode) {
    if (mode == 0) return min(max(i, 0), n - 1);
    if ((n & (n - 1)) == 0) {  // power-of-two extent: the modulo is a mask (two's complement)
        if (mode == 1) {
            const int m = i & (2 * n - 1);
            return m < n ? m : 2 * n - 1 - m;
        }
        return i & (n - 1);
    }
    if (mode == 1) {
        const int period = 2 * n;
        int m = i % period;
        m = m < 0 ? m + period : m;
        return m < n ? m : period - 1 - m;
    }
    int m = i % n;
    return m < 0 ? m + n : m;
}

struct TexRef {
    const uint8_t* base;
    int w, h, ch, srgb;
    int levels = 1;
};

// Mip level l of a texture as a texture of its own.
__device__ inline TexRef level_of(const TexRef& t, int l) {
    TexRef r = t;
    r.base = t.base + mip_offset(uint32_t(t.w), uint32_t(t.h), uint32_t(t.ch), uint32_t(l));
    r.w = int(mip_extent(uint32_t(t.w), uint32_t(l)));
    r.h = int(mip_extent(uint32_t(t.h), uint32_t(l)));
    r.levels = 1;
    return r;
}

// Sampler state of a sample: addressing per axis (0 clamp-to-edge, 1 mirrored repeat, 2 repeat)
// and nearest or bilinear filtering; the default is the REPEAT / bilinear sampler.
struct SamplerState {
    uint32_t wrap_s = 2u, wrap_t = 2u;
    bool nearest = false;      // magnification filter (and the filter of single-level textures)
    bool min_nearest = false;  // minification filter of textures with mip levels
};

template <class S>
__device__ inline SamplerState sampler_state(const S& s) {
    SamplerState r;
    r.wrap_s = s.wrap_s;
    r.wrap_t = s.wrap_t;
    r.nearest = s.mag_filter <= 2;
    r.min_nearest = s.min_filter <= 2;
    return r;
}

__device__ inline float4 texel(const TexRef& t, int x, int y, const float* lut) {
    const size_t o = size_t(y) * t.w + x;
    float4 c = make_float4(0.f, 0.f, 0.f, 1.f);
    if (t.ch == 4) {
        const uint32_t p = reinterpret_cast<const uint32_t*>(t.base)[o];
        const uint32_t r = p & 0xffu, g = (p >> 8) & 0xffu, b = (p >> 16) & 0xffu, a = p >> 24;
        c.x = t.srgb ? lut[r] : float(r) * (1.0f / 255.0f);
        c.y = t.srgb ? lut[g] : float(g) * (1.0f / 255.0f);
        c.z = t.srgb ? lut[b] : float(b) * (1.0f / 255.0f);
        c.w = float(a) * (1.0f / 255.0f);
    } else {
        const uint32_t p = reinterpret_cast<const uint16_t*>(t.base)[o];
        const uint32_t r = p & 0xffu, g = p >> 8;
        c.x = t.srgb ? lut[r] : float(r) * (1.0f / 255.0f);
        c.y = t.srgb ? lut[g] : float(g) * (1.0f / 255.0f);
    }
    return c;
}

__device__ inline float4 sample_tex(const TexRef& t, const SamplerState& s, float u, float v, const float* lut) {
    const uint32_t ws = s.wrap_s, wt = s.wrap_t;
    if (s.nearest) {
        const int x = wrap_coord(int(floorf(u * t.w)), t.w, ws), y = wrap_coord(int(floorf(v * t.h)), t.h, wt);
        return texel(t, x, y, lut);
    }
    const float x = u * float(t.w) - 0.5f, y = v * float(t.h) - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    const int x0 = int(fx0), y0 = int(fy0);
    const int xa = wrap_coord(x0, t.w, ws), xb = wrap_coord(x0 + 1, t.w, ws);
    const int ya = wrap_coord(y0, t.h, wt), yb = wrap_coord(y0 + 1, t.h, wt);
    const float4 c00 = texel(t, xa, ya, lut), c10 = texel(t, xb, ya, lut);
    const float4 c01 = texel(t, xa, yb, lut), c11 = texel(t, xb, yb, lut);
    const float w00 = (1 - ax) * (1 - ay), w10 = ax * (1 - ay), w01 = (1 - ax) * ay, w11 = ax * ay;
    float4 r;
    r.x = w00 * c00.x + w10 * c10.x + w01 * c01.x + w11 * c11.x;
    r.y = w00 * c00.y + w10 * c10.y + w01 * c01.y + w11 * c11.y;
    r.z = w00 * c00.z + w10 * c10.z + w01 * c01.z + w11 * c11.z;
    r.w = w00 * c00.w + w10 * c10.w + w01 * c01.w + w11 * c11.w;
    return r;
}

// Trilinear sampling (Vulkan's level-of-detail and mipmapping rules with mipmapMode LINEAR, no LOD
// bias or clamp): rho = the larger of |d(uv)/dx| and |d(uv)/dy| in level-0 texels, lambda =
// log2(rho); lambda <= 0 samples level 0 with the magnification filter, otherwise d = min(lambda,
// levels - 1) blends levels floor(d) and floor(d) + 1 (the last level when past it) by frac(d) with
// the minification filter.  Single-level textures keep the plain sample.
__device__ inline float4 sample_tex_lod(const TexRef& t, const SamplerState& s, float u, float v, float dudx, float dvdx,
                                        float dudy, float dvdy, const float* lut) {
    if (t.levels <= 1) return sample_tex(t, s, u, v, lut);
    const float W = float(t.w), H = float(t.h);
    const float rx = sqrtf(dudx * W * (dudx * W) + dvdx * H * (dvdx * H));
    const float ry = sqrtf(dudy * W * (dudy * W) + dvdy * H * (dvdy * H));
    const float lambda = log2f(fmaxf(rx, ry));
    if (!(lambda > 0.0f)) return sample_tex(level_of(t, 0), s, u, v, lut);
    SamplerState m = s;
    m.nearest = s.min_nearest;
    const float d = fminf(lambda, float(t.levels - 1));
    const float fl = floorf(d);
    const int dl = int(fl), dh = min(dl + 1, t.levels - 1);
    const float delta = d - fl;
    const float4 a = sample_tex(level_of(t, dl), m, u, v, lut);
    if (delta == 0.0f || dh == dl) return a;
    const float4 b = sample_tex(level_of(t, dh), m, u, v, lut);
    return make_float4(a.x + (b.x - a.x) * delta, a.y + (b.y - a.y) * delta, a.z + (b.z - a.z) * delta,
                       a.w + (b.w - a.w) * delta);
}

}  // namespace rsy
