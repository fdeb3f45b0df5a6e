// Output stage of a frame (Graphics.cpp:1908, 5835-5880): the shaded colour becomes the RGBA16F
// draw image, which is blitted with a linear filter to the swapchain extent into a B8G8R8A8_SRGB
// image (vkCmdBlitImage2: destination texel centres mapped to source coordinates, bilinear with
// clamp-to-edge, then the format's linear-to-sRGB encode).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.h"

namespace rsy {
namespace {

// RGBA32F -> RGBA16F, round to nearest even (one 8-byte store per pixel).
__global__ __launch_bounds__(256) void k_draw_image(const float4* color, uint2* draw, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = color[i];
    const __half2 rg = __floats2half2_rn(c.x, c.y), ba = __floats2half2_rn(c.z, c.w);
    draw[i] = make_uint2(*reinterpret_cast<const uint32_t*>(&rg), *reinterpret_cast<const uint32_t*>(&ba));
}

__device__ inline float4 load_half4(const uint2* img, int w, int x, int y) {
    const uint2 p = img[size_t(y) * w + x];
    const __half2 rg = *reinterpret_cast<const __half2*>(&p.x), ba = *reinterpret_cast<const __half2*>(&p.y);
    const float2 a = __half22float2(rg), b = __half22float2(ba);
    return make_float4(a.x, a.y, b.x, b.y);
}

// Linear to sRGB (the sRGB OETF of the Vulkan data-format specification) and 8-bit rounding.
__device__ inline uint32_t srgb8(float c) {
    c = fminf(fmaxf(c, 0.0f), 1.0f);
    const float s = c <= 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
    return uint32_t(rintf(s * 255.0f));
}

// One destination pixel of the blit: source coordinate (x + 0.5) * sw / dw - 0.5 on each axis,
// bilinear over the four nearest texels clamped to the image, then B8G8R8A8 with sRGB colour and
// linear alpha.
__global__ __launch_bounds__(256) void k_present(const uint2* draw, int sw, int sh, uint32_t* out, int dw, int dh) {
    const int x = int(blockIdx.x * 16 + (threadIdx.x & 15)), y = int(blockIdx.y * 16 + (threadIdx.x >> 4));
    if (x >= dw || y >= dh) return;
    const float u = (float(x) + 0.5f) * (float(sw) / float(dw)) - 0.5f;
    const float v = (float(y) + 0.5f) * (float(sh) / float(dh)) - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    const float au = u - fu, av = v - fv;
    const int x0 = min(max(int(fu), 0), sw - 1), x1 = min(max(int(fu) + 1, 0), sw - 1);
    const int y0 = min(max(int(fv), 0), sh - 1), y1 = min(max(int(fv) + 1, 0), sh - 1);
    const float4 c00 = load_half4(draw, sw, x0, y0), c10 = load_half4(draw, sw, x1, y0);
    const float4 c01 = load_half4(draw, sw, x0, y1), c11 = load_half4(draw, sw, x1, y1);
    const float w00 = (1.f - au) * (1.f - av), w10 = au * (1.f - av), w01 = (1.f - au) * av, w11 = au * av;
    const float r = w00 * c00.x + w10 * c10.x + w01 * c01.x + w11 * c11.x;
    const float g = w00 * c00.y + w10 * c10.y + w01 * c01.y + w11 * c11.y;
    const float b = w00 * c00.z + w10 * c10.z + w01 * c01.z + w11 * c11.z;
    const float a = w00 * c00.w + w10 * c10.w + w01 * c01.w + w11 * c11.w;
    const uint32_t a8 = uint32_t(rintf(fminf(fmaxf(a, 0.0f), 1.0f) * 255.0f));
    out[size_t(y) * dw + x] = srgb8(b) | (srgb8(g) << 8) | (srgb8(r) << 16) | (a8 << 24);
}

}  // namespace

hipError_t launch_present(const float4* color, uint2* draw, uint32_t sw, uint32_t sh, uint32_t* out, uint32_t dw,
                          uint32_t dh, hipStream_t s) {
    const uint64_t n = uint64_t(sw) * sh;
    if (n == 0 || dw == 0 || dh == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_draw_image, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, color, draw, n);
    hipLaunchKernelGGL(k_present, dim3((dw + 15) / 16, (dh + 15) / 16), dim3(256), 0, s, draw, int(sw), int(sh), out,
                       int(dw), int(dh));
    return hipGetLastError();
}

}  // namespace rsy
