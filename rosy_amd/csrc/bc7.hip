// BC7 block decode on the GPU (the format of the reference's images: Packager.cpp:216-308 writes
// every colour, normal, metallic-roughness and mixmap image as BC7, and Graphics.cpp:3620-3650
// uploads them as VK_FORMAT_BC7_SRGB_BLOCK / VK_FORMAT_BC7_UNORM_BLOCK).  The decode follows the
// published BC7 format: a 128-bit block per 4x4 texels, read from bit 0 upwards; the mode is the
// position of the lowest set bit of the first byte; then partition, rotation and index-selection
// fields, the endpoint colours channel by channel, alpha, p-bits, and the index sets (the anchor
// texel of each subset stores one index bit fewer, its top bit being 0).  Endpoints are widened to
// 8 bits by bit replication and interpolated as ((64 - w) e0 + w e1 + 32) >> 6.  A block with no
// mode bit decodes to transparent black.  Output is RGBA8 (sRGB decode stays with the sampler).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.h"

namespace rsy {
namespace {

// Per mode: subsets, partition bits, rotation bits, index-selection bits, colour bits, alpha bits,
// per-endpoint p-bits, shared p-bits, index bits, second index bits.
struct Bc7Mode {
    uint8_t ns, pb, rb, isb, cb, ab, epb, spb, ib, ib2;
};
__constant__ Bc7Mode kModes[8] = {
    {3, 4, 0, 0, 4, 0, 1, 0, 3, 0}, {2, 6, 0, 0, 6, 0, 0, 1, 3, 0}, {3, 6, 0, 0, 5, 0, 0, 0, 2, 0},
    {2, 6, 0, 0, 7, 0, 1, 0, 2, 0}, {1, 0, 2, 1, 5, 6, 0, 0, 2, 3}, {1, 0, 2, 0, 7, 8, 0, 0, 2, 2},
    {1, 0, 0, 0, 7, 7, 1, 0, 4, 0}, {2, 6, 0, 0, 5, 5, 1, 0, 2, 0}};

// Two-subset partitions: bit i set when texel i belongs to subset 1.
__constant__ uint16_t kPart2[64] = {
    0xCCCC, 0x8888, 0xEEEE, 0xECC8, 0xC880, 0xFEEC, 0xFEC8, 0xEC80, 0xC800, 0xFFEC, 0xFE80, 0xE800, 0xFFE8,
    0xFF00, 0xFFF0, 0xF000, 0xF710, 0x008E, 0x7100, 0x08CE, 0x008C, 0x7310, 0x3100, 0x8CCE, 0x088C, 0x3110,
    0x6666, 0x366C, 0x17E8, 0x0FF0, 0x718E, 0x399C, 0xAAAA, 0xF0F0, 0x5A5A, 0x33CC, 0x3C3C, 0x55AA, 0x9696,
    0xA55A, 0x73CE, 0x13C8, 0x324C, 0x3BDC, 0x6996, 0xC33C, 0x9966, 0x0660, 0x0272, 0x04E4, 0x4E40, 0x2720,
    0xC936, 0x936C, 0x39C6, 0x639C, 0x9336, 0x9CC6, 0x817E, 0xE718, 0xCCF0, 0x0FCC, 0x7744, 0xEE22};

// Three-subset partitions: two bits per texel (texel i in bits 2i, 2i+1).
__constant__ uint32_t kPart3[64] = {
    0xAA685050, 0x6A5A5040, 0x5A5A4200, 0x5450A0A8, 0xA5A50000, 0xA0A05050, 0x5555A0A0, 0x5A5A5050,
    0xAA550000, 0xAA555500, 0xAAAA5500, 0x90909090, 0x94949494, 0xA4A4A4A4, 0xA9A59450, 0x2A0A4250,
    0xA5945040, 0x0A425054, 0xA5A5A500, 0x55A0A0A0, 0xA8A85454, 0x6A6A4040, 0xA4A45000, 0x1A1A0500,
    0x0050A4A4, 0xAAA59090, 0x14696914, 0x69691400, 0xA08585A0, 0xAA821414, 0x50A4A450, 0x6A5A0200,
    0xA9A58000, 0x5090A0A8, 0xA8A09050, 0x24242424, 0x00AA5500, 0x24924924, 0x24499224, 0x50A50A50,
    0x500AA550, 0xAAAA4444, 0x66660000, 0xA5A0A5A0, 0x50A050A0, 0x69286928, 0x44AAAA44, 0x66666600,
    0xAA444444, 0x54A854A8, 0x95809580, 0x96969600, 0xA85454A8, 0x80959580, 0xAA141414, 0x96960000,
    0xAAAA1414, 0xA05050A0, 0xA0A5A5A0, 0x96000000, 0x40804080, 0xA9A8A9A8, 0xAAAAAA44, 0x2A4A5254};

// Anchor texel of subset 1 in two-subset partitions; of subsets 1 and 2 in three-subset ones.
__constant__ uint8_t kAnchor2[64] = {15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15,
                                     15, 2,  8,  2,  2,  8,  8,  15, 2,  8,  2,  2,  8,  8,  2,  2,
                                     15, 15, 6,  8,  2,  8,  15, 15, 2,  8,  2,  2,  2,  15, 15, 6,
                                     6,  2,  6,  8,  15, 15, 2,  2,  15, 15, 15, 15, 15, 2,  2,  15};
__constant__ uint8_t kAnchor3a[64] = {3,  3,  15, 15, 8,  3,  15, 15, 8,  8,  6,  6,  6,  5,  3,  3,
                                      3,  3,  8,  15, 3,  3,  6,  10, 5,  8,  8,  6,  8,  5,  15, 15,
                                      8,  15, 3,  5,  6,  10, 8,  15, 15, 3,  15, 5,  15, 15, 15, 15,
                                      3,  15, 5,  5,  5,  8,  5,  10, 5,  10, 8,  13, 15, 12, 3,  3};
__constant__ uint8_t kAnchor3b[64] = {15, 8,  8,  3,  15, 15, 3,  8,  15, 15, 15, 15, 15, 15, 15, 8,
                                      15, 8,  15, 3,  15, 8,  15, 8,  3,  15, 6,  10, 15, 15, 10, 8,
                                      15, 3,  15, 10, 10, 8,  9,  10, 6,  15, 8,  15, 3,  6,  6,  8,
                                      15, 3,  15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 3,  15, 15, 8};

__constant__ uint8_t kWeights2[4] = {0, 21, 43, 64};
__constant__ uint8_t kWeights3[8] = {0, 9, 18, 27, 37, 46, 55, 64};
__constant__ uint8_t kWeights4[16] = {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64};

// n <= 8 bits of the block from bit position `pos` (fields never cross more than one word boundary).
__device__ inline uint32_t bits_at(const uint32_t (&w)[4], uint32_t pos, uint32_t n) {
    if (n == 0) return 0u;
    const uint32_t i = pos >> 5, sft = pos & 31u;
    uint64_t v = uint64_t(w[i]) >> sft;
    if (sft + n > 32u && i < 3u) v |= uint64_t(w[i + 1]) << (32u - sft);
    return uint32_t(v) & ((1u << n) - 1u);
}

__device__ inline uint32_t widen(uint32_t v, uint32_t bits) {  // replicate the top bits downwards
    v <<= 8u - bits;
    return v | (v >> bits);
}

__device__ inline uint32_t weight(uint32_t bits, uint32_t i) {
    return bits == 2 ? kWeights2[i] : bits == 3 ? kWeights3[i] : kWeights4[i];
}

__device__ inline uint32_t lerp8(uint32_t e0, uint32_t e1, uint32_t w) { return ((64u - w) * e0 + w * e1 + 32u) >> 6; }

__device__ inline uint32_t subset_of(uint32_t ns, uint32_t part, uint32_t i) {
    if (ns == 2) return (kPart2[part] >> i) & 1u;
    if (ns == 3) return (kPart3[part] >> (2u * i)) & 3u;
    return 0u;
}

// Anchor texels before texel i (texel 0 always is one) and whether texel i is one.
__device__ inline void anchors_before(uint32_t ns, uint32_t part, uint32_t i, uint32_t& before, bool& is) {
    uint32_t a1 = 16u, a2 = 16u;  // 16: no such anchor
    if (ns == 2) a1 = kAnchor2[part];
    if (ns == 3) {
        a1 = kAnchor3a[part];
        a2 = kAnchor3b[part];
    }
    before = (i > 0 ? 1u : 0u) + (a1 < i ? 1u : 0u) + (a2 < i ? 1u : 0u);
    is = i == 0 || i == a1 || i == a2;
}

// Texel i (row i / 4, column i % 4) of a block as RGBA8, R in the low byte.  Every field's bit
// position follows from the mode: header, colour endpoints (channel, subset, endpoint), alpha
// endpoints, p-bits, then the index sets, where each anchor texel's index is one bit short.
__device__ uint32_t decode_texel(const uint4 blk, uint32_t i) {
    const uint32_t w[4] = {blk.x, blk.y, blk.z, blk.w};
    const uint32_t b0 = blk.x & 0xffu;
    if (b0 == 0u) return 0u;  // no mode bit: transparent black
    const uint32_t mode = uint32_t(__builtin_ctz(b0));
    const Bc7Mode M = kModes[mode];
    uint32_t pos = mode + 1u;
    const uint32_t part = bits_at(w, pos, M.pb);
    pos += M.pb;
    const uint32_t rot = bits_at(w, pos, M.rb);
    pos += M.rb;
    const uint32_t isel = bits_at(w, pos, M.isb);
    pos += M.isb;
    const uint32_t ns = M.ns, s = subset_of(ns, part, i);
    const uint32_t colour_at = pos, alpha_at = colour_at + 6u * ns * M.cb, pbit_at = alpha_at + 2u * ns * M.ab;
    const uint32_t index_at = pbit_at + (M.epb ? 2u * ns : M.spb ? ns : 0u);
    uint32_t e[2][4];
    const bool has_p = M.epb || M.spb;
    for (uint32_t k = 0; k < 2; ++k) {
        const uint32_t p = M.epb ? bits_at(w, pbit_at + 2u * s + k, 1) : M.spb ? bits_at(w, pbit_at + s, 1) : 0u;
        for (uint32_t c = 0; c < 3; ++c) {
            const uint32_t v = bits_at(w, colour_at + ((c * ns + s) * 2u + k) * M.cb, M.cb);
            e[k][c] = has_p ? widen((v << 1) | p, M.cb + 1u) : widen(v, M.cb);
        }
        if (M.ab) {
            const uint32_t v = bits_at(w, alpha_at + (s * 2u + k) * M.ab, M.ab);
            e[k][3] = has_p ? widen((v << 1) | p, M.ab + 1u) : widen(v, M.ab);
        } else {
            e[k][3] = 255u;
        }
    }
    uint32_t before;
    bool anchor;
    anchors_before(ns, part, i, before, anchor);
    const uint32_t idx = bits_at(w, index_at + i * M.ib - before, M.ib - (anchor ? 1u : 0u));
    uint32_t ci = idx, cb = M.ib, ai = idx, ab = M.ib;
    if (M.ib2) {
        // modes 4 and 5: a second index set (one subset: only texel 0 is an anchor) for alpha, or
        // for colour when the selection bit is set
        const uint32_t at2 = index_at + 16u * M.ib - 1u;
        const uint32_t idx2 = bits_at(w, at2 + i * M.ib2 - (i > 0 ? 1u : 0u), M.ib2 - (i == 0 ? 1u : 0u));
        if (isel) {
            ci = idx2;
            cb = M.ib2;
        } else {
            ai = idx2;
            ab = M.ib2;
        }
    }
    uint32_t r = lerp8(e[0][0], e[1][0], weight(cb, ci));
    uint32_t g = lerp8(e[0][1], e[1][1], weight(cb, ci));
    uint32_t b = lerp8(e[0][2], e[1][2], weight(cb, ci));
    uint32_t a = lerp8(e[0][3], e[1][3], weight(ab, ai));
    uint32_t t;
    if (rot == 1) { t = a; a = r; r = t; }
    else if (rot == 2) { t = a; a = g; g = t; }
    else if (rot == 3) { t = a; a = b; b = t; }
    return r | (g << 8) | (b << 16) | (a << 24);
}

// One thread per texel, 16 consecutive threads per block; texels outside the image (its width or
// height not a multiple of 4) are not written.
__global__ __launch_bounds__(256) void k_bc7_decode(const uint4* blocks, uint32_t width, uint32_t height,
                                                    uint32_t* rgba) {
    const uint32_t bw = (width + 3) / 4, bh = (height + 3) / 4;
    const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
    const uint32_t b = gid >> 4, i = gid & 15u;
    if (b >= bw * bh) return;
    const uint32_t x = (b % bw) * 4 + i % 4, y = (b / bw) * 4 + i / 4;
    const uint32_t texel = decode_texel(blocks[b], i);
    if (x < width && y < height) rgba[size_t(y) * width + x] = texel;
}

// One texel of mip level l + 1 from the 2x2 texels of level l above it (clamped to the level's
// edge when it is one texel wide or high; an odd level's last row or column is not read), with the
// reference's packager filter (Packager.cpp:216-308, NVTT's box filter): sRGB images are averaged
// as linear, alpha-premultiplied colour (decode, premultiply, average, un-premultiply where alpha
// is not 0, encode); other images average their stored values.  FP64, then rounded half up to 8
// bits.
__device__ inline double srgb_decode(double v) { return v <= 0.04045 ? v / 12.92 : pow((v + 0.055) / 1.055, 2.4); }
__device__ inline double srgb_encode(double x) { return x <= 0.0031308 ? 12.92 * x : 1.055 * pow(x, 1.0 / 2.4) - 0.055; }
__device__ inline uint32_t to8(double x) { return uint32_t(floor(fmin(fmax(x, 0.0), 1.0) * 255.0 + 0.5)); }

__global__ __launch_bounds__(256) void k_build_mip(const uint8_t* src, uint32_t sw, uint32_t sh, uint8_t* dst,
                                                   uint32_t dw, uint32_t dh, uint32_t ch, uint32_t srgb) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= dw * dh) return;
    const uint32_t x = i % dw, y = i / dw;
    const uint32_t xs[2] = {min(2 * x, sw - 1), min(2 * x + 1, sw - 1)};
    const uint32_t ys[2] = {min(2 * y, sh - 1), min(2 * y + 1, sh - 1)};
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = 0; j < 2; ++j)
        for (int k = 0; k < 2; ++k) {
            const uint8_t* t = src + (size_t(ys[j]) * sw + xs[k]) * ch;
            if (srgb && ch == 4) {
                const double a = t[3] / 255.0;
                for (int c = 0; c < 3; ++c) sum[c] += srgb_decode(t[c] / 255.0) * a;
                sum[3] += a;
            } else {
                for (uint32_t c = 0; c < ch; ++c) sum[c] += srgb ? srgb_decode(t[c] / 255.0) : t[c] / 255.0;
            }
        }
    uint8_t* o = dst + size_t(i) * ch;
    if (srgb && ch == 4) {
        const double a = sum[3] * 0.25;
        for (int c = 0; c < 3; ++c) {
            const double v = sum[c] * 0.25;
            o[c] = uint8_t(to8(srgb_encode(a > 0.0 ? v / a : v)));
        }
        o[3] = uint8_t(to8(a));
    } else {
        for (uint32_t c = 0; c < ch; ++c) o[c] = uint8_t(to8(srgb ? srgb_encode(sum[c] * 0.25) : sum[c] * 0.25));
    }
}

}  // namespace

hipError_t launch_build_mip(const uint8_t* src, uint32_t src_w, uint32_t src_h, uint8_t* dst, uint32_t channels,
                            uint32_t srgb, hipStream_t s) {
    const uint32_t dw = src_w >> 1 ? src_w >> 1 : 1u, dh = src_h >> 1 ? src_h >> 1 : 1u;
    hipLaunchKernelGGL(k_build_mip, dim3((dw * dh + 255) / 256), dim3(256), 0, s, src, src_w, src_h, dst, dw, dh,
                       channels, srgb);
    return hipGetLastError();
}

hipError_t launch_bc7_decode(const void* blocks, uint32_t width, uint32_t height, uint32_t* rgba, hipStream_t s) {
    const uint64_t n = uint64_t((width + 3) / 4) * ((height + 3) / 4) * 16;  // texels of whole blocks
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bc7_decode, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s,
                       static_cast<const uint4*>(blocks), width, height, rgba);
    return hipGetLastError();
}

}  // namespace rsy
