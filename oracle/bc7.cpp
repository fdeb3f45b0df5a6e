// CPU restatement of BC7 block decoding -- test infrastructure: the oracle the library's GPU
// decoder (rosy_amd/csrc/bc7.hip) is checked against; nothing in the product links it.
//
// The reference's images are BC7 (Packager.cpp:216-308 compresses every image with
// nvtt::Format_BC7; Graphics.cpp:3620-3650 uploads VK_FORMAT_BC7_SRGB_BLOCK for colour, metallic
// and mixmap images and VK_FORMAT_BC7_UNORM_BLOCK for normal maps).  NVTT is closed and is not
// used; this follows the published BC7 block format, written as a field walk over the 128 block
// bits held as a bit array, with the partition tables as 16-entry subset lists.
#include <cstdint>
#include <cstring>

#include "oracle.h"

namespace {

struct Mode {
    int subsets, partition_bits, rotation_bits, selector_bits, color_bits, alpha_bits, endpoint_pbits, shared_pbits,
        index_bits, index2_bits;
};
const Mode kMode[8] = {{3, 4, 0, 0, 4, 0, 1, 0, 3, 0}, {2, 6, 0, 0, 6, 0, 0, 1, 3, 0}, {3, 6, 0, 0, 5, 0, 0, 0, 2, 0},
                       {2, 6, 0, 0, 7, 0, 1, 0, 2, 0}, {1, 0, 2, 1, 5, 6, 0, 0, 2, 3}, {1, 0, 2, 0, 7, 8, 0, 0, 2, 2},
                       {1, 0, 0, 0, 7, 7, 1, 0, 4, 0}, {2, 6, 0, 0, 5, 5, 1, 0, 2, 0}};

const uint8_t kTwo[64][16] = {
    {0,0,1,1,0,0,1,1,0,0,1,1,0,0,1,1},{0,0,0,1,0,0,0,1,0,0,0,1,0,0,0,1},{0,1,1,1,0,1,1,1,0,1,1,1,0,1,1,1},{0,0,0,1,0,0,1,1,0,0,1,1,0,1,1,1},
    {0,0,0,0,0,0,0,1,0,0,0,1,0,0,1,1},{0,0,1,1,0,1,1,1,0,1,1,1,1,1,1,1},{0,0,0,1,0,0,1,1,0,1,1,1,1,1,1,1},{0,0,0,0,0,0,0,1,0,0,1,1,0,1,1,1},
    {0,0,0,0,0,0,0,0,0,0,0,1,0,0,1,1},{0,0,1,1,0,1,1,1,1,1,1,1,1,1,1,1},{0,0,0,0,0,0,0,1,0,1,1,1,1,1,1,1},{0,0,0,0,0,0,0,0,0,0,0,1,0,1,1,1},
    {0,0,0,1,0,1,1,1,1,1,1,1,1,1,1,1},{0,0,0,0,0,0,0,0,1,1,1,1,1,1,1,1},{0,0,0,0,1,1,1,1,1,1,1,1,1,1,1,1},{0,0,0,0,0,0,0,0,0,0,0,0,1,1,1,1},
    {0,0,0,0,1,0,0,0,1,1,1,0,1,1,1,1},{0,1,1,1,0,0,0,1,0,0,0,0,0,0,0,0},{0,0,0,0,0,0,0,0,1,0,0,0,1,1,1,0},{0,1,1,1,0,0,1,1,0,0,0,1,0,0,0,0},
    {0,0,1,1,0,0,0,1,0,0,0,0,0,0,0,0},{0,0,0,0,1,0,0,0,1,1,0,0,1,1,1,0},{0,0,0,0,0,0,0,0,1,0,0,0,1,1,0,0},{0,1,1,1,0,0,1,1,0,0,1,1,0,0,0,1},
    {0,0,1,1,0,0,0,1,0,0,0,1,0,0,0,0},{0,0,0,0,1,0,0,0,1,0,0,0,1,1,0,0},{0,1,1,0,0,1,1,0,0,1,1,0,0,1,1,0},{0,0,1,1,0,1,1,0,0,1,1,0,1,1,0,0},
    {0,0,0,1,0,1,1,1,1,1,1,0,1,0,0,0},{0,0,0,0,1,1,1,1,1,1,1,1,0,0,0,0},{0,1,1,1,0,0,0,1,1,0,0,0,1,1,1,0},{0,0,1,1,1,0,0,1,1,0,0,1,1,1,0,0},
    {0,1,0,1,0,1,0,1,0,1,0,1,0,1,0,1},{0,0,0,0,1,1,1,1,0,0,0,0,1,1,1,1},{0,1,0,1,1,0,1,0,0,1,0,1,1,0,1,0},{0,0,1,1,0,0,1,1,1,1,0,0,1,1,0,0},
    {0,0,1,1,1,1,0,0,0,0,1,1,1,1,0,0},{0,1,0,1,0,1,0,1,1,0,1,0,1,0,1,0},{0,1,1,0,1,0,0,1,0,1,1,0,1,0,0,1},{0,1,0,1,1,0,1,0,1,0,1,0,0,1,0,1},
    {0,1,1,1,0,0,1,1,1,1,0,0,1,1,1,0},{0,0,0,1,0,0,1,1,1,1,0,0,1,0,0,0},{0,0,1,1,0,0,1,0,0,1,0,0,1,1,0,0},{0,0,1,1,1,0,1,1,1,1,0,1,1,1,0,0},
    {0,1,1,0,1,0,0,1,1,0,0,1,0,1,1,0},{0,0,1,1,1,1,0,0,1,1,0,0,0,0,1,1},{0,1,1,0,0,1,1,0,1,0,0,1,1,0,0,1},{0,0,0,0,0,1,1,0,0,1,1,0,0,0,0,0},
    {0,1,0,0,1,1,1,0,0,1,0,0,0,0,0,0},{0,0,1,0,0,1,1,1,0,0,1,0,0,0,0,0},{0,0,0,0,0,0,1,0,0,1,1,1,0,0,1,0},{0,0,0,0,0,1,0,0,1,1,1,0,0,1,0,0},
    {0,1,1,0,1,1,0,0,1,0,0,1,0,0,1,1},{0,0,1,1,0,1,1,0,1,1,0,0,1,0,0,1},{0,1,1,0,0,0,1,1,1,0,0,1,1,1,0,0},{0,0,1,1,1,0,0,1,1,1,0,0,0,1,1,0},
    {0,1,1,0,1,1,0,0,1,1,0,0,1,0,0,1},{0,1,1,0,0,0,1,1,0,0,1,1,1,0,0,1},{0,1,1,1,1,1,1,0,1,0,0,0,0,0,0,1},{0,0,0,1,1,0,0,0,1,1,1,0,0,1,1,1},
    {0,0,0,0,1,1,1,1,0,0,1,1,0,0,1,1},{0,0,1,1,0,0,1,1,1,1,1,1,0,0,0,0},{0,0,1,0,0,0,1,0,1,1,1,0,1,1,1,0},{0,1,0,0,0,1,0,0,0,1,1,1,0,1,1,1}};

const uint8_t kThree[64][16] = {
    {0,0,1,1,0,0,1,1,0,2,2,1,2,2,2,2},{0,0,0,1,0,0,1,1,2,2,1,1,2,2,2,1},{0,0,0,0,2,0,0,1,2,2,1,1,2,2,1,1},{0,2,2,2,0,0,2,2,0,0,1,1,0,1,1,1},
    {0,0,0,0,0,0,0,0,1,1,2,2,1,1,2,2},{0,0,1,1,0,0,1,1,0,0,2,2,0,0,2,2},{0,0,2,2,0,0,2,2,1,1,1,1,1,1,1,1},{0,0,1,1,0,0,1,1,2,2,1,1,2,2,1,1},
    {0,0,0,0,0,0,0,0,1,1,1,1,2,2,2,2},{0,0,0,0,1,1,1,1,1,1,1,1,2,2,2,2},{0,0,0,0,1,1,1,1,2,2,2,2,2,2,2,2},{0,0,1,2,0,0,1,2,0,0,1,2,0,0,1,2},
    {0,1,1,2,0,1,1,2,0,1,1,2,0,1,1,2},{0,1,2,2,0,1,2,2,0,1,2,2,0,1,2,2},{0,0,1,1,0,1,1,2,1,1,2,2,1,2,2,2},{0,0,1,1,2,0,0,1,2,2,0,0,2,2,2,0},
    {0,0,0,1,0,0,1,1,0,1,1,2,1,1,2,2},{0,1,1,1,0,0,1,1,2,0,0,1,2,2,0,0},{0,0,0,0,1,1,2,2,1,1,2,2,1,1,2,2},{0,0,2,2,0,0,2,2,0,0,2,2,1,1,1,1},
    {0,1,1,1,0,1,1,1,0,2,2,2,0,2,2,2},{0,0,0,1,0,0,0,1,2,2,2,1,2,2,2,1},{0,0,0,0,0,0,1,1,0,1,2,2,0,1,2,2},{0,0,0,0,1,1,0,0,2,2,1,0,2,2,1,0},
    {0,1,2,2,0,1,2,2,0,0,1,1,0,0,0,0},{0,0,1,2,0,0,1,2,1,1,2,2,2,2,2,2},{0,1,1,0,1,2,2,1,1,2,2,1,0,1,1,0},{0,0,0,0,0,1,1,0,1,2,2,1,1,2,2,1},
    {0,0,2,2,1,1,0,2,1,1,0,2,0,0,2,2},{0,1,1,0,0,1,1,0,2,0,0,2,2,2,2,2},{0,0,1,1,0,1,2,2,0,1,2,2,0,0,1,1},{0,0,0,0,2,0,0,0,2,2,1,1,2,2,2,1},
    {0,0,0,0,0,0,0,2,1,1,2,2,1,2,2,2},{0,2,2,2,0,0,2,2,0,0,1,2,0,0,1,1},{0,0,1,1,0,0,1,2,0,0,2,2,0,2,2,2},{0,1,2,0,0,1,2,0,0,1,2,0,0,1,2,0},
    {0,0,0,0,1,1,1,1,2,2,2,2,0,0,0,0},{0,1,2,0,1,2,0,1,2,0,1,2,0,1,2,0},{0,1,2,0,2,0,1,2,1,2,0,1,0,1,2,0},{0,0,1,1,2,2,0,0,1,1,2,2,0,0,1,1},
    {0,0,1,1,1,1,2,2,2,2,0,0,0,0,1,1},{0,1,0,1,0,1,0,1,2,2,2,2,2,2,2,2},{0,0,0,0,0,0,0,0,2,1,2,1,2,1,2,1},{0,0,2,2,1,1,2,2,0,0,2,2,1,1,2,2},
    {0,0,2,2,0,0,1,1,0,0,2,2,0,0,1,1},{0,2,2,0,1,2,2,1,0,2,2,0,1,2,2,1},{0,1,0,1,2,2,2,2,2,2,2,2,0,1,0,1},{0,0,0,0,2,1,2,1,2,1,2,1,2,1,2,1},
    {0,1,0,1,0,1,0,1,0,1,0,1,2,2,2,2},{0,2,2,2,0,1,1,1,0,2,2,2,0,1,1,1},{0,0,0,2,1,1,1,2,0,0,0,2,1,1,1,2},{0,0,0,0,2,1,1,2,2,1,1,2,2,1,1,2},
    {0,2,2,2,0,1,1,1,0,1,1,1,0,2,2,2},{0,0,0,2,1,1,1,2,1,1,1,2,0,0,0,2},{0,1,1,0,0,1,1,0,0,1,1,0,2,2,2,2},{0,0,0,0,0,0,0,0,2,1,1,2,2,1,1,2},
    {0,1,1,0,0,1,1,0,2,2,2,2,2,2,2,2},{0,0,2,2,0,0,1,1,0,0,1,1,0,0,2,2},{0,0,2,2,1,1,2,2,1,1,2,2,0,0,2,2},{0,0,0,0,0,0,0,0,0,0,0,0,2,1,1,2},
    {0,0,0,2,0,0,0,1,0,0,0,2,0,0,0,1},{0,2,2,2,1,2,2,2,0,2,2,2,1,2,2,2},{0,1,0,1,2,2,2,2,2,2,2,2,2,2,2,2},{0,1,1,1,2,0,1,1,2,2,0,1,2,2,2,0}};

const uint8_t kFix2[64] = {15,15,15,15,15,15,15,15,15,15,15,15,15,15,15,15,15, 2, 8, 2, 2, 8, 8,15, 2, 8, 2, 2, 8, 8, 2, 2,
                           15,15, 6, 8, 2, 8,15,15, 2, 8, 2, 2, 2,15,15, 6, 6, 2, 6, 8,15,15, 2, 2,15,15,15,15,15, 2, 2,15};
const uint8_t kFix3a[64] = { 3, 3,15,15, 8, 3,15,15, 8, 8, 6, 6, 6, 5, 3, 3, 3, 3, 8,15, 3, 3, 6,10, 5, 8, 8, 6, 8, 5,15,15,
                             8,15, 3, 5, 6,10, 8,15,15, 3,15, 5,15,15,15,15, 3,15, 5, 5, 5, 8, 5,10, 5,10, 8,13,15,12, 3, 3};
const uint8_t kFix3b[64] = {15, 8, 8, 3,15,15, 3, 8,15,15,15,15,15,15,15, 8,15, 8,15, 3,15, 8,15, 8, 3,15, 6,10,15,15,10, 8,
                            15, 3,15,10,10, 8, 9,10, 6,15, 8,15, 3, 6, 6, 8,15, 3,15,15,15,15,15,15,15,15,15,15, 3,15,15, 8};

const int kW2[4] = {0, 21, 43, 64};
const int kW3[8] = {0, 9, 18, 27, 37, 46, 55, 64};
const int kW4[16] = {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64};

int weight(int bits, int index) { return bits == 2 ? kW2[index] : bits == 3 ? kW3[index] : kW4[index]; }

struct Bits {
    uint8_t bit[128];
    int at = 0;
    explicit Bits(const uint8_t block[16]) {
        for (int i = 0; i < 128; ++i) bit[i] = (block[i / 8] >> (i % 8)) & 1;
    }
    int take(int n) {  // n bits, the first taken the least significant
        int v = 0;
        for (int k = 0; k < n; ++k) v |= bit[at++] << k;
        return v;
    }
};

int expand(int value, int bits) {  // to 8 bits: the value's bits followed by its top bits again
    int v = value << (8 - bits);
    return v | (v >> bits);
}

int subset(int subsets, int partition, int texel) {
    return subsets == 3 ? kThree[partition][texel] : subsets == 2 ? kTwo[partition][texel] : 0;
}

bool anchor(int subsets, int partition, int texel) {
    if (texel == 0) return true;
    if (subsets == 2) return texel == kFix2[partition];
    if (subsets == 3) return texel == kFix3a[partition] || texel == kFix3b[partition];
    return false;
}

void decode(const uint8_t block[16], uint8_t out[16][4]) {
    int mode = -1;
    for (int b = 0; b < 8; ++b)
        if (block[0] & (1 << b)) {
            mode = b;
            break;
        }
    if (mode < 0) {  // reserved: transparent black
        std::memset(out, 0, 64);
        return;
    }
    Bits in(block);
    in.take(mode + 1);
    const Mode& m = kMode[mode];
    const int partition = in.take(m.partition_bits);
    const int rotation = in.take(m.rotation_bits);
    const int selector = in.take(m.selector_bits);
    int e[3][2][4] = {};  // subset, endpoint, channel (raw field values first)
    for (int c = 0; c < 3; ++c)
        for (int s = 0; s < m.subsets; ++s)
            for (int k = 0; k < 2; ++k) e[s][k][c] = in.take(m.color_bits);
    if (m.alpha_bits)
        for (int s = 0; s < m.subsets; ++s)
            for (int k = 0; k < 2; ++k) e[s][k][3] = in.take(m.alpha_bits);
    int pbit[3][2] = {};
    const bool has_p = m.endpoint_pbits || m.shared_pbits;
    for (int s = 0; s < m.subsets; ++s) {
        if (m.endpoint_pbits) {
            pbit[s][0] = in.take(1);
            pbit[s][1] = in.take(1);
        } else if (m.shared_pbits) {
            pbit[s][0] = pbit[s][1] = in.take(1);
        }
    }
    const int cbits = m.color_bits + (has_p ? 1 : 0);
    const int abits = m.alpha_bits ? m.alpha_bits + (has_p ? 1 : 0) : 0;
    for (int s = 0; s < m.subsets; ++s)
        for (int k = 0; k < 2; ++k) {
            for (int c = 0; c < 3; ++c)
                e[s][k][c] = expand(has_p ? (e[s][k][c] << 1) | pbit[s][k] : e[s][k][c], cbits);
            e[s][k][3] = m.alpha_bits ? expand(has_p ? (e[s][k][3] << 1) | pbit[s][k] : e[s][k][3], abits) : 255;
        }
    int first[16], second[16] = {};
    for (int t = 0; t < 16; ++t) first[t] = in.take(m.index_bits - (anchor(m.subsets, partition, t) ? 1 : 0));
    if (m.index2_bits)
        for (int t = 0; t < 16; ++t) second[t] = in.take(m.index2_bits - (t == 0 ? 1 : 0));
    for (int t = 0; t < 16; ++t) {
        const int s = subset(m.subsets, partition, t);
        int color_index = first[t], color_bits = m.index_bits, alpha_index = first[t], alpha_bits = m.index_bits;
        if (m.index2_bits) {
            // modes 4 and 5 carry a second index set: alpha's, or colour's when the selector is set
            if (selector) {
                color_index = second[t];
                color_bits = m.index2_bits;
            } else {
                alpha_index = second[t];
                alpha_bits = m.index2_bits;
            }
        }
        int px[4];
        for (int c = 0; c < 4; ++c) {
            const int w = c < 3 ? weight(color_bits, color_index) : weight(alpha_bits, alpha_index);
            px[c] = ((64 - w) * e[s][0][c] + w * e[s][1][c] + 32) >> 6;
        }
        if (rotation) {  // rotation r exchanges alpha with channel r - 1
            const int c = rotation - 1, tmp = px[c];
            px[c] = px[3];
            px[3] = tmp;
        }
        for (int c = 0; c < 4; ++c) out[t][c] = uint8_t(px[c]);
    }
}

}  // namespace

extern "C" {

int orc_bc7_decode(const uint8_t* blocks, uint32_t width, uint32_t height, uint8_t* rgba) {
    if (!blocks || !rgba || width == 0 || height == 0) return RSY_INVALID_ARGUMENT;
    const uint32_t bw = (width + 3) / 4, bh = (height + 3) / 4;
    for (uint32_t by = 0; by < bh; ++by)
        for (uint32_t bx = 0; bx < bw; ++bx) {
            uint8_t texels[16][4];
            decode(blocks + 16 * (size_t(by) * bw + bx), texels);
            for (int t = 0; t < 16; ++t) {
                const uint32_t x = bx * 4 + t % 4, y = by * 4 + t / 4;
                if (x < width && y < height) std::memcpy(rgba + 4 * (size_t(y) * width + x), texels[t], 4);
            }
        }
    return RSY_OK;
}

int orc_bc7_tables(uint8_t* two, uint8_t* three, uint8_t* anchors) {
    if (!two || !three || !anchors) return RSY_INVALID_ARGUMENT;
    std::memcpy(two, kTwo, sizeof(kTwo));
    std::memcpy(three, kThree, sizeof(kThree));
    std::memcpy(anchors, kFix2, 64);
    std::memcpy(anchors + 64, kFix3a, 64);
    std::memcpy(anchors + 128, kFix3b, 64);
    return RSY_OK;
}

}  // extern "C"
