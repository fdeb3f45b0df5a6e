// Neural material decoder: device pieces shared by the training, evaluation and shading kernels.
//
// A material's three maps (RGBA8 sRGB albedo, RG8 normal, RG8 metallic-roughness; the inputs of
// basic.slang:172-221) are replaced by an 8-channel latent grid of L x L cells per material and one
// MLP shared by every material:
//
//   features  f = [bilinear(latent, uv) (8), tri(ax), tri(ay), 1]       tri(a) = 1 - |2a - 1|
//   hidden    h = relu(W1 f[0..9] + b1)                                  32 units
//   output    o = clamp(W2 h + b2, 0, 1)                                 7 values
//
// o = (albedo r, g, b (linear), normal-map r, g (UNORM), metallic, roughness (linear)): the values
// the plain sampler returns for those maps (sRGB decoded before filtering).  The latent grid is
// addressed like a REPEAT texture of L x L texels (cell centres at (i + 0.5) / L); ax, ay are the
// bilinear weights, so tri() is continuous across cell borders.
//
// Inference runs the MLP for the 64 pixels of a wave on the matrix cores: gfx950's
// v_mfma_f32_32x32x16_f16 with FP16 operands and FP32 accumulation, hidden units as the M rows of
// layer 1 and the pixels as its N columns.  Layer 2 takes the layer-1 accumulators as its B operand
// in the order a lane holds them (its weights are stored in that K order), so nothing moves between
// the layers.  One K=16 step covers layer 1's 16 feature columns and two cover layer 2's 32 hidden
// units.  Weights are stored in FP16, or as OCP FP8 E4M3 codes widened to FP16 exactly when loaded;
// latents in FP16; activations enter the MFMA as FP16 hi/lo pairs, so the decoder is accurate to
// FP32 level for its stored parameters.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "frame_types.h"
#include "neural_types.h"
#include "texture.h"

namespace rsy {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// MFMA operands of the MLP, one entry of 8 K-values per lane.
// Layer 1 multiplies W1' = [W1 | b1 | 0] (32 x 16: the 10 inputs, the bias column, 5 zero columns)
// by the feature columns: w1[l] holds W1'[l % 32][8 (l / 32) + j], j = 0..7.
// Layer 2 multiplies W2 (rows 7..31 zero) by the hidden columns in two K-steps, K ordered as lane l
// holds its layer-1 accumulator: w2[s][l] holds W2[l % 32][nm_hidden(s, l / 32, j)].
// q1 / q2 hold the same operands as E4M3 codes, one byte each, for FP8 weights (fp8 = 1).
// scale[] turns the accumulators into the weights' units: 1 for FP16 weights; for FP8 the per-layer
// scale times 2^8, the factor between an E4M3 code's value and its exact FP16 widening.
struct NmWeights {
    h8 w1[64];
    h8 w2[2][64];
    uint2 q1[64];
    uint2 q2[2][64];
    float b2[8];
    float scale[2];
    uint32_t fp8;
    uint32_t pad;
};
static_assert(sizeof(NmWeights) == kNmWeightsBytes, "NmWeights layout");

// Hidden unit at layer-2 K-step s, lane half hb, element j: the layer-1 accumulator entry 8s + j of
// a lane of half hb holds hidden row 8 ((8s + j) / 4) + 4 hb + (8s + j) % 4.
__host__ __device__ constexpr uint32_t nm_hidden(uint32_t s, uint32_t hb, uint32_t j) {
    return 16u * s + 8u * (j / 4u) + 4u * hb + j % 4u;
}

__device__ inline uint32_t nm_lane() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Training / evaluation target: the plain sampler's values for the 7 decoded channels at (u, v)
// (REPEAT, bilinear: the sampler state every decoded material has).
__device__ inline void nm_target(const uint8_t* texels, const TexGPU* textures, const NmSlot& s, float u, float v,
                                 const float* lut, float out[kNmOut]) {
    const TexGPU ta = textures[s.albedo], tn = textures[s.normal], tm = textures[s.mr];
    const SamplerState repeat_linear{};
    const float4 a = sample_tex(TexRef{texels + ta.offset, int(ta.width), int(ta.height), int(ta.channels), int(ta.srgb)},
                                repeat_linear, u, v, lut);
    const float4 n = sample_tex(TexRef{texels + tn.offset, int(tn.width), int(tn.height), int(tn.channels), int(tn.srgb)},
                                repeat_linear, u, v, lut);
    const float4 m = sample_tex(TexRef{texels + tm.offset, int(tm.width), int(tm.height), int(tm.channels), int(tm.srgb)},
                                repeat_linear, u, v, lut);
    out[0] = a.x;
    out[1] = a.y;
    out[2] = a.z;
    out[3] = n.x;
    out[4] = n.y;
    out[5] = m.x;  // RG8 metallic-roughness: R = metallic, G = roughness
    out[6] = m.y;
}

// Inference features of one pixel from the FP16 latent grid of its material (L x L cells of
// 8 halves = 16 bytes; L a power of two), interpolated in FP32: f[0..7] latent channels,
// f[8], f[9] positional, f[10] = 1 (carries b1).
constexpr int kNmFeat = 11;
__device__ inline void nm_features(const uint4* grid, uint32_t L, float u, float v, float f[kNmFeat]) {
    const float x = u * float(L) - 0.5f, y = v * float(L) - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    const uint32_t m = L - 1;
    const uint32_t x0 = uint32_t(int(fx)) & m, y0 = uint32_t(int(fy)) & m;
    const uint32_t x1 = (x0 + 1) & m, y1 = (y0 + 1) & m;
    const uint4 q00 = grid[y0 * L + x0], q10 = grid[y0 * L + x1];
    const uint4 q01 = grid[y1 * L + x0], q11 = grid[y1 * L + x1];
    const float w00 = (1.f - ax) * (1.f - ay), w10 = ax * (1.f - ay), w01 = (1.f - ax) * ay, w11 = ax * ay;
    const uint32_t a00[4] = {q00.x, q00.y, q00.z, q00.w}, a10[4] = {q10.x, q10.y, q10.z, q10.w};
    const uint32_t a01[4] = {q01.x, q01.y, q01.z, q01.w}, a11[4] = {q11.x, q11.y, q11.z, q11.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const _Float16* p00 = reinterpret_cast<const _Float16*>(&a00[k]);
        const _Float16* p10 = reinterpret_cast<const _Float16*>(&a10[k]);
        const _Float16* p01 = reinterpret_cast<const _Float16*>(&a01[k]);
        const _Float16* p11 = reinterpret_cast<const _Float16*>(&a11[k]);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            f[2 * k + j] = w00 * float(p00[j]) + w10 * float(p10[j]) + w01 * float(p01[j]) + w11 * float(p11[j]);
    }
    f[8] = 1.f - fabsf(2.f * ax - 1.f);
    f[9] = 1.f - fabsf(2.f * ay - 1.f);
    f[10] = 1.f;
}

// FP32 values as unevaluated sums of two halves: hi = fp16(x), lo = fp16(x - hi) (|error| about
// 2^-22 |x|).  Products with FP16 weights are exact in the MFMA's FP32 accumulation, so feeding hi
// and lo as two B operands gives W x to FP32-level accuracy for the FP16 weights.
__device__ inline void nm_split8(const float (&x)[8], h8& hi, h8& lo) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        hi[k] = _Float16(x[k]);
        lo[k] = _Float16(x[k] - float(hi[k]));
    }
}

// Eight E4M3 codes widened to FP16: sign to bit 15, exponent and mantissa to bits 13..7.  Exact for
// every code (normal and subnormal); the FP16 value is the code's value times 2^-8.
__device__ inline h8 nm_widen_e4m3(uint2 q) {
    h8 r;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t b = ((k < 4 ? q.x : q.y) >> (8 * (k & 3))) & 0xffu;
        r[k] = __builtin_bit_cast(_Float16, uint16_t(((b & 0x80u) << 8) | ((b & 0x7fu) << 7)));
    }
    return r;
}

__device__ inline h8 nm_swap_halves(h8 v) {  // value of lane l ^ 32
    const int addr = int((nm_lane() ^ 32u) << 2);
    uint4 b = __builtin_bit_cast(uint4, v);
    b.x = uint32_t(__builtin_amdgcn_ds_bpermute(addr, int(b.x)));
    b.y = uint32_t(__builtin_amdgcn_ds_bpermute(addr, int(b.y)));
    b.z = uint32_t(__builtin_amdgcn_ds_bpermute(addr, int(b.z)));
    b.w = uint32_t(__builtin_amdgcn_ds_bpermute(addr, int(b.w)));
    return __builtin_bit_cast(h8, b);
}

__device__ inline float nm_swap_f(float v) {
    const int addr = int((nm_lane() ^ 32u) << 2);
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

// MLP of the 64 pixels of a wave, one pixel per lane (features f from nm_features); every lane of
// the wave must call it with the wave fully active.  Returns this lane's 7 clamped outputs.
//
// Pixel tile t (t = 0, 1) is lanes 32t .. 32t + 31.  The B operand of layer 1 for tile t, lane l is
// features 8 (l / 32) .. +7 of pixel 32t + l % 32: a lane holds its own pixel's features and swaps
// the other half with lane l ^ 32.  The layer-1 accumulator of lane l holds hidden rows
// 8 (r / 4) + 4 (l / 32) + r % 4 (r = 0..15) of pixel l % 32; after ReLU, entries 8s .. 8s + 7 are
// layer 2's B operand for K-step s.  Output rows 0..7 land in accumulator entries 0..3: rows 0-3 in
// lanes 0-31 and rows 4-7 in lanes 32-63 of the same column.  Every B operand goes in as an FP16
// hi / lo pair (nm_split8).
__device__ inline void nm_mlp_wave(const NmWeights* __restrict__ W, const float f[kNmFeat], float out[kNmOut]) {
    const uint32_t lane = nm_lane();
    const bool hi = lane >= 32u;
    h8 a1, a2[2];
    if (W->fp8) {  // uniform: one weight form per decoder
        a1 = nm_widen_e4m3(W->q1[lane]);
        a2[0] = nm_widen_e4m3(W->q2[0][lane]);
        a2[1] = nm_widen_e4m3(W->q2[1][lane]);
    } else {
        a1 = W->w1[lane];
        a2[0] = W->w2[0][lane];
        a2[1] = W->w2[1][lane];
    }
    // feature halves: 0..7 the latent channels; 8..15 the positional pair, the constant 1, zeros
    const float g0[8] = {f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7]};
    const float g1[8] = {f[8], f[9], f[10], 0.f, 0.f, 0.f, 0.f, 0.f};
    h8 q0h, q0l, q1h, q1l;
    nm_split8(g0, q0h, q0l);
    nm_split8(g1, q1h, q1l);
    // tile t wants half (l / 32) of pixel 32t + l % 32: the lane's own features for t == l / 32,
    // otherwise those of lane l ^ 32 (which sends the half this lane needs)
    const h8 own_h = hi ? q1h : q0h, own_l = hi ? q1l : q0l;
    const h8 far_h = nm_swap_halves(hi ? q0h : q1h), far_l = nm_swap_halves(hi ? q0l : q1l);
    const float s1 = W->scale[0], s2 = W->scale[1];
    float res[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const bool own = (t == 1) == hi;
        f32x16 acc = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, own ? own_h : far_h, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, own ? own_l : far_l, acc, 0, 0, 0);
        f32x16 acc2 = {};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            float h[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) h[j] = fmaxf(acc[8 * s + j] * s1, 0.f);
            h8 hh, hl;
            nm_split8(h, hh, hl);
            acc2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a2[s], hh, acc2, 0, 0, 0);
            acc2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a2[s], hl, acc2, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) res[t][r] = acc2[r];
    }
    // own pixel = tile (hi ? 1 : 0): rows 4*hi .. +3 here, the other four rows in lane l ^ 32
    float own_rows[4], far_rows[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        own_rows[r] = hi ? res[1][r] : res[0][r];
        far_rows[r] = nm_swap_f(hi ? res[0][r] : res[1][r]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float lo4 = hi ? far_rows[r] : own_rows[r];  // rows 0..3
        const float hi4 = hi ? own_rows[r] : far_rows[r];  // rows 4..7
        out[r] = lo4;
        if (r < 3) out[4 + r] = hi4;
    }
#pragma unroll
    for (int c = 0; c < kNmOut; ++c) out[c] = fminf(1.f, fmaxf(0.f, out[c] * s2 + W->b2[c]));
}

}  // namespace rsy
