// Material-decoder constants and records shared by host code and kernels.
#pragma once
#include <cstddef>
#include <cstdint>

namespace rsy {

constexpr int kNmCh = 8;       // latent channels
constexpr int kNmIn = 10;      // MLP inputs: 8 latent channels + 2 positional
constexpr int kNmHidden = 32;  // hidden ReLU units
constexpr int kNmOut = 7;      // albedo rgb, normal-map r g, metallic, roughness
// FP32 parameter vector: W1 [32][10] row-major, b1 [32], W2 [7][32] row-major, b2 [7]
constexpr int kNmW1 = 0;
constexpr int kNmB1 = kNmHidden * kNmIn;           // 320
constexpr int kNmW2 = kNmB1 + kNmHidden;           // 352
constexpr int kNmB2 = kNmW2 + kNmOut * kNmHidden;  // 576
constexpr int kNmParams = kNmB2 + kNmOut;          // 583
constexpr int kNmParamsPad = 584;                  // index 583 carries a loss sum where gradients are reduced

// Bytes of the MFMA-operand form of the MLP: the FP16 A operands of both layers (3 x 64 entries of
// 16 bytes), the same as E4M3 codes (3 x 64 entries of 8 bytes), FP32 b2, two FP32 layer scales,
// the weight-form flag and padding.
constexpr size_t kNmWeightsBytes = 3 * 64 * 16 + 3 * 64 * 8 + 8 * 4 + 2 * 4 + 2 * 4;

// One decoded material: the material index and its three source maps (texture table indices).
struct NmSlot {
    uint32_t material;
    uint32_t albedo, normal, mr;
};

}  // namespace rsy
