// Internal hook between the oracle's shading and decoder halves.  TEST INFRASTRUCTURE ONLY.
#pragma once
#include "oracle.h"

namespace orc {
// Decoder outputs (clamped to [0, 1]) of `material` at (u, v), in double precision.
void nm_decode_point(const orc_neural* nm, uint32_t material, double u, double v, double out[7]);
// The four latent cells (y * L + x within the material's grid) that nm_decode_point reads at (u, v).
void nm_cells(const orc_neural* nm, double u, double v, uint64_t cells[4]);
}  // namespace orc
