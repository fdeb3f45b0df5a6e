// FP64 restatement of the neural material decoder.  TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// The decoder has no counterpart in the reference; its definition is the one the HIP kernels
// implement (rosy_amd/csrc/neural.h):
//   f = [bilinear REPEAT fetch of the 8-channel latent grid at (u, v) (cell centres at
//        (i + 0.5) / L), 1 - |2 ax - 1|, 1 - |2 ay - 1|]     (ax, ay: bilinear weights)
//   h = relu(W1 f + b1) (32),  o = W2 h + b2 (7), clamped to [0, 1] when decoding;
// and the training objective is the mean over samples and the 7 channels of (o - target)^2, with
// the unclamped o.  Everything here is evaluated in double precision from the given float values.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "oracle.h"
#include "neural_oracle.h"

namespace orc {
namespace {

constexpr int kCh = 8, kIn = 10, kHid = 32, kOut = 7;
constexpr int kW1 = 0, kB1 = 320, kW2 = 352, kB2 = 576, kParams = 583;

struct Fetch {
    size_t corner[4];  // element offsets (channel 0) of the 4 cells inside the slot's grid
    double w[4];
    double ax, ay;
};

Fetch fetch(uint32_t L, double u, double v) {
    const double x = u * L - 0.5, y = v * L - 0.5;
    const double fx = std::floor(x), fy = std::floor(y);
    Fetch f;
    f.ax = x - fx;
    f.ay = y - fy;
    const long long Ll = L;
    const long long x0 = ((static_cast<long long>(fx) % Ll) + Ll) % Ll, y0 = ((static_cast<long long>(fy) % Ll) + Ll) % Ll;
    const long long x1 = (x0 + 1) % Ll, y1 = (y0 + 1) % Ll;
    f.corner[0] = size_t(y0 * Ll + x0) * kCh;
    f.corner[1] = size_t(y0 * Ll + x1) * kCh;
    f.corner[2] = size_t(y1 * Ll + x0) * kCh;
    f.corner[3] = size_t(y1 * Ll + x1) * kCh;
    f.w[0] = (1 - f.ax) * (1 - f.ay);
    f.w[1] = f.ax * (1 - f.ay);
    f.w[2] = (1 - f.ax) * f.ay;
    f.w[3] = f.ax * f.ay;
    return f;
}

void features(const float* grid, const Fetch& f, double in[kIn]) {
    for (int c = 0; c < kCh; ++c) {
        double s = 0;
        for (int k = 0; k < 4; ++k) s += f.w[k] * grid[f.corner[k] + c];
        in[c] = s;
    }
    in[8] = 1 - std::fabs(2 * f.ax - 1);
    in[9] = 1 - std::fabs(2 * f.ay - 1);
}

void mlp_forward(const float* p, const double in[kIn], double z[kHid], double h[kHid], double o[kOut]) {
    for (int j = 0; j < kHid; ++j) {
        double s = p[kB1 + j];
        for (int k = 0; k < kIn; ++k) s += double(p[kW1 + j * kIn + k]) * in[k];
        z[j] = s;
        h[j] = s > 0 ? s : 0;
    }
    for (int c = 0; c < kOut; ++c) {
        double s = p[kB2 + c];
        for (int j = 0; j < kHid; ++j) s += double(p[kW2 + c * kHid + j]) * h[j];
        o[c] = s;
    }
}

const float* slot_grid(const orc_neural* nm, uint32_t material) {
    return nm->latents + size_t(nm->slot_of[material]) * nm->latent_size * nm->latent_size * kCh;
}

}  // namespace

void nm_decode_point(const orc_neural* nm, uint32_t material, double u, double v, double out[7]) {
    double in[kIn], z[kHid], h[kHid];
    features(slot_grid(nm, material), fetch(nm->latent_size, u, v), in);
    mlp_forward(nm->mlp, in, z, h, out);
    for (int c = 0; c < kOut; ++c) out[c] = std::min(1.0, std::max(0.0, out[c]));
}

void nm_cells(const orc_neural* nm, double u, double v, uint64_t cells[4]) {
    const Fetch f = fetch(nm->latent_size, u, v);
    for (int k = 0; k < 4; ++k) cells[k] = f.corner[k] / kCh;
}

}  // namespace orc

extern "C" {

int orc_nm_decode(const orc_neural* nm, const uint32_t* material, const float* uv, uint32_t n, double* out) {
    if (!nm || !material || !uv || !out) return RSY_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < n; ++i) {
        if (material[i] >= nm->n_materials || nm->slot_of[material[i]] == 0xffffffffu) return RSY_INVALID_ARGUMENT;
        orc::nm_decode_point(nm, material[i], uv[2 * size_t(i)], uv[2 * size_t(i) + 1], out + 7 * size_t(i));
    }
    return RSY_OK;
}

int orc_nm_gradients(const orc_neural* nm, uint32_t n_slots, const uint32_t* material, const float* uv,
                     const float* target, uint32_t n, double* loss, double* grad_mlp, double* grad_latents,
                     double* min_abs_preact) {
    using namespace orc;
    if (!nm || !material || !uv || !target || n == 0) return RSY_INVALID_ARGUMENT;
    const uint32_t L = nm->latent_size;
    const size_t n_lat = size_t(n_slots) * L * L * kCh;
    std::vector<double> gm(kParams, 0.0), gl(n_lat, 0.0);
    const double norm = 1.0 / (double(n) * kOut);
    double total = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (material[i] >= nm->n_materials || nm->slot_of[material[i]] == 0xffffffffu) return RSY_INVALID_ARGUMENT;
        const size_t base = size_t(nm->slot_of[material[i]]) * L * L * kCh;
        const Fetch f = fetch(L, uv[2 * size_t(i)], uv[2 * size_t(i) + 1]);
        double in[kIn], z[kHid], h[kHid], o[kOut], dout[kOut], dh[kHid];
        features(nm->latents + base, f, in);
        mlp_forward(nm->mlp, in, z, h, o);
        double margin = std::numeric_limits<double>::infinity();
        for (int j = 0; j < kHid; ++j) margin = std::min(margin, std::fabs(z[j]));
        if (min_abs_preact) min_abs_preact[i] = margin;
        for (int c = 0; c < kOut; ++c) {
            const double d = o[c] - target[kOut * size_t(i) + c];
            total += d * d;
            dout[c] = 2 * d * norm;
            gm[kB2 + c] += dout[c];
            for (int j = 0; j < kHid; ++j) gm[kW2 + c * kHid + j] += dout[c] * h[j];
        }
        for (int j = 0; j < kHid; ++j) {
            double g = 0;
            for (int c = 0; c < kOut; ++c) g += double(nm->mlp[kW2 + c * kHid + j]) * dout[c];
            dh[j] = z[j] > 0 ? g : 0;
            gm[kB1 + j] += dh[j];
            for (int k = 0; k < kIn; ++k) gm[kW1 + j * kIn + k] += dh[j] * in[k];
        }
        for (int ch = 0; ch < kCh; ++ch) {
            double g = 0;
            for (int j = 0; j < kHid; ++j) g += double(nm->mlp[kW1 + j * kIn + ch]) * dh[j];
            for (int k = 0; k < 4; ++k) gl[base + f.corner[k] + ch] += f.w[k] * g;
        }
    }
    if (loss) *loss = total * norm;
    if (grad_mlp) std::copy(gm.begin(), gm.end(), grad_mlp);
    if (grad_latents) std::copy(gl.begin(), gl.end(), grad_latents);
    return RSY_OK;
}

}  // extern "C"
