// Material-decoder kernels: latent initialisation, training (forward, backward and Adam), FP16
// export, PSNR evaluation and point decoding.  The decoder itself is defined in neural.h.
//
// Training.  A step draws `batch` samples (u, v) and regresses the decoder onto the plain
// sampler's values there (mean squared error over the 7 channels), then applies one Adam update
// (lr with an optional cosine decay, beta1, beta2, eps) to the latents it touched and to the shared MLP.  Samples are drawn per
// workgroup inside a region of 2^log2R x 2^log2R latent cells of one material, positioned so that
// every bilinear footprint stays inside the region: the region's latents and their gradients live
// in LDS (gradients accumulate with LDS atomics, never with global ones), and the workgroup applies
// Adam to its region itself.  The regions of one step are distinct cells of one partition of every
// grid (the partition shifts by a random offset each step), so no two workgroups touch the same
// latent.  The MLP gradient is reduced per workgroup from per-sample records in LDS and summed over
// workgroups in a fixed order by one wave per parameter, which also applies Adam to it; a last
// one-thread kernel advances the device-side step counter (so a captured graph of steps replays as
// consecutive steps).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "frame_types.h"
#include "kernels.h"
#include "neural.h"

namespace rsy {
namespace {

constexpr int kMaxR = 32;   // largest region side (explicit mode: the whole grid)
constexpr int kRec = 81;    // per-sample record: features 10, hidden 32, dL/dout 7, dL/dhidden 32
constexpr int kRecIn = 0, kRecH = 10, kRecDo = 42, kRecDh = 49;

__device__ inline uint64_t mix64(uint64_t z) {  // SplitMix64 finaliser
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline float unit24(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }  // [0, 1)

// Learning-rate factor of step `step`: 0.5 (1 + cos(pi step / decay)) until `decay`, then 0.
__device__ inline float lr_factor(uint32_t step, uint32_t decay) {
    if (decay == 0) return 1.f;
    if (step >= decay) return 0.f;
    return 0.5f * (1.f + cospif(float(step) / float(decay)));
}

__device__ inline float block_sum(float v, float* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float s = 0.f;
    if (threadIdx.x == 0)
        for (int w = 0; w < int(blockDim.x >> 6); ++w) s += red[w];
    return s;
}

template <bool kExplicit>
__global__ __launch_bounds__(256) void k_nm_step(NmTrainLaunch T) {
    __shared__ float lat_s[kMaxR * kMaxR * kNmCh];
    __shared__ float grad_s[kMaxR * kMaxR * kNmCh];
    __shared__ float w_s[kNmParamsPad];
    __shared__ float rec_s[256 * kRec];
    __shared__ float red_s[4];
    const uint32_t tid = threadIdx.x;
    const uint32_t step = kExplicit ? 0u : *T.step;
    const uint32_t L = T.L, Lm = L - 1;
    const NmSlot* slots = static_cast<const NmSlot*>(T.slots);

    uint32_t slot, rx, ry, R;
    if (kExplicit) {
        slot = blockIdx.x;
        rx = ry = 0;
        R = L;
    } else {
        R = 1u << T.log2R;
        const uint32_t log2rx = T.log2L - T.log2R;
        const uint32_t rpm = 1u << (2 * log2rx);  // regions per grid
        const uint64_t n_regions = uint64_t(T.n_slots) * rpm;
        const uint64_t r = (uint64_t(step) * T.n_wg + blockIdx.x) % n_regions;
        slot = uint32_t(r % T.n_slots);
        const uint32_t reg = (uint32_t(r / T.n_slots) * 2654435761u) & (rpm - 1);  // odd multiplier: a bijection
        const uint64_t hs = mix64(T.seed ^ mix64(0x5bd1e995ull + step));
        const uint32_t ox = uint32_t(hs) & (R - 1), oy = uint32_t(hs >> 32) & (R - 1);
        rx = (reg & ((1u << log2rx) - 1)) * R + ox;
        ry = (reg >> log2rx) * R + oy;
    }
    const uint32_t n_el = R * R * kNmCh;
    const size_t grid_base = size_t(slot) * L * L * kNmCh;
    for (uint32_t i = tid; i < n_el; i += 256) {
        const uint32_t c = i & 7u, cell = i >> 3, cx = cell % R, cy = cell / R;
        const size_t g = grid_base + (size_t((ry + cy) & Lm) * L + ((rx + cx) & Lm)) * kNmCh + c;
        lat_s[i] = T.lat[g];
        grad_s[i] = 0.f;
    }
    for (uint32_t i = tid; i < uint32_t(kNmParamsPad); i += 256) w_s[i] = T.mlp[i];
    __syncthreads();

    uint32_t n_samples, s_base = 0;
    if (kExplicit) {
        s_base = T.ex_offsets[slot];
        n_samples = T.ex_offsets[slot + 1] - s_base;
    } else {
        n_samples = T.spw;
    }
    const NmSlot S = slots[slot];
    const uint64_t skey = mix64(T.seed + 0x632BE59BD9B4E019ull * (uint64_t(step) + 1)) + (uint64_t(blockIdx.x) << 32);
    float acc[3] = {0.f, 0.f, 0.f};
    float loss = 0.f;
    const float* W1 = w_s + kNmW1;
    const float* B1 = w_s + kNmB1;
    const float* W2 = w_s + kNmW2;
    const float* B2 = w_s + kNmB2;
    for (uint32_t base = 0; base < n_samples; base += 256) {
        const uint32_t q = base + tid;
        float* rec = rec_s + tid * kRec;
        if (q < n_samples) {
            float ax, ay, tgt[kNmOut];
            uint32_t i0, i1, j0, j1;
            if (kExplicit) {
                const float u = T.ex_uv[2 * size_t(s_base + q)], v = T.ex_uv[2 * size_t(s_base + q) + 1];
                const float x = u * float(L) - 0.5f, y = v * float(L) - 0.5f;
                const float fx = floorf(x), fy = floorf(y);
                ax = x - fx;
                ay = y - fy;
                i0 = uint32_t(int(fx)) & Lm;
                j0 = uint32_t(int(fy)) & Lm;
                i1 = (i0 + 1) & Lm;
                j1 = (j0 + 1) & Lm;
#pragma unroll
                for (int c = 0; c < kNmOut; ++c) tgt[c] = T.ex_target[kNmOut * size_t(s_base + q) + c];
            } else {
                const uint64_t h = mix64(skey + q);
                const float xl = unit24(uint32_t(h)) * float(R - 1), yl = unit24(uint32_t(h >> 32)) * float(R - 1);
                i0 = min(uint32_t(xl), R - 2);
                j0 = min(uint32_t(yl), R - 2);
                ax = xl - float(i0);
                ay = yl - float(j0);
                i1 = i0 + 1;
                j1 = j0 + 1;
                const float u = (float(rx) + xl + 0.5f) / float(L), v = (float(ry) + yl + 0.5f) / float(L);
                nm_target(T.texels, T.textures, S, u, v, T.srgb_lut, tgt);
            }
            const float w00 = (1.f - ax) * (1.f - ay), w10 = ax * (1.f - ay), w01 = (1.f - ax) * ay, w11 = ax * ay;
            const uint32_t c00 = (j0 * R + i0) * kNmCh, c10 = (j0 * R + i1) * kNmCh;
            const uint32_t c01 = (j1 * R + i0) * kNmCh, c11 = (j1 * R + i1) * kNmCh;
            float in[kNmIn];
#pragma unroll
            for (int c = 0; c < kNmCh; ++c)
                in[c] = w00 * lat_s[c00 + c] + w10 * lat_s[c10 + c] + w01 * lat_s[c01 + c] + w11 * lat_s[c11 + c];
            in[8] = 1.f - fabsf(2.f * ax - 1.f);
            in[9] = 1.f - fabsf(2.f * ay - 1.f);
            // forward: hidden activations go straight to the sample's LDS record
#pragma unroll 2
            for (int j = 0; j < kNmHidden; ++j) {
                float z = B1[j];
#pragma unroll
                for (int k = 0; k < kNmIn; ++k) z += W1[j * kNmIn + k] * in[k];
                rec[kRecH + j] = fmaxf(z, 0.f);
            }
            float dout[kNmOut];
#pragma unroll
            for (int c = 0; c < kNmOut; ++c) {
                float o = B2[c];
#pragma unroll 4
                for (int j = 0; j < kNmHidden; ++j) o += W2[c * kNmHidden + j] * rec[kRecH + j];
                const float d = o - tgt[c];
                loss += d * d;
                dout[c] = 2.f * d * T.inv_norm;
                rec[kRecDo + c] = dout[c];
            }
            // backward: dL/dhidden (ReLU mask), dL/dfeatures for the 8 latent channels
            float din[kNmCh] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
            for (int j = 0; j < kNmHidden; ++j) {
                float g = 0.f;
#pragma unroll
                for (int c = 0; c < kNmOut; ++c) g += W2[c * kNmHidden + j] * dout[c];
                const float dh = rec[kRecH + j] > 0.f ? g : 0.f;
                rec[kRecDh + j] = dh;
#pragma unroll
                for (int c = 0; c < kNmCh; ++c) din[c] += W1[j * kNmIn + c] * dh;
            }
#pragma unroll
            for (int c = 0; c < kNmCh; ++c) {
                atomicAdd(&grad_s[c00 + c], w00 * din[c]);
                atomicAdd(&grad_s[c10 + c], w10 * din[c]);
                atomicAdd(&grad_s[c01 + c], w01 * din[c]);
                atomicAdd(&grad_s[c11 + c], w11 * din[c]);
            }
#pragma unroll
            for (int k = 0; k < kNmIn; ++k) rec[kRecIn + k] = in[k];
        }
        __syncthreads();
        // MLP gradient components tid, tid + 256, tid + 512 over this chunk's samples
        const uint32_t cnt = min(256u, n_samples - base);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t g = tid + 256u * a;
            if (g >= uint32_t(kNmParams)) continue;
            float s = 0.f;
            if (g < uint32_t(kNmB1)) {
                const uint32_t j = g / kNmIn, k = g % kNmIn;
                for (uint32_t r = 0; r < cnt; ++r) s += rec_s[r * kRec + kRecDh + j] * rec_s[r * kRec + kRecIn + k];
            } else if (g < uint32_t(kNmW2)) {
                const uint32_t j = g - kNmB1;
                for (uint32_t r = 0; r < cnt; ++r) s += rec_s[r * kRec + kRecDh + j];
            } else if (g < uint32_t(kNmB2)) {
                const uint32_t c = (g - kNmW2) / kNmHidden, j = (g - kNmW2) % kNmHidden;
                for (uint32_t r = 0; r < cnt; ++r) s += rec_s[r * kRec + kRecDo + c] * rec_s[r * kRec + kRecH + j];
            } else {
                const uint32_t c = g - kNmB2;
                for (uint32_t r = 0; r < cnt; ++r) s += rec_s[r * kRec + kRecDo + c];
            }
            acc[a] += s;
        }
        __syncthreads();
    }
    const float loss_sum = block_sum(loss, red_s);
    // partial sums, parameter-major: partial[g * n_wg + workgroup]
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t g = tid + 256u * a;
        if (g < uint32_t(kNmParams)) T.partial[size_t(g) * T.n_wg + blockIdx.x] = acc[a];
    }
    if (tid == 0) T.partial[size_t(kNmParams) * T.n_wg + blockIdx.x] = loss_sum;

    // latents of the region: gradient out (explicit mode) or Adam
    const float t = float(step + 1);
    const float bc1 = 1.f - powf(T.beta1, t), bc2 = 1.f - powf(T.beta2, t);
    const float lr = T.lr_lat * lr_factor(step, T.decay_steps);
    for (uint32_t i = tid; i < n_el; i += 256) {
        const uint32_t c = i & 7u, cell = i >> 3, cx = cell % R, cy = cell / R;
        const size_t g = grid_base + (size_t((ry + cy) & Lm) * L + ((rx + cx) & Lm)) * kNmCh + c;
        const float gr = grad_s[i];
        if (kExplicit) {
            T.ex_grad_lat[g] = gr;
        } else {
            const float m = T.beta1 * T.lat_m[g] + (1.f - T.beta1) * gr;
            const float v = T.beta2 * T.lat_v[g] + (1.f - T.beta2) * gr * gr;
            T.lat_m[g] = m;
            T.lat_v[g] = v;
            T.lat[g] = lat_s[i] - lr * (m / bc1) / (sqrtf(v / bc2) + T.eps);
        }
    }
}

// One wave per MLP parameter (and one for the loss): sums that parameter's per-workgroup partials
// (lane-strided, then a fixed-order wave reduction), then Adam (or, in explicit mode, writes the
// summed gradient) and records the step's loss.
template <bool kExplicit>
__global__ __launch_bounds__(256) void k_nm_mlp(NmTrainLaunch T) {
    const uint32_t g = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (g >= uint32_t(kNmParamsPad)) return;
    const uint32_t step = kExplicit ? 0u : *T.step;
    const float* P = T.partial + size_t(g) * T.n_wg;
    float s = 0.f;
    for (uint32_t b = lane; b < T.n_wg; b += 64) s += P[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane != 0) return;
    if (kExplicit) {
        T.ex_grad_mlp[g] = s;
    } else if (g < uint32_t(kNmParams)) {
        const float t = float(step + 1);
        const float bc1 = 1.f - powf(T.beta1, t), bc2 = 1.f - powf(T.beta2, t);
        const float lr = T.lr_mlp * lr_factor(step, T.decay_steps);
        const float m = T.beta1 * T.mlp_m[g] + (1.f - T.beta1) * s;
        const float v = T.beta2 * T.mlp_v[g] + (1.f - T.beta2) * s * s;
        T.mlp_m[g] = m;
        T.mlp_v[g] = v;
        T.mlp[g] -= lr * (m / bc1) / (sqrtf(v / bc2) + T.eps);
    } else if (T.loss_out && step >= T.loss_base && step - T.loss_base < T.loss_cap) {
        T.loss_out[step - T.loss_base] = s * T.inv_norm;  // mean squared error of the step
    }
}

// Advances the device step counter once every reader of this step has finished.
__global__ void k_nm_advance(uint32_t* step) {
    if (threadIdx.x == 0) *step += 1u;
}

__global__ __launch_bounds__(256) void k_nm_init(NmInitLaunch I) {
    const uint64_t cells = uint64_t(I.n_slots) * I.L * I.L;
    const NmSlot* slots = static_cast<const NmSlot*>(I.slots);
    for (uint64_t cell = uint64_t(blockIdx.x) * 256 + threadIdx.x; cell < cells; cell += uint64_t(gridDim.x) * 256) {
        float* o = I.lat + cell * kNmCh;
        if (I.random) {
            const uint64_t h = mix64(I.seed ^ mix64(cell));
            const uint64_t h2 = mix64(h);
            const uint32_t r[8] = {uint32_t(h), uint32_t(h >> 32), uint32_t(h2), uint32_t(h2 >> 32),
                                   uint32_t(mix64(h2)), uint32_t(mix64(h2) >> 32), uint32_t(mix64(h2 + 1)),
                                   uint32_t(mix64(h2 + 1) >> 32)};
            for (int c = 0; c < kNmCh; ++c) o[c] = 0.2f * unit24(r[c]) - 0.1f;
        } else {
            const uint32_t slot = uint32_t(cell / (uint64_t(I.L) * I.L));
            const uint32_t local = uint32_t(cell % (uint64_t(I.L) * I.L));
            const float u = (float(local % I.L) + 0.5f) / float(I.L), v = (float(local / I.L) + 0.5f) / float(I.L);
            float tgt[kNmOut];
            nm_target(I.texels, I.textures, slots[slot], u, v, I.srgb_lut, tgt);
            for (int c = 0; c < kNmOut; ++c) o[c] = tgt[c];
            o[7] = 0.f;
        }
    }
}

__global__ __launch_bounds__(256) void k_nm_export(const float* lat, uint4* lat16, uint64_t cells) {
    for (uint64_t cell = uint64_t(blockIdx.x) * 256 + threadIdx.x; cell < cells; cell += uint64_t(gridDim.x) * 256) {
        const float4 a = reinterpret_cast<const float4*>(lat)[2 * cell];
        const float4 b = reinterpret_cast<const float4*>(lat)[2 * cell + 1];
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        const h2 p0 = h2{_Float16(a.x), _Float16(a.y)}, p1 = h2{_Float16(a.z), _Float16(a.w)};
        const h2 p2 = h2{_Float16(b.x), _Float16(b.y)}, p3 = h2{_Float16(b.z), _Float16(b.w)};
        lat16[cell] = make_uint4(__builtin_bit_cast(uint32_t, p0), __builtin_bit_cast(uint32_t, p1),
                                 __builtin_bit_cast(uint32_t, p2), __builtin_bit_cast(uint32_t, p3));
    }
}

// One block = a 16 x 16 texel tile of one slot's albedo map; all 256 lanes run the MLP.
__global__ __launch_bounds__(256) void k_nm_eval(NmEvalLaunch E) {
    __shared__ float red_s[4];
    uint32_t slot = 0;
    while (slot + 1 < E.n_slots && E.block_base[slot + 1] <= blockIdx.x) ++slot;
    const NmSlot S = static_cast<const NmSlot*>(E.slots)[slot];
    const TexGPU ta = E.textures[S.albedo];
    const uint32_t tiles_x = (ta.width + 15) / 16;
    const uint32_t tb = blockIdx.x - E.block_base[slot];
    const uint32_t tx = (tb % tiles_x) * 16 + (threadIdx.x & 15), ty = (tb / tiles_x) * 16 + (threadIdx.x >> 4);
    const bool valid = tx < ta.width && ty < ta.height;
    const float u = (float(tx) + 0.5f) / float(ta.width), v = (float(ty) + 0.5f) / float(ta.height);
    float f[kNmFeat];
    nm_features(E.lat16 + size_t(slot) * E.L * E.L, E.L, u, v, f);
    float dec[kNmOut];
    nm_mlp_wave(static_cast<const NmWeights*>(E.weights), f, dec);
    float err = 0.f;
    if (valid) {
        float tgt[kNmOut];
        nm_target(E.texels, E.textures, S, u, v, E.srgb_lut, tgt);
#pragma unroll
        for (int c = 0; c < kNmOut; ++c) err += (dec[c] - tgt[c]) * (dec[c] - tgt[c]);
    }
    const float s = block_sum(err, red_s);
    if (threadIdx.x == 0) E.partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_nm_points(NmPointLaunch P) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t ii = i < P.n ? i : P.n - 1;  // lanes past the end decode the last point (not stored)
    const uint32_t slot = P.slot[ii];
    const float u = P.uv[2 * size_t(ii)], v = P.uv[2 * size_t(ii) + 1];
    float o[kNmOut];
    if (P.targets) {
        nm_target(P.texels, P.textures, static_cast<const NmSlot*>(P.slots)[slot], u, v, P.srgb_lut, o);
    } else {
        float f[kNmFeat];
        nm_features(P.lat16 + size_t(slot) * P.L * P.L, P.L, u, v, f);
        nm_mlp_wave(static_cast<const NmWeights*>(P.weights), f, o);
    }
    if (i < P.n)
        for (int c = 0; c < kNmOut; ++c) P.out[kNmOut * size_t(i) + c] = o[c];
}

// MFMA-operand form of the MLP (see NmWeights): one wave, lane l writes its own operand entries.
// FP16: every weight rounded to FP16, scales 1.  FP8: each layer's weights (layer 1 with b1 folded
// in) divided by max|w| / 448 and rounded to OCP E4M3 (round to nearest even, the hardware
// conversion), stored as codes; inference widens them to FP16 exactly and the scale multiplies the
// layer's accumulator.
__device__ inline float layer_weight(const float* p, int layer, uint32_t i, uint32_t k) {
    if (layer == 0) return k < uint32_t(kNmIn) ? p[kNmW1 + i * kNmIn + k] : (k == uint32_t(kNmIn) ? p[kNmB1 + i] : 0.f);
    return i < uint32_t(kNmOut) ? p[kNmW2 + i * kNmHidden + k] : 0.f;
}

__global__ __launch_bounds__(64) void k_nm_export_weights(const float* p, NmWeights* W, uint32_t fp8) {
    const uint32_t l = threadIdx.x, i = l % 32, hb = l / 32;
    float scale[2] = {1.f, 1.f};
    if (fp8) {
        for (int layer = 0; layer < 2; ++layer) {
            float m = 0.f;
            for (uint32_t e = l; e < 32u * 32u; e += 64) m = fmaxf(m, fabsf(layer_weight(p, layer, e / 32, e % 32)));
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
            scale[layer] = m > 0.f ? m / 448.f : 1.f;
        }
    }
    // the 8 K-values of this lane's entry: layer-1 columns 8 hb + j; layer-2 hidden units nm_hidden(s, hb, j)
    auto put = [&](int layer, uint32_t s, h8& w, uint2& q) {
        uint32_t code[2] = {0u, 0u};
        for (uint32_t j = 0; j < 8; ++j) {
            const uint32_t k = layer == 0 ? 8 * hb + j : nm_hidden(s, hb, j);
            const float v = layer_weight(p, layer, i, k);
            if (fp8) {
                const uint32_t c = uint32_t(__builtin_amdgcn_cvt_pk_fp8_f32(v / scale[layer], 0.f, 0, false)) & 0xffu;
                code[j / 4] |= c << (8 * (j % 4));
                w[j] = _Float16(0.f);
            } else {
                w[j] = _Float16(v);
            }
        }
        q = make_uint2(code[0], code[1]);
    };
    put(0, 0, W->w1[l], W->q1[l]);
    put(1, 0, W->w2[0][l], W->q2[0][l]);
    put(1, 1, W->w2[1][l], W->q2[1][l]);
    if (l < 8) W->b2[l] = l < uint32_t(kNmOut) ? p[kNmB2 + l] : 0.f;
    // FP8: the widened codes carry a factor 2^-8 (nm_widen_e4m3)
    if (l < 2) W->scale[l] = fp8 ? scale[l] * 256.f : 1.f;
    if (l == 0) {
        W->fp8 = fp8 ? 1u : 0u;
        W->pad = 0u;
    }
}

// The inference MLP back in the parameter layout (FP32 values: operand value times layer scale).
__global__ __launch_bounds__(64) void k_nm_unpack_weights(const NmWeights* W, float* p) {
    const uint32_t l = threadIdx.x, i = l % 32, hb = l / 32;
    const bool fp8 = W->fp8 != 0;
    const h8 a1 = fp8 ? nm_widen_e4m3(W->q1[l]) : W->w1[l];
    const h8 a2[2] = {fp8 ? nm_widen_e4m3(W->q2[0][l]) : W->w2[0][l], fp8 ? nm_widen_e4m3(W->q2[1][l]) : W->w2[1][l]};
    for (uint32_t j = 0; j < 8; ++j) {
        const uint32_t k = 8 * hb + j;
        const float v = float(a1[j]) * W->scale[0];
        if (k < uint32_t(kNmIn)) p[kNmW1 + i * kNmIn + k] = v;
        else if (k == uint32_t(kNmIn)) p[kNmB1 + i] = v;
    }
    for (uint32_t s = 0; s < 2; ++s)
        for (uint32_t j = 0; j < 8; ++j)
            if (i < uint32_t(kNmOut)) p[kNmW2 + i * kNmHidden + nm_hidden(s, hb, j)] = float(a2[s][j]) * W->scale[1];
    if (l < uint32_t(kNmOut)) p[kNmB2 + l] = W->b2[l];
}

}  // namespace

hipError_t launch_nm_export_weights(const float* mlp, void* weights, uint32_t fp8, hipStream_t s) {
    hipLaunchKernelGGL(k_nm_export_weights, dim3(1), dim3(64), 0, s, mlp, static_cast<NmWeights*>(weights), fp8);
    return hipGetLastError();
}

hipError_t launch_nm_unpack_weights(const void* weights, float* mlp, hipStream_t s) {
    hipLaunchKernelGGL(k_nm_unpack_weights, dim3(1), dim3(64), 0, s, static_cast<const NmWeights*>(weights), mlp);
    return hipGetLastError();
}

hipError_t launch_nm_init(const NmInitLaunch& l, hipStream_t s) {
    const uint64_t cells = uint64_t(l.n_slots) * l.L * l.L;
    const uint32_t blocks = uint32_t(std::min<uint64_t>((cells + 255) / 256, 65536));
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nm_init, dim3(blocks), dim3(256), 0, s, l);
    return hipGetLastError();
}

hipError_t launch_nm_step(const NmTrainLaunch& l, bool explicit_batch, hipStream_t s) {
    if (explicit_batch) {
        if (l.L > uint32_t(kMaxR)) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_nm_step<true>, dim3(l.n_slots), dim3(256), 0, s, l);
        hipLaunchKernelGGL(k_nm_mlp<true>, dim3(kNmParamsPad / 4), dim3(256), 0, s, l);
    } else {
        if ((1u << l.log2R) > uint32_t(kMaxR) || l.log2R > l.log2L || l.log2R < 1) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_nm_step<false>, dim3(l.n_wg), dim3(256), 0, s, l);
        hipLaunchKernelGGL(k_nm_mlp<false>, dim3(kNmParamsPad / 4), dim3(256), 0, s, l);
        hipLaunchKernelGGL(k_nm_advance, dim3(1), dim3(64), 0, s, l.step);
    }
    return hipGetLastError();
}

hipError_t launch_nm_export(const float* lat, uint4* lat16, uint64_t cells, hipStream_t s) {
    const uint32_t blocks = uint32_t(std::min<uint64_t>((cells + 255) / 256, 65536));
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nm_export, dim3(blocks), dim3(256), 0, s, lat, lat16, cells);
    return hipGetLastError();
}

hipError_t launch_nm_eval(const NmEvalLaunch& l, uint32_t n_blocks, hipStream_t s) {
    if (n_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nm_eval, dim3(n_blocks), dim3(256), 0, s, l);
    return hipGetLastError();
}

hipError_t launch_nm_points(const NmPointLaunch& l, hipStream_t s) {
    if (l.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nm_points, dim3((l.n + 255) / 256), dim3(256), 0, s, l);
    return hipGetLastError();
}

}  // namespace rsy
