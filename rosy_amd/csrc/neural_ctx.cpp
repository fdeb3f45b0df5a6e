// Host side of the neural material decoder behind the C ABI: eligibility, parameter buffers,
// the training loop (HIP graphs of Adam steps), FP16 export, PSNR evaluation and the active set.
// The decoder has no counterpart in the reference; the plain maps it replaces are sampled as in
// basic.slang:172-221 (albedo, normal map, metallic-roughness).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "context.h"
#include "kernels.h"
#include "neural_types.h"
#include "rosy_amd.h"

namespace rsy {

struct NeuralState {
    rsy_neural_config cfg{};
    uint32_t L = 0, log2L = 0, log2R = 0;
    std::vector<NmSlot> slots;      // eligible materials, in material order
    std::vector<uint32_t> slot_of;  // per material: slot or 0xFFFFFFFF
    std::vector<uint32_t> active;   // per material
    std::vector<float> psnr;        // per material (NaN: not eligible / not evaluated)
    std::vector<uint32_t> block_base;  // evaluation blocks per slot (prefix, n_slots + 1)
    DevBuf slots_d, slot_of_d, active_d, block_base_d;
    DevBuf lat, lat_m, lat_v, mlp, mlp_m, mlp_v, partial, step_d, loss_d, lat16, weights_d, eval_partial;
    DevBuf tmp[5];
    uint32_t n_wg = 0, spw = 0;
    uint32_t steps = 0;
    bool evaluated = false;
    float train_ms = 0.f, eval_ms = 0.f, last_loss = NAN;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ~NeuralState() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
    uint64_t n_latents() const { return uint64_t(slots.size()) * L * L * kNmCh; }
};

void destroy_neural(NeuralState* n) { delete n; }

bool neural_ready(const rsy_ctx* c) { return c->neural && c->neural->evaluated; }

void neural_shade_args(const rsy_ctx* c, ShadeLaunch& sh) {
    const NeuralState* n = c->neural;
    sh.nm_slot = n->slot_of_d.as<uint32_t>();
    sh.nm_active = n->active_d.as<uint32_t>();
    sh.nm_lat16 = n->lat16.as<uint4>();
    sh.nm_weights = n->weights_d.p;
    sh.nm_L = n->L;
    sh.neural = 1;
}

namespace {

bool pow2(uint32_t v) { return v && !(v & (v - 1)); }

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float uniform(uint64_t& s, float lo, float hi) { return lo + (hi - lo) * float(double(splitmix(s) >> 11) * 0x1.0p-53); }

// A material is decodable when its three maps are the formats the decoder reproduces, single-level
// (the decoder reproduces level 0 and has no level of detail), and every one of them is sampled
// bilinearly with REPEAT addressing (the sampler state the training targets use): opaque
// four-channel albedo, two-channel normal map, two-channel metallic-roughness (mixmap when present,
// as basic.slang:198-221 prefers it).
bool eligible(const rsy_ctx* c, uint32_t m, NmSlot& out) {
    const rsy_gpu_material& md = c->h_materials[m];
    auto usable = [&](uint32_t img, uint32_t smp, uint32_t channels) {
        if (img == RSY_ABSENT || img >= c->n_textures || c->h_tex_table[img].channels != channels) return false;
        if (c->h_tex_table[img].levels > 1) return false;
        if (smp < c->n_samplers) {
            const rsy_sampler& s = c->h_samplers[smp];
            if (s.mag_filter <= 2 || s.wrap_s != 2 || s.wrap_t != 2) return false;
        }
        return true;
    };
    const uint32_t mix = md.mixmap_sampled_image_index;
    const bool use_mix = mix != RSY_ABSENT && mix < c->n_textures;
    const uint32_t mr = use_mix ? mix : md.metallic_sampled_image_index;
    const uint32_t mrs = use_mix ? md.mixmap_sampler_index : md.metallic_sampler_index;
    if (!usable(md.color_sampled_image_index, md.color_sampler_index, 4) || !c->tex_opaque[md.color_sampled_image_index])
        return false;
    if (!usable(md.normal_sampled_image_index, md.normal_sampler_index, 2)) return false;
    if (!usable(mr, mrs, 2)) return false;
    out = NmSlot{m, md.color_sampled_image_index, md.normal_sampled_image_index, mr};
    return true;
}

// Starting MLP.  From maps: hidden pairs (2c, 2c + 1) = relu(+f_c), relu(-f_c) and output
// c = h_2c - h_2c+1, so the decoder starts as the bilinear interpolation of latent channels 0..6
// (which hold the plain values at the cell centres); units 14..31 start with small random input
// weights and no output weight.  Random: He-uniform layers, output bias 0.5.
std::vector<float> initial_mlp(uint32_t mode, uint64_t seed) {
    std::vector<float> p(kNmParamsPad, 0.f);
    uint64_t s = seed ^ 0x6A09E667F3BCC909ull;
    if (mode == RSY_NEURAL_INIT_FROM_MAPS) {
        for (int c = 0; c < kNmOut; ++c) {
            p[kNmW1 + (2 * c) * kNmIn + c] = 1.f;
            p[kNmW1 + (2 * c + 1) * kNmIn + c] = -1.f;
            p[kNmW2 + c * kNmHidden + 2 * c] = 1.f;
            p[kNmW2 + c * kNmHidden + 2 * c + 1] = -1.f;
        }
        for (int j = 2 * kNmOut; j < kNmHidden; ++j)
            for (int k = 0; k < kNmIn; ++k) p[kNmW1 + j * kNmIn + k] = uniform(s, -0.1f, 0.1f);
    } else {
        const float b1 = std::sqrt(6.f / kNmIn), b2 = std::sqrt(6.f / kNmHidden) * 0.5f;
        for (int i = 0; i < kNmHidden * kNmIn; ++i) p[kNmW1 + i] = uniform(s, -b1, b1);
        for (int i = 0; i < kNmOut * kNmHidden; ++i) p[kNmW2 + i] = uniform(s, -b2, b2);
        for (int c = 0; c < kNmOut; ++c) p[kNmB2 + c] = 0.5f;
    }
    return p;
}

NmTrainLaunch train_args(rsy_ctx* c, NeuralState* n) {
    NmTrainLaunch T{};
    T.lat = n->lat.as<float>();
    T.lat_m = n->lat_m.as<float>();
    T.lat_v = n->lat_v.as<float>();
    T.mlp = n->mlp.as<float>();
    T.mlp_m = n->mlp_m.as<float>();
    T.mlp_v = n->mlp_v.as<float>();
    T.partial = n->partial.as<float>();
    T.step = n->step_d.as<uint32_t>();
    T.slots = n->slots_d.p;
    T.n_slots = uint32_t(n->slots.size());
    T.textures = c->tex_table.as<TexGPU>();
    T.texels = c->texels.as<uint8_t>();
    T.srgb_lut = c->srgb_lut.as<float>();
    T.L = n->L;
    T.log2L = n->log2L;
    T.log2R = n->log2R;
    T.n_wg = n->n_wg;
    T.spw = n->spw;
    T.inv_norm = 1.f / (float(n->n_wg) * float(n->spw) * float(kNmOut));
    T.lr_mlp = n->cfg.lr_mlp;
    T.lr_lat = n->cfg.lr_latent;
    T.beta1 = n->cfg.beta1;
    T.beta2 = n->cfg.beta2;
    T.eps = n->cfg.eps;
    T.decay_steps = n->cfg.decay_steps;
    T.seed = n->cfg.seed;
    return T;
}

rsy_result need_neural(rsy_ctx* c) {
    if (!c) return RSY_INVALID_ARGUMENT;
    if (!c->neural) return fail(c, RSY_INVALID_STATE, "rsy_neural_init has not been called");
    return RSY_OK;
}

}  // namespace
}  // namespace rsy

using namespace rsy;

extern "C" {

void rsy_neural_default_config(rsy_neural_config* cfg) {
    if (!cfg) return;
    *cfg = rsy_neural_config{};
    cfg->latent_size = 256;
    cfg->batch = 65536;
    cfg->init = RSY_NEURAL_INIT_FROM_MAPS;
    cfg->region = 16;
    cfg->lr_mlp = 1e-3f;
    cfg->lr_latent = 1e-3f;
    cfg->beta1 = 0.9f;
    cfg->beta2 = 0.999f;
    cfg->eps = 1e-8f;
    cfg->psnr_threshold = 40.f;
    cfg->seed = 1;
}

rsy_result rsy_neural_init(rsy_ctx* c, const rsy_neural_config* cfg) {
    if (!c || !cfg) return RSY_INVALID_ARGUMENT;
    if (!pow2(cfg->latent_size) || cfg->latent_size < 4 || cfg->latent_size > 4096)
        return fail(c, RSY_INVALID_ARGUMENT, "latent_size must be a power of two in [4, 4096]");
    if (cfg->batch == 0 || cfg->batch % 256 != 0) return fail(c, RSY_INVALID_ARGUMENT, "batch must be a positive multiple of 256");
    if (!pow2(cfg->region) || cfg->region < 2 || cfg->region > 32 || cfg->region > cfg->latent_size)
        return fail(c, RSY_INVALID_ARGUMENT, "region must be a power of two in [2, min(32, latent_size)]");
    if (cfg->init > RSY_NEURAL_INIT_RANDOM) return fail(c, RSY_INVALID_ARGUMENT, "unknown init mode");
    if (cfg->weights > RSY_NEURAL_WEIGHTS_FP8) return fail(c, RSY_INVALID_ARGUMENT, "unknown weight format");
    if (!(cfg->lr_mlp >= 0.f) || !(cfg->lr_latent >= 0.f) || !(cfg->beta1 >= 0.f && cfg->beta1 < 1.f) ||
        !(cfg->beta2 >= 0.f && cfg->beta2 < 1.f) || !(cfg->eps > 0.f))
        return fail(c, RSY_INVALID_ARGUMENT, "invalid Adam hyper-parameters");
    if (c->h_materials.empty()) return fail(c, RSY_INVALID_STATE, "rsy_set_materials has not been called");
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    destroy_neural(c->neural);
    c->neural = nullptr;
    invalidate_graph(c);
    NeuralState* n = new (std::nothrow) NeuralState();
    if (!n) return RSY_ALLOCATION_FAILURE;
    c->neural = n;  // owned by the context from here (freed on error paths by the next init/destroy)
    n->cfg = *cfg;
    n->L = cfg->latent_size;
    n->log2L = ilog2(n->L);
    n->log2R = ilog2(cfg->region);
    const uint32_t nm = c->n_materials;
    n->slot_of.assign(nm, 0xffffffffu);
    n->active.assign(nm, 0u);
    n->psnr.assign(nm, NAN);
    for (uint32_t m = 0; m < nm; ++m) {
        NmSlot s;
        if (eligible(c, m, s)) {
            n->slot_of[m] = uint32_t(n->slots.size());
            n->slots.push_back(s);
        }
    }
    const uint32_t ns = uint32_t(n->slots.size());
    const uint64_t cells = uint64_t(ns) * n->L * n->L;
    const uint64_t rpm = uint64_t(n->L / cfg->region) * (n->L / cfg->region);
    const uint64_t n_regions = std::max<uint64_t>(1, uint64_t(ns) * rpm);
    n->n_wg = uint32_t(std::min<uint64_t>(cfg->batch / 256, n_regions));
    n->spw = (cfg->batch + n->n_wg - 1) / n->n_wg;
    // evaluation blocks: 16 x 16 texel tiles of each slot's albedo map
    n->block_base.assign(ns + 1, 0);
    for (uint32_t i = 0; i < ns; ++i) {
        const TexGPU& t = c->h_tex_table[n->slots[i].albedo];
        n->block_base[i + 1] = n->block_base[i] + ((t.width + 15) / 16) * ((t.height + 15) / 16);
    }
    hipStream_t s = c->stream;
    HIP_TRY(c, hipEventCreate(&n->ev0), "event");
    HIP_TRY(c, hipEventCreate(&n->ev1), "event");
    HIP_TRY(c, n->slots_d.ensure(std::max<size_t>(1, ns) * sizeof(NmSlot)), "allocate decoder");
    HIP_TRY(c, n->slot_of_d.ensure(std::max<size_t>(1, nm) * 4), "allocate decoder");
    HIP_TRY(c, n->active_d.ensure(std::max<size_t>(1, nm) * 4), "allocate decoder");
    HIP_TRY(c, n->block_base_d.ensure((ns + 1) * 4), "allocate decoder");
    const size_t lat_bytes = std::max<uint64_t>(1, cells) * kNmCh * 4;
    HIP_TRY(c, n->lat.ensure(lat_bytes), "allocate latents");
    HIP_TRY(c, n->lat_m.ensure(lat_bytes), "allocate latents");
    HIP_TRY(c, n->lat_v.ensure(lat_bytes), "allocate latents");
    HIP_TRY(c, n->lat16.ensure(std::max<uint64_t>(1, cells) * 16), "allocate latents");
    HIP_TRY(c, n->mlp.ensure(kNmParamsPad * 4), "allocate decoder");
    HIP_TRY(c, n->mlp_m.ensure(kNmParamsPad * 4), "allocate decoder");
    HIP_TRY(c, n->mlp_v.ensure(kNmParamsPad * 4), "allocate decoder");
    HIP_TRY(c, n->partial.ensure(size_t(std::max<uint32_t>(n->n_wg, std::max<uint32_t>(1, ns))) * kNmParamsPad * 4),
            "allocate decoder");
    HIP_TRY(c, n->step_d.ensure(4), "allocate decoder");
    HIP_TRY(c, n->weights_d.ensure(kNmWeightsBytes), "allocate decoder");
    HIP_TRY(c, n->eval_partial.ensure(std::max<uint32_t>(1, n->block_base[ns]) * 4), "allocate decoder");
    if (ns) HIP_TRY(c, hipMemcpy(n->slots_d.p, n->slots.data(), ns * sizeof(NmSlot), hipMemcpyHostToDevice), "upload");
    if (nm) {
        HIP_TRY(c, hipMemcpy(n->slot_of_d.p, n->slot_of.data(), nm * 4, hipMemcpyHostToDevice), "upload");
        HIP_TRY(c, hipMemcpy(n->active_d.p, n->active.data(), nm * 4, hipMemcpyHostToDevice), "upload");
    }
    HIP_TRY(c, hipMemcpy(n->block_base_d.p, n->block_base.data(), (ns + 1) * 4, hipMemcpyHostToDevice), "upload");
    const std::vector<float> p = initial_mlp(cfg->init, cfg->seed);
    HIP_TRY(c, hipMemcpy(n->mlp.p, p.data(), kNmParamsPad * 4, hipMemcpyHostToDevice), "upload");
    HIP_TRY(c, hipMemsetAsync(n->mlp_m.p, 0, kNmParamsPad * 4, s), "clear");
    HIP_TRY(c, hipMemsetAsync(n->mlp_v.p, 0, kNmParamsPad * 4, s), "clear");
    HIP_TRY(c, hipMemsetAsync(n->lat_m.p, 0, lat_bytes, s), "clear");
    HIP_TRY(c, hipMemsetAsync(n->lat_v.p, 0, lat_bytes, s), "clear");
    HIP_TRY(c, hipMemsetAsync(n->step_d.p, 0, 4, s), "clear");
    NmInitLaunch I{};
    I.lat = n->lat.as<float>();
    I.n_slots = ns;
    I.L = n->L;
    I.random = cfg->init == RSY_NEURAL_INIT_RANDOM ? 1u : 0u;
    I.seed = cfg->seed;
    I.slots = n->slots_d.p;
    I.textures = c->tex_table.as<TexGPU>();
    I.texels = c->texels.as<uint8_t>();
    I.srgb_lut = c->srgb_lut.as<float>();
    HIP_TRY(c, launch_nm_init(I, s), "initialise latents");
    HIP_TRY(c, hipStreamSynchronize(s), "synchronize");
    return RSY_OK;
}

rsy_result rsy_neural_train(rsy_ctx* c, uint32_t n_steps, float* loss_history) {
    rsy_result r = need_neural(c);
    if (r != RSY_OK) return r;
    NeuralState* n = c->neural;
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    n->evaluated = false;  // exported parameters are stale from here
    if (n_steps == 0 || n->slots.empty()) return RSY_OK;
    HIP_TRY(c, n->loss_d.ensure(size_t(n_steps) * 4), "allocate loss history");
    NmTrainLaunch T = train_args(c, n);
    T.loss_out = n->loss_d.as<float>();
    T.loss_base = n->steps;
    T.loss_cap = n_steps;
    hipStream_t s = c->stream;
    // a graph of up to 64 consecutive steps (the step counter lives on the device), replayed
    const uint32_t chunk = std::min<uint32_t>(n_steps, 64);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIP_TRY(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    hipError_t e = hipSuccess;
    for (uint32_t i = 0; i < chunk && e == hipSuccess; ++i) e = launch_nm_step(T, false, s);
    hipError_t e2 = hipStreamEndCapture(s, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(c, e != hipSuccess ? e : e2, "capture training steps");
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(c, e, "instantiate training graph");
    HIP_TRY(c, hipEventRecord(n->ev0, s), "event");
    uint32_t done = 0;
    for (; done + chunk <= n_steps && e == hipSuccess; done += chunk) e = hipGraphLaunch(exec, s);
    for (; done < n_steps && e == hipSuccess; ++done) e = launch_nm_step(T, false, s);
    if (e == hipSuccess) e = hipEventRecord(n->ev1, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipGraphExecDestroy(exec);
    if (e != hipSuccess) return hip_fail(c, e, "training steps");
    HIP_TRY(c, hipEventElapsedTime(&n->train_ms, n->ev0, n->ev1), "event");
    n->steps += n_steps;
    std::vector<float> loss(n_steps);
    HIP_TRY(c, hipMemcpy(loss.data(), n->loss_d.p, size_t(n_steps) * 4, hipMemcpyDeviceToHost), "read loss");
    n->last_loss = loss.back();
    if (loss_history) std::memcpy(loss_history, loss.data(), size_t(n_steps) * 4);
    return RSY_OK;
}

rsy_result rsy_neural_evaluate(rsy_ctx* c, float* psnr, uint32_t n_materials) {
    rsy_result r = need_neural(c);
    if (r != RSY_OK) return r;
    NeuralState* n = c->neural;
    if (psnr && n_materials != c->n_materials) return fail(c, RSY_INVALID_ARGUMENT, "psnr array size != material count");
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    hipStream_t s = c->stream;
    HIP_TRY(c, hipStreamSynchronize(s), "synchronize");
    const uint32_t ns = uint32_t(n->slots.size());
    HIP_TRY(c, hipEventRecord(n->ev0, s), "event");
    HIP_TRY(c, launch_nm_export(n->lat.as<float>(), n->lat16.as<uint4>(), uint64_t(ns) * n->L * n->L, s), "export latents");
    HIP_TRY(c, launch_nm_export_weights(n->mlp.as<float>(), n->weights_d.p, n->cfg.weights == RSY_NEURAL_WEIGHTS_FP8 ? 1u : 0u, s),
            "export weights");
    NmEvalLaunch E{};
    E.lat16 = n->lat16.as<uint4>();
    E.weights = n->weights_d.p;
    E.L = n->L;
    E.slots = n->slots_d.p;
    E.n_slots = ns;
    E.block_base = n->block_base_d.as<uint32_t>();
    E.textures = c->tex_table.as<TexGPU>();
    E.texels = c->texels.as<uint8_t>();
    E.srgb_lut = c->srgb_lut.as<float>();
    E.partial = n->eval_partial.as<float>();
    const uint32_t blocks = n->block_base[ns];
    HIP_TRY(c, launch_nm_eval(E, blocks, s), "evaluate decoder");
    HIP_TRY(c, hipEventRecord(n->ev1, s), "event");
    std::vector<float> part(blocks);
    if (blocks) HIP_TRY(c, hipMemcpyAsync(part.data(), n->eval_partial.p, size_t(blocks) * 4, hipMemcpyDeviceToHost, s), "read");
    HIP_TRY(c, hipStreamSynchronize(s), "synchronize");
    HIP_TRY(c, hipEventElapsedTime(&n->eval_ms, n->ev0, n->ev1), "event");
    for (uint32_t i = 0; i < ns; ++i) {
        double sum = 0.0;
        for (uint32_t b = n->block_base[i]; b < n->block_base[i + 1]; ++b) sum += part[b];
        const TexGPU& t = c->h_tex_table[n->slots[i].albedo];
        const double mse = sum / (double(t.width) * t.height * kNmOut);
        const double p = mse > 0.0 ? 10.0 * std::log10(1.0 / mse) : INFINITY;
        const uint32_t m = n->slots[i].material;
        n->psnr[m] = float(p);
        n->active[m] = p >= double(n->cfg.psnr_threshold) ? 1u : 0u;
    }
    if (c->n_materials)
        HIP_TRY(c, hipMemcpy(n->active_d.p, n->active.data(), size_t(c->n_materials) * 4, hipMemcpyHostToDevice), "upload");
    n->evaluated = true;
    if (psnr) std::memcpy(psnr, n->psnr.data(), size_t(n_materials) * 4);
    return RSY_OK;
}

rsy_result rsy_neural_get_stats(rsy_ctx* c, rsy_neural_stats* out) {
    if (!out) return RSY_INVALID_ARGUMENT;
    rsy_result r = need_neural(c);
    if (r != RSY_OK) return r;
    const NeuralState* n = c->neural;
    rsy_neural_stats st{};
    st.n_materials = c->n_materials;
    st.n_eligible = uint32_t(n->slots.size());
    st.n_active = uint32_t(std::count(n->active.begin(), n->active.end(), 1u));
    st.steps = n->steps;
    st.train_ms = n->train_ms;
    st.eval_ms = n->eval_ms;
    st.last_loss = n->last_loss;
    *out = st;
    return RSY_OK;
}

rsy_result rsy_neural_set_active(rsy_ctx* c, const uint32_t* active, uint32_t n_materials) {
    rsy_result r = need_neural(c);
    if (r != RSY_OK) return r;
    NeuralState* n = c->neural;
    if (!active || n_materials != c->n_materials) return fail(c, RSY_INVALID_ARGUMENT, "active array size != material count");
    if (!n->evaluated) return fail(c, RSY_INVALID_STATE, "rsy_neural_evaluate has not been called");
    for (uint32_t m = 0; m < n_materials; ++m)
        if (active[m] > 1u || (active[m] && n->slot_of[m] == 0xffffffffu))
            return fail(c, RSY_INVALID_ARGUMENT, "only eligible materials can be decoded");
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    n->active.assign(active, active + n_materials);
    if (n_materials) HIP_TRY(c, hipMemcpy(n->active_d.p, active, size_t(n_materials) * 4, hipMemcpyHostToDevice), "upload");
    return RSY_OK;
}

rsy_result rsy_neural_get_params(rsy_ctx* c, float* latents, uint64_t n_latents, float* mlp, uint32_t n_mlp) {
    rsy_result r = need_neural(c);
    if (r != RSY_OK) return r;
    NeuralState* n = c->neural;
    if ((latents && n_latents != n->n_latents()) || (mlp && n_mlp != uint32_t(kNmParams)))
        return fail(c, RSY_INVALID_ARGUMENT, "parameter array size mismatch");
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    if (latents && n_latents) HIP_TRY(c, hipMemcpy(latents, n->lat.p, n_latents * 4, hipMemcpyDeviceToHost), "read latents");
    if (mlp) HIP_TRY(c, hipMemcpy(mlp, n->mlp.p, kNmParams * 4, hipMemcpyDeviceToHost), "read mlp");
    return RSY_OK;
}

rsy_result rsy_neural_set_params(rsy_ctx* c, const float* latents, uint64_t n_latents, const float* mlp, uint32_t n_mlp) {
    rsy_result r = need_neural(c);
    if (r != RSY_OK) return r;
    NeuralState* n = c->neural;
    if ((latents && n_latents != n->n_latents()) || (mlp && n_mlp != uint32_t(kNmParams)))
        return fail(c, RSY_INVALID_ARGUMENT, "parameter array size mismatch");
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    if (latents && n_latents) HIP_TRY(c, hipMemcpy(n->lat.p, latents, n_latents * 4, hipMemcpyHostToDevice), "write latents");
    if (mlp) HIP_TRY(c, hipMemcpy(n->mlp.p, mlp, kNmParams * 4, hipMemcpyHostToDevice), "write mlp");
    // training resumes from these parameters with fresh Adam moments
    const size_t lat_bytes = std::max<uint64_t>(1, n->n_latents()) * 4;
    HIP_TRY(c, hipMemset(n->lat_m.p, 0, lat_bytes), "clear");
    HIP_TRY(c, hipMemset(n->lat_v.p, 0, lat_bytes), "clear");
    HIP_TRY(c, hipMemset(n->mlp_m.p, 0, kNmParamsPad * 4), "clear");
    HIP_TRY(c, hipMemset(n->mlp_v.p, 0, kNmParamsPad * 4), "clear");
    HIP_TRY(c, hipMemset(n->step_d.p, 0, 4), "clear");
    n->steps = 0;
    n->evaluated = false;
    return RSY_OK;
}

rsy_result rsy_neural_gradients(rsy_ctx* c, const uint32_t* material, const float* uv, const float* target, uint32_t n_pts,
                                float* loss, float* grad_mlp, float* grad_latents, uint64_t n_latents) {
    rsy_result r = need_neural(c);
    if (r != RSY_OK) return r;
    NeuralState* n = c->neural;
    if (n_pts == 0 || !material || !uv || !target) return fail(c, RSY_INVALID_ARGUMENT, "empty or null batch");
    if (n->L > 32) return fail(c, RSY_INVALID_ARGUMENT, "gradients of the whole grid need latent_size <= 32");
    if (grad_latents && n_latents != n->n_latents()) return fail(c, RSY_INVALID_ARGUMENT, "latent gradient size mismatch");
    const uint32_t ns = uint32_t(n->slots.size());
    std::vector<uint32_t> order(n_pts), offsets(ns + 1, 0);
    for (uint32_t i = 0; i < n_pts; ++i) {
        if (material[i] >= c->n_materials || n->slot_of[material[i]] == 0xffffffffu)
            return fail(c, RSY_INVALID_ARGUMENT, "sample of a material without a latent grid");
        ++offsets[n->slot_of[material[i]] + 1];
    }
    std::partial_sum(offsets.begin(), offsets.end(), offsets.begin());
    std::vector<uint32_t> fill(offsets.begin(), offsets.end() - 1);
    std::vector<float> suv(size_t(n_pts) * 2), stgt(size_t(n_pts) * kNmOut);
    for (uint32_t i = 0; i < n_pts; ++i) {
        const uint32_t k = fill[n->slot_of[material[i]]]++;
        suv[2 * size_t(k)] = uv[2 * size_t(i)];
        suv[2 * size_t(k) + 1] = uv[2 * size_t(i) + 1];
        std::memcpy(&stgt[kNmOut * size_t(k)], &target[kNmOut * size_t(i)], kNmOut * 4);
    }
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    hipStream_t s = c->stream;
    HIP_TRY(c, hipStreamSynchronize(s), "synchronize");
    HIP_TRY(c, n->tmp[0].ensure(suv.size() * 4), "allocate");
    HIP_TRY(c, n->tmp[1].ensure(stgt.size() * 4), "allocate");
    HIP_TRY(c, n->tmp[2].ensure(offsets.size() * 4), "allocate");
    HIP_TRY(c, n->tmp[3].ensure(std::max<uint64_t>(1, n->n_latents()) * 4), "allocate");
    HIP_TRY(c, n->tmp[4].ensure(kNmParamsPad * 4), "allocate");
    HIP_TRY(c, hipMemcpy(n->tmp[0].p, suv.data(), suv.size() * 4, hipMemcpyHostToDevice), "upload");
    HIP_TRY(c, hipMemcpy(n->tmp[1].p, stgt.data(), stgt.size() * 4, hipMemcpyHostToDevice), "upload");
    HIP_TRY(c, hipMemcpy(n->tmp[2].p, offsets.data(), offsets.size() * 4, hipMemcpyHostToDevice), "upload");
    NmTrainLaunch T = train_args(c, n);
    T.n_wg = ns;
    T.inv_norm = 1.f / (float(n_pts) * float(kNmOut));
    T.ex_uv = n->tmp[0].as<float>();
    T.ex_target = n->tmp[1].as<float>();
    T.ex_offsets = n->tmp[2].as<uint32_t>();
    T.ex_grad_lat = n->tmp[3].as<float>();
    T.ex_grad_mlp = n->tmp[4].as<float>();
    HIP_TRY(c, launch_nm_step(T, true, s), "gradient kernels");
    HIP_TRY(c, hipStreamSynchronize(s), "synchronize");
    std::vector<float> g(kNmParamsPad);
    HIP_TRY(c, hipMemcpy(g.data(), n->tmp[4].p, kNmParamsPad * 4, hipMemcpyDeviceToHost), "read gradients");
    if (grad_mlp) std::memcpy(grad_mlp, g.data(), kNmParams * 4);
    if (loss) *loss = g[kNmParams] * T.inv_norm;
    if (grad_latents && n_latents)
        HIP_TRY(c, hipMemcpy(grad_latents, n->tmp[3].p, n_latents * 4, hipMemcpyDeviceToHost), "read gradients");
    return RSY_OK;
}

rsy_result rsy_neural_set_weights(rsy_ctx* c, uint32_t weights) {
    rsy_result r = need_neural(c);
    if (r != RSY_OK) return r;
    if (weights > RSY_NEURAL_WEIGHTS_FP8) return fail(c, RSY_INVALID_ARGUMENT, "unknown weight format");
    c->neural->cfg.weights = weights;
    c->neural->evaluated = false;
    return RSY_OK;
}

rsy_result rsy_neural_get_inference_mlp(rsy_ctx* c, float* mlp, uint32_t n_mlp) {
    rsy_result r = need_neural(c);
    if (r != RSY_OK) return r;
    NeuralState* n = c->neural;
    if (!mlp || n_mlp != uint32_t(kNmParams)) return fail(c, RSY_INVALID_ARGUMENT, "mlp array must hold 583 values");
    if (!n->evaluated) return fail(c, RSY_INVALID_STATE, "rsy_neural_evaluate has not been called");
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    HIP_TRY(c, n->tmp[4].ensure(kNmParamsPad * 4), "allocate");
    HIP_TRY(c, launch_nm_unpack_weights(n->weights_d.p, n->tmp[4].as<float>(), c->stream), "unpack weights");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    HIP_TRY(c, hipMemcpy(mlp, n->tmp[4].p, kNmParams * 4, hipMemcpyDeviceToHost), "read mlp");
    return RSY_OK;
}

rsy_result rsy_neural_decode(rsy_ctx* c, const uint32_t* material, const float* uv, uint32_t n_pts, uint32_t targets,
                             float* out) {
    rsy_result r = need_neural(c);
    if (r != RSY_OK) return r;
    NeuralState* n = c->neural;
    if (!material || !uv || !out) return fail(c, RSY_INVALID_ARGUMENT, "null argument");
    if (!targets && !n->evaluated) return fail(c, RSY_INVALID_STATE, "rsy_neural_evaluate has not been called");
    if (n_pts == 0) return RSY_OK;
    std::vector<uint32_t> slot(n_pts);
    for (uint32_t i = 0; i < n_pts; ++i) {
        if (material[i] >= c->n_materials || n->slot_of[material[i]] == 0xffffffffu)
            return fail(c, RSY_INVALID_ARGUMENT, "point of a material without a latent grid");
        slot[i] = n->slot_of[material[i]];
    }
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    hipStream_t s = c->stream;
    HIP_TRY(c, hipStreamSynchronize(s), "synchronize");
    HIP_TRY(c, n->tmp[0].ensure(size_t(n_pts) * 8), "allocate");
    HIP_TRY(c, n->tmp[1].ensure(size_t(n_pts) * kNmOut * 4), "allocate");
    HIP_TRY(c, n->tmp[2].ensure(size_t(n_pts) * 4), "allocate");
    HIP_TRY(c, hipMemcpy(n->tmp[0].p, uv, size_t(n_pts) * 8, hipMemcpyHostToDevice), "upload");
    HIP_TRY(c, hipMemcpy(n->tmp[2].p, slot.data(), size_t(n_pts) * 4, hipMemcpyHostToDevice), "upload");
    NmPointLaunch P{};
    P.lat16 = n->lat16.as<uint4>();
    P.weights = n->weights_d.p;
    P.L = n->L;
    P.slots = n->slots_d.p;
    P.slot = n->tmp[2].as<uint32_t>();
    P.uv = n->tmp[0].as<float>();
    P.n = n_pts;
    P.targets = targets ? 1u : 0u;
    P.textures = c->tex_table.as<TexGPU>();
    P.texels = c->texels.as<uint8_t>();
    P.srgb_lut = c->srgb_lut.as<float>();
    P.out = n->tmp[1].as<float>();
    HIP_TRY(c, launch_nm_points(P, s), "decode kernel");
    HIP_TRY(c, hipStreamSynchronize(s), "synchronize");
    HIP_TRY(c, hipMemcpy(out, n->tmp[1].p, size_t(n_pts) * kNmOut * 4, hipMemcpyDeviceToHost), "read");
    return RSY_OK;
}

}  // extern "C"
