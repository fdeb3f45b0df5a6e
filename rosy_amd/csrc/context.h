// Context internals shared by the frame orchestration (context.cpp) and the material decoder's
// host side (neural_ctx.cpp).  Not part of the C ABI: callers see rsy_ctx as an opaque handle.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "frame_types.h"
#include "kernels.h"
#include "rosy_amd.h"

namespace rsy {

constexpr uint32_t kMaxViewExtent = 8192;
constexpr uint32_t kMaxViews = RSY_MAX_VIEWS;

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), cap(o.cap) {
        o.p = nullptr;
        o.cap = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            cap = o.cap;
            o.p = nullptr;
            o.cap = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    // Grows (never shrinks); contents are not preserved.
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        release();
        const size_t want = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipHostMalloc(&p, std::max<size_t>(bytes, 256), hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = std::max<size_t>(bytes, 256);
        return hipSuccess;
    }
};

struct FrameViewDesc {
    bool shadow;
    uint32_t width, height;
    float vp[16];
    // camera only
    float camera_position[4];
    // shadow only
    float split_far;
};

struct Layout {
    // key
    std::vector<FrameViewDesc> views;
    uint32_t tile = 0;
    uint64_t lists_version = 0;
    uint32_t shade = 0;  // rsy_shade_mode
    // derived
    std::vector<ViewGPU> vg;
    std::vector<BlockEntry> vblocks, tblocks;
    uint32_t setup_tpt = 1;  // triangles per setup thread (kSetupPairBlocks)
    uint32_t total_bins = 0;
    uint64_t total_xv = 0;
    uint32_t rec_total = 0, clip_total = 0, big_total = 0, cull_total = 0;
    uint64_t want_pairs = 0;  // estimate of the frame's (triangle, bin) pairs
    uint32_t raster_grid = 0; // raster workgroups: the last frame's work items plus slack (0: estimate)
    uint32_t n_shadow = 0, n_camera = 0, n_blend = 0;
    uint32_t first_blend = 0, blend_bins = 0;  // first blend view, bins of all blend views
    bool valid = false;
};

inline bool same_views(const std::vector<FrameViewDesc>& a, const std::vector<FrameViewDesc>& b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].shadow != b[i].shadow || a[i].width != b[i].width || a[i].height != b[i].height) return false;
    return true;
}

inline uint32_t ilog2(uint32_t v) {
    uint32_t r = 0;
    while ((1u << (r + 1)) <= v) ++r;
    return r;
}

// Device addresses of one frame's working memory (everything its kernels use between the inputs
// and the outputs) and of its outputs.  The ordinary frame points them into the context's own
// buffers; the allocation-free frame carves the working memory from one static device array and
// takes the outputs from the caller.
struct FrameMem {
    ViewGPU* views = nullptr;
    float4* xv = nullptr;
    TriRec* recs = nullptr;
    uint32_t* bin_tiles = nullptr;
    ShadeRec* shade_recs = nullptr;
    ClipEntry* clipq = nullptr;
    uint32_t* counters = nullptr;
    uint32_t* draw_culled = nullptr;
    uint32_t* bin_counts = nullptr;
    uint32_t* bin_offsets = nullptr;
    uint32_t* pair_slots = nullptr;
    uint4* big_pairs = nullptr;
    uint32_t* block_valid = nullptr;
    uint32_t* totals = nullptr;
    void* scan_partial = nullptr;
    BlockEntry* vblocks = nullptr;
    BlockEntry* tblocks = nullptr;
    uint32_t* lists = nullptr;
    uint32_t list_cap = 0;
    uint4* items = nullptr;
    ShadeView* shade_views = nullptr;
    ShadeCascade* shade_cascades = nullptr;
    ShadeView* blend_views = nullptr;
    // the blended pass's fragments
    uint4* frags = nullptr;
    uint32_t* frag_keys = nullptr;
    float4* frag_color = nullptr;
    uint8_t* frag_mask = nullptr;
    uint32_t frag_cap = 0;
    // outputs: per camera view, per cascade, per blend view (in view order within each kind)
    uint64_t* vis[kMaxViews] = {};
    float4* color[kMaxViews] = {};
    float4* sample_color[kMaxViews] = {};
    float* shadow_factor[kMaxViews] = {};
    float* shadow_map[kMaxViews] = {};
    uint2* blend_list[kMaxViews] = {};
};

// The allocation-free frame's working memory: a static device array of kStaticScratchBytes per
// GPU (static_scratch.hip).  static_scratch_init looks up the calling device's copy (once per
// device; rsy_create calls it); static_scratch returns it, or null before the lookup.
constexpr size_t kStaticScratchBytes = size_t(128) << 20;
hipError_t static_scratch_init(int device);
void* static_scratch(int device);

struct NeuralState;  // material decoder (neural_ctx.cpp)
void destroy_neural(NeuralState* n);

}  // namespace rsy

struct rsy_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side_stream = nullptr;  // the blended fragments beside the opaque shading (joined each frame)
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_stage[8] = {};  // [5] end of shading, [6]/[7] around rsy_present
    bool present_timed = false;
    std::string err;
    uint32_t debug_flags = 0;

    // geometry
    rsy::DevBuf vertices, px, py, pz, indices;
    uint64_t n_vertices = 0, n_indices = 0;
    std::vector<rsy_mesh_range> meshes;
    std::vector<float> mesh_bounds;  // per mesh: object-space min xyz, max xyz of its vertices
    bool has_geometry = false;
    // objects + surfaces
    rsy::DevBuf objects;
    std::vector<rsy_object_record> h_objects;
    std::vector<rsy_draw> surfaces;
    bool has_objects = false;
    // materials
    rsy::DevBuf materials, tex_table, texels, samplers, srgb_lut;
    uint32_t n_materials = 0, n_textures = 0, n_samplers = 0;
    std::vector<uint32_t> tex_channels;
    std::vector<rsy_gpu_material> h_materials;
    std::vector<rsy::TexGPU> h_tex_table;
    std::vector<rsy_sampler> h_samplers;
    std::vector<uint8_t> tex_opaque;  // four-channel texture whose alpha is 255 everywhere
    rsy::DevBuf tex_staging;          // BC7 blocks of one level on their way to the decoder
    bool any_mips = false;            // some texture has more than one level (LOD sampling)
    // draw lists: 0 = shadow (all), 1 = opaque (!blended), 2 = blended (the blended pass)
    std::vector<rsy::DrawGPU> lists[3];
    std::vector<uint32_t> list_index[3];  // surface indices
    uint32_t list_tris[3] = {0, 0, 0};
    uint64_t list_verts[3] = {0, 0, 0};
    rsy::DevBuf lists_d[3];
    uint64_t lists_version = 0;
    bool lists_dirty = true;
    // per-frame uniforms
    rsy_scene_data sd{};
    rsy_raster_state rs{};
    bool has_scene = false;
    rsy::DevBuf scene_d;

    // frame buffers
    rsy::Layout layout;
    rsy::FrameMem mem;            // the ordinary frame's addresses (into the buffers below)
    rsy::Layout static_layout;    // the allocation-free frame's layout (host tables only)
    std::vector<rsy::FrameViewDesc> static_views;
    rsy::DevBuf views_d, xv, recs, bin_tiles, shade_recs, clipq, counters, bin_counts, bin_offsets, pair_slots, big_pairs, block_valid, totals,
        lists_mem, vblocks_d, tblocks_d, shade_views_d, shade_cascades_d, items,
        draw_culled, scan_partial;
    uint32_t list_cap = 0;
    uint32_t rec_scale = 1;  // grows on overflow
    std::vector<rsy::DevBuf> vis_out, color_out, shadow_factor_out, shadow_maps;
    std::vector<rsy::DevBuf> blend_lists;  // per blend view: per pixel (first fragment, count)
    rsy::DevBuf frags, frag_keys, frag_color, frag_mask;  // the blended pass's fragments
    std::vector<rsy::DevBuf> sample_color;  // per camera view, 4 samples with a blended pass: colour planes
    uint32_t frag_cap = 0;
    rsy::DevBuf blend_shade_views_d;
    std::vector<rsy::ShadeView> h_blend_views;
    rsy::PinnedBuf staging, readback_counters;
    std::vector<rsy::ShadeView> h_shade_views;
    std::vector<rsy::ShadeCascade> h_shade_cascades;
    uint32_t shade_blocks = 0;

    // last frame (for re-running after an overflow)
    std::vector<rsy::FrameViewDesc> last_views;
    uint32_t last_shade = 0;
    bool frame_pending = false;
    bool frame_validated = false;
    rsy_frame_stats stats{};
    uint32_t autotuned_tile = 0;
    // output stage (rsy_present): the draw image of the presented view and the swapchain image
    rsy::DevBuf draw_image, present_image;
    uint32_t draw_w = 0, draw_h = 0, present_w = 0, present_h = 0;

    // graph cache
    hipGraphExec_t graph_exec = nullptr;
    bool graph_valid = false;

    hipError_t last = hipSuccess;

    // material decoder state (null until rsy_neural_init)
    rsy::NeuralState* neural = nullptr;
    ~rsy_ctx() { rsy::destroy_neural(neural); }
};

namespace rsy {

inline rsy_result fail(rsy_ctx* c, rsy_result r, const std::string& msg) {
    if (c) c->err = msg;
    return r;
}

inline rsy_result hip_fail(rsy_ctx* c, hipError_t e, const char* what) {
    if (c) c->err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? RSY_ALLOCATION_FAILURE : RSY_GRAPHICS_FRAME_FAILURE;
}

#define HIP_TRY(ctx, expr, what)                    \
    do {                                            \
        hipError_t e_ = (expr);                     \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, what); \
    } while (0)

inline void invalidate_graph(rsy_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
    c->graph_valid = false;
}

// Material decoder hooks used by the frame path (neural_ctx.cpp).
bool neural_ready(const rsy_ctx* c);                     // evaluated: RSY_SHADE_NEURAL may render
void neural_shade_args(const rsy_ctx* c, ShadeLaunch& sh);  // decoder inputs of the shading kernel

}  // namespace rsy
