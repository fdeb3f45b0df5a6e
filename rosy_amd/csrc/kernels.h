// Launch descriptors for the frame kernels (host orchestration -> kernels).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "frame_types.h"
#include "rosy_amd.h"

namespace rsy {

// Work of one 256-thread transform or setup block (a BlockEntry): two vertices per thread, or one
// triangle (setup moves about 90 bytes per triangle and keeps HBM busy at one per thread; two per
// thread measured slower).
constexpr uint32_t kTransformPerThread = 2;
constexpr uint32_t kTransformBlockVerts = 256 * kTransformPerThread;
constexpr uint32_t kTransformViews = 8;  // consecutive views of one draw list per transform block
constexpr uint32_t kSetupBlockTris = 256;
// A frame whose setup would take more blocks than this runs two triangles per thread (512 per block).
constexpr uint32_t kSetupPairBlocks = 8192;
// A view's setup blocks clear its output when each block's share is at most this many bytes; a view
// with fewer, larger shares is cleared by a fill launch instead
constexpr uint64_t kSetupClearBytes = 128u << 10;

// The three draw lists of a frame, indexed by ViewGPU::list.
struct DrawLists {
    const DrawGPU* l[3];  // shadow, opaque, blended
    __host__ __device__ const DrawGPU* of(uint32_t list) const { return l[list]; }
};

struct GeometryLaunch {
    const ViewGPU* views;
    uint32_t n_views;
    const DrawGPU* draws0;  // shadow list
    const DrawGPU* draws1;  // opaque list
    const DrawGPU* draws2;  // blended list
    const float* objects;   // packed 164-byte object records
    const BlockEntry* vblocks;
    uint32_t n_vblocks;
    const BlockEntry* tblocks;
    uint32_t n_tblocks;
    uint32_t setup_tpt;     // triangles per setup thread (1 or 2): a block covers 256 * setup_tpt
    const float *px, *py, *pz;
    const uint32_t* indices;
    float4* xv;
    TriRec* recs;
    uint32_t* bin_tiles;    // per record slot: its bins for the scatter (pack_bins), or kNoBins
    ShadeRec* shade_recs;   // per record slot of a camera or blend view: its shading record
    ClipEntry* clipq;
    uint32_t* counters;
    uint32_t* bin_counts;
    uint32_t* pair_slots;   // per record slot: its slot inside each of its (at most 4) bins, as two
                            // uint2 arrays of slot_count entries (bins 0-1, then bins 2-3)
    uint32_t slot_count;    // record slots of the frame (all views)
    uint4* big_pairs;       // (bin, slot, record) of records with more bins
    uint32_t* block_valid;  // records set up per setup block
    uint32_t* draw_culled;  // per (view, draw): 1 when the draw's box is outside the view
    uint32_t shift;  // log2(tile)
    uint32_t clip_blocks;
    uint32_t total_bins;
};

struct BinLaunch {
    const ViewGPU* views;
    uint32_t n_views;
    uint32_t total_bins;
    uint32_t* bin_counts;
    uint32_t* bin_offsets;
    const uint32_t* pair_slots;
    uint32_t slot_count;
    const uint4* big_pairs;
    const uint32_t* block_valid;
    const uint32_t* bin_tiles;
    uint32_t n_tblocks;
    uint32_t* totals;  // [0] (record, bin) pairs, [1] records set up, [2] overflow bits (lists, raster grid),
                       // [3] raster work items, [4] first mid-class item, [5] empty and split bins,
                       // [6] first split item, [7] first light-class item
    const TriRec* recs;
    const uint32_t* counters;
    uint32_t* lists;
    uint32_t list_cap;
    uint32_t shift;
    uint32_t scatter_blocks;
    void* scan_partial;     // kScanPartialBytes per scan block
    uint4* items;           // raster work items (bin, first list entry, records, split): split bins' first
    uint32_t items_cap;     // items capacity
    uint32_t raster_grid;   // workgroups of the raster launch (more items than these: totals[2] bit 1)
};

struct RasterLaunch {
    const ViewGPU* views;
    uint32_t n_views;
    uint32_t total_bins;
    uint32_t tile;
    const uint32_t* lists;
    uint32_t list_cap;
    const TriRec* recs;
    uint64_t rec_cap_total;
    const uint4* items;
    uint32_t items_cap;
    const uint32_t* totals;  // [3] work items
    uint32_t samples;  // of the camera views (1 or 4)
    uint32_t grid;     // workgroups: one per work item (the last frame's count, with slack)
};

// The blended pass over the bins of the blend views (one launch).
struct BlendLaunch {
    const ViewGPU* views;
    uint32_t n_views;
    uint32_t first_blend;  // first blend view; blend views come last and own the last bins
    uint32_t blend_bins;
    uint32_t tile;
    const uint32_t* bin_counts;
    const uint32_t* bin_offsets;
    const uint32_t* lists;
    uint32_t list_cap;
    const TriRec* recs;
    uint64_t rec_cap_total;
    uint32_t* counters;     // per-view counters (fragment count, overflow flags)
    uint4* frags;           // per fragment: key, x | y << 16, blend view - first_blend, colour index
    uint32_t* frag_keys;    // a bin's lists when they exceed the LDS budget
    float4* frag_color;     // per fragment: shaded colour (discarded: kDiscarded in w)
    uint8_t* frag_mask;     // 4 samples: per fragment, the samples it covers and passes the depth test at
    uint32_t frag_cap;
    uint32_t samples;       // samples per pixel of the camera views (1 or 4)
};

// Per-pixel shading inputs of one camera view.
struct ShadeView {
    uint32_t width, height;
    uint32_t block_base;  // first shading block of this view
    uint32_t blocks_x;
    const unsigned long long* vis;
    const TriRec* recs;    // the raster view's record slots, indexed by triangle id
    const ShadeRec* shade_recs;  // the raster view's shading records, indexed by triangle id
    uint32_t xv_base;      // first transformed vertex of the raster view
    uint32_t samples;      // 1, or 4: vis holds one plane per sample and colour is their resolve
    float4* color;
    float4* sample_color;  // 4 samples: per-sample colour planes for the blended pass (nullable)
    float* shadow_factor;  // nullable debug output
    float view_projection[16];
    float camera_position[4];
};

constexpr uint32_t kMaxShadeCascades = RSY_MAX_VIEWS;  // the shading kernels keep a frame's cascades in LDS

struct ShadeCascade {
    uint32_t size;
    float split_far;
    const float* map;
    float m[16];
};

struct ShadeLaunch {
    const ShadeView* views;  // device array
    uint32_t n_views;
    uint32_t total_blocks;
    const ShadeCascade* cascades;  // device array
    uint32_t n_cascades;
    uint32_t pcf_mode;
    const void* scene;  // rsy_scene_data on device
    const DrawGPU* draws;
    uint32_t n_draws;
    const uint32_t* indices;
    const void* vertices;  // rsy_position
    const float* objects;
    const void* materials;  // rsy_gpu_material
    uint32_t n_materials;
    const TexGPU* textures;
    uint32_t n_textures;
    const uint8_t* texels;
    const void* samplers;  // rsy_sampler
    uint32_t n_samplers;
    float clear_color[4];
    const float* srgb_lut;      // 256 entries
    const float4* xv;           // transformed vertices of all views
    uint32_t samples;           // samples per pixel of the camera views (1 or 4)
    uint32_t lod;               // 1: some texture has mip levels (trilinear sampling at the pixel's LOD)
    // material decoder (shade mode NEURAL only)
    const uint32_t* nm_slot;    // per material: latent grid slot or 0xFFFFFFFF
    const uint32_t* nm_active;  // per material: 1 = decoded by the network, 0 = plain textures
    const uint4* nm_lat16;      // FP16 latent grids, L * L cells of 8 halves per slot
    const void* nm_weights;     // NmWeights
    uint32_t nm_L;
    uint32_t neural;            // 1: launch the decoder-capable shading kernel
};

// One material-decoder training step (or, in explicit mode, the exact gradient of a given batch).
struct NmTrainLaunch {
    float* lat;      // [slots][L][L][8] FP32 latents
    float* lat_m;    // Adam moments of the latents
    float* lat_v;
    float* mlp;      // [584] FP32 MLP parameters (W1, b1, W2, b2)
    float* mlp_m;
    float* mlp_v;
    float* partial;  // [n_wg][584]: per-workgroup MLP gradient sums, loss sum at index 583
    uint32_t* step;  // device step counter: completed Adam steps
    float* loss_out; // loss per step, index step - loss_base (nullable)
    uint32_t loss_base, loss_cap;
    const void* slots;  // NmSlot
    uint32_t n_slots;
    const TexGPU* textures;
    const uint8_t* texels;
    const float* srgb_lut;
    uint32_t L, log2L;
    uint32_t log2R;   // region side 2^log2R cells
    uint32_t n_wg;    // workgroups per step
    uint32_t spw;     // samples per workgroup
    float inv_norm;   // 1 / (batch * 7)
    float lr_mlp, lr_lat, beta1, beta2, eps;
    uint32_t decay_steps;  // cosine decay of both learning rates to 0 over this many steps (0: constant)
    uint64_t seed;
    // explicit-batch mode: one workgroup per slot over its whole grid
    const float* ex_uv;          // [n][2]
    const float* ex_target;      // [n][7]
    const uint32_t* ex_offsets;  // [slots + 1]
    float* ex_grad_lat;          // [slots][L][L][8]
    float* ex_grad_mlp;          // [584]: MLP gradient, loss at 583
};

struct NmInitLaunch {
    float* lat;
    uint32_t n_slots, L;
    uint32_t random;  // 0: latents from the plain maps (channels 0..6), 1: uniform noise
    uint64_t seed;
    const void* slots;
    const TexGPU* textures;
    const uint8_t* texels;
    const float* srgb_lut;
};

// PSNR evaluation: every texel centre of each slot's albedo map, decoded vs plain.
struct NmEvalLaunch {
    const uint4* lat16;
    const void* weights;  // NmWeights
    uint32_t L;
    const void* slots;
    uint32_t n_slots;
    const uint32_t* block_base;  // [slots + 1] first 16x16-texel block of each slot
    const TexGPU* textures;
    const uint8_t* texels;
    const float* srgb_lut;
    float* partial;  // per block: sum of squared errors over its texels and 7 channels
};

// Decode (or sample the plain maps) at explicit (slot, uv) points.
struct NmPointLaunch {
    const uint4* lat16;
    const void* weights;
    uint32_t L;
    const void* slots;
    const uint32_t* slot;  // [n]
    const float* uv;       // [n][2]
    uint32_t n;
    uint32_t targets;      // 1: plain-map targets instead of the decoder
    const TexGPU* textures;
    const uint8_t* texels;
    const float* srgb_lut;
    float* out;  // [n][7]
};

// texture upload: BC7 blocks of a width x height image into RGBA8 texels; mip level l + 1 from level l
hipError_t launch_bc7_decode(const void* blocks, uint32_t width, uint32_t height, uint32_t* rgba, hipStream_t s);
hipError_t launch_build_mip(const uint8_t* src, uint32_t src_w, uint32_t src_h, uint8_t* dst, uint32_t channels,
                            uint32_t srgb, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
hipError_t launch_geometry(const GeometryLaunch& g, hipStream_t s);
hipError_t launch_binning(const BinLaunch& b, hipStream_t s);
constexpr uint32_t kScanBlockBins = 1024;    // bins per scan block
constexpr uint32_t kScanPartialBytes = 32;  // per scan block
hipError_t launch_raster(const RasterLaunch& r, hipStream_t s);
uint32_t raster_items_cap(uint32_t total_bins, uint32_t list_cap);  // work-item table entries
// output stage: colour -> RGBA16F draw image (sw x sh) -> linear blit to B8G8R8A8_SRGB (dw x dh)
hipError_t launch_present(const float4* color, uint2* draw, uint32_t sw, uint32_t sh, uint32_t* out, uint32_t dw,
                          uint32_t dh, hipStream_t s);
hipError_t launch_shade(const ShadeLaunch& l, hipStream_t s);
// the blended pass: fragment lists and their shading (needs the visibility buffers and shadow maps),
// then the blending over the shaded colour
hipError_t launch_blend_fragments(const BlendLaunch& b, const ShadeLaunch& l, hipStream_t s);
hipError_t launch_blend_resolve(const BlendLaunch& b, const ShadeLaunch& l, hipStream_t s);
hipError_t launch_nm_init(const NmInitLaunch& l, hipStream_t s);
hipError_t launch_nm_step(const NmTrainLaunch& l, bool explicit_batch, hipStream_t s);
hipError_t launch_nm_export(const float* lat, uint4* lat16, uint64_t cells, hipStream_t s);
hipError_t launch_nm_export_weights(const float* mlp, void* weights, uint32_t fp8, hipStream_t s);
hipError_t launch_nm_unpack_weights(const void* weights, float* mlp, hipStream_t s);
hipError_t launch_nm_eval(const NmEvalLaunch& l, uint32_t n_blocks, hipStream_t s);
hipError_t launch_nm_points(const NmPointLaunch& l, hipStream_t s);

}  // namespace rsy
