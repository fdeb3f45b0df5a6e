// Launch descriptors for the frame kernels (host orchestration -> kernels).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "frame_types.h"

namespace rsy {

struct GeometryLaunch {
    const ViewGPU* views;
    uint32_t n_views;
    const DrawGPU* draws0;  // shadow list
    const DrawGPU* draws1;  // opaque list
    const float* objects;   // packed 164-byte object records
    const BlockEntry* vblocks;
    uint32_t n_vblocks;
    const BlockEntry* tblocks;
    uint32_t n_tblocks;
    const float *px, *py, *pz;
    const uint32_t* indices;
    float4* xv;
    TriRec* recs;
    uint2* clipq;
    uint32_t* counters;
    uint32_t* bin_counts;
    uint32_t* pair_slots;   // per record slot: its slot inside each of its (at most 4) bins
    uint4* big_pairs;       // (bin, slot, record) of records with more bins
    uint32_t* block_valid;  // records set up per setup block
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
    const uint4* big_pairs;
    const uint32_t* block_valid;
    uint32_t n_tblocks;
    uint32_t* totals;  // [0] (record, bin) pairs, [1] records set up, [2] bin lists overflowed
    const TriRec* recs;
    const uint32_t* counters;
    uint32_t* lists;
    uint32_t list_cap;
    uint32_t shift;
    uint32_t scatter_blocks;
};

struct RasterLaunch {
    const ViewGPU* views;
    uint32_t n_views;
    uint32_t total_bins;
    uint32_t tile;
    const uint32_t* bin_counts;
    const uint32_t* bin_offsets;
    const uint32_t* lists;
    uint32_t list_cap;
    const TriRec* recs;
    uint64_t rec_cap_total;
};

// Per-pixel shading inputs of one camera view.
struct ShadeView {
    uint32_t width, height;
    uint32_t block_base;  // first shading block of this view
    uint32_t blocks_x;
    const unsigned long long* vis;
    float4* color;
    float* shadow_factor;  // nullable debug output
    float view_projection[16];
    float camera_position[4];
};

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
    const uint32_t* xv_bases;   // per shading view: first transformed vertex of its raster view
};

hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
hipError_t launch_geometry(const GeometryLaunch& g, hipStream_t s);
hipError_t launch_binning(const BinLaunch& b, hipStream_t s);
hipError_t launch_raster(const RasterLaunch& r, hipStream_t s);
hipError_t launch_shade(const ShadeLaunch& l, hipStream_t s);

}  // namespace rsy
