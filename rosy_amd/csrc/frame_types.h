// Device-side records shared by the host orchestration (context.cpp) and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rsy {

// Global-memory view of a device pointer that was read from memory (a record or LDS).  The compiler
// cannot tell such a pointer from an LDS or scratch one and would use flat instructions, whose
// waits also hold back the scalar loads issued beside them.
// (The host pass only type-checks device code, and sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using GlobalPtr = __attribute__((address_space(1))) T*;
#else
template <class T> using GlobalPtr = T*;
#endif
template <class T> __device__ inline GlobalPtr<T> gmem(T* p) { return (GlobalPtr<T>)p; }

// One triangle after clip, snap, cull and setup: 48 bytes, three 16-byte loads.
struct alignas(16) TriRec {
    int32_t X0, Y0, X1, Y1;  // snapped vertices, 8 sub-pixel bits
    int32_t X2, Y2;
    float z0;                // depth plane at (X0, Y0) (depth bias folded in for shadow views)
    float dzdx;              // per sub-pixel unit
    float dzdy;
    uint32_t id;             // triangle id in the view's draw list | kRecSwapped
    uint32_t bbx;            // px0 | px1 << 16 (inclusive pixel bounds)
    uint32_t bby;            // py0 | py1 << 16
};
static_assert(sizeof(TriRec) == 48, "TriRec is 48 bytes");

// Setup stores every record clockwise in pixel space (y down): a triangle whose snapped vertices
// were counter-clockwise has vertices 1 and 2 exchanged, and its id carries this bit.
constexpr uint32_t kRecSwapped = 0x80000000u;
constexpr uint32_t kRecIdMask = 0x7fffffffu;

// What shading needs of a camera or blend view's triangle, written by setup into the triangle's slot
// (32 bytes, two 16-byte loads): the three vertices' indices into the global vertex buffer, the
// draw index in the view's list (| kShadeNoRecord when the triangle was clipped and has no raster
// record in its slot), then the vertices' clip w and the draw's material index.
struct ShadeRec {
    uint32_t v0, v1, v2;
    uint32_t draw;
    float w0, w1, w2;
    uint32_t material;
};
static_assert(sizeof(ShadeRec) == 32, "ShadeRec is 32 bytes");
constexpr uint32_t kShadeNoRecord = 0x80000000u;

// A triangle queued for clipping (64 bytes, four 16-byte words): its three clip-space vertices as
// the transform wrote them and its triangle id, so that the clipper reads one entry instead of
// walking draw, indices and vertices.
struct ClipEntry {
    float4 v[3];
    uint32_t id, pad[3];
};
static_assert(sizeof(ClipEntry) == 64, "ClipEntry is 64 bytes");

// Per-draw data of one draw list (shadow or opaque), in list order.
struct DrawGPU {
    uint32_t first_index;  // mesh.index_offset + start_index
    uint32_t tri_count;
    uint32_t tri_base;     // prefix over the list
    uint32_t vtx_base;     // first vertex of the mesh in the global vertex buffer
    uint32_t vtx_count;
    uint32_t xv_off;       // offset of this draw's transformed vertices inside a view's block
    uint32_t object;       // graphic_objects_offset + graphics_object_index
    uint32_t material;     // material index or 0xFFFFFFFF
    float bmin[3], bmax[3];  // object-space box of the mesh's vertices (view culling)
    uint32_t flags;        // kDraw* bits of the draw's material, resolved on the host
};
// DrawGPU::flags: some sampler of the material clamps to the edge (texture coordinates are then
// used as they are; REPEAT and MIRRORED_REPEAT samplers take them with whole periods removed).
constexpr uint32_t kDrawClampSampler = 1u;

// Per-view counters: records appended by the clipper, queued clipped triangles, entries of the
// big-pair list, overflow flags, clipped-triangle statistic.
// The first blend view's kCtrFrags slot counts the blended pass's fragments of the frame.
enum : uint32_t {
    kCtrRecords = 0, kCtrClip = 1, kCtrBig = 2, kCtrOverflow = 3, kCtrClipped = 4, kCtrFrags = 5, kCtrPerView = 8
};
// kCtrOverflow bits: 1 clip queue, 2 records, 4 big pairs, 8 blended fragments
constexpr uint32_t kOverflowFrags = 8u;

struct ViewGPU {
    uint32_t width, height;
    uint32_t tile, tiles_x, tiles_y;
    uint32_t tile_base;     // first bin of this view in the global bin arrays
    uint32_t is_shadow;
    uint32_t list;          // 0 = shadow list, 1 = opaque list, 2 = blended list
    uint32_t n_draws;
    uint32_t xv_base;       // first transformed vertex of this view
    uint32_t rec_base, rec_cap;  // triangle slots [rec_base, rec_base + tri_count), then clipper output
    uint32_t tri_count;
    uint32_t clip_base, clip_cap;
    uint32_t big_base, big_cap;  // (bin, slot, record) entries of records with many bins
    uint32_t cull, reverse, bias;
    float bias_const, bias_clamp, bias_slope;
    uint32_t cull_base;     // first per-draw culling flag of this view
    uint32_t blend;         // 1: the blended list of camera view `source`, drawn by the blended pass
    uint32_t source;
    uint32_t samples;       // 1, or 4 for the camera and blend views of a 4-sample frame
    uint32_t setup_clear;   // 1: the view's setup blocks clear its output; 0: a fill launch does (or none: blend)
    uint64_t out;           // device address: u64 visibility (camera; one plane per sample), u32 depth
                            // bits (shadow), or for a blend view its pixels' fragment lists (uint2:
                            // first fragment, count)
    float view_projection[16];
};

// Work of one block of the transform or setup kernel: its block size of items of one draw from
// `start`, with the addresses the block needs resolved on the host (the layout's draw and view
// tables are fixed), so that the block's first data loads wait on no draw or view record.  A setup
// block serves one view; a transform block serves `xv_first` consecutive views from `view`, all
// drawing the same list.
struct BlockEntry {
    uint32_t view, draw, start;
    uint32_t list;       // the view's draw list
    uint32_t first;      // transform: the block's first vertex in the vertex streams; setup: its first index word
    uint32_t count;      // vertices or triangles of the block
    uint32_t out_first;  // transform: its first transformed vertex inside a view's block; setup: its first record slot
    uint32_t xv_first;   // setup: the draw's first transformed vertex in the view (indices are mesh-local);
                         // transform: the number of views
    uint32_t cull;       // setup: the (view, draw) culling flag; transform: the draw (each view adds its cull_base)
    uint32_t object;     // transform: the draw's object record; setup: the triangle id of its first triangle
    uint32_t vtx_base;   // setup: the mesh's first vertex (shading records hold global vertex indices)
    uint32_t material;   // setup: the draw's material (shading records)
};
static_assert(sizeof(BlockEntry) == 48, "BlockEntry is 48 bytes");

struct TexGPU {
    uint64_t offset;  // bytes into the texel pool (level 0; the levels follow one another)
    uint32_t width, height;
    uint32_t channels;  // 2 or 4
    uint32_t srgb;
    uint32_t levels;    // mip levels (1: level 0 only)
    uint32_t pad;
};

// Size of mip level l of a width x height image.
__host__ __device__ inline uint32_t mip_extent(uint32_t e, uint32_t l) { return e >> l ? e >> l : 1u; }

// Bytes into a texture's texels of mip level l.
__host__ __device__ inline uint64_t mip_offset(uint32_t w, uint32_t h, uint32_t ch, uint32_t l) {
    uint64_t o = 0;
    for (uint32_t k = 0; k < l; ++k) o += uint64_t(mip_extent(w, k)) * mip_extent(h, k) * ch;
    return o;
}

// The standard locations of a 4-sample pixel, (0.375, 0.125), (0.875, 0.375), (0.125, 0.625) and
// (0.625, 0.875), in 1/256 pixel.
__host__ __device__ constexpr int32_t msaa_x(int s) { return s == 0 ? 96 : s == 1 ? 224 : s == 2 ? 32 : 160; }
__host__ __device__ constexpr int32_t msaa_y(int s) { return 32 + 64 * s; }

}  // namespace rsy
