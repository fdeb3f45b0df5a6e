// Geometry, binning and bin-raster kernels of the frame path (gfx950, wave64).
//
// Replaces the reference's shadow and main raster passes (shadow.slang:6-20, basic.slang:100-116
// and the fixed-function state of Graphics.cpp:4988-5045, 5195-5292, 5553-5653) with a software
// pipeline.  Device-scope atomics are the scarce resource on MI355X, so every stage is laid out to
// issue few of them:
//   k_transform  per (view, vertex) clip = (VP * W) * p from SoA x/y/z (12 B in, 16 B out, coalesced);
//                                   the block's VP * W is formed once in LDS (fmaf chain, k order); the
//                                   grid also zeroes the frame's counters and bin counts
//   k_setup      per (view, tri)    clip test, snap, cull, bbox, depth plane, bias; the record goes
//                                   to the triangle's own slot (no append); bin counts use one device
//                                   atomic per distinct bin per block (LDS hash table), records that
//                                   span many bins are spread over the wave; each block also clears
//                                   its share of the view's output
//   k_clip       per clipped tri    near/far/guard-band Sutherland-Hodgman + fan, one thread each
//   k_scan_*     all bins           exclusive prefix sum of the bin counts in two launches (block
//                                   partials, then every block's offsets from them) and the raster
//                                   work items
//   k_scatter    per record         bin lists into the slots setup handed out (no atomics)
//   k_raster     per work item      one workgroup per non-empty bin (or per 128 records of a crowded
//                                   bin), the bin's depth/ID tile in LDS: every record's rows become
//                                   tasks spread over the workgroup; LDS atomic max resolves depth and
//                                   the tile goes to HBM once (crowded bins merge with device atomic
//                                   max)
// Every floating-point step is compiled without contraction; the CPU oracle restates the same
// rules, so visibility and depth are bit-exact by construction.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bin_raster.h"
#include "fixed_point.h"
#include "frame_types.h"
#include "kernels.h"

namespace rsy {
namespace {

// Snap, cull and set up one triangle.  Same rules, in the same operation order, as the oracle.
__device__ bool setup_tri(const float4 v[3], uint32_t id, const ViewGPU& V, TriRec& r) {
    const float hw = 0.5f * float(V.width), hh = 0.5f * float(V.height);
    int32_t X[3], Y[3];
    float Z[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) snap(v[k], hw, hh, X[k], Y[k], Z[k]);
    const int64_t area2 = int64_t(X[1] - X[0]) * (Y[2] - Y[0]) - int64_t(X[2] - X[0]) * (Y[1] - Y[0]);
    if (area2 == 0) return false;
    if (V.cull) {
        const bool front = V.reverse ? (area2 < 0) : (area2 > 0);
        if (front) return false;
    }
    if (area2 > 0) {
        int32_t t = X[1]; X[1] = X[2]; X[2] = t;
        t = Y[1]; Y[1] = Y[2]; Y[2] = t;
        const float z = Z[1]; Z[1] = Z[2]; Z[2] = z;
    }
    const int32_t minx = min(X[0], min(X[1], X[2])), maxx = max(X[0], max(X[1], X[2]));
    const int32_t miny = min(Y[0], min(Y[1], Y[2])), maxy = max(Y[0], max(Y[1], Y[2]));
    // pixels with a sample point inside the vertices' box: the centre (128, 128), or with 4 samples
    // anywhere in [32, 224] on each axis
    const int32_t lo = V.samples == 4 ? 32 : 128, hi = V.samples == 4 ? 224 : 128;
    const int32_t px0 = max(0, (minx - hi + 255) >> 8);
    const int32_t px1 = min(int32_t(V.width) - 1, (maxx - lo) >> 8);
    const int32_t py0 = max(0, (miny - hi + 255) >> 8);
    const int32_t py1 = min(int32_t(V.height) - 1, (maxy - lo) >> 8);
    if (px0 > px1 || py0 > py1) return false;
    const float fx1 = float(X[1] - X[0]), fy1 = float(Y[1] - Y[0]);
    const float fx2 = float(X[2] - X[0]), fy2 = float(Y[2] - Y[0]);
    const float dz1 = Z[1] - Z[0], dz2 = Z[2] - Z[0];
    const float t0 = fx1 * fy2, t1 = fx2 * fy1;
    const float den = t0 - t1;
    float dzdx = 0.f, dzdy = 0.f;
    if (den != 0.f) {
        const float a0 = dz1 * fy2, a1 = dz2 * fy1;
        const float b0 = dz2 * fx1, b1 = dz1 * fx2;
        dzdx = (a0 - a1) / den;
        dzdy = (b0 - b1) / den;
    }
    float z0 = Z[0];
    if (V.is_shadow && V.bias) {
        const float m = fmaxf(fabsf(dzdx), fabsf(dzdy)) * 256.0f;
        const float zmax = fmaxf(fabsf(Z[0]), fmaxf(fabsf(Z[1]), fabsf(Z[2])));
        const int eb = int((fbits(zmax) >> 23) & 0xffu);
        const int e = eb == 0 ? -126 : eb - 127;
        const float rr = ldexpf(1.0f, e - 23);
        const float ms = m * V.bias_slope, rc = rr * V.bias_const;
        float o = ms + rc;
        if (V.bias_clamp > 0.f) o = fminf(o, V.bias_clamp);
        else if (V.bias_clamp < 0.f) o = fmaxf(o, V.bias_clamp);
        z0 = z0 + o;
    }
    r.X0 = X[0]; r.Y0 = Y[0]; r.X1 = X[1]; r.Y1 = Y[1]; r.X2 = X[2]; r.Y2 = Y[2];
    r.z0 = z0;
    r.dzdx = dzdx;
    r.dzdy = dzdy;
    r.id = id | (area2 > 0 ? kRecSwapped : 0u);
    r.bbx = uint32_t(px0) | (uint32_t(px1) << 16);
    r.bby = uint32_t(py0) | (uint32_t(py1) << 16);
    return true;
}

__device__ inline float4 transform(const float* m, float x, float y, float z) {
    float4 c;
    c.x = fmaf(m[8], z, fmaf(m[4], y, fmaf(m[0], x, m[12])));
    c.y = fmaf(m[9], z, fmaf(m[5], y, fmaf(m[1], x, m[13])));
    c.z = fmaf(m[10], z, fmaf(m[6], y, fmaf(m[2], x, m[14])));
    c.w = fmaf(m[11], z, fmaf(m[7], y, fmaf(m[3], x, m[15])));
    return c;
}

__device__ inline uint64_t lanemask_lt() { return (uint64_t(1) << lane_id()) - 1; }

// Reserve `want` slots of a counter for every lane with want = 1, one atomic per wave.
__device__ inline uint32_t wave_append(bool want, uint32_t* counter) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return 0;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (int(lane_id()) == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
    base = __shfl(base, leader);
    return base + uint32_t(__popcll(mask & lanemask_lt()));
}

// A record slot's bins for the scatter, in one word: the first bin's column and row (13 bits each)
// and the width and height in bins less one (2 bits each; a record binned this way covers at most
// kSmallBins bins).  kNoBins: the slot has no record, or its record goes through the big-pair list.
constexpr uint32_t kNoBins = 0xffffffffu;
__device__ inline uint32_t pack_bins(int tx0, int ty0, int bw, int bh) {
    return uint32_t(tx0) | (uint32_t(ty0) << 13) | (uint32_t(bw - 1) << 26) | (uint32_t(bh - 1) << 28);
}

__device__ inline void rec_tiles(const TriRec& r, uint32_t shift, int& tx0, int& ty0, int& tx1, int& ty1) {
    tx0 = int(r.bbx & 0xffffu) >> shift;
    tx1 = int(r.bbx >> 16) >> shift;
    ty0 = int(r.bby & 0xffffu) >> shift;
    ty1 = int(r.bby >> 16) >> shift;
}

constexpr int kSmallBins = 4;  // records touching at most this many bins are binned in the block's LDS table

}  // namespace

// ---------------------------------------------------------------------------------------------

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Zeroes two ranges (the per-view counters and the bin counts) in one launch; graph-safe.
__global__ void k_clear2(uint32_t* a, uint32_t na, uint32_t* b, uint32_t nb) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < na + nb; i += gridDim.x * blockDim.x) {
        if (i < na) a[i] = 0u;
        else b[i - na] = 0u;
    }
}

// Draw culling: a draw is skipped when its object-space box provably contributes nothing to the
// view.  Every test carries a margin for the rounding of the transform (|error| <= a few ulp of
// sum |m_rj p_j| per coordinate; the margin takes 2^-18 of that sum); a vertex of the mesh is a
// convex combination of the box's corners and its own sum is no larger, so a distance that is
// negative (positive) beyond the margin at every corner is negative (positive) at the vertex's FP32
// clip position too.  Two cases:
//  - every corner outside one clip plane (near, far, guard band): every triangle of the draw is
//    rejected by the per-triangle test;
//  - every corner inside all six clip planes and beyond one side of the view (x < -w, x > w,
//    y < -w or y > w): no triangle is clipped, and every vertex snaps onto or beyond that edge of
//    the image (x / w > 1 rounds to >= 1, so X >= width * 256; x / w < -1 gives X <= 0), while
//    the sample points lie strictly inside (1/256 to 255/256 of a pixel): no triangle's box holds a
//    sample point, so none produces a record.
// Skipping the draw changes nothing.  corner_outside gives corner k's bits (k = 0..7: x from bit 0,
// y from bit 1, z from bit 2): bit `plane` (0..5) when the corner is outside that clip plane, bit
// 6 + side when it is beyond side `side` of the view, and kCornerInside when it is inside all six
// clip planes, each by the margin.
constexpr uint32_t kCornerInside = 1u << 10;

__device__ inline uint32_t corner_outside(const float* m, const float* bmin, const float* bmax, int k) {
    const float x = (k & 1) ? bmax[0] : bmin[0], y = (k & 2) ? bmax[1] : bmin[1];
    const float z = (k & 4) ? bmax[2] : bmin[2];
    const float4 c = transform(m, x, y, z);
    float sr[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
        sr[r] = fabsf(m[r] * x) + fabsf(m[4 + r] * y) + fabsf(m[8 + r] * z) + fabsf(m[12 + r]);
    uint32_t bits = kCornerInside;
#pragma unroll
    for (int plane = 0; plane < 6; ++plane) {
        float bound;
        switch (plane) {
            case 0: bound = sr[3] + sr[2]; break;
            case 1: bound = sr[2]; break;
            case 2: case 3: bound = sr[0] + 4.0f * sr[3]; break;
            default: bound = sr[1] + 4.0f * sr[3]; break;
        }
        const float d = plane_dist(c, plane), margin = bound * 0x1.0p-18f;
        if (d < -margin) bits |= 1u << plane;
        if (!(d > margin)) bits &= ~kCornerInside;
    }
    const float side[4] = {c.x + c.w, c.w - c.x, c.y + c.w, c.w - c.y};
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (side[q] < -(sr[q >> 1] + sr[3]) * 0x1.0p-18f) bits |= 1u << (6 + q);
    return bits;
}

// True when the bits of the eight corners in lanes 8 * g .. 8 * g + 7 of the wave show the box
// outside the view.
__device__ inline bool box_outside(uint32_t bits, uint32_t g) {
    uint32_t all = 0;
#pragma unroll
    for (int b = 0; b < 11; ++b)
        all |= (((__builtin_amdgcn_ballot_w64((bits >> b) & 1u) >> (8u * g)) & 0xffull) == 0xffull) ? 1u << b : 0u;
    return (all & 0x3fu) != 0u || ((all & kCornerInside) != 0u && (all & 0x3c0u) != 0u);
}

// clip = (VP * W) * p per (view, vertex).  Every block covers kTransformBlockVerts vertices of one
// draw, kTransformPerThread per thread (their loads issued together), for up to kTransformViews
// consecutive views that draw the same list: the positions are read once (coalesced 12 B from the
// SoA x/y/z streams) and each view's clip positions are written (coalesced 16 B).  Each view's
// MVP = views[v].view_projection * objects[draw.object].transform is formed in LDS by the fmaf chain
// in k order while the vertex loads are in flight.  A draw whose box is outside a view is skipped
// for that view (eight lanes test one box corner each), and the draw's first block records that
// for k_setup.
// Also zeroes the frame's counters and bin counts (`clear` words at `clear_a` then at `clear_b`), spread
// over the grid, before the setup kernel counts into them.
__global__ __launch_bounds__(256) void k_transform(const BlockEntry* blocks, const ViewGPU* views, DrawLists lists,
                                                   const float* objects, const float* px, const float* py,
                                                   const float* pz, float4* xv, uint32_t* draw_culled,
                                                   uint32_t* clear_a, uint32_t na, uint32_t* clear_b, uint32_t nb) {
    __shared__ float m[kTransformViews][16];
    __shared__ uint32_t culled[kTransformViews];
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < na + nb; i += gridDim.x * 256) {
        if (i < na) clear_a[i] = 0u;
        else clear_b[i - na] = 0u;
    }
    const BlockEntry e = blocks[blockIdx.x];
    const uint32_t nv = e.xv_first;
    const DrawGPU& dr = lists.of(e.list)[e.draw];  // its box only
    constexpr int R = int(kTransformPerThread);
    float x[R], y[R], z[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const uint32_t v = uint32_t(j) * 256u + threadIdx.x;
        x[j] = y[j] = z[j] = 0.f;
        if (v < e.count) {
            const uint32_t g = e.first + v;
            x[j] = px[g];
            y[j] = py[g];
            z[j] = pz[g];
        }
    }
    if (threadIdx.x < 16u * nv) {
        const uint32_t k = threadIdx.x >> 4;
        const int col = (threadIdx.x >> 2) & 3, row = threadIdx.x & 3;
        const float* a = views[e.view + k].view_projection;
        const float* b = objects + size_t(e.object) * 41u;  // 164-byte packed records
        float acc = a[row] * b[col * 4 + 0];
        acc = fmaf(a[4 + row], b[col * 4 + 1], acc);
        acc = fmaf(a[8 + row], b[col * 4 + 2], acc);
        acc = fmaf(a[12 + row], b[col * 4 + 3], acc);
        m[k][threadIdx.x & 15u] = acc;
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // wave 0: lanes 8k .. 8k + 7 test view k's box corners
        const uint32_t k = threadIdx.x >> 3;
        const bool outside = box_outside(k < nv ? corner_outside(m[k], dr.bmin, dr.bmax, int(threadIdx.x & 7u)) : 0u, k);
        if ((threadIdx.x & 7u) == 0 && k < nv) {
            culled[k] = outside ? 1u : 0u;
            if (e.start == 0) draw_culled[views[e.view + k].cull_base + e.cull] = outside ? 1u : 0u;
        }
    }
    __syncthreads();
    for (uint32_t k = 0; k < nv; ++k) {
        if (culled[k]) continue;
        float4* out = xv + size_t(views[e.view + k].xv_base) + e.out_first;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const uint32_t v = uint32_t(j) * 256u + threadIdx.x;
            if (v < e.count) out[v] = transform(m[k], x[j], y[j], z[j]);
        }
    }
}

// Bins of a record with at most kSmallBins bins go through the block's LDS table in k_setup; the
// others (and every clipper record) are binned by the whole wave: a wave prefix over their bin
// counts numbers the wave's (record, bin) pairs, one atomic reserves that many entries of the view's
// big-pair list, and every lane takes pairs in turn (its record found by a binary search over the
// prefix with lane shuffles), with one global atomic per pair for the slot inside the bin.
__device__ void bin_wide(bool active, const TriRec& r, uint32_t ri, const ViewGPU& V, uint32_t shift,
                         uint32_t* bin_counts, uint4* big_pairs, uint32_t* ctr) {
    int tx0 = 0, ty0 = 0, tx1 = -1, ty1 = -1;
    if (active) rec_tiles(r, shift, tx0, ty0, tx1, ty1);
    const int bw = tx1 - tx0 + 1;
    const uint32_t bn = active ? uint32_t(bw * (ty1 - ty0 + 1)) : 0u;
    const uint32_t lane = lane_id();
    const uint32_t incl = wave_inclusive_sum(bn);
    const uint32_t total = uint32_t(__builtin_amdgcn_readlane(int(incl), 63));
    if (total == 0) return;
    const uint32_t excl = incl - bn;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&ctr[kCtrBig], total);
    base = uint32_t(__builtin_amdgcn_readlane(int(base), 0));
    // kPairBatch pairs per lane at a time: their slot atomics are all in flight before the first
    // result is needed
    constexpr int kPairBatch = 4;
    for (uint32_t p0 = 0; p0 < total; p0 += 64u * kPairBatch) {
        uint32_t key[kPairBatch], local[kPairBatch], rid[kPairBatch];
#pragma unroll
        for (int j = 0; j < kPairBatch; ++j) {
            const uint32_t p = p0 + 64u * uint32_t(j) + lane;
            // owner: the last lane whose exclusive prefix is <= p (lanes without pairs share the
            // prefix of the next lane that has some, and come before it)
            int lo = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
                const uint32_t e = uint32_t(__shfl(int(excl), lo + step));
                if (e <= p) lo += step;
            }
            const uint32_t o_excl = uint32_t(__shfl(int(excl), lo));
            const int o_tx0 = __shfl(tx0, lo), o_ty0 = __shfl(ty0, lo), o_bw = __shfl(bw, lo);
            rid[j] = uint32_t(__shfl(int(ri), lo));
            key[j] = 0u;
            local[j] = 0u;
            if (p < total) {
                const int k = int(p - o_excl);
                key[j] = V.tile_base + uint32_t((o_ty0 + k / o_bw) * int(V.tiles_x) + o_tx0 + k % o_bw);
                local[j] = atomicAdd(&bin_counts[key[j]], 1u);
            }
        }
#pragma unroll
        for (int j = 0; j < kPairBatch; ++j) {
            const uint32_t p = p0 + 64u * uint32_t(j) + lane;
            if (p < total) {
                if (base + p < V.big_cap) big_pairs[V.big_base + base + p] = make_uint4(key[j], local[j], rid[j], 0u);
                else atomicOr(&ctr[kCtrOverflow], 4u);
            }
        }
    }
}

constexpr uint32_t kEmpty = 0xffffffffu;

// A block's work entry, read as three 16-byte words.
__device__ inline BlockEntry load_entry(const BlockEntry* blocks, uint32_t b) {
    const uint4* p = reinterpret_cast<const uint4*>(blocks + b);
    const uint4 w0 = p[0], w1 = p[1], w2 = p[2];
    BlockEntry e;
    e.view = w0.x; e.draw = w0.y; e.start = w0.z; e.list = w0.w;
    e.first = w1.x; e.count = w1.y; e.out_first = w1.z; e.xv_first = w1.w;
    e.cull = w2.x; e.object = w2.y; e.vtx_base = w2.z; e.material = w2.w;
    return e;
}

// One setup block: 256 threads, kTpt triangles each (a block covers 256 * kTpt triangles of one
// view and draw).  Two per thread pay on frames with many blocks per resident slot: each thread
// keeps two triangles' loads in flight, at fewer waves per SIMD.  The LDS bin table holds up to
// 256 * kTpt * kSmallBins distinct keys, so a probe always finds its key or an empty entry.
template <int kTpt>
// (the arguments that a block needs before its culling decision come first: the first 64 bytes of
// the arguments are loaded with the kernel's start, the rest when first used)
__global__ __launch_bounds__(256) void k_setup(const BlockEntry* blocks, const ViewGPU* views,
                                               const uint32_t* draw_culled, const uint32_t* indices,
                                               const float4* xv, uint32_t* bin_tiles, uint32_t* block_valid,
                                               TriRec* recs, ShadeRec* shade_recs, ClipEntry* clipq,
                                               uint32_t* counters, uint32_t* bin_counts, uint32_t* pair_slots,
                                               uint32_t slot_count, uint4* big_pairs, uint32_t shift) {
    constexpr uint32_t kHash = 256u * kTpt * kSmallBins;
    constexpr uint32_t kHashShift = kTpt == 1 ? 22u : 21u;  // 10- or 11-bit multiplicative hash
    static_assert(kHash == (1u << (32u - kHashShift)), "hash bits");
    __shared__ uint32_t hkey[kHash];
    __shared__ uint32_t hcount[kHash];  // per key: records of the block, then the block's first slot
    __shared__ uint32_t wave_valid[4];
    for (uint32_t i = threadIdx.x; i < kHash; i += 256) {
        hkey[i] = kEmpty;
        hcount[i] = 0u;
    }
    const BlockEntry e = load_entry(blocks, blockIdx.x);
    uint32_t* ctr = counters + e.view * kCtrPerView;
    // the thread's triangles of the block (t[j]) and their indices, requested first, then the view
    // record and the draw's culling flag
    uint32_t t[kTpt], slot[kTpt];
    bool valid[kTpt];
    uint32_t vi[kTpt][3];
#pragma unroll
    for (int j = 0; j < kTpt; ++j) {
        t[j] = threadIdx.x + 256u * uint32_t(j);
        valid[j] = t[j] < e.count;
        slot[j] = e.out_first + t[j];  // every triangle owns one record slot
        vi[j][0] = vi[j][1] = vi[j][2] = 0u;
        if (valid[j]) {
            const uint32_t* ip = indices + e.first + 3u * t[j];
            vi[j][0] = ip[0];
            vi[j][1] = ip[1];
            vi[j][2] = ip[2];
        }
    }
    // this block's share of the view's output words, cleared (for a view whose blocks' shares are
    // small; the host clears the others with a fill launch): k_raster then stores every bin that
    // holds records, and merges the items of a split bin into its cleared words.  A view's words are
    // split over its setup blocks in proportion to their record slots, which tile the view's
    // triangles (the stores overlap the block's loads instead of taking a pass of their own)
    const uint32_t culled = draw_culled[e.cull];  // requested before the clear
    const ViewGPU V = views[e.view];
    // word boundary of slot s: min(units, floor(s * units / tri_count)) in FP64, units at the view's
    // last slot -- non-decreasing in s, so the blocks' ranges tile the view's words.  The range is
    // formed before the branch (the empty asm takes it as an input), so that the view fields it
    // reads are requested with the others instead of after the branch's own wait.
    const uint64_t units = uint64_t(V.width) * V.height * (V.is_shadow ? 1u : V.samples);
    const double per_slot = double(units) / double(max(V.tri_count, 1u));
    const uint32_t s0 = e.out_first - V.rec_base, s1 = s0 + e.count;
    const uint64_t u0 = min(units, uint64_t(double(s0) * per_slot));
    const uint64_t u1 = s1 >= V.tri_count ? units : min(units, uint64_t(double(s1) * per_slot));
    const uint64_t out = V.out;
    asm volatile("" ::"s"(u0), "s"(u1), "s"(out));
    if (V.setup_clear) {
        if (V.is_shadow)
            for (uint64_t u = u0 + threadIdx.x; u < u1; u += 256u) reinterpret_cast<uint32_t*>(out)[u] = 0u;
        else
            for (uint64_t u = u0 + threadIdx.x; u < u1; u += 256u) reinterpret_cast<unsigned long long*>(out)[u] = 0ull;
    }
    if (culled) {  // the whole draw is outside the view: empty slots only
#pragma unroll
        for (int j = 0; j < kTpt; ++j)
            if (valid[j]) bin_tiles[slot[j]] = kNoBins;
        if (threadIdx.x == 0) block_valid[blockIdx.x] = 0u;
        return;
    }
    float4 c[kTpt][3];
#pragma unroll
    for (int j = 0; j < kTpt; ++j)
        if (valid[j]) {
            const float4* base = xv + size_t(e.xv_first);
            c[j][0] = base[vi[j][0]];
            c[j][1] = base[vi[j][1]];
            c[j][2] = base[vi[j][2]];
        }
    TriRec r[kTpt];
    bool has[kTpt], small[kTpt];
    int tx0[kTpt], ty0[kTpt], bw[kTpt], n[kTpt];
#pragma unroll
    for (int j = 0; j < kTpt; ++j) {
        int status = 0;  // 0 reject, 1 accept, 2 clip
        if (valid[j]) {
            bool reject = false;
#pragma unroll
            for (int plane = 0; plane < 6; ++plane) {
                const int outside = int(plane_dist(c[j][0], plane) < 0.f) + int(plane_dist(c[j][1], plane) < 0.f) +
                                    int(plane_dist(c[j][2], plane) < 0.f);
                reject |= outside == 3;
            }
            status = reject ? 0 : (all_inside(c[j][0], c[j][1], c[j][2]) ? 1 : 2);
        }
        has[j] = status == 1 && setup_tri(c[j], e.object + t[j], V, r[j]);
        if (valid[j]) {
            if (has[j]) recs[slot[j]] = r[j];
            // a camera or blend view's triangle that a pixel can show (set up into a raster record,
            // or queued for clipping): its shading record; culled and boxless triangles need none
            if (!V.is_shadow && (has[j] || status == 2)) {
                uint4* sr = reinterpret_cast<uint4*>(shade_recs + slot[j]);
                sr[0] = make_uint4(e.vtx_base + vi[j][0], e.vtx_base + vi[j][1], e.vtx_base + vi[j][2],
                                   e.draw | (has[j] ? 0u : kShadeNoRecord));
                sr[1] = make_uint4(__float_as_uint(c[j][0].w), __float_as_uint(c[j][1].w), __float_as_uint(c[j][2].w),
                                   e.material);
            }
        }
        const uint32_t cslot = wave_append(status == 2, &ctr[kCtrClip]);
        if (status == 2) {
            if (cslot < V.clip_cap) {
                ClipEntry& q = clipq[V.clip_base + cslot];
                q.v[0] = c[j][0];
                q.v[1] = c[j][1];
                q.v[2] = c[j][2];
                q.id = e.object + t[j];
            } else {
                atomicOr(&ctr[kCtrOverflow], 1u);
            }
        }
        int x0 = 0, y0 = 0, x1 = -1, y1 = -1;
        if (has[j]) rec_tiles(r[j], shift, x0, y0, x1, y1);
        tx0[j] = x0;
        ty0[j] = y0;
        bw[j] = x1 - x0 + 1;
        n[j] = has[j] ? bw[j] * (y1 - y0 + 1) : 0;
        small[j] = has[j] && n[j] <= kSmallBins;
        if (valid[j]) bin_tiles[slot[j]] = small[j] ? pack_bins(x0, y0, bw[j], y1 - y0 + 1) : kNoBins;
    }
    __syncthreads();  // table cleared
    // small records: rank within the block for each bin, through the LDS table
    uint32_t hidx[kTpt][kSmallBins], rank[kTpt][kSmallBins];
#pragma unroll
    for (int j = 0; j < kTpt; ++j)
#pragma unroll
        for (int k = 0; k < kSmallBins; ++k) {
            hidx[j][k] = 0u;
            rank[j][k] = 0u;
            if (small[j] && k < n[j]) {
                const uint32_t key = V.tile_base + uint32_t((ty0[j] + k / bw[j]) * int(V.tiles_x) + tx0[j] + k % bw[j]);
                uint32_t h = (key * 2654435761u) >> kHashShift;
                while (true) {
                    const uint32_t prev = atomicCAS(&hkey[h], kEmpty, key);
                    if (prev == kEmpty || prev == key) break;
                    h = (h + 1) & (kHash - 1);
                }
                hidx[j][k] = h;
                rank[j][k] = atomicAdd(&hcount[h], 1u);
            }
        }
    __syncthreads();
    // one global atomic per distinct bin of the block: the block's first slot in that bin.  A thread
    // owns kHash / 256 table entries; their atomics are all issued before the first result is
    // stored, so that the thread waits for one round trip instead of one per entry
    {
        constexpr uint32_t kOwn = kHash / 256u;
        uint32_t key[kOwn], got[kOwn];
#pragma unroll
        for (uint32_t q = 0; q < kOwn; ++q) {
            key[q] = hkey[threadIdx.x + 256u * q];
            got[q] = hcount[threadIdx.x + 256u * q];
        }
#pragma unroll
        for (uint32_t q = 0; q < kOwn; ++q)
            if (key[q] != kEmpty) got[q] = atomicAdd(&bin_counts[key[q]], got[q]);
#pragma unroll
        for (uint32_t q = 0; q < kOwn; ++q)
            if (key[q] != kEmpty) hcount[threadIdx.x + 256u * q] = got[q];
    }
    __syncthreads();
    uint32_t nvalid = 0;
#pragma unroll
    for (int j = 0; j < kTpt; ++j) {
        if (small[j]) {  // slots in bins 0-1, and in bins 2-3 only for a record with more than two bins
            uint2* s01 = reinterpret_cast<uint2*>(pair_slots);
            s01[slot[j]] = make_uint2(hcount[hidx[j][0]] + rank[j][0], n[j] > 1 ? hcount[hidx[j][1]] + rank[j][1] : 0u);
            if (n[j] > 2)
                s01[slot_count + slot[j]] =
                    make_uint2(hcount[hidx[j][2]] + rank[j][2], n[j] > 3 ? hcount[hidx[j][3]] + rank[j][3] : 0u);
        }
        // records with many bins: the whole wave spreads their bins
        bin_wide(has[j] && !small[j], r[j], slot[j], V, shift, bin_counts, big_pairs, ctr);
        nvalid += uint32_t(__popcll(__ballot(has[j])));
    }
    if (lane_id() == 0) wave_valid[threadIdx.x >> 6] = nvalid;
    __syncthreads();
    if (threadIdx.x == 0) block_valid[blockIdx.x] = wave_valid[0] + wave_valid[1] + wave_valid[2] + wave_valid[3];
}

// One thread per queued triangle: Sutherland-Hodgman against the near, far and guard-band planes
// in that order (intersections from the inside vertex), the polygon kept in the thread's slice of
// LDS, then the fan (v0, vk, vk+1) set up as records; record slots are handed out per wave and
// the fan triangles are binned by the whole wave.
constexpr int kClipThreads = 128;
// A triangle cut by 6 planes has at most 9 vertices in exact arithmetic.  In FP32 a distance that
// rounds to the other side near a corner can add a crossing: a vertex past the bound is dropped (the
// oracle's clipper holds the same bound), so the polygon never leaves its thread's slice.
constexpr int kClipMax = 9;

__global__ __launch_bounds__(kClipThreads) void k_clip(const ViewGPU* views, const ClipEntry* clipq, TriRec* recs,
                                                       uint32_t* counters, uint32_t* bin_counts, uint4* big_pairs,
                                                       uint32_t shift) {
    __shared__ float4 poly[kClipThreads][2][kClipMax];
    const uint32_t view = blockIdx.y;
    const ViewGPU V = views[view];
    uint32_t* ctr = counters + view * kCtrPerView;
    const uint32_t n = min(ctr[kCtrClip], V.clip_cap);
    for (uint32_t base = blockIdx.x * kClipThreads; base < n; base += gridDim.x * kClipThreads) {
        const uint32_t q = base + threadIdx.x;
        const bool active = q < n;
        float4* A = poly[threadIdx.x][0];
        float4* B = poly[threadIdx.x][1];
        int cnt = 0;
        uint32_t id = 0;
        if (active) {
            const ClipEntry& ent = clipq[V.clip_base + q];
            A[0] = ent.v[0];
            A[1] = ent.v[1];
            A[2] = ent.v[2];
            id = ent.id;
            cnt = 3;
            for (int plane = 0; plane < 6 && cnt >= 3; ++plane) {
                int m = 0;
                for (int i = 0; i < cnt; ++i) {
                    const float4 cur = A[i];
                    const float4 nxt = A[i + 1 < cnt ? i + 1 : 0];
                    const float dc = plane_dist(cur, plane), dn = plane_dist(nxt, plane);
                    const bool ic = dc >= 0.f, in_ = dn >= 0.f;
                    if (ic && m < kClipMax) B[m++] = cur;
                    if (ic != in_ && m < kClipMax) {
                        const float4 pi = ic ? cur : nxt;
                        const float4 po = ic ? nxt : cur;
                        const float di = ic ? dc : dn, dout = ic ? dn : dc;
                        const float tt = di / (di - dout);
                        float4 qv;
                        qv.x = pi.x + tt * (po.x - pi.x);
                        qv.y = pi.y + tt * (po.y - pi.y);
                        qv.z = pi.z + tt * (po.z - pi.z);
                        qv.w = pi.w + tt * (po.w - pi.w);
                        B[m++] = qv;
                    }
                }
                cnt = m;
                float4* t = A;
                A = B;
                B = t;
            }
            bool ok = cnt >= 3;
            for (int i = 0; i < cnt && ok; ++i) ok = A[i].w > 0.f;
            if (!ok) cnt = 0;
        }
        const uint32_t nclipped = uint32_t(__popcll(__ballot(active)));
        if (lane_id() == 0 && nclipped) atomicAdd(&ctr[kCtrClipped], nclipped);
        // fan triangle k = (v0, vk, vk+1): the ones that set up into records (bit k - 1), then one
        // slot range for the wave's records, then each record set up again, stored and binned
        uint32_t fan = 0;
        for (int k = 1; k + 1 < cnt; ++k) {
            TriRec r;
            const float4 tv[3] = {A[0], A[k], A[k + 1]};
            if (setup_tri(tv, id, V, r)) fan |= 1u << (k - 1);
        }
        const uint32_t nrec = uint32_t(__popc(fan));
        const uint32_t incl = wave_inclusive_sum(nrec);
        const uint32_t wave_total = uint32_t(__builtin_amdgcn_readlane(int(incl), 63));
        uint32_t slot = 0;
        if (lane_id() == 0 && wave_total) slot = atomicAdd(&ctr[kCtrRecords], wave_total);
        slot = uint32_t(__builtin_amdgcn_readlane(int(slot), 0)) + incl - nrec;
        for (int k = 1; k <= kClipMax - 2; ++k) {
            if (__ballot((fan >> (k - 1)) != 0u) == 0ull) break;  // no lane has fan triangle k or later
            TriRec r;
            bool has = false;
            if ((fan >> (k - 1)) & 1u) {
                const float4 tv[3] = {A[0], A[k], A[k + 1]};
                has = setup_tri(tv, id, V, r);
            }
            uint32_t ri = 0;
            if (has) {
                if (V.tri_count + slot < V.rec_cap) {
                    ri = V.rec_base + V.tri_count + slot;
                    recs[ri] = r;
                } else {
                    atomicOr(&ctr[kCtrOverflow], 2u);
                    has = false;
                }
                ++slot;
            }
            bin_wide(has, r, ri, V, shift, bin_counts, big_pairs, ctr);
        }
    }
}

// Bin scan geometry: each scan block covers kScanPerThread bins per thread.
constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanPerThread = 4;

// Raster work items: a bin with 1..kItemRecs records is one item (its tile is stored whole); a
// larger bin is split into one item per kItemRecs records, and its items merge into the output
// with device atomic max into the cleared pixels.  Max is order independent, so the result is the
// same whichever item finishes first.  An empty bin has no item: its pixels keep the clear (made by
// the view's setup blocks, or by a fill launch for a view with few setup blocks).  One table holds every item as (bin, first list entry, records, split),
// ordered so that the longest items start first and the short ones fill the end of the launch:
// whole bins with at least kHeavyRecs records, then those with at least kMidRecs, then the split
// bins' items (each a share of a crowded bin), then the remaining whole bins.
constexpr uint32_t kHeavyRecs = 128;
constexpr uint32_t kMidRecs = 48;

// N exclusive prefix sums over the workgroup at once (one pair of barriers for all of them): v[j]
// becomes its exclusive prefix and total[j] the workgroup's sum.  lds_waves holds 16 words per value.
// The wave totals are scanned by the first row of 16 lanes of every wave (one LDS read per value)
// and the wave's base and the block total read from that row's lanes.
template <int N>
__device__ inline void block_exclusive_scan64_n(uint64_t (&v)[N], uint64_t* lds_waves, uint64_t (&total)[N]) {
    uint64_t x[N];
#pragma unroll
    for (int j = 0; j < N; ++j) x[j] = wave_inclusive_sum64(v[j]);
    const uint32_t wave = threadIdx.x / 64, nwaves = blockDim.x / 64;
    if (lane_id() == 63)
#pragma unroll
        for (int j = 0; j < N; ++j) lds_waves[j * 16 + wave] = x[j];
    __syncthreads();
    const uint32_t lane = lane_id();
#pragma unroll
    for (int j = 0; j < N; ++j) {
        uint64_t t = lane < nwaves ? lds_waves[j * 16 + lane] : 0ull;
        t += dpp_move64<0x111, 0xf>(t);  // row_shr:1, 2, 4, 8: lanes 0..15 hold the waves' inclusive prefix
        t += dpp_move64<0x112, 0xf>(t);
        t += dpp_move64<0x114, 0xf>(t);
        t += dpp_move64<0x118, 0xf>(t);
        const uint64_t wave_off = wave == 0 ? 0ull : readlane64(t, wave - 1);
        total[j] = readlane64(t, nwaves - 1);
        v[j] = wave_off + x[j] - v[j];
    }
    __syncthreads();
}

// Bin offsets (exclusive prefix of the counts) and the work-item table, in two launches: block
// partial sums, then each block's final scan (which scans the partials itself for its base).
// Each block covers kScanChunk bins with two 64-bit block scans: one of (record count | split
// items << 32) -- the counts of a frame stay below 2^32, so the low half never carries -- and one
// of four 16-bit bin classes (empty or split bins, heavy, mid and light whole bins), each at most
// kScanChunk per block.
constexpr uint32_t kScanChunk = kScanThreads * kScanPerThread;
static_assert(kScanChunk <= 65535, "16-bit bin class fields");
static_assert(kScanChunk == kScanBlockBins, "scan partial buffer sizing");

struct ScanPartial {
    uint64_t a;      // records | split items << 32
    uint32_t f[4];   // empty or split bins, heavy, mid, light whole bins
    uint32_t valid;  // records set up by this block's share of the setup blocks (a frame statistic)
    uint32_t pad;
};
static_assert(sizeof(ScanPartial) == kScanPartialBytes, "scan partial size");

enum BinClass : uint32_t { kClassZero = 0, kClassHeavy = 1, kClassMid = 2, kClassLight = 3, kClassSplit = 4 };

__device__ inline uint32_t bin_class(uint32_t v) {
    if (v > kItemRecs) return kClassSplit;
    if (v == 0) return kClassZero;
    return v >= kHeavyRecs ? kClassHeavy : (v >= kMidRecs ? kClassMid : kClassLight);
}

__device__ inline void bin_fields(uint32_t v, uint64_t& a, uint64_t& b) {
    const uint32_t cls = bin_class(v);
    if (cls == kClassSplit) {
        a += uint64_t(v) | (uint64_t((v + kItemRecs - 1) / kItemRecs) << 32);
        b += 1u;  // counted with the empty bins: no whole-bin item
    } else {
        a += v;
        b += uint64_t(1) << (16u * cls);
    }
}

// The frame's totals (block 0, thread 0 of the scan) from the sums over all bins: a (records |
// split items << 32), f01 and f23 (the class counts as 32-bit halves) and the set-up records.
__device__ inline void scan_write_totals(uint64_t a, uint64_t f01, uint64_t f23, uint32_t valid, uint32_t list_cap,
                                         uint32_t items_cap, uint32_t raster_grid, uint32_t* totals) {
    const uint32_t n_split = uint32_t(a >> 32);
    const uint32_t n_heavy = uint32_t(f01 >> 32), n_mid = uint32_t(f23);
    const uint32_t pairs = uint32_t(a), n_zero = uint32_t(f01), n_light = uint32_t(f23 >> 32);
    const uint64_t n_items = uint64_t(n_heavy) + n_mid + n_split + n_light;
    totals[0] = pairs;
    totals[1] = valid;
    // bit 0: bin lists or item table too small; bit 1: more items than raster workgroups
    totals[2] = ((pairs > list_cap || n_items > items_cap) ? 1u : 0u) | (n_items > raster_grid ? 2u : 0u);
    totals[3] = uint32_t(min<uint64_t>(n_items, items_cap));
    totals[4] = n_heavy;                    // first mid item
    totals[5] = n_zero;                     // empty and split bins (a statistic)
    totals[6] = n_heavy + n_mid;            // first split item
    totals[7] = n_heavy + n_mid + n_split;  // first light item
}

// A thread's P consecutive bins from i0: list offsets and work items, from the
// block's base over the earlier bins (base), the thread's exclusive prefix inside the block (ea:
// records | split items << 32, eb: 16-bit class fields) and the frame's class totals.
template <uint32_t P>
__device__ inline void scan_write_bins(const uint32_t (&v)[P], uint32_t i0, uint32_t n, const ScanPartial& base,
                                       uint64_t ea_local, uint64_t eb, uint32_t n_heavy, uint32_t n_mid,
                                       uint32_t n_split, uint32_t* offsets, uint4* items, uint32_t items_cap) {
    const uint64_t ea = base.a + ea_local;
    uint32_t run = uint32_t(ea);
    uint64_t r_split = uint64_t(n_heavy) + n_mid + uint32_t(ea >> 32);
    // next item slot of each whole-bin class: the class's first item + earlier blocks + this block
    uint64_t r_heavy = base.f[1] + ((eb >> 16) & 0xffffu);
    uint64_t r_mid = uint64_t(n_heavy) + base.f[2] + ((eb >> 32) & 0xffffu);
    uint64_t r_light = uint64_t(n_heavy) + n_mid + n_split + base.f[3] + ((eb >> 48) & 0xffffu);
#pragma unroll
    for (uint32_t k = 0; k < P; ++k)
        if (i0 + k < n) {
            offsets[i0 + k] = run;
            const uint32_t cls = bin_class(v[k]);
            if (cls == kClassSplit) {
                const uint32_t ni = (v[k] + kItemRecs - 1) / kItemRecs;
                for (uint32_t c = 0; c < ni; ++c, ++r_split)
                    if (r_split < items_cap)
                        items[r_split] = make_uint4(i0 + k, run + c * kItemRecs, min(kItemRecs, v[k] - c * kItemRecs), 1u);
            } else if (cls != kClassZero) {
                const uint64_t slot = cls == kClassHeavy ? r_heavy++ : (cls == kClassMid ? r_mid++ : r_light++);
                if (slot < items_cap) items[slot] = make_uint4(i0 + k, run, v[k], 0u);
            }
            run += v[k];
        }
}

// Block sums of the bin fields; each block also sums an equal share of the setup blocks' record
// counts, so that no block walks the whole table.
__global__ __launch_bounds__(kScanThreads) void k_scan_partial(const uint32_t* counts, uint32_t n, ScanPartial* partial,
                                                               const uint32_t* block_valid, uint32_t n_blocks) {
    __shared__ uint64_t lds64[3 * 16];
    const uint32_t i0 = blockIdx.x * kScanChunk + threadIdx.x * kScanPerThread;
    uint64_t a = 0, b = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanPerThread; ++k)
        if (i0 + k < n) bin_fields(counts[i0 + k], a, b);
    const uint32_t share = (n_blocks + gridDim.x - 1) / gridDim.x;
    const uint32_t v0 = blockIdx.x * share, v1 = min(n_blocks, v0 + share);
    uint32_t vs = 0;
    for (uint32_t i = v0 + threadIdx.x; i < v1; i += kScanThreads) vs += block_valid[i];
    uint64_t sv[3] = {a, b, vs}, st[3];
    block_exclusive_scan64_n<3>(sv, lds64, st);
    if (threadIdx.x == 0) {
        ScanPartial p{};
        p.a = st[0];
        p.valid = uint32_t(st[2]);
        for (int k = 0; k < 4; ++k) p.f[k] = uint32_t(st[1] >> (16 * k)) & 0xffffu;
        partial[blockIdx.x] = p;
    }
}

// Every block scans the block partials itself (there are few: one per kScanChunk bins) for its
// base and the totals; block 0 writes the totals.  Then the block's bins get their list offsets,
// raster work items.
__global__ __launch_bounds__(kScanThreads) void k_scan_final(const uint32_t* counts, uint32_t n, uint32_t* offsets,
                                                     const ScanPartial* partial, uint32_t n_parts,
                                                     uint32_t list_cap, uint32_t* totals, uint4* items,
                                                     uint32_t items_cap, uint32_t raster_grid) {
    __shared__ uint64_t lds64[4 * 16];
    __shared__ ScanPartial base_lds;
    // this block's counts are requested before the scan of the partials
    const uint32_t i0 = blockIdx.x * kScanChunk + threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread];
#pragma unroll
    for (uint32_t k = 0; k < kScanPerThread; ++k) v[k] = i0 + k < n ? counts[i0 + k] : 0u;
    uint64_t carry_a = 0, carry_01 = 0, carry_23 = 0;
    uint32_t valid = 0;
    for (uint32_t at = 0; at < n_parts; at += kScanThreads) {
        const uint32_t i = at + threadIdx.x;
        ScanPartial p{};
        if (i < n_parts) p = partial[i];
        // class counts of a block fit 16 bits, their sums over blocks 32 bits: two classes per scan
        const uint64_t v01 = uint64_t(p.f[0]) | (uint64_t(p.f[1]) << 32);
        const uint64_t v23 = uint64_t(p.f[2]) | (uint64_t(p.f[3]) << 32);
        uint64_t sv[4] = {p.a, v01, v23, p.valid}, st[4];
        block_exclusive_scan64_n<4>(sv, lds64, st);
        const uint64_t ea = carry_a + sv[0], e01 = carry_01 + sv[1], e23 = carry_23 + sv[2];
        if (i == blockIdx.x) {
            ScanPartial q{};
            q.a = ea;
            q.f[0] = uint32_t(e01);
            q.f[1] = uint32_t(e01 >> 32);
            q.f[2] = uint32_t(e23);
            q.f[3] = uint32_t(e23 >> 32);
            base_lds = q;
        }
        carry_a += st[0];
        carry_01 += st[1];
        carry_23 += st[2];
        valid += uint32_t(st[3]);  // the frame's set-up records
    }
    const uint32_t n_split = uint32_t(carry_a >> 32);
    const uint32_t n_heavy = uint32_t(carry_01 >> 32), n_mid = uint32_t(carry_23);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        scan_write_totals(carry_a, carry_01, carry_23, valid, list_cap, items_cap, raster_grid, totals);
    __syncthreads();
    uint64_t a = 0, b = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanPerThread; ++k)
        if (i0 + k < n) bin_fields(v[k], a, b);
    uint64_t sv[2] = {a, b}, st[2];
    const ScanPartial base = base_lds;
    block_exclusive_scan64_n<2>(sv, lds64, st);
    scan_write_bins<kScanPerThread>(v, i0, n, base, sv[0], sv[1], n_heavy, n_mid, n_split, offsets, items, items_cap);
}

// The same scan in one launch for frames of at most kDirectScanBins bins: kDirectThreads threads
// per block, one bin per thread.  Every block reduces the counts of all bins itself (thread t takes
// bins t, t + kDirectThreads, ..., their loads in flight together): those before its own bins give
// its base, all of them the class totals.  No block waits for another.  Blocks of 512 threads for
// frames of at most kSmallScanBins bins (their block scans finish sooner, and every thread still
// has one batch of loads), of 1024 above.
constexpr uint32_t kDirectScanBins = 65536;  // per-thread class counts stay below 2^16
constexpr uint32_t kSmallScanBins = 16384;
constexpr uint32_t kDirectBatch = 16;
template <uint32_t kDirectThreads>
__global__ __launch_bounds__(kDirectThreads) void k_scan_direct(const uint32_t* counts, uint32_t n, uint32_t* offsets,
                                                                const uint32_t* block_valid, uint32_t n_blocks,
                                                                uint32_t list_cap, uint32_t* totals, uint4* items,
                                                                uint32_t items_cap, uint32_t raster_grid) {
    __shared__ uint64_t lds64[7 * 16];
    const uint32_t i0 = blockIdx.x * kDirectThreads + threadIdx.x;
    const uint32_t lo = blockIdx.x * kDirectThreads;
    uint32_t v[1];
    v[0] = i0 < n ? counts[i0] : 0u;
    uint64_t ba = 0, bb = 0, ta = 0, tb = 0;  // before this block's bins / all bins
    for (uint32_t j = threadIdx.x; j < n; j += kDirectBatch * kDirectThreads) {
        uint32_t c[kDirectBatch];
#pragma unroll
        for (uint32_t k = 0; k < kDirectBatch; ++k) c[k] = j + k * kDirectThreads < n ? counts[j + k * kDirectThreads] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < kDirectBatch; ++k) {
            const uint32_t i = j + k * kDirectThreads;
            if (i < n) {
                uint64_t a1 = 0, b1 = 0;
                bin_fields(c[k], a1, b1);
                ta += a1;
                tb += b1;
                if (i < lo) {
                    ba += a1;
                    bb += b1;
                }
            }
        }
    }
    uint32_t vs = 0;
    if (blockIdx.x == 0)  // the set-up records, kDirectBatch loads in flight together
        for (uint32_t j = threadIdx.x; j < n_blocks; j += kDirectBatch * kDirectThreads) {
            uint32_t c[kDirectBatch];
#pragma unroll
            for (uint32_t k = 0; k < kDirectBatch; ++k)
                c[k] = j + k * kDirectThreads < n_blocks ? block_valid[j + k * kDirectThreads] : 0u;
#pragma unroll
            for (uint32_t k = 0; k < kDirectBatch; ++k) vs += c[k];
        }
    // the four 16-bit class fields as two pairs of 32-bit halves for the sums over the block
    const uint64_t b01 = (bb & 0xffffu) | (((bb >> 16) & 0xffffu) << 32), b23 = ((bb >> 32) & 0xffffu) | ((bb >> 48) << 32);
    const uint64_t t01 = (tb & 0xffffu) | (((tb >> 16) & 0xffffu) << 32), t23 = ((tb >> 32) & 0xffffu) | ((tb >> 48) << 32);
    uint64_t sr[7] = {ba, b01, b23, ta, t01, t23, vs}, tot[7];
    block_exclusive_scan64_n<7>(sr, lds64, tot);
    ScanPartial base{};
    base.a = tot[0];
    base.f[0] = uint32_t(tot[1]);
    base.f[1] = uint32_t(tot[1] >> 32);
    base.f[2] = uint32_t(tot[2]);
    base.f[3] = uint32_t(tot[2] >> 32);
    const uint32_t n_split = uint32_t(tot[3] >> 32);
    const uint32_t n_heavy = uint32_t(tot[4] >> 32), n_mid = uint32_t(tot[5]);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        scan_write_totals(tot[3], tot[4], tot[5], uint32_t(tot[6]), list_cap, items_cap, raster_grid, totals);
    uint64_t a = 0, b = 0;
    if (i0 < n) bin_fields(v[0], a, b);
    uint64_t sv[2] = {a, b}, st[2];
    block_exclusive_scan64_n<2>(sv, lds64, st);
    scan_write_bins<1>(v, i0, n, base, sv[0], sv[1], n_heavy, n_mid, n_split, offsets, items, items_cap);
}

// Bin list entries of B record slots of a view (i0, i0 + stride, ...: those below the view's
// triangle count; records with more than kSmallBins bins go through the big-pair list): every
// slot's bin word is requested first, then every record's bin slots and bin offsets, then the
// entries are written (a store ahead of a load would otherwise hold the next slot's loads back).
// B = 2 pays when a thread has more than one slot.
template <int B>
__device__ inline void scatter_slots(const ViewGPU& V, const uint32_t* bin_tiles, const uint32_t* bin_offsets,
                                     const uint32_t* pair_slots, uint32_t slot_count, uint32_t* lists,
                                     uint32_t list_cap, uint32_t i0, uint32_t stride) {
    uint32_t w[B];
#pragma unroll
    for (int j = 0; j < B; ++j) {
        const uint32_t i = i0 + uint32_t(j) * stride;
        w[j] = i < V.tri_count ? bin_tiles[V.rec_base + i] : kNoBins;
    }
    uint32_t nb[B], base[B][kSmallBins];
    uint4 sl[B];
#pragma unroll
    for (int j = 0; j < B; ++j) {
        nb[j] = 0u;
        sl[j] = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int k = 0; k < kSmallBins; ++k) base[j][k] = 0u;
        if (w[j] != kNoBins) {
            const int tx0 = int(w[j] & 0x1fffu), ty0 = int((w[j] >> 13) & 0x1fffu);
            const int bw = int((w[j] >> 26) & 3u) + 1, bh = int((w[j] >> 28) & 3u) + 1;
            nb[j] = uint32_t(bw * bh);
            const uint2* s01 = reinterpret_cast<const uint2*>(pair_slots);
            const uint32_t ri = V.rec_base + i0 + uint32_t(j) * stride;
            const uint2 a = s01[ri];
            const uint2 b = nb[j] > 2u ? s01[slot_count + ri] : make_uint2(0u, 0u);
            sl[j] = make_uint4(a.x, a.y, b.x, b.y);
#pragma unroll
            for (int k = 0; k < kSmallBins; ++k) {
                const uint32_t key = V.tile_base + uint32_t((ty0 + k / bw) * int(V.tiles_x) + tx0 + k % bw);
                base[j][k] = uint32_t(k) < nb[j] ? bin_offsets[key] : 0u;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < B; ++j) {
        const uint32_t ri = V.rec_base + i0 + uint32_t(j) * stride;
        const uint32_t slots[kSmallBins] = {sl[j].x, sl[j].y, sl[j].z, sl[j].w};
#pragma unroll
        for (int k = 0; k < kSmallBins; ++k) {
            const uint32_t pos = base[j][k] + slots[k];
            if (uint32_t(k) < nb[j] && pos < list_cap) lists[pos] = ri;
        }
    }
}

// One (bin, slot, record) entry of the big-pair list into its bin's list.
__device__ inline void scatter_big_pair(const uint4 bp, const uint32_t* bin_offsets, uint32_t* lists, uint32_t list_cap) {
    const uint32_t pos = bin_offsets[bp.x] + bp.y;
    if (pos < list_cap) lists[pos] = bp.z;
}

// Bin lists without atomics: every (record, bin) pair already holds its slot inside the bin (from
// k_setup's block table, or the big-pair list), so the entry goes to bin_offsets[bin] + slot.
__global__ __launch_bounds__(256) void k_scatter(const ViewGPU* views, const uint32_t* bin_tiles,
                                                 const uint32_t* counters, const uint32_t* bin_offsets,
                                                 const uint32_t* pair_slots, uint32_t slot_count,
                                                 const uint4* big_pairs, uint32_t* lists, uint32_t list_cap) {
    const uint32_t view = blockIdx.y;
    const ViewGPU V = views[view];
    const uint32_t nbig = min(counters[view * kCtrPerView + kCtrBig], V.big_cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    if (V.tri_count > 2u * stride) {  // more than two record slots per thread: four at a time
        for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < V.tri_count; i0 += 4u * stride)
            scatter_slots<4>(V, bin_tiles, bin_offsets, pair_slots, slot_count, lists, list_cap, i0, stride);
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nbig; i += stride)
            scatter_big_pair(big_pairs[V.big_base + i], bin_offsets, lists, list_cap);
    } else if (V.tri_count > stride) {  // two record slots per thread at a time, then the big pairs
        for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < V.tri_count; i0 += 2u * stride)
            scatter_slots<2>(V, bin_tiles, bin_offsets, pair_slots, slot_count, lists, list_cap, i0, stride);
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nbig; i += stride)
            scatter_big_pair(big_pairs[V.big_base + i], bin_offsets, lists, list_cap);
    } else {  // the record slots and then the big pairs, one grid-stride sequence
        const uint32_t n = V.tri_count + nbig;
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
            if (i < V.tri_count)
                scatter_slots<1>(V, bin_tiles, bin_offsets, pair_slots, slot_count, lists, list_cap, i, stride);
            else scatter_big_pair(big_pairs[V.big_base + (i - V.tri_count)], bin_offsets, lists, list_cap);
        }
    }
}

// Bin raster: one workgroup per work item (a whole non-empty bin, or up to kItemRecs records of a
// crowded bin), the bin's tile in LDS (rows padded to T + 1 words so that lanes on the same column of
// different rows fall in different banks): 64-bit visibility words (depth bits above triangle id
// + 1: nearer depth wins under reverse-Z, the later triangle wins ties) in a camera view, 32-bit
// depth words in a shadow view.  A whole bin stores its tile; the items of a split bin merge into
// the cleared bin with global atomic max.  Workgroups are dealt round-robin over the 8 XCDs, and
// consecutive items are neighbouring bins, so the dense regions of a view spread over every XCD.
template <int T, int N>
__global__ __launch_bounds__(256) void k_raster(const ViewGPU* views, uint32_t n_views, const uint32_t* lists,
                                                uint32_t list_cap, const TriRec* recs, uint64_t rec_cap_total,
                                                const uint4* items, const uint32_t* totals) {
    constexpr int S = T + 1;  // padded row stride (words)
    __shared__ unsigned long long tile[N * T * S];  // camera views: one plane per sample
    __shared__ TriRec chunk_recs[kChunkRecs];
    __shared__ uint8_t owner[kChunkRecs * T];  // task (row) -> record of the chunk
    __shared__ uint32_t row_start[kChunkRecs];  // exclusive prefix of row counts
    __shared__ uint32_t scan_lds[4];

    // One workgroup per item.  The grid is sized from the last frame's item count (the count of
    // this frame is known only on the device): the bin scan flags a frame with more items than
    // workgroups (totals[2] bit 1), which is re-run with a grid that fits.  The item is loaded
    // beside the count (blockIdx.x < gridDim.x <= the table's capacity).
    const uint32_t n_items = totals[3];
    const uint4 it = items[blockIdx.x];
    // (the test reads the item too, which keeps its load beside the count's instead of after the
    // branch; an item's record count never reaches 2^31)
    if (blockIdx.x >= n_items || int32_t(it.z) < 0) return;
    // this thread's record of the item: the list entry, then the record itself as three 16-byte
    // words (X0 Y0 X1 Y1 | X2 Y2 z0 dzdx | dzdy id bbx bby); the view lookup and the tile clear
    // overlap these loads
    const uint32_t off = it.y;
    const uint32_t end = uint32_t(min(uint64_t(off) + it.z, uint64_t(list_cap)));
    const uint32_t n = end > off ? end - off : 0u;  // at most kItemRecs
    const uint32_t ri = threadIdx.x < n ? lists[off + threadIdx.x] : 0xffffffffu;
    const uint32_t bin = it.x;
    uint32_t view = 0;
    while (view + 1 < n_views && views[view + 1].tile_base <= bin) ++view;
    const ViewGPU V = views[view];
    if (V.blend) return;  // blend views are drawn by the blended pass (k_blend)
    const uint32_t local_bin = bin - V.tile_base;
    const int ox = int(local_bin % V.tiles_x) * T, oy = int(local_bin / V.tiles_x) * T;
    // view fields copied out: the output stores could alias the view table for the compiler
    const int width = int(V.width), height = int(V.height);
    const uint64_t out_addr = V.out;
    const bool split = it.w != 0u;
    // unconditional loads (record 0 stands in for a missing entry), so that the first wait for
    // them comes after the tile clear
    const bool have = ri < rec_cap_total;
    const uint4* src = reinterpret_cast<const uint4*>(recs + (have ? ri : 0u));
    const uint4 w0 = src[0], w1 = src[1], w2 = src[2];

    if (V.is_shadow) {
        uint32_t* t32 = reinterpret_cast<uint32_t*>(tile);
        for (int p = threadIdx.x; p < T * S; p += 256) t32[p] = 0u;
        raster_item<T>(ShadowFrag<T>{t32}, chunk_recs, owner, row_start, scan_lds, have, w0, w1, w2, n, ox, oy);
        __syncthreads();  // tile complete (and cleared, for an item without records)
        uint32_t* out = reinterpret_cast<uint32_t*>(out_addr);
        for (int p = threadIdx.x; p < T * T; p += 256) {
            const int x = ox + p % T, y = oy + p / T;
            if (x < width && y < height) {
                const uint32_t w = t32[(p / T) * S + p % T];
                const size_t o = size_t(y) * width + x;
                if (!split) out[o] = w;
                else if (w != 0u) atomicMax(out + o, w);
            }
        }
    } else {
        for (int p = threadIdx.x; p < N * T * S; p += 256) tile[p] = 0ull;
        if constexpr (N == 4)
            raster_item<T>(CameraFragMS<T>{tile}, chunk_recs, owner, row_start, scan_lds, have, w0, w1, w2, n, ox, oy);
        else
            raster_item<T>(CameraFrag<T>{tile}, chunk_recs, owner, row_start, scan_lds, have, w0, w1, w2, n, ox, oy);
        __syncthreads();
        unsigned long long* out = reinterpret_cast<unsigned long long*>(out_addr);
        const size_t plane = size_t(width) * height;
        for (int p = threadIdx.x; p < T * T; p += 256) {
            const int x = ox + p % T, y = oy + p / T;
            if (x < width && y < height) {
                const size_t o = size_t(y) * width + x;
#pragma unroll
                for (int s = 0; s < N; ++s) {
                    const unsigned long long w = tile[s * T * S + (p / T) * S + p % T];
                    if (!split) out[s * plane + o] = w;
                    else if (w != 0ull) atomicMax(out + s * plane + o, w);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Launchers

hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint32_t blocks = min((n + 255) / 256, 4096u);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_geometry(const GeometryLaunch& g, hipStream_t s) {
    if (g.n_views == 0) return hipSuccess;
    const DrawLists lists{{g.draws0, g.draws1, g.draws2}};
    // the transform clears the counters and bin counts; without vertices a clear of its own
    const uint32_t nclear = g.n_views * kCtrPerView + g.total_bins;
    if (g.n_vblocks)
        hipLaunchKernelGGL(k_transform, dim3(g.n_vblocks), dim3(256), 0, s, g.vblocks, g.views, lists,
                           g.objects, g.px, g.py, g.pz, g.xv, g.draw_culled, g.counters, g.n_views * kCtrPerView,
                           g.bin_counts, g.total_bins);
    else
        hipLaunchKernelGGL(k_clear2, dim3(min((nclear + 255) / 256, 2048u)), dim3(256), 0, s, g.counters,
                           g.n_views * kCtrPerView, g.bin_counts, g.total_bins);
    if (g.n_tblocks && g.setup_tpt == 2)
        hipLaunchKernelGGL(k_setup<2>, dim3(g.n_tblocks), dim3(256), 0, s, g.tblocks, g.views, g.draw_culled, g.indices, g.xv,
                           g.bin_tiles, g.block_valid, g.recs, g.shade_recs, g.clipq, g.counters, g.bin_counts,
                           g.pair_slots, g.slot_count, g.big_pairs, g.shift);
    else if (g.n_tblocks)
        hipLaunchKernelGGL(k_setup<1>, dim3(g.n_tblocks), dim3(256), 0, s, g.tblocks, g.views, g.draw_culled, g.indices, g.xv,
                           g.bin_tiles, g.block_valid, g.recs, g.shade_recs, g.clipq, g.counters, g.bin_counts,
                           g.pair_slots, g.slot_count, g.big_pairs, g.shift);
    hipLaunchKernelGGL(k_clip, dim3(g.clip_blocks, g.n_views), dim3(kClipThreads), 0, s, g.views, g.clipq, g.recs,
                       g.counters, g.bin_counts, g.big_pairs, g.shift);
    return hipGetLastError();
}

hipError_t launch_binning(const BinLaunch& b, hipStream_t s) {
    const uint32_t parts = (b.total_bins + kScanChunk - 1) / kScanChunk;
    ScanPartial* partial = reinterpret_cast<ScanPartial*>(b.scan_partial);
    if (b.total_bins <= kDirectScanBins) {
#define RSY_SCAN_DIRECT(NT)                                                                                        \
    hipLaunchKernelGGL(k_scan_direct<NT>, dim3((b.total_bins + NT - 1) / NT), dim3(NT), 0, s, b.bin_counts,       \
                       b.total_bins, b.bin_offsets, b.block_valid, b.n_tblocks, b.list_cap, b.totals, b.items,    \
                       b.items_cap, b.raster_grid)
        if (b.total_bins <= kSmallScanBins) RSY_SCAN_DIRECT(512);
        else RSY_SCAN_DIRECT(1024);
#undef RSY_SCAN_DIRECT
    } else {
        hipLaunchKernelGGL(k_scan_partial, dim3(parts), dim3(kScanThreads), 0, s, b.bin_counts, b.total_bins, partial,
                           b.block_valid, b.n_tblocks);
        hipLaunchKernelGGL(k_scan_final, dim3(parts), dim3(kScanThreads), 0, s, b.bin_counts, b.total_bins,
                           b.bin_offsets, partial, parts, b.list_cap, b.totals, b.items, b.items_cap,
                           b.raster_grid);
    }
    hipLaunchKernelGGL(k_scatter, dim3(b.scatter_blocks, b.n_views), dim3(256), 0, s, b.views, b.bin_tiles, b.counters,
                       b.bin_offsets, b.pair_slots, b.slot_count, b.big_pairs, b.lists, b.list_cap);
    return hipGetLastError();
}

hipError_t launch_raster(const RasterLaunch& r, hipStream_t s) {
    const uint32_t grid = min(r.grid, r.items_cap);
    if (grid == 0) return hipSuccess;
#define RSY_RASTER(T, N)                                                                                         \
    hipLaunchKernelGGL((k_raster<T, N>), dim3(grid), dim3(256), 0, s, r.views, r.n_views, r.lists, r.list_cap, \
                       r.recs, r.rec_cap_total, r.items, r.totals)
    if (r.samples == 1) {
        if (r.tile == 8) RSY_RASTER(8, 1);
        else if (r.tile == 16) RSY_RASTER(16, 1);
        else if (r.tile == 32) RSY_RASTER(32, 1);
        else if (r.tile == 64) RSY_RASTER(64, 1);
        else return hipErrorInvalidValue;
    } else if (r.tile == 32 && r.samples == 4) {
        RSY_RASTER(32, 4);
    } else {
        return hipErrorInvalidValue;  // 4-sample frames use 32-pixel bins
    }
#undef RSY_RASTER
    return hipGetLastError();
}

uint32_t raster_items_cap(uint32_t total_bins, uint32_t list_cap) {
    return total_bins + (list_cap + kItemRecs - 1) / kItemRecs;
}

}  // namespace rsy
