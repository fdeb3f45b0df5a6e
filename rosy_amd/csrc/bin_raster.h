// Bin-raster building blocks shared by the raster kernels (raster.hip) and the fused blended pass
// (shade.hip): the workgroup scan, the per-record raster flags, the fragment operations on an LDS
// tile, one row of a record, and one work item's records through the tile.  Device code only.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "frame_types.h"

namespace rsy {

__device__ inline uint32_t lane_id() { return __lane_id(); }

// Inclusive prefix sums over the 64 lanes of a wave (every lane active): DPP row shifts by 1, 2, 4
// and 8 give each row of 16 lanes its own prefix, then row broadcasts add the totals of the rows
// before (lane 15 into rows 1 and 3, lane 31 into rows 2 and 3).  VALU moves only, no LDS permutes.
template <int kCtrl, int kRowMask>
__device__ inline uint32_t dpp_move(uint32_t x) {
    return uint32_t(__builtin_amdgcn_update_dpp(0, int(x), kCtrl, kRowMask, 0xf, false));
}

__device__ inline uint32_t wave_inclusive_sum(uint32_t x) {
    x += dpp_move<0x111, 0xf>(x);  // row_shr:1
    x += dpp_move<0x112, 0xf>(x);  // row_shr:2
    x += dpp_move<0x114, 0xf>(x);  // row_shr:4
    x += dpp_move<0x118, 0xf>(x);  // row_shr:8
    x += dpp_move<0x142, 0xa>(x);  // row_bcast:15
    x += dpp_move<0x143, 0xc>(x);  // row_bcast:31
    return x;
}

template <int kCtrl, int kRowMask>
__device__ inline uint64_t dpp_move64(uint64_t x) {
    return uint64_t(dpp_move<kCtrl, kRowMask>(uint32_t(x))) | (uint64_t(dpp_move<kCtrl, kRowMask>(uint32_t(x >> 32))) << 32);
}

// Lane `lane` (wave-uniform) of a 64-bit value.
__device__ inline uint64_t readlane64(uint64_t x, uint32_t lane) {
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(x)), int(lane)));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(x >> 32)), int(lane)));
    return uint64_t(lo) | (uint64_t(hi) << 32);
}

__device__ inline uint64_t wave_inclusive_sum64(uint64_t x) {
    x += dpp_move64<0x111, 0xf>(x);
    x += dpp_move64<0x112, 0xf>(x);
    x += dpp_move64<0x114, 0xf>(x);
    x += dpp_move64<0x118, 0xf>(x);
    x += dpp_move64<0x142, 0xa>(x);
    x += dpp_move64<0x143, 0xc>(x);
    return x;
}

__device__ inline uint32_t fbits(float f) { return __float_as_uint(f); }

constexpr uint32_t kItemRecs = 256;   // records per raster work item (one per thread)
constexpr uint32_t kChunkRecs = 128;  // records staged in LDS at a time

// Exclusive prefix sum over the workgroup (wave scans by DPP, wave totals through LDS).
// Every thread of the block must call it; it synchronises the block.  kTailSync = false leaves out
// the closing barrier, for a caller that synchronises again before lds_waves is next written.
template <bool kTailSync = true>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_waves, uint32_t& total) {
    const uint32_t x = wave_inclusive_sum(v);
    const uint32_t wave = threadIdx.x / 64, nwaves = blockDim.x / 64;
    if (lane_id() == 63) lds_waves[wave] = x;
    __syncthreads();
    uint32_t wave_off = 0, sum = 0;
    for (uint32_t w = 0; w < nwaves; ++w) {
        const uint32_t s = lds_waves[w];
        if (w < wave) wave_off += s;
        sum += s;
    }
    if constexpr (kTailSync) __syncthreads();
    total = sum;
    return wave_off + x - v;
}

// Per-record raster flags, computed once when a record is staged for a bin: bit k (k < 3) is set when
// edge k is not a top-left edge (its value must exceed 0: F_k = E_k - 1), bit 3 when the edge
// values need 64 bits.
constexpr uint32_t kMetaWide = 8u;

__device__ inline uint32_t record_meta(int32_t X0, int32_t Y0, int32_t X1, int32_t Y1, int32_t X2, int32_t Y2) {
    const int32_t X[3] = {X0, X1, X2}, Y[3] = {Y0, Y1, Y2};
    uint32_t meta = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int f = k == 2 ? 0 : k + 1;
        const int32_t A = Y[f] - Y[k], B = X[k] - X[f];
        const bool tl = A > 0 || (A == 0 && B > 0);
        meta |= tl ? 0u : (1u << k);
    }
    // the whole triangle's box (sub-pixel units) bounds every |E| over it: under 64 pixels on each
    // side every |E| < 2^30 at pixel centres inside the box, and < 2^31 at the pixels and sample
    // points a record's box reaches (at most 192 sub-pixel units beyond the vertices)
    const int32_t wx = max(X[0], max(X[1], X[2])) - min(X[0], min(X[1], X[2]));
    const int32_t wy = max(Y[0], max(Y[1], Y[2])) - min(Y[0], min(Y[1], Y[2]));
    if (wx >= 64 * 256 || wy >= 64 * 256) meta |= kMetaWide;
    return meta;
}

// Fragment operations of the bin raster on the LDS tile (rows padded to T + 1 words).  Each takes
// a covered pixel's tile index, its 31-bit depth bits and the record's key (triangle id + 1).
template <int T>
struct ShadowFrag {  // shadow view: largest depth (GEQUAL under reverse-Z)
    static constexpr int kStride = T + 1, kSamples = 1;
    uint32_t* tile;
    __device__ void operator()(int i, uint32_t zb, uint32_t) const { atomicMax(&tile[i], zb); }
};
template <int T>
struct CameraFrag {  // camera view: depth bits above key, the largest word wins
    static constexpr int kStride = T + 1, kSamples = 1;
    unsigned long long* tile;
    __device__ void operator()(int i, uint32_t zb, uint32_t key) const {
        atomicMax(&tile[i], ((unsigned long long)zb << 32) | key);
    }
};
template <int T>
struct CountFrag {  // blend view: count the fragments that pass the depth test against the opaque depth
    static constexpr int kStride = T + 1, kSamples = 1;
    uint32_t* count;
    const uint32_t* opaque;  // depth bits of the camera view
    __device__ void operator()(int i, uint32_t zb, uint32_t) const {
        if (zb >= opaque[i]) atomicAdd(&count[i], 1u);
    }
};
template <int T>
struct FillFrag {  // blend view: append the key of each passing fragment to its pixel's list
    static constexpr int kStride = T + 1, kSamples = 1;
    uint32_t* cursor;        // next free entry of the pixel's list
    const uint32_t* opaque;
    uint32_t* keys;          // the bin's lists (LDS or global memory)
    __device__ void operator()(int i, uint32_t zb, uint32_t key) const {
        if (zb >= opaque[i]) keys[atomicAdd(&cursor[i], 1u)] = key;
    }
};

// The same operations on 4-sample tiles: one call per covered pixel with the mask of covered samples
// and each sample's depth bits; sample planes lie T * (T + 1) words apart.
template <int T>
struct CameraFragMS {
    static constexpr int kStride = T + 1, kSamples = 4, kPlane = T * (T + 1);
    unsigned long long* tile;
    __device__ void operator()(int i, uint32_t mask, const uint32_t (&zb)[4], uint32_t key) const {
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (mask & (1u << s)) atomicMax(&tile[s * kPlane + i], ((unsigned long long)zb[s] << 32) | key);
    }
};
template <int T>
__device__ inline uint32_t passing_samples(const uint32_t* opaque, int i, uint32_t mask, const uint32_t (&zb)[4]) {
    uint32_t m = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if ((mask & (1u << s)) && zb[s] >= opaque[s * T * (T + 1) + i]) m |= 1u << s;
    return m;
}
template <int T>
struct CountFragMS {  // blend view: count the primitives with a sample passing the depth test
    static constexpr int kStride = T + 1, kSamples = 4;
    uint32_t* count;
    const uint32_t* opaque;  // depth bits of the camera view's sample planes
    __device__ void operator()(int i, uint32_t mask, const uint32_t (&zb)[4], uint32_t) const {
        if (passing_samples<T>(opaque, i, mask, zb)) atomicAdd(&count[i], 1u);
    }
};
template <int T>
struct FillFragMS {  // blend view: append key | passing-sample mask << 28 to the pixel's list
    static constexpr int kStride = T + 1, kSamples = 4;
    uint32_t* cursor;
    const uint32_t* opaque;
    uint32_t* keys;
    __device__ void operator()(int i, uint32_t mask, const uint32_t (&zb)[4], uint32_t key) const {
        const uint32_t m = passing_samples<T>(opaque, i, mask, zb);
        if (m) keys[atomicAdd(&cursor[i], 1u)] = key | (m << 28);
    }
};

// One row of a record inside the bin: E_k(x) for the three edges advances by 256 * A_k per pixel
// (exact integer steps; EdgeT = int32_t unless the record is wide), depth is the FP32 plane rule of
// setup, and covered pixels merge into the LDS row with atomic max (64-bit depth | id + 1 words in
// a camera view, 32-bit depth words in a shadow view).  The top-left rule is folded into the edge
// value: F_k = E_k - 1 + tl_k, so "E_k > 0, or E_k == 0 on a top-left edge" is F_k >= 0 for
// integer E_k, and a pixel is covered when (F_0 | F_1 | F_2) has a clear sign bit -- one test, no
// branches per edge.
template <typename EdgeT, class Op>
__device__ inline void raster_row(const TriRec& r, uint32_t meta, int ox, int oy, int y, int bx0, int bx1,
                                  const Op& op) {
    const int32_t X[3] = {r.X0, r.X1, r.X2}, Y[3] = {r.Y0, r.Y1, r.Y2};
    const int32_t Py = (oy + y) * 256 + 128;
    int32_t Px = (ox + bx0) * 256 + 128;
    EdgeT f[3], step[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int n = k == 2 ? 0 : k + 1;
        const int32_t A = Y[n] - Y[k], B = X[k] - X[n];
        f[k] = EdgeT(A) * EdgeT(Px - X[k]) + EdgeT(B) * EdgeT(Py - Y[k]) - EdgeT((meta >> k) & 1u);
        step[k] = EdgeT(A) * EdgeT(256);
    }
    const float fy = float(Py - r.Y0);
    const uint32_t key = (r.id & kRecIdMask) + 1u;
    const int row = y * Op::kStride;
    for (int x = bx0; x <= bx1; ++x, Px += 256) {
        if ((f[0] | f[1] | f[2]) >= 0) {
            float z = fmaf(r.dzdy, fy, fmaf(r.dzdx, float(Px - r.X0), r.z0));
            z = fminf(fmaxf(z, 0.0f), 1.0f);
            op(row + x, fbits(z) & 0x7fffffffu, key);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) f[k] += step[k];
    }
}

// One row of a record at the 4 sample locations of each pixel: the centre's edge values plus a fixed
// offset per sample and edge (exact integers), the depth plane at each covered sample.
template <typename EdgeT, class Op>
__device__ inline void raster_row_ms(const TriRec& r, uint32_t meta, int ox, int oy, int y, int bx0, int bx1,
                                     const Op& op) {
    const int32_t X[3] = {r.X0, r.X1, r.X2}, Y[3] = {r.Y0, r.Y1, r.Y2};
    const int32_t Py = (oy + y) * 256 + 128;
    int32_t Px = (ox + bx0) * 256 + 128;
    EdgeT f[3], step[3], d[4][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int n = k == 2 ? 0 : k + 1;
        const int32_t A = Y[n] - Y[k], B = X[k] - X[n];
        f[k] = EdgeT(A) * EdgeT(Px - X[k]) + EdgeT(B) * EdgeT(Py - Y[k]) - EdgeT((meta >> k) & 1u);
        step[k] = EdgeT(A) * EdgeT(256);
#pragma unroll
        for (int s = 0; s < 4; ++s) d[s][k] = EdgeT(A) * EdgeT(msaa_x(s) - 128) + EdgeT(B) * EdgeT(msaa_y(s) - 128);
    }
    float fy[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) fy[s] = float(Py + msaa_y(s) - 128 - r.Y0);
    const uint32_t key = (r.id & kRecIdMask) + 1u;
    const int row = y * Op::kStride;
    for (int x = bx0; x <= bx1; ++x, Px += 256) {
        uint32_t mask = 0, zb[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            zb[s] = 0u;
            if (((f[0] + d[s][0]) | (f[1] + d[s][1]) | (f[2] + d[s][2])) >= 0) {
                float z = fmaf(r.dzdy, fy[s], fmaf(r.dzdx, float(Px + msaa_x(s) - 128 - r.X0), r.z0));
                z = fminf(fmaxf(z, 0.0f), 1.0f);
                zb[s] = fbits(z) & 0x7fffffffu;
                mask |= 1u << s;
            }
        }
        if (mask) op(row + x, mask, zb, key);
#pragma unroll
        for (int k = 0; k < 3; ++k) f[k] += step[k];
    }
}

// The records of one work item through the bin's LDS tile.  Every thread loads one record of the
// item (its box clipped to the bin, its raster flags computed) in one round of loads; the records
// then go through LDS kChunkRecs at a time, each chunk staged by its own threads.  A staged
// record's rows become tasks (a prefix sum numbers them and each record writes its index into the
// task-owner table), and every thread takes tasks in turn, so small and large triangles spread
// evenly over the workgroup.
template <int T, class Op>
__device__ inline void raster_item(const Op& op, TriRec* chunk_recs, uint8_t* owner, uint32_t* row_start,
                                   uint32_t* scan_lds, bool have, uint4 w0, uint4 w1, uint4 w2, uint32_t n, int ox,
                                   int oy) {
    static_assert(kItemRecs == 256 && kItemRecs % kChunkRecs == 0, "one record per thread");
    uint32_t my_rows = 0;
    if (have) {
        const int bx0 = max(int(w2.z & 0xffffu), ox) - ox, bx1 = min(int(w2.z >> 16), ox + T - 1) - ox;
        const int by0 = max(int(w2.w & 0xffffu), oy) - oy, by1 = min(int(w2.w >> 16), oy + T - 1) - oy;
        if (bx0 <= bx1 && by0 <= by1) {
            my_rows = uint32_t(by1 - by0 + 1);
            // the staged copy keeps the clipped box in bbx and the raster flags in bby
            w2.z = uint32_t(bx0) | (uint32_t(bx1) << 8) | (uint32_t(by0) << 16);
            w2.w = record_meta(int32_t(w0.x), int32_t(w0.y), int32_t(w0.z), int32_t(w0.w), int32_t(w1.x),
                               int32_t(w1.y));
        }
    }
    const uint32_t slot = threadIdx.x % kChunkRecs;
    for (uint32_t c = 0; c * kChunkRecs < n; ++c) {
        const bool stager = threadIdx.x / kChunkRecs == c;
        const uint32_t rows = stager ? my_rows : 0u;
        if (rows) {
            uint4* dst = reinterpret_cast<uint4*>(chunk_recs + slot);
            dst[0] = w0;
            dst[1] = w1;
            dst[2] = w2;
        }
        uint32_t total_rows;
        // synchronises; the barrier after the owner table comes before the next chunk's scan writes
        const uint32_t first = block_exclusive_scan<false>(rows, scan_lds, total_rows);
        if (stager) row_start[slot] = first;
        for (uint32_t j = 0; j < rows; ++j) owner[first + j] = uint8_t(slot);
        __syncthreads();
        for (uint32_t task = threadIdx.x; task < total_rows; task += 256) {
            const uint32_t k = owner[task];
            const TriRec q = chunk_recs[k];
            const int bx0 = int(q.bbx & 0xffu), bx1 = int((q.bbx >> 8) & 0xffu);
            const int y = int((q.bbx >> 16) & 0xffu) + int(task - row_start[k]);
            if constexpr (Op::kSamples == 4) {
                if (q.bby & kMetaWide)
                    raster_row_ms<int64_t>(q, q.bby, ox, oy, y, bx0, bx1, op);
                else
                    raster_row_ms<int32_t>(q, q.bby, ox, oy, y, bx0, bx1, op);
            } else {
                if (q.bby & kMetaWide)
                    raster_row<int64_t>(q, q.bby, ox, oy, y, bx0, bx1, op);
                else
                    raster_row<int32_t>(q, q.bby, ox, oy, y, bx0, bx1, op);
            }
        }
        __syncthreads();
    }
}

}  // namespace rsy
