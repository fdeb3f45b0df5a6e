// Geometry, binning and tile-raster kernels of the frame path (gfx950, wave64).
//
// Replaces the reference's shadow and main raster passes (shadow.slang:6-20, basic.slang:100-116
// and the fixed-function state of Graphics.cpp:4988-5045, 5195-5292, 5553-5653) with a software
// pipeline:
//   k_mvp        per (view, draw)   M = VP * W                       (fmaf chain, k order)
//   k_transform  per (view, vertex) clip = M * p from SoA x/y/z      (12 B in, 16 B out, coalesced)
//   k_setup      per (view, tri)    clip test, snap, cull, bbox, depth plane, bias; records are
//                                   compacted with a wave ballot and each record's bins are counted
//                                   (wave-cooperative for triangles covering many bins)
//   k_clip       per clipped tri    near/far/guard-band Sutherland-Hodgman + fan, one wave each
//   k_scan_*     all bins           exclusive prefix sum of the bin counts
//   k_scatter    per record         bin lists (triangle record indices)
//   k_raster     per (view, bin)    LDS-staged records; small triangles walk their pixels with LDS
//                                   atomic max, large ones are tested pixel-parallel in registers;
//                                   the bin is written once (the clear never touches HBM)
// Every floating-point step is compiled without contraction; the CPU oracle restates the same
// rules, so visibility and depth are bit-exact by construction.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fixed_point.h"
#include "frame_types.h"
#include "kernels.h"

namespace rsy {
namespace {

constexpr int kSmallPixels = 24;  // records with at most this many bbox pixels in a bin walk them serially

__device__ inline uint32_t lane_id() { return __lane_id(); }

__device__ inline uint32_t fbits(float f) { return __float_as_uint(f); }

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
    const int32_t px0 = max(0, (minx - 128 + 255) >> 8);
    const int32_t px1 = min(int32_t(V.width) - 1, (maxx - 128) >> 8);
    const int32_t py0 = max(0, (miny - 128 + 255) >> 8);
    const int32_t py1 = min(int32_t(V.height) - 1, (maxy - 128) >> 8);
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
    r.id = id;
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

__device__ inline void rec_tiles(const TriRec& r, uint32_t shift, int& tx0, int& ty0, int& tx1, int& ty1) {
    tx0 = int(r.bbx & 0xffffu) >> shift;
    tx1 = int(r.bbx >> 16) >> shift;
    ty0 = int(r.bby & 0xffffu) >> shift;
    ty1 = int(r.bby >> 16) >> shift;
}

// Bin counting for one record per active lane: lanes with few bins loop themselves, records with
// many bins are spread over the whole wave.  Must be called by a converged wave.
__device__ void count_bins(bool active, const TriRec& r, const ViewGPU& V, uint32_t shift, uint32_t* bin_counts,
                           uint32_t* pair_counter) {
    int tx0 = 0, ty0 = 0, tx1 = -1, ty1 = -1;
    if (active) rec_tiles(r, shift, tx0, ty0, tx1, ty1);
    const int n = active ? (tx1 - tx0 + 1) * (ty1 - ty0 + 1) : 0;
    const bool small = active && n <= 4;
    uint32_t* counts = bin_counts + V.tile_base;
    if (small)
        for (int ty = ty0; ty <= ty1; ++ty)
            for (int tx = tx0; tx <= tx1; ++tx) atomicAdd(&counts[ty * int(V.tiles_x) + tx], 1u);
    uint64_t big = __ballot(active && !small);
    while (big) {
        const int l = __ffsll((unsigned long long)big) - 1;
        big &= big - 1;
        const int btx0 = __shfl(tx0, l), bty0 = __shfl(ty0, l), btx1 = __shfl(tx1, l), bty1 = __shfl(ty1, l);
        const int bw = btx1 - btx0 + 1, bn = bw * (bty1 - bty0 + 1);
        for (int k = int(lane_id()); k < bn; k += 64)
            atomicAdd(&counts[(bty0 + k / bw) * int(V.tiles_x) + btx0 + k % bw], 1u);
    }
    // wave-aggregated pair statistic
    uint32_t s = uint32_t(n);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane_id() == 0 && s) atomicAdd(pair_counter, s);
}

}  // namespace

// ---------------------------------------------------------------------------------------------

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// MVP per (view, draw): views[v].view_projection * objects[draw.object].transform.
__global__ void k_mvp(const ViewGPU* views, uint32_t n_views, const DrawGPU* draws0, const DrawGPU* draws1,
                      const float* objects, float* mvps) {
    const uint32_t v = blockIdx.y;
    if (v >= n_views) return;
    const ViewGPU& V = views[v];
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= V.n_draws) return;
    const DrawGPU& dr = (V.list ? draws1 : draws0)[d];
    const float* a = V.view_projection;
    const float* b = objects + size_t(dr.object) * 41u;  // 164-byte packed records
    float* c = mvps + size_t(V.mvp_base + d) * 16u;
#pragma unroll
    for (int col = 0; col < 4; ++col)
#pragma unroll
        for (int row = 0; row < 4; ++row) {
            float s = a[row] * b[col * 4 + 0];
            s = fmaf(a[4 + row], b[col * 4 + 1], s);
            s = fmaf(a[8 + row], b[col * 4 + 2], s);
            s = fmaf(a[12 + row], b[col * 4 + 3], s);
            c[col * 4 + row] = s;
        }
}

__global__ __launch_bounds__(256) void k_transform(const BlockEntry* blocks, const ViewGPU* views,
                                                   const DrawGPU* draws0, const DrawGPU* draws1,
                                                   const float* mvps, const float* px, const float* py,
                                                   const float* pz, float4* xv) {
    const BlockEntry e = blocks[blockIdx.x];
    const ViewGPU& V = views[e.view];
    const DrawGPU& dr = (V.list ? draws1 : draws0)[e.draw];
    const uint32_t v = e.start + threadIdx.x;
    if (v >= dr.vtx_count) return;
    const float* m = mvps + size_t(V.mvp_base + e.draw) * 16u;
    const uint32_t g = dr.vtx_base + v;
    xv[size_t(V.xv_base) + dr.xv_off + v] = transform(m, px[g], py[g], pz[g]);
}

__global__ __launch_bounds__(256) void k_setup(const BlockEntry* blocks, const ViewGPU* views, const DrawGPU* draws0,
                                               const DrawGPU* draws1, const uint32_t* indices, const float4* xv,
                                               TriRec* recs, uint2* clipq, uint32_t* counters,
                                               uint32_t* bin_counts, uint32_t shift) {
    const BlockEntry e = blocks[blockIdx.x];
    const ViewGPU& V = views[e.view];
    const DrawGPU& dr = (V.list ? draws1 : draws0)[e.draw];
    uint32_t* ctr = counters + e.view * kCtrPerView;
    const uint32_t t = e.start + threadIdx.x;
    const bool valid = t < dr.tri_count;
    int status = 0;  // 0 reject, 1 accept, 2 clip
    float4 c[3];
    if (valid) {
        const uint32_t* ip = indices + dr.first_index + 3u * t;
        const float4* base = xv + size_t(V.xv_base) + dr.xv_off;
        c[0] = base[ip[0]];
        c[1] = base[ip[1]];
        c[2] = base[ip[2]];
        bool reject = false;
#pragma unroll
        for (int plane = 0; plane < 6; ++plane) {
            const int outside = int(plane_dist(c[0], plane) < 0.f) + int(plane_dist(c[1], plane) < 0.f) +
                                int(plane_dist(c[2], plane) < 0.f);
            reject |= outside == 3;
        }
        status = reject ? 0 : (all_inside(c[0], c[1], c[2]) ? 1 : 2);
    }
    TriRec r;
    bool has = false;
    if (status == 1) has = setup_tri(c, dr.tri_base + t, V, r);

    const uint32_t cslot = wave_append(status == 2, &ctr[kCtrClip]);
    if (status == 2) {
        if (cslot < V.clip_cap) clipq[V.clip_base + cslot] = make_uint2(e.draw, t);
        else atomicOr(&ctr[kCtrOverflow], 1u);
    }
    const uint32_t rslot = wave_append(has, &ctr[kCtrRecords]);
    if (has) {
        if (rslot < V.rec_cap) {
            recs[V.rec_base + rslot] = r;
        } else {
            atomicOr(&ctr[kCtrOverflow], 2u);
            has = false;
        }
    }
    count_bins(has, r, V, shift, bin_counts, &ctr[kCtrPairs]);
}

// One wave per queued triangle: lane 0 clips and sets up the fan; the wave counts the bins.
__global__ __launch_bounds__(256) void k_clip(const ViewGPU* views, const DrawGPU* draws0, const DrawGPU* draws1,
                                              const uint32_t* indices, const float4* xv, const uint2* clipq,
                                              TriRec* recs, uint32_t* counters, uint32_t* bin_counts,
                                              uint32_t shift) {
    const uint32_t view = blockIdx.y;
    const ViewGPU& V = views[view];
    uint32_t* ctr = counters + view * kCtrPerView;
    const uint32_t n = min(ctr[kCtrClip], V.clip_cap);
    const uint32_t waves_per_block = blockDim.x / 64;
    const uint32_t wid = blockIdx.x * waves_per_block + threadIdx.x / 64;
    const uint32_t stride = gridDim.x * waves_per_block;
    for (uint32_t q = wid; q < n; q += stride) {
        TriRec out[7];
        int nout = 0;
        if (lane_id() == 0) {
            const uint2 ent = clipq[V.clip_base + q];
            const DrawGPU& dr = (V.list ? draws1 : draws0)[ent.x];
            const uint32_t* ip = indices + dr.first_index + 3u * ent.y;
            const float4* base = xv + size_t(V.xv_base) + dr.xv_off;
            float4 a[12], b[12];
            a[0] = base[ip[0]];
            a[1] = base[ip[1]];
            a[2] = base[ip[2]];
            int cnt = 3;
            for (int plane = 0; plane < 6 && cnt >= 3; ++plane) {
                int m = 0;
                for (int i = 0; i < cnt; ++i) {
                    const float4 cur = a[i];
                    const float4 nxt = a[(i + 1) % cnt];
                    const float dc = plane_dist(cur, plane), dn = plane_dist(nxt, plane);
                    const bool ic = dc >= 0.f, in_ = dn >= 0.f;
                    if (ic) b[m++] = cur;
                    if (ic != in_) {
                        const float4 pi = ic ? cur : nxt;
                        const float4 po = ic ? nxt : cur;
                        const float di = ic ? dc : dn, dout = ic ? dn : dc;
                        const float tt = di / (di - dout);
                        float4 qv;
                        qv.x = pi.x + tt * (po.x - pi.x);
                        qv.y = pi.y + tt * (po.y - pi.y);
                        qv.z = pi.z + tt * (po.z - pi.z);
                        qv.w = pi.w + tt * (po.w - pi.w);
                        b[m++] = qv;
                    }
                }
                cnt = m;
                for (int i = 0; i < cnt; ++i) a[i] = b[i];
            }
            bool ok = cnt >= 3;
            for (int i = 0; i < cnt && ok; ++i) ok = a[i].w > 0.f;
            if (ok) {
                for (int k = 1; k + 1 < cnt && nout < 7; ++k) {
                    const float4 tv[3] = {a[0], a[k], a[k + 1]};
                    TriRec r;
                    if (setup_tri(tv, dr.tri_base + ent.y, V, r)) {
                        const uint32_t slot = atomicAdd(&ctr[kCtrRecords], 1u);
                        if (slot < V.rec_cap) {
                            recs[V.rec_base + slot] = r;
                            out[nout++] = r;
                        } else {
                            atomicOr(&ctr[kCtrOverflow], 2u);
                        }
                    }
                }
            }
            atomicAdd(&ctr[kCtrClipped], 1u);
        }
        nout = __shfl(nout, 0);
        for (int k = 0; k < nout; ++k) {
            TriRec r;
            r.bbx = __shfl(out[k].bbx, 0);
            r.bby = __shfl(out[k].bby, 0);
            count_bins(lane_id() == 0, r, V, shift, bin_counts, &ctr[kCtrPairs]);
        }
    }
}

// Exclusive scan of n bin counts: per-block sums, a one-block scan of the sums, then the local scan.
constexpr uint32_t kScanPerThread = 8;
constexpr uint32_t kScanBlock = 256 * kScanPerThread;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_waves, uint32_t& total) {
    // inclusive scan within the wave
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (int(lane_id()) >= off) x += y;
    }
    const uint32_t wave = threadIdx.x / 64, nwaves = blockDim.x / 64;
    if (lane_id() == 63) lds_waves[wave] = x;
    __syncthreads();
    uint32_t wave_off = 0, sum = 0;
    for (uint32_t w = 0; w < nwaves; ++w) {
        const uint32_t s = lds_waves[w];
        if (w < wave) wave_off += s;
        sum += s;
    }
    __syncthreads();
    total = sum;
    return wave_off + x - v;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* counts, uint32_t n, uint32_t* block_sums) {
    __shared__ uint32_t lds[4];
    const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPerThread;
    uint32_t s = 0;
    for (uint32_t k = 0; k < kScanPerThread; ++k)
        if (base + k < n) s += counts[base + k];
    uint32_t total;
    block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_scan_top(uint32_t* block_sums, uint32_t nb, uint32_t* total_out) {
    __shared__ uint32_t lds[4];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 256) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? block_sums[i] : 0;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, lds, total);
        if (i < nb) block_sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(256) void k_scan_down(const uint32_t* counts, uint32_t n, const uint32_t* block_sums,
                                                   uint32_t* offsets, uint32_t* cursor) {
    __shared__ uint32_t lds[4];
    const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread];
    uint32_t s = 0;
    for (uint32_t k = 0; k < kScanPerThread; ++k) {
        v[k] = base + k < n ? counts[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, lds, total) + block_sums[blockIdx.x];
    for (uint32_t k = 0; k < kScanPerThread; ++k)
        if (base + k < n) {
            offsets[base + k] = run;
            cursor[base + k] = run;
            run += v[k];
        }
}

// Bin lists: every record index is appended to each bin its pixel bounds touch.
__global__ __launch_bounds__(256) void k_scatter(const ViewGPU* views, const TriRec* recs, const uint32_t* counters,
                                                 uint32_t* cursor, uint32_t* lists, uint32_t list_cap,
                                                 uint32_t shift) {
    const uint32_t view = blockIdx.y;
    const ViewGPU& V = views[view];
    const uint32_t n = min(counters[view * kCtrPerView + kCtrRecords], V.rec_cap);
    uint32_t* cur = cursor + V.tile_base;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < n;
        int tx0 = 0, ty0 = 0, tx1 = -1, ty1 = -1;
        const uint32_t ri = V.rec_base + i;
        if (active) {
            const uint2 bb = *reinterpret_cast<const uint2*>(&recs[ri].bbx);
            TriRec r;
            r.bbx = bb.x;
            r.bby = bb.y;
            rec_tiles(r, shift, tx0, ty0, tx1, ty1);
        }
        const int nb = active ? (tx1 - tx0 + 1) * (ty1 - ty0 + 1) : 0;
        const bool small = active && nb <= 4;
        if (small)
            for (int ty = ty0; ty <= ty1; ++ty)
                for (int tx = tx0; tx <= tx1; ++tx) {
                    const uint32_t slot = atomicAdd(&cur[ty * int(V.tiles_x) + tx], 1u);
                    if (slot < list_cap) lists[slot] = ri;
                }
        uint64_t big = __ballot(active && !small);
        while (big) {
            const int l = __ffsll((unsigned long long)big) - 1;
            big &= big - 1;
            const int btx0 = __shfl(tx0, l), bty0 = __shfl(ty0, l), btx1 = __shfl(tx1, l), bty1 = __shfl(ty1, l);
            const uint32_t bri = __shfl(ri, l);
            const int bw = btx1 - btx0 + 1, bn = bw * (bty1 - bty0 + 1);
            for (int k = int(lane_id()); k < bn; k += 64) {
                const uint32_t slot = atomicAdd(&cur[(bty0 + k / bw) * int(V.tiles_x) + btx0 + k % bw], 1u);
                if (slot < list_cap) lists[slot] = bri;
            }
        }
    }
}

// Tile raster.  One workgroup per bin; blocks are mapped so that neighbouring bins run on the same
// XCD (the hardware deals workgroups round-robin over the 8 XCDs, each with its own L2).
template <int T>
__global__ __launch_bounds__(256) void k_raster(const ViewGPU* views, uint32_t n_views, uint32_t total_bins,
                                                const uint32_t* bin_counts, const uint32_t* bin_offsets,
                                                const uint32_t* lists, uint32_t list_cap, const TriRec* recs,
                                                uint32_t rec_cap_total) {
    constexpr int NT = (T * T < 256) ? T * T : 256;  // threads per workgroup
    constexpr int PPT = T * T / NT;                  // pixels per thread
    __shared__ TriRec big_recs[NT];
    __shared__ unsigned long long best_lds[T * T];
    __shared__ uint32_t big_n;

    const uint32_t chunk = (total_bins + 7) / 8;
    const uint32_t bin = (blockIdx.x % 8) * chunk + blockIdx.x / 8;
    if (bin >= total_bins) return;
    uint32_t view = 0;
    while (view + 1 < n_views && views[view + 1].tile_base <= bin) ++view;
    const ViewGPU& V = views[view];
    const uint32_t local_bin = bin - V.tile_base;
    const int tx = int(local_bin % V.tiles_x), ty = int(local_bin / V.tiles_x);
    const int ox = tx * T, oy = ty * T;
    const bool shadow = V.is_shadow != 0;

    for (int p = threadIdx.x; p < T * T; p += NT) best_lds[p] = 0ull;
    if (threadIdx.x == 0) big_n = 0;
    unsigned long long best[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) best[k] = 0ull;
    __syncthreads();

    const uint32_t count = bin_counts[bin];
    const uint32_t off = bin_offsets[bin];
    const uint32_t end = min(off + count, list_cap);
    for (uint32_t base = off; base < end; base += NT) {
        const uint32_t i = base + threadIdx.x;
        if (i < end) {
            const uint32_t ri = lists[i];
            if (ri < rec_cap_total) {
                const TriRec r = recs[ri];
                const int bx0 = max(int(r.bbx & 0xffffu), ox), bx1 = min(int(r.bbx >> 16), ox + T - 1);
                const int by0 = max(int(r.bby & 0xffffu), oy), by1 = min(int(r.bby >> 16), oy + T - 1);
                const int area = (bx1 - bx0 + 1) * (by1 - by0 + 1);
                if (bx0 <= bx1 && by0 <= by1) {
                    if (area <= kSmallPixels) {
                        int32_t A[3], B[3], X[3], Y[3];
                        X[0] = r.X0; X[1] = r.X1; X[2] = r.X2;
                        Y[0] = r.Y0; Y[1] = r.Y1; Y[2] = r.Y2;
                        bool tl[3];
#pragma unroll
                        for (int e = 0; e < 3; ++e) {
                            const int f = e == 2 ? 0 : e + 1;
                            A[e] = Y[f] - Y[e];
                            B[e] = X[e] - X[f];
                            tl[e] = A[e] > 0 || (A[e] == 0 && B[e] > 0);
                        }
                        for (int y = by0; y <= by1; ++y) {
                            const int32_t Py = y * 256 + 128;
                            for (int x = bx0; x <= bx1; ++x) {
                                const int32_t Px = x * 256 + 128;
                                bool inside = true;
#pragma unroll
                                for (int e = 0; e < 3; ++e) {
                                    const int64_t ev = int64_t(A[e]) * (Px - X[e]) + int64_t(B[e]) * (Py - Y[e]);
                                    inside &= ev > 0 || (ev == 0 && tl[e]);
                                }
                                if (inside) {
                                    float z = fmaf(r.dzdy, float(Py - r.Y0), fmaf(r.dzdx, float(Px - r.X0), r.z0));
                                    z = fminf(fmaxf(z, 0.0f), 1.0f);
                                    const uint32_t zb = fbits(z) & 0x7fffffffu;
                                    const unsigned long long w =
                                        shadow ? (unsigned long long)zb
                                               : ((unsigned long long)zb << 32) | (unsigned long long)(r.id + 1u);
                                    atomicMax(&best_lds[(y - oy) * T + (x - ox)], w);
                                }
                            }
                        }
                    } else {
                        const uint32_t slot = atomicAdd(&big_n, 1u);
                        big_recs[slot] = r;
                    }
                }
            }
        }
        __syncthreads();
        const uint32_t nbig = big_n;
        for (uint32_t j = 0; j < nbig; ++j) {
            const TriRec r = big_recs[j];
            int32_t A[3], B[3], X[3], Y[3];
            X[0] = r.X0; X[1] = r.X1; X[2] = r.X2;
            Y[0] = r.Y0; Y[1] = r.Y1; Y[2] = r.Y2;
            bool tl[3];
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const int f = e == 2 ? 0 : e + 1;
                A[e] = Y[f] - Y[e];
                B[e] = X[e] - X[f];
                tl[e] = A[e] > 0 || (A[e] == 0 && B[e] > 0);
            }
            const int bx0 = int(r.bbx & 0xffffu), bx1 = int(r.bbx >> 16);
            const int by0 = int(r.bby & 0xffffu), by1 = int(r.bby >> 16);
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                const int p = threadIdx.x + k * NT;
                const int x = ox + p % T, y = oy + p / T;
                if (x < bx0 || x > bx1 || y < by0 || y > by1) continue;
                const int32_t Px = x * 256 + 128, Py = y * 256 + 128;
                bool inside = true;
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    const int64_t ev = int64_t(A[e]) * (Px - X[e]) + int64_t(B[e]) * (Py - Y[e]);
                    inside &= ev > 0 || (ev == 0 && tl[e]);
                }
                if (!inside) continue;
                float z = fmaf(r.dzdy, float(Py - r.Y0), fmaf(r.dzdx, float(Px - r.X0), r.z0));
                z = fminf(fmaxf(z, 0.0f), 1.0f);
                const uint32_t zb = fbits(z) & 0x7fffffffu;
                const unsigned long long w =
                    shadow ? (unsigned long long)zb : ((unsigned long long)zb << 32) | (unsigned long long)(r.id + 1u);
                best[k] = w > best[k] ? w : best[k];
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) big_n = 0;
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int p = threadIdx.x + k * NT;
        const int x = ox + p % T, y = oy + p / T;
        const unsigned long long a = best_lds[p];
        const unsigned long long w = a > best[k] ? a : best[k];
        if (x < int(V.width) && y < int(V.height)) {
            const size_t o = size_t(y) * V.width + x;
            if (shadow) reinterpret_cast<uint32_t*>(V.out)[o] = uint32_t(w);
            else reinterpret_cast<unsigned long long*>(V.out)[o] = w;
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
    hipLaunchKernelGGL(k_mvp, dim3((g.max_draws + 63) / 64, g.n_views), dim3(64), 0, s, g.views, g.n_views, g.draws0,
                       g.draws1, g.objects, g.mvps);
    if (g.n_vblocks)
        hipLaunchKernelGGL(k_transform, dim3(g.n_vblocks), dim3(256), 0, s, g.vblocks, g.views, g.draws0, g.draws1,
                           g.mvps, g.px, g.py, g.pz, g.xv);
    if (g.n_tblocks)
        hipLaunchKernelGGL(k_setup, dim3(g.n_tblocks), dim3(256), 0, s, g.tblocks, g.views, g.draws0, g.draws1,
                           g.indices, g.xv, g.recs, g.clipq, g.counters, g.bin_counts, g.shift);
    hipLaunchKernelGGL(k_clip, dim3(g.clip_blocks, g.n_views), dim3(256), 0, s, g.views, g.draws0, g.draws1,
                       g.indices, g.xv, g.clipq, g.recs, g.counters, g.bin_counts, g.shift);
    return hipGetLastError();
}

hipError_t launch_binning(const BinLaunch& b, hipStream_t s) {
    const uint32_t nb = (b.total_bins + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, s, b.bin_counts, b.total_bins, b.block_sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, s, b.block_sums, nb, b.total_pairs);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(256), 0, s, b.bin_counts, b.total_bins, b.block_sums,
                       b.bin_offsets, b.cursor);
    hipLaunchKernelGGL(k_scatter, dim3(b.scatter_blocks, b.n_views), dim3(256), 0, s, b.views, b.recs, b.counters,
                       b.cursor, b.lists, b.list_cap, b.shift);
    return hipGetLastError();
}

hipError_t launch_raster(const RasterLaunch& r, hipStream_t s) {
    const uint32_t grid = ((r.total_bins + 7) / 8) * 8;
    switch (r.tile) {
        case 8:
            hipLaunchKernelGGL(k_raster<8>, dim3(grid), dim3(64), 0, s, r.views, r.n_views, r.total_bins,
                               r.bin_counts, r.bin_offsets, r.lists, r.list_cap, r.recs, r.rec_cap_total);
            break;
        case 16:
            hipLaunchKernelGGL(k_raster<16>, dim3(grid), dim3(256), 0, s, r.views, r.n_views, r.total_bins,
                               r.bin_counts, r.bin_offsets, r.lists, r.list_cap, r.recs, r.rec_cap_total);
            break;
        case 32:
            hipLaunchKernelGGL(k_raster<32>, dim3(grid), dim3(256), 0, s, r.views, r.n_views, r.total_bins,
                               r.bin_counts, r.bin_offsets, r.lists, r.list_cap, r.recs, r.rec_cap_total);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace rsy
