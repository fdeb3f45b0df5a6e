// Host orchestration behind the C ABI: device buffers, draw lists, frame layout, kernel sequence,
// HIP-graph capture, overflow recovery and readback.
//
// Stands in for graphics_device (src/Engine/Graphics.cpp:443-6111): set_asset (3525-4689),
// set_graphic_objects (4691-4921), update_graphic_objects (4923-4927), update (6016-6061) and
// render (4934-6014).  One context owns one HIP stream; every call on a context is serialised by the
// caller, as the reference's single-threaded main loop does (Engine.cpp:175-224).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "context.h"
#include "frame_types.h"
#include "host_math.h"
#include "kernels.h"
#include "rosy_amd.h"

using namespace rsy;

namespace {

// Orders a frame's setup blocks so that the blocks of different views that read the same indices
// (the same `first`) run on the same XCD, one after another, and share its L2: workgroups are dealt
// to the 8 XCDs round-robin, so within each group of 8 chunks the order is (view rank, chunk) and a
// chunk's blocks are 8 ids apart.  The order changes no result: every block writes its own slots,
// and bin counts are sums.
void interleave_views(std::vector<BlockEntry>& blocks) {
    constexpr uint32_t kXcds = 8;
    std::vector<uint32_t> chunk(blocks.size()), rank(blocks.size());
    {
        std::vector<std::pair<uint32_t, uint32_t>> order;  // (first, block)
        order.reserve(blocks.size());
        for (uint32_t i = 0; i < blocks.size(); ++i) order.push_back({blocks[i].first, i});
        std::stable_sort(order.begin(), order.end());
        uint32_t id = 0;
        for (size_t k = 0; k < order.size(); ++k) {
            if (k > 0 && order[k].first != order[k - 1].first) ++id;
            chunk[order[k].second] = id;
            rank[order[k].second] = (k > 0 && order[k].first == order[k - 1].first) ? rank[order[k - 1].second] + 1 : 0;
        }
    }
    std::vector<uint32_t> idx(blocks.size());
    for (uint32_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
        const uint32_t ga = chunk[a] / kXcds, gb = chunk[b] / kXcds;
        if (ga != gb) return ga < gb;
        if (rank[a] != rank[b]) return rank[a] < rank[b];
        return chunk[a] < chunk[b];
    });
    std::vector<BlockEntry> out;
    out.reserve(blocks.size());
    for (uint32_t i : idx) out.push_back(blocks[i]);
    blocks.swap(out);
}

// kDraw* flags of a draw with material index `material` under the context's materials and samplers.
uint32_t draw_flags(const rsy_ctx* c, uint32_t material) {
    uint32_t flags = 0;
    if (material < c->h_materials.size()) {
        const rsy_gpu_material& md = c->h_materials[material];
        const uint32_t si[4] = {md.color_sampler_index, md.normal_sampler_index, md.metallic_sampler_index,
                                md.mixmap_sampler_index};
        for (uint32_t k : si)
            if (k < c->h_samplers.size() && (c->h_samplers[k].wrap_s == 0 || c->h_samplers[k].wrap_t == 0))
                flags |= kDrawClampSampler;
    }
    return flags;
}

// Rebuilds the per-list draw tables (Graphics.cpp:4716-4735: shadow = every surface,
// opaque = !blended, blended = blended) from the surfaces and the mesh table.
rsy_result rebuild_lists(rsy_ctx* c) {
    for (int l = 0; l < 3; ++l) {
        c->lists[l].clear();
        c->list_index[l].clear();
        c->list_tris[l] = 0;
        c->list_verts[l] = 0;
    }
    for (uint32_t i = 0; i < c->surfaces.size(); ++i) {
        const rsy_draw& s = c->surfaces[i];
        if (s.mesh_index >= c->meshes.size()) return fail(c, RSY_INVALID_ARGUMENT, "surface mesh_index out of range");
        const rsy_mesh_range& m = c->meshes[s.mesh_index];
        const uint64_t obj = s.graphic_objects_offset + s.graphics_object_index;
        if (obj >= c->h_objects.size()) return fail(c, RSY_INVALID_ARGUMENT, "surface object index out of range");
        if (uint64_t(s.start_index) + s.index_count > m.num_indices)
            return fail(c, RSY_INVALID_ARGUMENT, "surface index range exceeds its mesh");
        DrawGPU d{};
        d.first_index = m.index_offset + s.start_index;
        d.tri_count = s.index_count / 3;
        d.vtx_base = uint32_t(m.vertex_buffer_offset / sizeof(rsy_position));
        d.vtx_count = m.num_vertices;
        d.object = uint32_t(obj);
        d.material = s.material_index < 0xFFFFFFFFull ? uint32_t(s.material_index) : 0xFFFFFFFFu;
        d.flags = draw_flags(c, d.material);
        for (int k = 0; k < 3; ++k) {
            d.bmin[k] = c->mesh_bounds[size_t(s.mesh_index) * 6 + k];
            d.bmax[k] = c->mesh_bounds[size_t(s.mesh_index) * 6 + 3 + k];
        }
        const int lists_of[2] = {0, s.blended ? 2 : 1};
        for (int l : lists_of) {
            d.tri_base = c->list_tris[l];
            d.xv_off = uint32_t(c->list_verts[l]);
            c->lists[l].push_back(d);
            c->list_index[l].push_back(i);
            c->list_tris[l] += d.tri_count;
            c->list_verts[l] += d.vtx_count;
        }
    }
    for (int l = 0; l < 3; ++l)
        if (c->list_tris[l] >= 0x7fffffffu || c->list_verts[l] >= 0x7fffffffull)
            return fail(c, RSY_OVERFLOW, "draw list too large");
    for (int l = 0; l < 3; ++l) {
        const size_t bytes = std::max<size_t>(1, c->lists[l].size()) * sizeof(DrawGPU);
        HIP_TRY(c, c->lists_d[l].ensure(bytes), "allocate draw list");
        if (!c->lists[l].empty())
            HIP_TRY(c, hipMemcpy(c->lists_d[l].p, c->lists[l].data(), c->lists[l].size() * sizeof(DrawGPU), hipMemcpyHostToDevice),
                    "upload draw list");
    }
    c->lists_version++;
    c->lists_dirty = false;
    invalidate_graph(c);
    return RSY_OK;
}

rsy_result resolve_views(rsy_ctx* c, const rsy_frame_desc* f, std::vector<FrameViewDesc>& views, uint32_t& shade) {
    views.clear();
    shade = f ? f->shade_mode : RSY_SHADE_PLAIN;
    if (shade > RSY_SHADE_NEURAL) return fail(c, RSY_INVALID_ARGUMENT, "unknown shade mode");
    if (shade == RSY_SHADE_NEURAL && !neural_ready(c))
        return fail(c, RSY_INVALID_STATE, "material decoder not trained and evaluated (rsy_neural_evaluate)");
    // shadow cascades first (the reference records the shadow pass before the main pass)
    if (f && f->n_cascades) {
        if (!f->cascades) return fail(c, RSY_INVALID_ARGUMENT, "cascades == NULL");
        for (uint32_t i = 0; i < f->n_cascades; ++i) {
            FrameViewDesc v{};
            v.shadow = true;
            v.width = v.height = f->cascades[i].size;
            std::memcpy(v.vp, f->cascades[i].light_view_projection, sizeof(v.vp));
            v.split_far = f->cascades[i].split_far;
            views.push_back(v);
        }
    } else {
        FrameViewDesc v{};
        v.shadow = true;
        v.width = v.height = (f && f->default_shadow_size) ? f->default_shadow_size : 2048u;
        std::memcpy(v.vp, c->sd.shadow_projection_near, sizeof(v.vp));
        v.split_far = INFINITY;
        views.push_back(v);
    }
    if (f && f->n_camera_views) {
        if (!f->camera_views) return fail(c, RSY_INVALID_ARGUMENT, "camera_views == NULL");
        for (uint32_t i = 0; i < f->n_camera_views; ++i) {
            FrameViewDesc v{};
            v.shadow = false;
            v.width = f->camera_views[i].width;
            v.height = f->camera_views[i].height;
            std::memcpy(v.vp, f->camera_views[i].view_projection, sizeof(v.vp));
            std::memcpy(v.camera_position, f->camera_views[i].camera_position, sizeof(v.camera_position));
            views.push_back(v);
        }
    } else {
        FrameViewDesc v{};
        v.shadow = false;
        v.width = uint32_t(c->sd.draw_extent[0]);
        v.height = uint32_t(c->sd.draw_extent[1]);
        std::memcpy(v.vp, c->sd.view_projection, sizeof(v.vp));
        std::memcpy(v.camera_position, c->sd.camera_position, sizeof(v.camera_position));
        views.push_back(v);
    }
    if (views.size() > kMaxViews) return fail(c, RSY_INVALID_ARGUMENT, "too many views");
    for (const auto& v : views)
        if (v.width == 0 || v.height == 0 || v.width > kMaxViewExtent || v.height > kMaxViewExtent)
            return fail(c, RSY_INVALID_ARGUMENT, "view extent must be in [1, 8192]");
    return RSY_OK;
}

uint32_t msaa_of(const rsy_ctx* c) { return c->rs.msaa_samples == 4 ? 4u : 1u; }

uint32_t choose_tile(const rsy_ctx* c) {
    if (msaa_of(c) == 4) return 32u;  // the 4-sample tile of a 64-pixel bin exceeds the LDS budget
    if (c->rs.tile_size) return c->rs.tile_size;  // 8, 16, 32 or 64 (rsy_update checks it)
    return c->autotuned_tile ? c->autotuned_tile : 32u;
}

// Computes the bin, vertex, record and clip-queue layout of a frame into L: host tables only.
// Record, clip-queue and bin-pair capacities are scaled by rec_scale (it doubles on overflow).
rsy_result plan_layout(rsy_ctx* c, const std::vector<FrameViewDesc>& views, uint32_t tile, uint32_t shade,
                       uint32_t rec_scale, Layout& L) {
    L.valid = false;
    L.views = views;
    L.tile = tile;
    L.lists_version = c->lists_version;
    L.shade = shade;
    L.vg.clear();
    L.vblocks.clear();
    L.tblocks.clear();
    L.total_bins = 0;
    L.blend_bins = 0;
    L.total_xv = 0;
    L.rec_total = L.clip_total = L.big_total = 0;
    L.cull_total = 0;
    L.n_shadow = L.n_camera = L.n_blend = 0;
    const uint32_t shift = ilog2(tile);
    // the requested views, then one blend view per camera view when the blended list is drawn
    // (the blended list drawn over that camera view)
    std::vector<uint32_t> order(views.size());
    for (uint32_t vi = 0; vi < views.size(); ++vi) order[vi] = vi;
    const bool blend = shade != RSY_SHADE_NONE && !c->lists[2].empty();
    if (blend && tile < 16)
        return fail(c, RSY_INVALID_ARGUMENT, "the blended pass collects fragments in 16x16 sub-tiles: bins of 16 pixels or more");
    if (blend && msaa_of(c) == 4 && c->list_tris[2] >= (1u << 28))
        return fail(c, RSY_INVALID_ARGUMENT, "4-sample blended pass: the blended list must hold fewer than 2^28 triangles");
    if (blend)
        for (uint32_t vi = 0; vi < views.size(); ++vi)
            if (!views[vi].shadow) order.push_back(vi);
    L.first_blend = uint32_t(views.size());
    const bool min_cap = (c->debug_flags & RSY_DEBUG_MIN_CAPACITY) != 0;
    for (uint32_t oi = 0; oi < order.size(); ++oi) {
        const uint32_t vi = uint32_t(L.vg.size());
        const FrameViewDesc& f = views[order[oi]];
        const bool is_blend = oi >= views.size();
        ViewGPU g{};
        g.width = f.width;
        g.height = f.height;
        g.tile = tile;
        g.tiles_x = (f.width + tile - 1) >> shift;
        g.tiles_y = (f.height + tile - 1) >> shift;
        g.tile_base = L.total_bins;
        g.is_shadow = f.shadow ? 1 : 0;
        g.list = f.shadow ? 0 : (is_blend ? 2 : 1);
        g.blend = is_blend ? 1 : 0;
        g.source = order[oi];
        g.samples = f.shadow ? 1u : msaa_of(c);
        const auto& list = c->lists[g.list];
        g.n_draws = uint32_t(list.size());
        g.cull_base = L.cull_total;
        L.cull_total += g.n_draws;
        g.xv_base = uint32_t(L.total_xv);
        const uint32_t tris = c->list_tris[g.list];
        // capacities in 64 bits: rec_scale doubles on every record overflow
        const uint64_t clip_cap = uint64_t(min_cap ? 256u : std::max<uint32_t>(4096u, tris / 8)) * rec_scale;
        const uint64_t rec_cap = uint64_t(tris) + 2 * clip_cap;
        // a triangle spans about (32 / tile)^2 times as many bins below 32-pixel bins
        const uint64_t small_bins = tile < 32 ? uint64_t(32 / tile) * (32 / tile) : 1u;
        const uint64_t big_cap =
            uint64_t(min_cap ? 4096u : std::max<uint32_t>(65536u, tris / 4)) * (min_cap ? 1u : small_bins) * rec_scale;
        if (rec_cap >= 0xffffffffull || uint64_t(L.rec_total) + rec_cap >= 0xffffffffull)
            return fail(c, RSY_OVERFLOW, "record buffer too large");
        if (uint64_t(L.clip_total) + clip_cap >= 0xffffffffull) return fail(c, RSY_OVERFLOW, "clip queue too large");
        if (uint64_t(L.big_total) + big_cap >= 0xffffffffull) return fail(c, RSY_OVERFLOW, "bin pair buffer too large");
        g.rec_base = L.rec_total;
        g.tri_count = tris;
        g.rec_cap = uint32_t(rec_cap);
        g.clip_base = L.clip_total;
        g.clip_cap = uint32_t(clip_cap);
        g.big_base = L.big_total;
        g.big_cap = uint32_t(big_cap);
        // the blended pass draws both faces (VK_CULL_MODE_NONE, Graphics.cpp:5675)
        g.cull = c->rs.cull_enabled && !is_blend ? 1 : 0;
        g.reverse = c->rs.reverse_winding_order_enabled ? 1 : 0;
        g.bias = c->rs.depth_bias_enabled ? 1 : 0;
        g.bias_const = c->rs.depth_bias_constant;
        g.bias_clamp = c->rs.depth_bias_clamp;
        g.bias_slope = c->rs.depth_bias_slope_factor;
        std::memcpy(g.view_projection, f.vp, sizeof(f.vp));
        L.total_bins += g.tiles_x * g.tiles_y;
        if (is_blend) L.blend_bins += g.tiles_x * g.tiles_y;
        L.total_xv += c->list_verts[g.list];
        L.rec_total += g.rec_cap;
        L.clip_total += g.clip_cap;
        L.big_total += g.big_cap;
        for (uint32_t d = 0; d < g.n_draws; ++d) {
            const DrawGPU& dr = list[d];
            for (uint32_t s = 0; s < dr.tri_count; s += kSetupBlockTris)
                L.tblocks.push_back({vi, d, s, g.list, dr.first_index + 3u * s, std::min(kSetupBlockTris, dr.tri_count - s),
                                     g.rec_base + dr.tri_base + s, g.xv_base + dr.xv_off, g.cull_base + d, dr.tri_base + s,
                                     dr.vtx_base, dr.material});
        }
        if (is_blend) ++L.n_blend;
        else if (f.shadow) ++L.n_shadow;
        else ++L.n_camera;
        L.vg.push_back(g);
    }
    if (L.total_xv >= 0xffffffffull) return fail(c, RSY_OVERFLOW, "transformed vertex buffer too large");
    // transform blocks: a chunk of a draw's vertices for a run of up to kTransformViews consecutive
    // views of the same draw list (the shadow cascades, the camera views), so that the positions are
    // read once for all of them
    for (uint32_t v0 = 0; v0 < L.vg.size();) {
        const uint32_t list_id = L.vg[v0].list;
        uint32_t nv = 1;
        while (v0 + nv < L.vg.size() && nv < kTransformViews && L.vg[v0 + nv].list == list_id) ++nv;
        const auto& list = c->lists[list_id];
        for (uint32_t d = 0; d < uint32_t(list.size()); ++d) {
            const DrawGPU& dr = list[d];
            for (uint32_t s = 0; s < dr.vtx_count; s += kTransformBlockVerts)
                L.vblocks.push_back({v0, d, s, list_id, dr.vtx_base + s, std::min(kTransformBlockVerts, dr.vtx_count - s),
                                     dr.xv_off + s, nv, d, dr.object, 0u, 0u});
        }
        v0 += nv;
    }
    // a frame with many setup blocks runs two triangles per thread: each draw's chunks in pairs
    L.setup_tpt = L.tblocks.size() > kSetupPairBlocks ? 2u : 1u;
    if (L.setup_tpt == 2) {
        std::vector<BlockEntry> paired;
        paired.reserve(L.tblocks.size() / 2 + 1);
        for (size_t i = 0; i < L.tblocks.size(); ++i) {
            BlockEntry b = L.tblocks[i];
            if (i + 1 < L.tblocks.size() && L.tblocks[i + 1].view == b.view && L.tblocks[i + 1].draw == b.draw &&
                L.tblocks[i + 1].start == b.start + kSetupBlockTris) {
                b.count += L.tblocks[i + 1].count;  // the second chunk's slots, indices and ids follow the first's
                ++i;
            }
            paired.push_back(b);
        }
        L.tblocks.swap(paired);
    }
    interleave_views(L.tblocks);  // (the transform keeps view order: its output writes stream per view)
    // output clears: in the view's setup blocks when each block's share is small, else a fill launch
    std::vector<uint32_t> view_blocks(L.vg.size(), 0u);
    for (const BlockEntry& b : L.tblocks) ++view_blocks[b.view];
    for (size_t v = 0; v < L.vg.size(); ++v) {
        ViewGPU& g = L.vg[v];
        const uint64_t bytes = uint64_t(g.width) * g.height * (g.is_shadow ? 4u : 8u * g.samples);
        g.setup_clear = !g.blend && view_blocks[v] && bytes <= uint64_t(view_blocks[v]) * kSetupClearBytes ? 1u : 0u;
    }
    // bin lists: an estimate of the frame's (triangle, bin) pairs
    uint64_t want_pairs = 0;
    const uint64_t per_tri = tile < 32 ? 3u * (32 / tile) * (32 / tile) : 3u;
    for (const auto& g : L.vg) want_pairs += uint64_t(c->list_tris[g.list]) * per_tri + g.tiles_x * g.tiles_y;
    want_pairs *= rec_scale;
    if (min_cap) want_pairs = std::min<uint64_t>(want_pairs, 4096);
    L.want_pairs = std::min<uint64_t>(want_pairs, 0xfffffff0ull);
    L.raster_grid = 0;
    return RSY_OK;
}

// Points each view's output word at its buffer in m (visibility, shadow map or blend list).
void bind_outputs(Layout& L, const FrameMem& m) {
    uint32_t ci = 0, si = 0, bi = 0;
    for (auto& g : L.vg) {
        if (g.blend) g.out = reinterpret_cast<uint64_t>(m.blend_list[bi++]);
        else if (g.is_shadow) g.out = reinterpret_cast<uint64_t>(m.shadow_map[si++]);
        else g.out = reinterpret_cast<uint64_t>(m.vis[ci++]);
    }
}

// Computes the frame's layout and (re)allocates the context's frame buffers for it.
rsy_result build_layout(rsy_ctx* c, const std::vector<FrameViewDesc>& views, uint32_t tile, uint32_t shade) {
    Layout& L = c->layout;
    const bool reuse = L.valid && same_views(L.views, views) && L.tile == tile && L.lists_version == c->lists_version &&
                       L.shade == shade;
    if (reuse) {
        L.views = views;  // matrices may change without a new layout
        for (size_t i = 0; i < views.size(); ++i) std::memcpy(L.vg[i].view_projection, views[i].vp, sizeof(views[i].vp));
        for (size_t i = views.size(); i < L.vg.size(); ++i)  // blend views follow their camera view
            std::memcpy(L.vg[i].view_projection, views[L.vg[i].source].vp, sizeof(views[0].vp));
        return RSY_OK;
    }
    invalidate_graph(c);
    rsy_result r = plan_layout(c, views, tile, shade, c->rec_scale, L);
    if (r != RSY_OK) return r;
    // buffers
    const size_t nv = L.vg.size();
    HIP_TRY(c, c->views_d.ensure(nv * sizeof(ViewGPU)), "allocate views");
    HIP_TRY(c, c->xv.ensure(size_t(std::max<uint64_t>(1, L.total_xv)) * sizeof(float4)), "allocate vertices");
    HIP_TRY(c, c->recs.ensure(size_t(std::max<uint32_t>(1, L.rec_total)) * sizeof(TriRec)), "allocate records");
    HIP_TRY(c, c->clipq.ensure(size_t(std::max<uint32_t>(1, L.clip_total)) * sizeof(ClipEntry)), "allocate clip queue");
    HIP_TRY(c, c->counters.ensure(nv * kCtrPerView * sizeof(uint32_t)), "allocate counters");
    HIP_TRY(c, c->draw_culled.ensure(size_t(std::max<uint32_t>(1, L.cull_total)) * 4), "allocate culling flags");
    HIP_TRY(c, c->bin_counts.ensure(size_t(L.total_bins) * 4), "allocate bins");
    HIP_TRY(c, c->bin_offsets.ensure(size_t(L.total_bins) * 4), "allocate bins");
    HIP_TRY(c, c->pair_slots.ensure(size_t(std::max<uint32_t>(1, L.rec_total)) * 16), "allocate bin slots");
    HIP_TRY(c, c->bin_tiles.ensure(size_t(std::max<uint32_t>(1, L.rec_total)) * 4), "allocate record bins");
    HIP_TRY(c, c->shade_recs.ensure(size_t(std::max<uint32_t>(1, L.rec_total)) * sizeof(ShadeRec)), "allocate shading records");
    HIP_TRY(c, c->big_pairs.ensure(size_t(std::max<uint32_t>(1, L.big_total)) * 16), "allocate bin pairs");
    HIP_TRY(c, c->block_valid.ensure(std::max<size_t>(1, L.tblocks.size()) * 4), "allocate statistics");
    HIP_TRY(c, c->totals.ensure(32), "allocate statistics");
    HIP_TRY(c, c->scan_partial.ensure((size_t(L.total_bins) / kScanBlockBins + 1) * kScanPartialBytes), "allocate bins");
    HIP_TRY(c, c->vblocks_d.ensure(std::max<size_t>(1, L.vblocks.size()) * sizeof(BlockEntry)), "allocate blocks");
    HIP_TRY(c, c->tblocks_d.ensure(std::max<size_t>(1, L.tblocks.size()) * sizeof(BlockEntry)), "allocate blocks");
    if (!L.vblocks.empty())
        HIP_TRY(c, hipMemcpy(c->vblocks_d.p, L.vblocks.data(), L.vblocks.size() * sizeof(BlockEntry), hipMemcpyHostToDevice), "upload blocks");
    if (!L.tblocks.empty())
        HIP_TRY(c, hipMemcpy(c->tblocks_d.p, L.tblocks.data(), L.tblocks.size() * sizeof(BlockEntry), hipMemcpyHostToDevice), "upload blocks");
    // bin lists: start from the estimate, grow on overflow
    if (L.want_pairs > c->list_cap) {
        HIP_TRY(c, c->lists_mem.ensure(size_t(L.want_pairs) * 4), "allocate bin lists");
        c->list_cap = uint32_t(L.want_pairs);
    }
    HIP_TRY(c, c->items.ensure(size_t(raster_items_cap(L.total_bins, c->list_cap)) * sizeof(uint4)), "allocate work items");
    // outputs
    c->vis_out.resize(L.n_camera);
    c->color_out.resize(L.n_camera);
    c->shadow_factor_out.resize(L.n_camera);
    c->sample_color.resize(L.n_camera);
    c->shadow_maps.resize(L.n_shadow);
    c->blend_lists.resize(L.n_blend);
    FrameMem& m = c->mem;
    m = FrameMem{};
    uint32_t ci = 0, si = 0, bi = 0;
    uint64_t blend_px = 0;
    for (auto& g : L.vg) {
        const size_t px = size_t(g.width) * g.height;
        if (g.blend) {
            HIP_TRY(c, c->blend_lists[bi].ensure(px * 8), "allocate blended fragment lists");
            m.blend_list[bi] = c->blend_lists[bi].as<uint2>();
            blend_px += px;
            ++bi;
        } else if (g.is_shadow) {
            HIP_TRY(c, c->shadow_maps[si].ensure(px * 4), "allocate shadow map");
            m.shadow_map[si] = c->shadow_maps[si].as<float>();
            ++si;
        } else {
            HIP_TRY(c, c->vis_out[ci].ensure(px * 8 * g.samples), "allocate visibility buffer");
            HIP_TRY(c, c->color_out[ci].ensure(px * 16), "allocate colour buffer");
            if (g.samples == 4 && L.n_blend) {
                HIP_TRY(c, c->sample_color[ci].ensure(px * 16 * 4), "allocate sample colours");
                m.sample_color[ci] = c->sample_color[ci].as<float4>();
            }
            HIP_TRY(c, c->shadow_factor_out[ci].ensure(px * 4), "allocate shadow factor");
            m.vis[ci] = c->vis_out[ci].as<uint64_t>();
            m.color[ci] = c->color_out[ci].as<float4>();
            m.shadow_factor[ci] = c->shadow_factor_out[ci].as<float>();
            ++ci;
        }
    }
    if (L.n_blend) {
        // blended fragments: start at two per pixel of the camera views (grows on overflow)
        const uint64_t first = (c->debug_flags & RSY_DEBUG_MIN_CAPACITY) ? 4096 : 2 * blend_px + 4096;
        const uint64_t want = std::min<uint64_t>(std::max<uint64_t>(c->frag_cap, first), 0xfffffff0ull);
        HIP_TRY(c, c->frags.ensure(size_t(want) * 16), "allocate blended fragments");
        HIP_TRY(c, c->frag_keys.ensure(size_t(want) * 4), "allocate blended fragment keys");
        HIP_TRY(c, c->frag_color.ensure(size_t(want) * 16), "allocate blended fragment colours");
        HIP_TRY(c, c->frag_mask.ensure(size_t(want)), "allocate blended fragment samples");
        c->frag_cap = uint32_t(want);
    }
    HIP_TRY(c, c->readback_counters.ensure(nv * kCtrPerView * 4 + 64), "allocate readback");
    m.views = c->views_d.as<ViewGPU>();
    m.xv = c->xv.as<float4>();
    m.recs = c->recs.as<TriRec>();
    m.bin_tiles = c->bin_tiles.as<uint32_t>();
    m.shade_recs = c->shade_recs.as<ShadeRec>();
    m.clipq = c->clipq.as<ClipEntry>();
    m.counters = c->counters.as<uint32_t>();
    m.draw_culled = c->draw_culled.as<uint32_t>();
    m.bin_counts = c->bin_counts.as<uint32_t>();
    m.bin_offsets = c->bin_offsets.as<uint32_t>();
    m.pair_slots = c->pair_slots.as<uint32_t>();
    m.big_pairs = c->big_pairs.as<uint4>();
    m.block_valid = c->block_valid.as<uint32_t>();
    m.totals = c->totals.as<uint32_t>();
    m.scan_partial = c->scan_partial.p;
    m.vblocks = c->vblocks_d.as<BlockEntry>();
    m.tblocks = c->tblocks_d.as<BlockEntry>();
    m.lists = c->lists_mem.as<uint32_t>();
    m.list_cap = c->list_cap;
    m.items = c->items.as<uint4>();
    m.frags = c->frags.as<uint4>();
    m.frag_keys = c->frag_keys.as<uint32_t>();
    m.frag_color = c->frag_color.as<float4>();
    m.frag_mask = c->frag_mask.as<uint8_t>();
    m.frag_cap = c->frag_cap;
    const size_t nsv = L.n_camera, nsc = L.n_shadow, nbv = L.n_blend;
    HIP_TRY(c, c->shade_views_d.ensure(std::max<size_t>(1, nsv) * sizeof(ShadeView) + 64), "allocate shade views");
    HIP_TRY(c, c->shade_cascades_d.ensure(std::max<size_t>(1, nsc) * sizeof(ShadeCascade) + 64), "allocate cascades");
    HIP_TRY(c, c->blend_shade_views_d.ensure(std::max<size_t>(1, nbv) * sizeof(ShadeView) + 64), "allocate blend views");
    const size_t staging = ((nv * sizeof(ViewGPU) + 255) & ~size_t(255)) + ((nsv * sizeof(ShadeView) + 255) & ~size_t(255)) +
                           ((nsc * sizeof(ShadeCascade) + 255) & ~size_t(255)) + nbv * sizeof(ShadeView) + 256;
    HIP_TRY(c, c->staging.ensure(staging), "allocate staging");
    m.shade_views = c->shade_views_d.as<ShadeView>();
    m.shade_cascades = c->shade_cascades_d.as<ShadeCascade>();
    m.blend_views = c->blend_shade_views_d.as<ShadeView>();
    bind_outputs(L, m);
    L.valid = true;
    return RSY_OK;
}

// Per-frame constants (view records, shading views and cascades) go to the device with async copies
// on the context stream: through `staging` (pinned, at least the tables' size) when given, else
// straight from host memory.  They are not part of the captured frame graph: a replayed graph reads
// whatever the last upload left in device memory.
rsy_result upload_frame(rsy_ctx* c, const Layout& L, const FrameMem& m, void* staging, uint32_t debug_flags) {
    hipStream_t s = c->stream;
    const uint32_t nv = uint32_t(L.vg.size());
    c->h_shade_views.clear();
    c->h_shade_cascades.clear();
    c->h_blend_views.clear();
    uint32_t camera_of[kMaxViews] = {};  // shading view index of each camera view
    uint32_t blocks = 0, ci = 0;
    for (uint32_t i = 0; i < nv; ++i) {
        const ViewGPU& vg = L.vg[i];
        if (vg.blend) {
            // the blended pass of camera view vg.source: its records and transformed vertices, shaded
            // into that view's colour buffer
            ShadeView sv = c->h_shade_views[camera_of[vg.source]];
            sv.vis = nullptr;
            sv.recs = m.recs + vg.rec_base;
            sv.shade_recs = m.shade_recs + vg.rec_base;
            sv.xv_base = vg.xv_base;
            sv.shadow_factor = nullptr;
            c->h_blend_views.push_back(sv);
        } else if (vg.is_shadow) {
            ShadeCascade sc{};
            sc.size = vg.width;
            sc.split_far = L.views[i].split_far;
            sc.map = reinterpret_cast<const float*>(vg.out);
            std::memcpy(sc.m, vg.view_projection, sizeof(sc.m));
            c->h_shade_cascades.push_back(sc);
        } else {
            ShadeView sv{};
            sv.width = vg.width;
            sv.height = vg.height;
            sv.blocks_x = (vg.width + 15) / 16;
            sv.block_base = blocks;
            blocks += sv.blocks_x * ((vg.height + 15) / 16);
            sv.vis = reinterpret_cast<const unsigned long long*>(vg.out);
            sv.recs = m.recs + vg.rec_base;
            sv.shade_recs = m.shade_recs + vg.rec_base;
            sv.xv_base = vg.xv_base;
            sv.samples = vg.samples;
            sv.color = m.color[ci];
            sv.sample_color = vg.samples == 4 && L.n_blend ? m.sample_color[ci] : nullptr;
            sv.shadow_factor = (debug_flags & RSY_DEBUG_SHADOW_FACTOR) ? m.shadow_factor[ci] : nullptr;
            std::memcpy(sv.view_projection, vg.view_projection, sizeof(sv.view_projection));
            std::memcpy(sv.camera_position, L.views[i].camera_position, sizeof(sv.camera_position));
            camera_of[i] = uint32_t(c->h_shade_views.size());
            c->h_shade_views.push_back(sv);
            ++ci;
        }
    }
    c->shade_blocks = blocks;
    const size_t nsv = c->h_shade_views.size(), nsc = c->h_shade_cascades.size(), nbv = c->h_blend_views.size();
    struct Table {
        void* dst;
        const void* src;
        size_t bytes;
    };
    const Table tables[4] = {{m.views, L.vg.data(), nv * sizeof(ViewGPU)},
                             {m.shade_views, c->h_shade_views.data(), nsv * sizeof(ShadeView)},
                             {m.shade_cascades, c->h_shade_cascades.data(), nsc * sizeof(ShadeCascade)},
                             {m.blend_views, c->h_blend_views.data(), nbv * sizeof(ShadeView)}};
    size_t at = 0;
    for (const Table& t : tables) {
        if (!t.bytes) continue;
        const void* src = t.src;
        if (staging) {
            uint8_t* st = static_cast<uint8_t*>(staging) + at;
            std::memcpy(st, t.src, t.bytes);
            src = st;
            at += (t.bytes + 255) & ~size_t(255);
        }
        HIP_TRY(c, hipMemcpyAsync(t.dst, src, t.bytes, hipMemcpyHostToDevice, s), "upload frame constants");
    }
    return RSY_OK;
}

// Enqueues the kernels of one frame on the context stream (graph-capturable: kernels only).  With
// `bookkeeping` (the ordinary frame, not while capturing): HIP events around the frame and, with
// RSY_DEBUG_STAGE_TIMING, between its stages, and the counters copied to the pinned readback buffer.
rsy_result enqueue_frame(rsy_ctx* c, const Layout& L, const FrameMem& m, bool capturing, bool bookkeeping) {
    hipStream_t s = c->stream;
    const uint32_t nv = uint32_t(L.vg.size());
    const uint32_t shift = ilog2(L.tile);
    const bool events = bookkeeping && !capturing;
    const bool timing = events && (c->debug_flags & RSY_DEBUG_STAGE_TIMING);
    if (events) HIP_TRY(c, hipEventRecord(c->ev_begin, s), "event");

    GeometryLaunch g{};
    g.views = m.views;
    g.n_views = nv;
    g.draws0 = c->lists_d[0].as<DrawGPU>();
    g.draws1 = c->lists_d[1].as<DrawGPU>();
    g.draws2 = c->lists_d[2].as<DrawGPU>();
    g.objects = c->objects.as<float>();
    g.vblocks = m.vblocks;
    g.n_vblocks = uint32_t(L.vblocks.size());
    g.tblocks = m.tblocks;
    g.n_tblocks = uint32_t(L.tblocks.size());
    g.setup_tpt = L.setup_tpt;
    g.px = c->px.as<float>();
    g.py = c->py.as<float>();
    g.pz = c->pz.as<float>();
    g.indices = c->indices.as<uint32_t>();
    g.xv = m.xv;
    g.recs = m.recs;
    g.bin_tiles = m.bin_tiles;
    g.shade_recs = m.shade_recs;
    g.clipq = m.clipq;
    g.counters = m.counters;
    g.bin_counts = m.bin_counts;
    g.pair_slots = m.pair_slots;
    g.slot_count = L.rec_total;
    g.big_pairs = m.big_pairs;
    g.block_valid = m.block_valid;
    g.draw_culled = m.draw_culled;
    g.shift = shift;
    g.clip_blocks = 128;
    g.total_bins = L.total_bins;
    HIP_TRY(c, launch_geometry(g, s), "geometry kernels");
    // the outputs that the setup blocks do not clear (views with few setup blocks for their size)
    for (const ViewGPU& v : L.vg)
        if (!v.blend && !v.setup_clear) {
            const uint64_t words = uint64_t(v.width) * v.height * (v.is_shadow ? 1u : 2u * v.samples);
            HIP_TRY(c, launch_fill_u32(reinterpret_cast<uint32_t*>(v.out), uint32_t(words), 0u, s), "clear output");
        }
    if (timing) HIP_TRY(c, hipEventRecord(c->ev_stage[0], s), "event");

    BinLaunch b{};
    b.views = g.views;
    b.n_views = nv;
    b.total_bins = L.total_bins;
    b.bin_counts = g.bin_counts;
    b.bin_offsets = m.bin_offsets;
    b.pair_slots = m.pair_slots;
    b.slot_count = L.rec_total;
    b.big_pairs = m.big_pairs;
    b.block_valid = m.block_valid;
    b.n_tblocks = uint32_t(L.tblocks.size());
    b.totals = m.totals;
    b.recs = g.recs;
    b.bin_tiles = g.bin_tiles;
    b.counters = g.counters;
    b.lists = m.lists;
    b.list_cap = m.list_cap;
    b.shift = shift;
    uint32_t max_tris = 0;
    for (const auto& v : L.vg) max_tris = std::max(max_tris, v.tri_count);
    b.scatter_blocks = std::max<uint32_t>(64, std::min<uint32_t>(4096, (max_tris + 1023) / 1024));
    b.scan_partial = m.scan_partial;
    b.items = m.items;
    b.items_cap = raster_items_cap(L.total_bins, m.list_cap);
    // one raster workgroup per work item: as many as the last frame of this layout had (with slack),
    // or an estimate (every bin, and a split item per 256 estimated pairs); the scan flags a frame
    // with more items, and finish_frame re-runs it with a grid that fits
    uint32_t grid = L.raster_grid ? L.raster_grid
                                  : uint32_t(std::min<uint64_t>(uint64_t(L.total_bins) + L.want_pairs / 256 + 256,
                                                                0xffffffffull));
    if ((c->debug_flags & RSY_DEBUG_MIN_CAPACITY) && !L.raster_grid) grid = std::min(grid, 64u);  // the re-run path
    b.raster_grid = std::min(grid, b.items_cap);
    HIP_TRY(c, launch_binning(b, s), "binning kernels");
    if (timing) HIP_TRY(c, hipEventRecord(c->ev_stage[1], s), "event");

    RasterLaunch r{};
    r.views = g.views;
    r.n_views = nv;
    r.total_bins = L.total_bins;
    r.tile = L.tile;
    r.lists = b.lists;
    r.list_cap = m.list_cap;
    r.recs = g.recs;
    r.rec_cap_total = L.rec_total;
    r.items = b.items;
    r.items_cap = b.items_cap;
    r.totals = b.totals;
    r.samples = msaa_of(c);
    r.grid = b.raster_grid;
    HIP_TRY(c, launch_raster(r, s), "raster kernel");
    if (timing) HIP_TRY(c, hipEventRecord(c->ev_stage[2], s), "event");

    if (L.shade && L.n_camera) {
        ShadeLaunch sh{};
        sh.views = m.shade_views;
        sh.n_views = uint32_t(c->h_shade_views.size());
        sh.total_blocks = c->shade_blocks;
        sh.cascades = m.shade_cascades;
        sh.n_cascades = uint32_t(c->h_shade_cascades.size());
        sh.pcf_mode = c->rs.pcf_mode;
        sh.scene = c->scene_d.p;
        sh.draws = c->lists_d[1].as<DrawGPU>();
        sh.n_draws = uint32_t(c->lists[1].size());
        sh.indices = g.indices;
        sh.vertices = c->vertices.p;
        sh.objects = g.objects;
        sh.materials = c->materials.p;
        sh.n_materials = c->n_materials;
        sh.textures = c->tex_table.as<TexGPU>();
        sh.n_textures = c->n_textures;
        sh.texels = c->texels.as<uint8_t>();
        sh.samplers = c->samplers.p;
        sh.n_samplers = c->n_samplers;
        std::memcpy(sh.clear_color, c->rs.clear_color, sizeof(sh.clear_color));
        sh.srgb_lut = c->srgb_lut.as<float>();
        sh.xv = g.xv;
        sh.samples = msaa_of(c);
        sh.lod = c->any_mips ? 1u : 0u;
        if (L.shade == RSY_SHADE_NEURAL) neural_shade_args(c, sh);
        // the blended pass: the blended list of every camera view in primitive order
        BlendLaunch bl{};
        ShadeLaunch sb = sh;
        if (L.n_blend) {
            bl.views = g.views;
            bl.n_views = nv;
            bl.first_blend = L.first_blend;
            bl.blend_bins = L.blend_bins;
            bl.tile = L.tile;
            bl.bin_counts = g.bin_counts;
            bl.bin_offsets = b.bin_offsets;
            bl.lists = b.lists;
            bl.list_cap = m.list_cap;
            bl.recs = g.recs;
            bl.rec_cap_total = L.rec_total;
            bl.counters = g.counters;
            bl.frags = m.frags;
            bl.frag_keys = m.frag_keys;
            bl.frag_color = m.frag_color;
            bl.frag_mask = m.frag_mask;
            bl.frag_cap = m.frag_cap;
            bl.samples = msaa_of(c);
            sb.views = m.blend_views;
            sb.n_views = uint32_t(c->h_blend_views.size());
            sb.draws = c->lists_d[2].as<DrawGPU>();
            sb.n_draws = uint32_t(c->lists[2].size());
            sb.neural = 0;
        }
        // The blended fragments need only the visibility buffers and shadow maps, so outside stage
        // timing they are collected and shaded on the side stream while the opaque shading runs;
        // the blending waits for both.
        const bool side = L.n_blend && !timing;
        if (side) {
            HIP_TRY(c, hipEventRecord(c->ev_fork, s), "event");
            HIP_TRY(c, hipStreamWaitEvent(c->side_stream, c->ev_fork, 0), "fork");
            HIP_TRY(c, launch_blend_fragments(bl, sb, c->side_stream), "blended fragment kernels");
            HIP_TRY(c, hipEventRecord(c->ev_join, c->side_stream), "event");
        }
        if (timing) HIP_TRY(c, hipEventRecord(c->ev_stage[3], s), "event");
        HIP_TRY(c, launch_shade(sh, s), "shade kernel");
        if (timing) HIP_TRY(c, hipEventRecord(c->ev_stage[5], s), "event");
        if (L.n_blend) {
            if (side) HIP_TRY(c, hipStreamWaitEvent(s, c->ev_join, 0), "join");
            else HIP_TRY(c, launch_blend_fragments(bl, sb, s), "blended fragment kernels");
            HIP_TRY(c, launch_blend_resolve(bl, sb, s), "blend kernel");
        }
    } else if (timing) {
        HIP_TRY(c, hipEventRecord(c->ev_stage[3], s), "event");
        HIP_TRY(c, hipEventRecord(c->ev_stage[5], s), "event");
    }
    if (timing) HIP_TRY(c, hipEventRecord(c->ev_stage[4], s), "event");
    if (events) {
        HIP_TRY(c, hipMemcpyAsync(c->readback_counters.p, m.counters, nv * kCtrPerView * 4, hipMemcpyDeviceToHost, s),
                "readback");
        HIP_TRY(c, hipMemcpyAsync(static_cast<uint8_t*>(c->readback_counters.p) + nv * kCtrPerView * 4, m.totals, 16,
                                  hipMemcpyDeviceToHost, s), "readback");
        HIP_TRY(c, hipEventRecord(c->ev_end, s), "event");
    }
    return RSY_OK;
}

rsy_result ready_to_render(rsy_ctx* c) {
    if (!c->has_geometry) return fail(c, RSY_INVALID_STATE, "rsy_set_geometry has not been called");
    if (!c->has_objects) return fail(c, RSY_INVALID_STATE, "rsy_set_graphic_objects has not been called");
    if (!c->has_scene) return fail(c, RSY_INVALID_STATE, "rsy_update has not been called");
    if (c->lists_dirty) {
        rsy_result r = rebuild_lists(c);
        if (r != RSY_OK) return r;
    }
    return RSY_OK;
}

// Raster workgroups for the frames after one with `items` work items: those, a quarter more
// and 256 (a moving view changes its non-empty bins gradually).  Workgroups past the frame's items
// exit at once but still take dispatch slots, so the grid is not sized for every bin of the views.
uint32_t raster_grid_for(uint32_t items) {
    return uint32_t(std::min<uint64_t>(uint64_t(items) + items / 4 + 256, 0xffffffffull));
}

// Reads the last frame's counters (after a synchronize) and reports whether it overflowed.
bool collect_stats(rsy_ctx* c) {
    const Layout& L = c->layout;
    const uint32_t nv = uint32_t(L.vg.size());
    const uint32_t* ctr = static_cast<const uint32_t*>(c->readback_counters.p);
    const uint32_t* totals = ctr + nv * kCtrPerView;
    bool overflow = totals[2] != 0;
    rsy_frame_stats st{};
    st.tile_pairs = totals[0];
    st.triangles_setup = totals[1];
    for (uint32_t v = 0; v < nv; ++v) {
        const uint32_t* vc = ctr + v * kCtrPerView;
        overflow |= vc[kCtrOverflow] != 0;
        st.triangles_setup += vc[kCtrRecords];
        st.triangles_clipped += vc[kCtrClipped];
        const uint32_t list = L.vg[v].list;
        st.triangle_count += c->list_tris[list];
        st.draw_call_count += c->lists[list].size();
    }
    st.tile_size = L.tile;
    st.overflowed = overflow ? 1 : 0;
    // the raster grid of the next frames of this layout: this frame's work items and some slack
    if (!overflow && totals[3]) c->layout.raster_grid = raster_grid_for(totals[3]);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev_begin, c->ev_end) == hipSuccess) st.gpu_ms = ms;
    if (c->debug_flags & RSY_DEBUG_STAGE_TIMING) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, c->ev_begin, c->ev_stage[0]) == hipSuccess) st.stage_ms[0] = t;
        if (hipEventElapsedTime(&t, c->ev_stage[0], c->ev_stage[1]) == hipSuccess) st.stage_ms[1] = t;
        if (hipEventElapsedTime(&t, c->ev_stage[1], c->ev_stage[2]) == hipSuccess) st.stage_ms[2] = t;
        if (hipEventElapsedTime(&t, c->ev_stage[3], c->ev_stage[5]) == hipSuccess) st.stage_ms[3] = t;
        if (hipEventElapsedTime(&t, c->ev_stage[5], c->ev_stage[4]) == hipSuccess) st.stage_ms[4] = t;
    }
    c->stats = st;
    return overflow;
}

rsy_result render_now(rsy_ctx* c, const std::vector<FrameViewDesc>& views, uint32_t shade) {
    rsy_result r = build_layout(c, views, choose_tile(c), shade);
    if (r != RSY_OK) return r;
    r = upload_frame(c, c->layout, c->mem, c->staging.p, c->debug_flags);
    if (r != RSY_OK) return r;
    r = enqueue_frame(c, c->layout, c->mem, false, true);
    if (r != RSY_OK) return r;
    c->last_views = views;
    c->last_shade = shade;
    c->frame_pending = true;
    return RSY_OK;
}

// Waits for the stream; if the last frame overflowed a bin or record buffer, grows the buffers and
// renders it again (bounded number of attempts).
rsy_result finish_frame(rsy_ctx* c) {
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    if (!c->frame_pending) return RSY_OK;
    for (int attempt = 0; attempt < 8; ++attempt) {
        if (!collect_stats(c)) {
            c->stats.overflowed = uint32_t(attempt);  // re-runs this frame needed
            c->frame_pending = false;
            return RSY_OK;
        }
        // grow: bin list capacity to the measured pair count, record/clip capacity x2
        const uint32_t nv = uint32_t(c->layout.vg.size());
        const uint32_t total_pairs = static_cast<const uint32_t*>(c->readback_counters.p)[nv * kCtrPerView];
        const uint32_t over_bits = static_cast<const uint32_t*>(c->readback_counters.p)[nv * kCtrPerView + 2];
        const bool list_over = (over_bits & 1u) != 0;
        if (over_bits & 2u) {  // more raster work items than workgroups: a grid for this frame's count
            c->layout.raster_grid =
                raster_grid_for(static_cast<const uint32_t*>(c->readback_counters.p)[nv * kCtrPerView + 3]);
            invalidate_graph(c);
        }
        bool rec_over = false, frags_over = false;
        for (uint32_t v = 0; v < nv; ++v) {
            const uint32_t ov = static_cast<const uint32_t*>(c->readback_counters.p)[v * kCtrPerView + kCtrOverflow];
            rec_over |= (ov & ~kOverflowFrags) != 0;
            frags_over |= (ov & kOverflowFrags) != 0;
        }
        if (frags_over) {
            // the blended fragments of the frame, measured by the collection's counter
            const uint32_t n = static_cast<const uint32_t*>(
                c->readback_counters.p)[c->layout.first_blend * kCtrPerView + kCtrFrags];
            const uint64_t want = std::min<uint64_t>(uint64_t(n) + n / 4 + 4096, 0xfffffff0ull);
            if (want <= c->frag_cap) return fail(c, RSY_OVERFLOW, "blended fragments exceed 2^32");
            c->frag_cap = uint32_t(want);
            c->layout.valid = false;
        }
        if (rec_over) {
            if (c->rec_scale >= (1u << 20)) return fail(c, RSY_OVERFLOW, "record capacity cannot grow further");
            c->rec_scale *= 2;
            c->layout.valid = false;
        }
        if (list_over) {
            const uint64_t want = std::min<uint64_t>(uint64_t(total_pairs) + total_pairs / 4 + 1024, 0xfffffff0ull);
            if (want <= c->list_cap) return fail(c, RSY_OVERFLOW, "bin lists exceed 2^32 entries");
            HIP_TRY(c, c->lists_mem.ensure(size_t(want) * 4), "grow bin lists");
            c->list_cap = uint32_t(want);
            c->layout.valid = false;  // the frame's addresses are rebound to the grown lists
        }
        rsy_result r = render_now(c, c->last_views, c->last_shade);
        if (r != RSY_OK) return r;
        HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    }
    return fail(c, RSY_OVERFLOW, "frame still overflows after growing its buffers");
}

std::vector<float> make_srgb_lut() {
    std::vector<float> lut(256);
    for (int i = 0; i < 256; ++i) {
        const double v = i / 255.0;
        lut[i] = float(v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4));
    }
    return lut;
}

}  // namespace

extern "C" {

uint32_t rsy_abi_version(void) { return RSY_ABI_VERSION; }

const char* rsy_last_error(const rsy_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

rsy_result rsy_create(const rsy_config* cfg, rsy_ctx** out_ctx) {
    if (!out_ctx) return RSY_INVALID_ARGUMENT;
    *out_ctx = nullptr;
    if (cfg && cfg->abi_version != RSY_ABI_VERSION) return RSY_INVALID_ARGUMENT;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return RSY_GRAPHICS_INIT_FAILURE;
    const int dev = cfg ? cfg->device : 0;
    if (dev < 0 || dev >= n) return RSY_INVALID_ARGUMENT;
    rsy_ctx* c = new (std::nothrow) rsy_ctx();
    if (!c) return RSY_ALLOCATION_FAILURE;
    c->device = dev;
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev_begin) != hipSuccess || hipEventCreate(&c->ev_end) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
        delete c;
        return RSY_GRAPHICS_INIT_FAILURE;
    }
    for (auto& e : c->ev_stage)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return RSY_GRAPHICS_INIT_FAILURE;
        }
    // the allocation-free frame's static scratch: this device's copy, looked up once
    if (static_scratch_init(dev) != hipSuccess) {
        delete c;
        return RSY_GRAPHICS_INIT_FAILURE;
    }
    const std::vector<float> lut = make_srgb_lut();
    if (c->srgb_lut.ensure(256 * 4) != hipSuccess ||
        hipMemcpy(c->srgb_lut.p, lut.data(), 256 * 4, hipMemcpyHostToDevice) != hipSuccess ||
        c->scene_d.ensure(sizeof(rsy_scene_data)) != hipSuccess) {
        delete c;
        return RSY_ALLOCATION_FAILURE;
    }
    c->rs.tile_size = 0;
    c->rs.clear_color[1] = 0.05f;
    c->rs.clear_color[2] = 0.1f;
    c->rs.clear_color[3] = 1.f;
    *out_ctx = c;
    return RSY_OK;
}

void rsy_destroy(rsy_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    invalidate_graph(ctx);
    if (ctx->side_stream) (void)hipStreamSynchronize(ctx->side_stream);
    if (ctx->ev_begin) (void)hipEventDestroy(ctx->ev_begin);
    if (ctx->ev_end) (void)hipEventDestroy(ctx->ev_end);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    for (auto& e : ctx->ev_stage)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->side_stream) (void)hipStreamDestroy(ctx->side_stream);
    delete ctx;
}

rsy_result rsy_set_geometry(rsy_ctx* c, const rsy_position* vertices, uint64_t n_vertices, const uint32_t* indices,
                            uint64_t n_indices, const rsy_mesh_range* meshes, uint32_t n_meshes) {
    if (!c) return RSY_INVALID_ARGUMENT;
    if ((n_vertices && !vertices) || (n_indices && !indices) || (n_meshes && !meshes))
        return fail(c, RSY_INVALID_ARGUMENT, "null geometry pointer");
    if (n_vertices >= 0xffffffffull || n_indices >= 0xffffffffull)
        return fail(c, RSY_OVERFLOW, "geometry exceeds 2^32 vertices or indices");
    for (uint32_t m = 0; m < n_meshes; ++m) {
        const rsy_mesh_range& r = meshes[m];
        if (r.vertex_buffer_offset % sizeof(rsy_position) != 0)
            return fail(c, RSY_INVALID_ARGUMENT, "vertex_buffer_offset must be a multiple of 64");
        const uint64_t vb = r.vertex_buffer_offset / sizeof(rsy_position);
        if (vb + r.num_vertices > n_vertices || uint64_t(r.index_offset) + r.num_indices > n_indices)
            return fail(c, RSY_INVALID_ARGUMENT, "mesh range outside the buffers");
        for (uint64_t i = r.index_offset; i < uint64_t(r.index_offset) + r.num_indices; ++i)
            if (indices[i] >= r.num_vertices) return fail(c, RSY_INVALID_ARGUMENT, "index outside its mesh");
    }
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    const size_t nv = std::max<uint64_t>(1, n_vertices);
    HIP_TRY(c, c->vertices.ensure(nv * sizeof(rsy_position)), "allocate vertices");
    HIP_TRY(c, c->px.ensure(nv * 4), "allocate positions");
    HIP_TRY(c, c->py.ensure(nv * 4), "allocate positions");
    HIP_TRY(c, c->pz.ensure(nv * 4), "allocate positions");
    HIP_TRY(c, c->indices.ensure(std::max<uint64_t>(1, n_indices) * 4), "allocate indices");
    if (n_vertices) {
        HIP_TRY(c, hipMemcpy(c->vertices.p, vertices, n_vertices * sizeof(rsy_position), hipMemcpyHostToDevice), "upload vertices");
        std::vector<float> sx(n_vertices), sy(n_vertices), sz(n_vertices);
        for (uint64_t i = 0; i < n_vertices; ++i) {
            sx[i] = vertices[i].vertex[0];
            sy[i] = vertices[i].vertex[1];
            sz[i] = vertices[i].vertex[2];
        }
        HIP_TRY(c, hipMemcpy(c->px.p, sx.data(), n_vertices * 4, hipMemcpyHostToDevice), "upload positions");
        HIP_TRY(c, hipMemcpy(c->py.p, sy.data(), n_vertices * 4, hipMemcpyHostToDevice), "upload positions");
        HIP_TRY(c, hipMemcpy(c->pz.p, sz.data(), n_vertices * 4, hipMemcpyHostToDevice), "upload positions");
    }
    if (n_indices) HIP_TRY(c, hipMemcpy(c->indices.p, indices, n_indices * 4, hipMemcpyHostToDevice), "upload indices");
    c->n_vertices = n_vertices;
    c->n_indices = n_indices;
    c->meshes.assign(meshes, meshes + n_meshes);
    c->mesh_bounds.assign(size_t(n_meshes) * 6, 0.f);
    for (uint32_t m = 0; m < n_meshes; ++m) {
        const uint64_t vb = meshes[m].vertex_buffer_offset / sizeof(rsy_position);
        float* b = &c->mesh_bounds[size_t(m) * 6];
        for (int k = 0; k < 3; ++k) {
            b[k] = INFINITY;
            b[3 + k] = -INFINITY;
        }
        for (uint64_t v = vb; v < vb + meshes[m].num_vertices; ++v)
            for (int k = 0; k < 3; ++k) {
                b[k] = std::min(b[k], vertices[v].vertex[k]);
                b[3 + k] = std::max(b[3 + k], vertices[v].vertex[k]);
            }
        if (meshes[m].num_vertices == 0)
            for (int k = 0; k < 6; ++k) b[k] = 0.f;  // empty mesh: any box (no triangles to cull)
    }
    c->has_geometry = true;
    c->lists_dirty = true;
    c->layout.valid = false;
    invalidate_graph(c);
    return RSY_OK;
}

rsy_result rsy_set_materials(rsy_ctx* c, const rsy_gpu_material* materials, uint32_t n_materials,
                             const rsy_texture_desc* textures, uint32_t n_textures, const rsy_sampler* samplers,
                             uint32_t n_samplers) {
    if (!c) return RSY_INVALID_ARGUMENT;
    if ((n_materials && !materials) || (n_textures && !textures) || (n_samplers && !samplers))
        return fail(c, RSY_INVALID_ARGUMENT, "null material pointer");
    std::vector<TexGPU> table(n_textures);
    uint64_t pool = 0, staging = 0;
    for (uint32_t t = 0; t < n_textures; ++t) {
        const rsy_texture_desc& d = textures[t];
        if (d.width == 0 || d.height == 0 || d.width > 16384 || d.height > 16384 || !d.pixels || d.format > RSY_TEX_BC7_UNORM)
            return fail(c, RSY_INVALID_ARGUMENT, "invalid texture description");
        const uint32_t full = ilog2(std::max(d.width, d.height)) + 1;  // levels down to 1 x 1
        const bool build = d.mip_levels == RSY_MIPS_BUILD;
        const uint32_t given = build ? 1u : std::max(1u, d.mip_levels);
        if (given > full) return fail(c, RSY_INVALID_ARGUMENT, "more mip levels than the image has");
        const bool bc7 = d.format == RSY_TEX_BC7_SRGB || d.format == RSY_TEX_BC7_UNORM;
        const uint32_t ch = (d.format == RSY_TEX_RG8_UNORM || d.format == RSY_TEX_RG8_SRGB) ? 2 : 4;
        const uint32_t srgb = d.format == RSY_TEX_RGBA8_SRGB || d.format == RSY_TEX_RG8_SRGB || d.format == RSY_TEX_BC7_SRGB;
        table[t] = {pool, d.width, d.height, ch, srgb, build ? full : given, 0u};
        pool += (mip_offset(d.width, d.height, ch, table[t].levels) + 255) & ~uint64_t(255);
        if (bc7)
            for (uint32_t l = 0; l < given; ++l)
                staging += uint64_t((mip_extent(d.width, l) + 3) / 4) * ((mip_extent(d.height, l) + 3) / 4) * 16;
    }
    auto check_img = [&](uint32_t i) { return i == RSY_ABSENT || i < n_textures; };
    for (uint32_t m = 0; m < n_materials; ++m) {
        const rsy_gpu_material& g = materials[m];
        if (!check_img(g.color_sampled_image_index) || !check_img(g.normal_sampled_image_index) ||
            !check_img(g.metallic_sampled_image_index) || !check_img(g.mixmap_sampled_image_index))
            return fail(c, RSY_INVALID_ARGUMENT, "material image index out of range");
    }
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    HIP_TRY(c, c->materials.ensure(std::max<size_t>(1, n_materials) * sizeof(rsy_gpu_material)), "allocate materials");
    HIP_TRY(c, c->tex_table.ensure(std::max<size_t>(1, n_textures) * sizeof(TexGPU)), "allocate textures");
    HIP_TRY(c, c->texels.ensure(std::max<uint64_t>(256, pool)), "allocate texels");
    HIP_TRY(c, c->samplers.ensure(std::max<size_t>(1, n_samplers) * sizeof(rsy_sampler)), "allocate samplers");
    if (n_materials) HIP_TRY(c, hipMemcpy(c->materials.p, materials, n_materials * sizeof(rsy_gpu_material), hipMemcpyHostToDevice), "upload materials");
    if (n_textures) HIP_TRY(c, hipMemcpy(c->tex_table.p, table.data(), n_textures * sizeof(TexGPU), hipMemcpyHostToDevice), "upload textures");
    if (staging) HIP_TRY(c, c->tex_staging.ensure(staging), "allocate texture staging");
    uint64_t staged = 0;  // every BC7 image's blocks go to their own part of the staging buffer
    for (uint32_t t = 0; t < n_textures; ++t) {
        // the given levels (BC7 decoded on the GPU), then the built ones, each from the one above
        const rsy_texture_desc& d = textures[t];
        const TexGPU& g = table[t];
        const bool bc7 = d.format == RSY_TEX_BC7_SRGB || d.format == RSY_TEX_BC7_UNORM;
        const uint32_t given = d.mip_levels == RSY_MIPS_BUILD ? 1u : std::max(1u, d.mip_levels);
        const uint8_t* src = static_cast<const uint8_t*>(d.pixels);
        uint8_t* base = c->texels.as<uint8_t>() + g.offset;
        if (bc7) {
            uint64_t n = 0;
            for (uint32_t l = 0; l < given; ++l)
                n += uint64_t((mip_extent(g.width, l) + 3) / 4) * ((mip_extent(g.height, l) + 3) / 4) * 16;
            HIP_TRY(c, hipMemcpyAsync(c->tex_staging.as<uint8_t>() + staged, src, n, hipMemcpyHostToDevice, c->stream),
                    "upload BC7 blocks");
        }
        for (uint32_t l = 0; l < given; ++l) {
            const uint32_t w = mip_extent(g.width, l), h = mip_extent(g.height, l);
            uint8_t* dst = base + mip_offset(g.width, g.height, g.channels, l);
            if (bc7) {
                const uint64_t n = uint64_t((w + 3) / 4) * ((h + 3) / 4) * 16;
                HIP_TRY(c, launch_bc7_decode(c->tex_staging.as<uint8_t>() + staged, w, h, reinterpret_cast<uint32_t*>(dst),
                                             c->stream), "BC7 decode");
                staged += n;
            } else {
                const uint64_t n = uint64_t(w) * h * g.channels;
                HIP_TRY(c, hipMemcpy(dst, src, n, hipMemcpyHostToDevice), "upload texels");
                src += n;
            }
        }
        for (uint32_t l = given; l < g.levels; ++l)
            HIP_TRY(c, launch_build_mip(base + mip_offset(g.width, g.height, g.channels, l - 1), mip_extent(g.width, l - 1),
                                        mip_extent(g.height, l - 1), base + mip_offset(g.width, g.height, g.channels, l),
                                        g.channels, g.srgb, c->stream), "build mip level");
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    if (n_samplers) HIP_TRY(c, hipMemcpy(c->samplers.p, samplers, n_samplers * sizeof(rsy_sampler), hipMemcpyHostToDevice), "upload samplers");
    c->n_materials = n_materials;
    c->n_textures = n_textures;
    c->n_samplers = n_samplers;
    c->tex_channels.clear();
    for (auto& t : table) c->tex_channels.push_back(t.channels);
    c->h_materials.assign(materials, materials + n_materials);
    c->h_tex_table = table;
    c->h_samplers.assign(samplers, samplers + n_samplers);
    // the draws' sampler flags follow the materials: refreshed in place, so that the allocation-free
    // render path keeps working after a material change
    if (!c->lists_dirty)
        for (int l = 0; l < 3; ++l) {
            for (DrawGPU& d : c->lists[l]) d.flags = draw_flags(c, d.material);
            if (!c->lists[l].empty())
                HIP_TRY(c, hipMemcpy(c->lists_d[l].p, c->lists[l].data(), c->lists[l].size() * sizeof(DrawGPU),
                                     hipMemcpyHostToDevice), "upload draw list");
        }
    c->tex_opaque.assign(n_textures, 0);
    c->any_mips = false;
    std::vector<uint8_t> decoded;
    for (uint32_t t = 0; t < n_textures; ++t) {
        c->any_mips |= table[t].levels > 1;
        if (table[t].channels != 4) continue;
        const uint64_t n = uint64_t(table[t].width) * table[t].height;
        const uint8_t* px = static_cast<const uint8_t*>(textures[t].pixels);
        if (textures[t].format == RSY_TEX_BC7_SRGB || textures[t].format == RSY_TEX_BC7_UNORM) {
            decoded.resize(n * 4);  // level 0 as decoded
            HIP_TRY(c, hipMemcpy(decoded.data(), c->texels.as<uint8_t>() + table[t].offset, n * 4, hipMemcpyDeviceToHost),
                    "read decoded texels");
            px = decoded.data();
        }
        bool opaque = true;
        for (uint64_t i = 0; i < n && opaque; ++i) opaque = px[4 * i + 3] == 255;
        c->tex_opaque[t] = opaque ? 1 : 0;
    }
    // a decoder trained on the previous materials no longer applies
    destroy_neural(c->neural);
    c->neural = nullptr;
    invalidate_graph(c);
    return RSY_OK;
}

rsy_result rsy_set_graphic_objects(rsy_ctx* c, const rsy_object_record* objects, uint32_t n_objects,
                                   const rsy_draw* surfaces, uint32_t n_surfaces) {
    if (!c) return RSY_INVALID_ARGUMENT;
    // Graphics.cpp:4693-4697: an empty object set is an invalid argument
    if (n_objects == 0 || !objects || (n_surfaces && !surfaces))
        return fail(c, RSY_INVALID_ARGUMENT, "no graphic objects");
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    HIP_TRY(c, c->objects.ensure(size_t(n_objects) * sizeof(rsy_object_record)), "allocate objects");
    HIP_TRY(c, hipMemcpy(c->objects.p, objects, size_t(n_objects) * sizeof(rsy_object_record), hipMemcpyHostToDevice), "upload objects");
    c->h_objects.assign(objects, objects + n_objects);
    c->surfaces.assign(surfaces, surfaces + n_surfaces);
    c->has_objects = true;
    c->lists_dirty = true;
    c->layout.valid = false;
    invalidate_graph(c);
    if (c->has_geometry) return rebuild_lists(c);
    return RSY_OK;
}

rsy_result rsy_update_graphic_objects(rsy_ctx* c, uint64_t offset, const rsy_object_record* objects, uint32_t n_objects) {
    if (!c) return RSY_INVALID_ARGUMENT;
    if (!c->has_objects) return fail(c, RSY_INVALID_STATE, "rsy_set_graphic_objects has not been called");
    if (n_objects && !objects) return fail(c, RSY_INVALID_ARGUMENT, "null objects");
    if (offset + n_objects > c->h_objects.size()) return fail(c, RSY_INVALID_ARGUMENT, "object update out of range");
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    std::copy(objects, objects + n_objects, c->h_objects.begin() + offset);
    if (n_objects)
        HIP_TRY(c, hipMemcpy(c->objects.as<uint8_t>() + offset * sizeof(rsy_object_record), objects,
                             size_t(n_objects) * sizeof(rsy_object_record), hipMemcpyHostToDevice), "upload objects");
    return RSY_OK;
}

rsy_result rsy_update(rsy_ctx* c, const rsy_scene_data* sd, const rsy_raster_state* state) {
    if (!c || !sd) return RSY_INVALID_ARGUMENT;
    if (state) {
        if (state->tile_size != 0 && state->tile_size != 8 && state->tile_size != 16 && state->tile_size != 32 &&
            state->tile_size != 64)
            return fail(c, RSY_INVALID_ARGUMENT, "tile_size must be 0, 8, 16, 32 or 64");
        if (state->pcf_mode > RSY_PCF_GRID_3X3) return fail(c, RSY_INVALID_ARGUMENT, "unknown pcf mode");
        if (state->msaa_samples > 1 && state->msaa_samples != 4)
            return fail(c, RSY_INVALID_ARGUMENT, "msaa_samples must be 0, 1 or 4");
        if (state->msaa_samples == 4 && state->tile_size == 64)
            return fail(c, RSY_INVALID_ARGUMENT, "4-sample frames use 32-pixel bins");
        const bool raster_changed = state->cull_enabled != c->rs.cull_enabled ||
                                    state->reverse_winding_order_enabled != c->rs.reverse_winding_order_enabled ||
                                    state->depth_bias_enabled != c->rs.depth_bias_enabled ||
                                    state->depth_bias_constant != c->rs.depth_bias_constant ||
                                    state->depth_bias_clamp != c->rs.depth_bias_clamp ||
                                    state->depth_bias_slope_factor != c->rs.depth_bias_slope_factor ||
                                    state->tile_size != c->rs.tile_size || state->pcf_mode != c->rs.pcf_mode ||
                                    (state->msaa_samples == 4) != (c->rs.msaa_samples == 4) ||
                                    std::memcmp(state->clear_color, c->rs.clear_color, sizeof(c->rs.clear_color)) != 0;
        c->rs = *state;
        if (raster_changed) c->layout.valid = false;
    }
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    c->sd = *sd;
    HIP_TRY(c, hipMemcpy(c->scene_d.p, sd, sizeof(rsy_scene_data), hipMemcpyHostToDevice), "upload scene data");
    c->has_scene = true;
    return RSY_OK;
}

rsy_result rsy_render(rsy_ctx* c, const rsy_frame_desc* frame) {
    if (!c) return RSY_INVALID_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    rsy_result r = ready_to_render(c);
    if (r != RSY_OK) return r;
    if (c->frame_pending) {
        r = finish_frame(c);
        if (r != RSY_OK) return r;
    }
    std::vector<FrameViewDesc> views;
    uint32_t shade = 0;
    r = resolve_views(c, frame, views, shade);
    if (r != RSY_OK) return r;
    return render_now(c, views, shade);
}

rsy_result rsy_synchronize(rsy_ctx* c) {
    if (!c) return RSY_INVALID_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    return finish_frame(c);
}

rsy_result rsy_render_graph(rsy_ctx* c, const rsy_frame_desc* frame, uint32_t n_replays) {
    if (!c) return RSY_INVALID_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    // one eager frame sizes every buffer (and re-runs on overflow)
    rsy_result r = rsy_render(c, frame);
    if (r != RSY_OK) return r;
    r = finish_frame(c);
    if (r != RSY_OK) return r;
    if (!c->graph_valid) {
        invalidate_graph(c);
        hipGraph_t graph = nullptr;
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
        r = enqueue_frame(c, c->layout, c->mem, true, true);
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (r != RSY_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return r;
        }
        if (e != hipSuccess) return hip_fail(c, e, "end capture");
        e = hipGraphInstantiate(&c->graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return hip_fail(c, e, "instantiate graph");
        c->graph_valid = true;
    }
    HIP_TRY(c, hipEventRecord(c->ev_begin, c->stream), "event");
    for (uint32_t i = 0; i < n_replays; ++i) HIP_TRY(c, hipGraphLaunch(c->graph_exec, c->stream), "graph launch");
    const uint32_t nv = uint32_t(c->layout.vg.size());
    HIP_TRY(c, hipMemcpyAsync(c->readback_counters.p, c->counters.p, nv * kCtrPerView * 4, hipMemcpyDeviceToHost, c->stream), "readback");
    HIP_TRY(c, hipMemcpyAsync(static_cast<uint8_t*>(c->readback_counters.p) + nv * kCtrPerView * 4, c->totals.p, 16,
                              hipMemcpyDeviceToHost, c->stream), "readback");
    HIP_TRY(c, hipEventRecord(c->ev_end, c->stream), "event");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    const bool over = collect_stats(c);
    if (n_replays) c->stats.gpu_ms /= float(n_replays);
    if (over) return fail(c, RSY_OVERFLOW, "graph replay overflowed its buffers");
    return RSY_OK;
}

// The allocation-free frame (no counterpart in the reference): rsy_render's kernels with their
// working memory carved from the static scratch array and the caller's output buffers.
rsy_result rsy_render_noalloc(rsy_ctx* c, const rsy_frame_desc* frame, const rsy_frame_targets* t) {
    if (!c) return RSY_INVALID_ARGUMENT;
    if (!t) return fail(c, RSY_INVALID_ARGUMENT, "targets == NULL");
    if (!c->has_geometry) return fail(c, RSY_INVALID_STATE, "rsy_set_geometry has not been called");
    if (!c->has_objects) return fail(c, RSY_INVALID_STATE, "rsy_set_graphic_objects has not been called");
    if (!c->has_scene) return fail(c, RSY_INVALID_STATE, "rsy_update has not been called");
    if (c->lists_dirty)
        return fail(c, RSY_INVALID_STATE, "draw lists not built: call rsy_set_graphic_objects after rsy_set_geometry");
    if (msaa_of(c) != 1) return fail(c, RSY_INVALID_STATE, "the allocation-free frame is single-sampled (msaa_samples 1)");
    std::vector<FrameViewDesc>& views = c->static_views;
    uint32_t shade = 0;
    rsy_result r = resolve_views(c, frame, views, shade);
    if (r != RSY_OK) return r;
    if (shade != RSY_SHADE_NONE && !c->lists[2].empty())
        return fail(c, RSY_INVALID_STATE, "the allocation-free frame has no blended pass (the blended list is not empty)");
    uint32_t n_cam = 0, n_casc = 0;
    for (const auto& v : views) (v.shadow ? n_casc : n_cam)++;
    if (t->n_camera_views != n_cam || t->n_cascades != n_casc)
        return fail(c, RSY_INVALID_ARGUMENT, "targets do not match the frame's camera views and cascades");
    if (!t->visibility || !t->shadow_maps || (shade != RSY_SHADE_NONE && !t->color))
        return fail(c, RSY_INVALID_ARGUMENT, "null target array");
    for (uint32_t i = 0; i < n_cam; ++i)
        if (!t->visibility[i] || (shade != RSY_SHADE_NONE && !t->color[i]))
            return fail(c, RSY_INVALID_ARGUMENT, "null camera view target");
    for (uint32_t i = 0; i < n_casc; ++i)
        if (!t->shadow_maps[i]) return fail(c, RSY_INVALID_ARGUMENT, "null shadow map target");
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    uint8_t* const base = static_cast<uint8_t*>(static_scratch(c->device));
    if (!base) return fail(c, RSY_INVALID_STATE, "static scratch of this device not looked up");

    Layout& L = c->static_layout;
    r = plan_layout(c, views, choose_tile(c), shade, 1u, L);
    if (r != RSY_OK) return r;
    // carve the working memory, 256-byte aligned pieces in a fixed order
    const size_t nv = L.vg.size();
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t o = at;
        at += (std::max<size_t>(bytes, 4) + 255) & ~size_t(255);
        return base + o;
    };
    FrameMem m{};
    m.views = reinterpret_cast<ViewGPU*>(take(nv * sizeof(ViewGPU)));
    m.counters = reinterpret_cast<uint32_t*>(take(nv * kCtrPerView * 4));
    m.totals = reinterpret_cast<uint32_t*>(take(32));
    m.shade_views = reinterpret_cast<ShadeView*>(take(L.n_camera * sizeof(ShadeView)));
    m.shade_cascades = reinterpret_cast<ShadeCascade*>(take(L.n_shadow * sizeof(ShadeCascade)));
    m.vblocks = reinterpret_cast<BlockEntry*>(take(L.vblocks.size() * sizeof(BlockEntry)));
    m.tblocks = reinterpret_cast<BlockEntry*>(take(L.tblocks.size() * sizeof(BlockEntry)));
    m.block_valid = reinterpret_cast<uint32_t*>(take(L.tblocks.size() * 4));
    m.draw_culled = reinterpret_cast<uint32_t*>(take(size_t(L.cull_total) * 4));
    m.bin_counts = reinterpret_cast<uint32_t*>(take(size_t(L.total_bins) * 4));
    m.bin_offsets = reinterpret_cast<uint32_t*>(take(size_t(L.total_bins) * 4));
    m.scan_partial = take((size_t(L.total_bins) / kScanBlockBins + 1) * kScanPartialBytes);
    m.xv = reinterpret_cast<float4*>(take(size_t(L.total_xv) * sizeof(float4)));
    m.recs = reinterpret_cast<TriRec*>(take(size_t(L.rec_total) * sizeof(TriRec)));
    m.bin_tiles = reinterpret_cast<uint32_t*>(take(size_t(L.rec_total) * 4));
    m.shade_recs = reinterpret_cast<ShadeRec*>(take(size_t(L.rec_total) * sizeof(ShadeRec)));
    m.pair_slots = reinterpret_cast<uint32_t*>(take(size_t(L.rec_total) * 16));
    m.clipq = reinterpret_cast<ClipEntry*>(take(size_t(L.clip_total) * sizeof(ClipEntry)));
    m.big_pairs = reinterpret_cast<uint4*>(take(size_t(L.big_total) * 16));
    // the bin lists and work items take the rest: at least the frame's estimated pairs
    const auto need_for = [&](uint64_t pairs) {  // with both pieces' alignment slack
        return uint64_t(at) + pairs * 4 + uint64_t(raster_items_cap(L.total_bins, uint32_t(pairs))) * sizeof(uint4) + 512;
    };
    if (need_for(L.want_pairs) > kStaticScratchBytes) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "the frame needs %.1f MiB of working memory, more than the %zu MiB static scratch",
                      double(need_for(L.want_pairs)) / double(1 << 20), kStaticScratchBytes >> 20);
        return fail(c, RSY_OVERFLOW, msg);
    }
    // the largest capacity that fits (need_for grows with it; the estimate fits, as checked)
    uint64_t lo = L.want_pairs, hi = std::min<uint64_t>((kStaticScratchBytes - at) / 4, 0xfffffff0ull);
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        if (need_for(mid) <= kStaticScratchBytes) lo = mid;
        else hi = mid - 1;
    }
    m.list_cap = uint32_t(lo);
    m.lists = reinterpret_cast<uint32_t*>(take(size_t(m.list_cap) * 4));
    m.items = reinterpret_cast<uint4*>(take(size_t(raster_items_cap(L.total_bins, m.list_cap)) * sizeof(uint4)));
    if (at > kStaticScratchBytes) return fail(c, RSY_OVERFLOW, "static scratch layout exceeds its size");
    // the caller's outputs
    for (uint32_t i = 0; i < n_cam; ++i) {
        m.vis[i] = t->visibility[i];
        m.color[i] = shade != RSY_SHADE_NONE ? reinterpret_cast<float4*>(t->color[i]) : nullptr;
    }
    for (uint32_t i = 0; i < n_casc; ++i) m.shadow_map[i] = t->shadow_maps[i];
    bind_outputs(L, m);
    // a frame from rsy_render may still be in flight on the stream: the tables are written behind it
    if (!L.vblocks.empty())
        HIP_TRY(c, hipMemcpyAsync(m.vblocks, L.vblocks.data(), L.vblocks.size() * sizeof(BlockEntry), hipMemcpyHostToDevice,
                                  c->stream), "upload blocks");
    if (!L.tblocks.empty())
        HIP_TRY(c, hipMemcpyAsync(m.tblocks, L.tblocks.data(), L.tblocks.size() * sizeof(BlockEntry), hipMemcpyHostToDevice,
                                  c->stream), "upload blocks");
    r = upload_frame(c, L, m, nullptr, 0u);
    if (r != RSY_OK) return r;
    // at most two passes: the second when the frame had more raster work items than the estimated
    // grid (nothing is allocated for it)
    for (int pass = 0; pass < 2; ++pass) {
        r = enqueue_frame(c, L, m, false, false);
        if (r != RSY_OK) return r;
        uint32_t ctr[kMaxViews * kCtrPerView + 8] = {};
        HIP_TRY(c, hipMemcpyAsync(ctr, m.counters, nv * kCtrPerView * 4, hipMemcpyDeviceToHost, c->stream), "readback");
        HIP_TRY(c, hipMemcpyAsync(ctr + nv * kCtrPerView, m.totals, 16, hipMemcpyDeviceToHost, c->stream), "readback");
        HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
        const uint32_t bits = ctr[nv * kCtrPerView + 2];
        bool over = (bits & 1u) != 0;
        for (size_t v = 0; v < nv; ++v) over |= ctr[v * kCtrPerView + kCtrOverflow] != 0;
        if (over) return fail(c, RSY_OVERFLOW, "the frame overflowed the working memory the static scratch holds");
        if (!(bits & 2u)) return RSY_OK;
        L.raster_grid = raster_grid_for(ctr[nv * kCtrPerView + 3]);
    }
    return fail(c, RSY_OVERFLOW, "the frame's raster work items outnumber the workgroups");
}

static rsy_result camera_index(rsy_ctx* c, uint32_t view, uint32_t& idx) {
    if (!c->layout.valid) return fail(c, RSY_INVALID_STATE, "no frame rendered");
    if (view >= c->layout.n_camera) return fail(c, RSY_INVALID_ARGUMENT, "camera view index out of range");
    idx = view;
    return RSY_OK;
}

static const ViewGPU* camera_vg(rsy_ctx* c, uint32_t view) {
    uint32_t k = 0;
    for (const auto& g : c->layout.vg)
        if (!g.is_shadow && k++ == view) return &g;
    return nullptr;
}

rsy_result rsy_read_visibility(rsy_ctx* c, uint32_t view, uint64_t* out, uint64_t n_pixels) {
    if (!c || !out) return RSY_INVALID_ARGUMENT;
    rsy_result r = rsy_synchronize(c);
    if (r != RSY_OK) return r;
    uint32_t idx;
    if ((r = camera_index(c, view, idx)) != RSY_OK) return r;
    const ViewGPU* g = camera_vg(c, view);
    if (n_pixels != uint64_t(g->width) * g->height * g->samples)
        return fail(c, RSY_INVALID_ARGUMENT, "pixel count mismatch (width * height * samples)");
    HIP_TRY(c, hipMemcpy(out, c->vis_out[idx].p, n_pixels * 8, hipMemcpyDeviceToHost), "read visibility");
    return RSY_OK;
}

rsy_result rsy_read_color(rsy_ctx* c, uint32_t view, float* out, uint64_t n_pixels) {
    if (!c || !out) return RSY_INVALID_ARGUMENT;
    rsy_result r = rsy_synchronize(c);
    if (r != RSY_OK) return r;
    uint32_t idx;
    if ((r = camera_index(c, view, idx)) != RSY_OK) return r;
    if (!c->layout.shade) return fail(c, RSY_INVALID_STATE, "last frame was not shaded");
    const ViewGPU* g = camera_vg(c, view);
    if (n_pixels != uint64_t(g->width) * g->height) return fail(c, RSY_INVALID_ARGUMENT, "pixel count mismatch");
    HIP_TRY(c, hipMemcpy(out, c->color_out[idx].p, n_pixels * 16, hipMemcpyDeviceToHost), "read colour");
    return RSY_OK;
}

rsy_result rsy_present(rsy_ctx* c, uint32_t view, uint32_t width, uint32_t height) {
    if (!c) return RSY_INVALID_ARGUMENT;
    rsy_result r = rsy_synchronize(c);
    if (r != RSY_OK) return r;
    uint32_t idx;
    if ((r = camera_index(c, view, idx)) != RSY_OK) return r;
    if (!c->layout.shade) return fail(c, RSY_INVALID_STATE, "last frame was not shaded");
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return fail(c, RSY_INVALID_ARGUMENT, "present extent must be 1..16384 on each side");
    const ViewGPU* g = camera_vg(c, view);
    HIP_TRY(c, c->draw_image.ensure(size_t(g->width) * g->height * 8), "allocate draw image");
    HIP_TRY(c, c->present_image.ensure(size_t(width) * height * 4), "allocate present image");
    const bool timing = (c->debug_flags & RSY_DEBUG_STAGE_TIMING) != 0;
    if (timing) HIP_TRY(c, hipEventRecord(c->ev_stage[6], c->stream), "event");
    HIP_TRY(c, launch_present(c->color_out[idx].as<float4>(), c->draw_image.as<uint2>(), g->width, g->height,
                              c->present_image.as<uint32_t>(), width, height, c->stream), "present kernels");
    if (timing) HIP_TRY(c, hipEventRecord(c->ev_stage[7], c->stream), "event");
    c->present_timed = timing;
    c->draw_w = g->width;
    c->draw_h = g->height;
    c->present_w = width;
    c->present_h = height;
    return RSY_OK;
}

rsy_result rsy_read_draw_image(rsy_ctx* c, uint16_t* out, uint64_t n_pixels) {
    if (!c || !out) return RSY_INVALID_ARGUMENT;
    if (c->draw_w == 0) return fail(c, RSY_INVALID_STATE, "nothing presented");
    if (n_pixels != uint64_t(c->draw_w) * c->draw_h) return fail(c, RSY_INVALID_ARGUMENT, "pixel count mismatch");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    HIP_TRY(c, hipMemcpy(out, c->draw_image.p, n_pixels * 8, hipMemcpyDeviceToHost), "read draw image");
    return RSY_OK;
}

rsy_result rsy_read_present(rsy_ctx* c, uint8_t* out, uint64_t n_pixels) {
    if (!c || !out) return RSY_INVALID_ARGUMENT;
    if (c->present_w == 0) return fail(c, RSY_INVALID_STATE, "nothing presented");
    if (n_pixels != uint64_t(c->present_w) * c->present_h) return fail(c, RSY_INVALID_ARGUMENT, "pixel count mismatch");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    HIP_TRY(c, hipMemcpy(out, c->present_image.p, n_pixels * 4, hipMemcpyDeviceToHost), "read present image");
    return RSY_OK;
}

rsy_result rsy_read_shadow_factor(rsy_ctx* c, uint32_t view, float* out, uint64_t n_pixels) {
    if (!c || !out) return RSY_INVALID_ARGUMENT;
    rsy_result r = rsy_synchronize(c);
    if (r != RSY_OK) return r;
    uint32_t idx;
    if ((r = camera_index(c, view, idx)) != RSY_OK) return r;
    if (!(c->debug_flags & RSY_DEBUG_SHADOW_FACTOR)) return fail(c, RSY_INVALID_STATE, "shadow factor output disabled");
    const ViewGPU* g = camera_vg(c, view);
    if (n_pixels != uint64_t(g->width) * g->height) return fail(c, RSY_INVALID_ARGUMENT, "pixel count mismatch");
    HIP_TRY(c, hipMemcpy(out, c->shadow_factor_out[idx].p, n_pixels * 4, hipMemcpyDeviceToHost), "read shadow factor");
    return RSY_OK;
}

rsy_result rsy_read_shadow_map(rsy_ctx* c, uint32_t cascade, float* out, uint64_t n_texels) {
    if (!c || !out) return RSY_INVALID_ARGUMENT;
    rsy_result r = rsy_synchronize(c);
    if (r != RSY_OK) return r;
    if (!c->layout.valid || cascade >= c->layout.n_shadow) return fail(c, RSY_INVALID_ARGUMENT, "cascade index out of range");
    const ViewGPU& g = c->layout.vg[cascade];  // shadow views come first
    if (n_texels != uint64_t(g.width) * g.height) return fail(c, RSY_INVALID_ARGUMENT, "texel count mismatch");
    HIP_TRY(c, hipMemcpy(out, c->shadow_maps[cascade].p, n_texels * 4, hipMemcpyDeviceToHost), "read shadow map");
    return RSY_OK;
}

rsy_result rsy_read_texture(rsy_ctx* c, uint32_t texture, uint32_t level, uint8_t* out, uint64_t n_bytes) {
    if (!c || !out) return RSY_INVALID_ARGUMENT;
    if (texture >= c->n_textures) return fail(c, RSY_INVALID_ARGUMENT, "texture index out of range");
    const TexGPU& g = c->h_tex_table[texture];
    if (level >= g.levels) return fail(c, RSY_INVALID_ARGUMENT, "mip level out of range");
    const uint64_t n = uint64_t(mip_extent(g.width, level)) * mip_extent(g.height, level) * g.channels;
    if (n_bytes != n) return fail(c, RSY_INVALID_ARGUMENT, "byte count mismatch");
    HIP_TRY(c, hipSetDevice(c->device), "set device");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "synchronize");
    HIP_TRY(c, hipMemcpy(out, c->texels.as<uint8_t>() + g.offset + mip_offset(g.width, g.height, g.channels, level), n,
                         hipMemcpyDeviceToHost), "read texture");
    return RSY_OK;
}

rsy_result rsy_get_frame_stats(rsy_ctx* c, rsy_frame_stats* out) {
    if (!c || !out) return RSY_INVALID_ARGUMENT;
    rsy_result r = rsy_synchronize(c);
    if (r != RSY_OK) return r;
    if (c->present_timed) {
        float t = 0.f;
        HIP_TRY(c, hipEventSynchronize(c->ev_stage[7]), "synchronize");
        if (hipEventElapsedTime(&t, c->ev_stage[6], c->ev_stage[7]) == hipSuccess) c->stats.stage_ms[5] = t;
    }
    *out = c->stats;
    return RSY_OK;
}

rsy_result rsy_get_device_outputs(rsy_ctx* c, uint32_t view, const void** vis, const void** color) {
    if (!c) return RSY_INVALID_ARGUMENT;
    uint32_t idx;
    rsy_result r = camera_index(c, view, idx);
    if (r != RSY_OK) return r;
    if (vis) *vis = c->vis_out[idx].p;
    if (color) *color = c->color_out[idx].p;
    return RSY_OK;
}

rsy_result rsy_set_debug_flags(rsy_ctx* c, uint32_t flags) {
    if (!c) return RSY_INVALID_ARGUMENT;
    if (flags != c->debug_flags) invalidate_graph(c);
    c->debug_flags = flags;
    return RSY_OK;
}

rsy_result rsy_autotune(rsy_ctx* c, const rsy_frame_desc* frame, uint32_t n_iters, uint32_t* chosen_tile) {
    if (!c) return RSY_INVALID_ARGUMENT;
    const uint32_t saved = c->rs.tile_size;
    float best_ms = 1e30f;
    uint32_t best = 32;
    for (uint32_t t : {8u, 16u, 32u, 64u}) {
        if (t != 32 && msaa_of(c) == 4) continue;  // 4-sample frames use 32-pixel bins
        if (t == 8 && !c->lists[2].empty() && (!frame || frame->shade_mode != RSY_SHADE_NONE))
            continue;  // the blended pass needs bins of 16 pixels or more
        c->rs.tile_size = t;
        c->layout.valid = false;
        rsy_result r = rsy_render_graph(c, frame, std::max<uint32_t>(1, n_iters));
        if (r != RSY_OK) {
            c->rs.tile_size = saved;
            return r;
        }
        if (c->stats.gpu_ms < best_ms) {
            best_ms = c->stats.gpu_ms;
            best = t;
        }
    }
    c->rs.tile_size = saved;
    c->autotuned_tile = best;
    c->layout.valid = false;
    invalidate_graph(c);
    if (chosen_tile) *chosen_tile = best;
    return RSY_OK;
}

}  // extern "C"
