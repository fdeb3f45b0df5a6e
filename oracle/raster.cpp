// CPU oracle, raster half.  TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Restates the shadow and main raster passes of the reference (shadow.slang:6-20,
// basic.slang:100-116, fixed-function state Graphics.cpp:4988-5045, 5195-5292, 5553-5653) with the
// fixed-point rules listed in oracle.h.  Built with -ffp-contract=off so that every FP32 operation
// rounds exactly where the source says.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "fixed_point.h"
#include "oracle.h"

namespace {

using orc::C4;
using orc::plane_dist;
using orc::transform;
using orc::mat_mul_fma;

struct Rec {
    int32_t X[3], Y[3];
    float z0, dzdx, dzdy;
    uint32_t id;
    int32_t px0, px1, py0, py1;
};

uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

struct Params {
    uint32_t width, height;
    bool shadow, cull, reverse, bias;
    float bias_const, bias_clamp, bias_slope;
    int32_t sx = 128, sy = 128;  // the sample point inside a pixel, in 1/256 pixel (centre: 128, 128)
};

// Snap, cull and set up one (sub-)triangle; returns false when it produces no fragments.
bool setup_tri(const C4 v[3], uint32_t id, const Params& p, Rec& r) {
    int32_t X[3], Y[3];
    float Z[3];
    for (int k = 0; k < 3; ++k) orc::snap(v[k], p.width, p.height, X[k], Y[k], Z[k]);
    const int64_t area2 = int64_t(X[1] - X[0]) * (Y[2] - Y[0]) - int64_t(X[2] - X[0]) * (Y[1] - Y[0]);
    if (area2 == 0) return false;
    if (p.cull) {
        const bool front = p.reverse ? (area2 < 0) : (area2 > 0);
        if (front) return false;  // cull mode FRONT
    }
    if (area2 > 0) {
        std::swap(X[1], X[2]);
        std::swap(Y[1], Y[2]);
        std::swap(Z[1], Z[2]);
    }
    const int32_t minx = std::min({X[0], X[1], X[2]}), maxx = std::max({X[0], X[1], X[2]});
    const int32_t miny = std::min({Y[0], Y[1], Y[2]}), maxy = std::max({Y[0], Y[1], Y[2]});
    // pixels whose sample point lies inside the vertices' box
    r.px0 = std::max<int32_t>(0, (minx - p.sx + 255) >> 8);
    r.px1 = std::min<int32_t>(int32_t(p.width) - 1, (maxx - p.sx) >> 8);
    r.py0 = std::max<int32_t>(0, (miny - p.sy + 255) >> 8);
    r.py1 = std::min<int32_t>(int32_t(p.height) - 1, (maxy - p.sy) >> 8);
    if (r.px0 > r.px1 || r.py0 > r.py1) return false;
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
    if (p.shadow && p.bias) {
        const float m = std::fmax(std::fabs(dzdx), std::fabs(dzdy)) * 256.0f;
        const float zmax = std::fmax(std::fabs(Z[0]), std::fmax(std::fabs(Z[1]), std::fabs(Z[2])));
        const int eb = int((fbits(zmax) >> 23) & 0xffu);
        const int e = eb == 0 ? -126 : eb - 127;
        const float rr = std::ldexp(1.0f, e - 23);
        const float ms = m * p.bias_slope, rc = rr * p.bias_const;
        float o = ms + rc;
        if (p.bias_clamp > 0.f) o = std::fmin(o, p.bias_clamp);
        else if (p.bias_clamp < 0.f) o = std::fmax(o, p.bias_clamp);
        z0 = z0 + o;
    }
    for (int k = 0; k < 3; ++k) {
        r.X[k] = X[k];
        r.Y[k] = Y[k];
    }
    r.z0 = z0;
    r.dzdx = dzdx;
    r.dzdy = dzdy;
    r.id = id;
    return true;
}

// A clipped polygon keeps at most kClipMax vertices: the exact bound for a triangle cut by six
// planes.  In FP32 a distance that rounds to the other side near a corner can add a crossing; a
// vertex past the bound is dropped, as the GPU clipper does (the same polygon on both sides).
constexpr int kClipMax = 9;

// Clip one triangle (some vertex outside some plane) and set up the fan.
void clip_and_setup(const C4 in[3], uint32_t id, const Params& p, std::vector<Rec>& out) {
    C4 a[kClipMax], b[kClipMax];
    int n = 3;
    for (int k = 0; k < 3; ++k) a[k] = in[k];
    for (int plane = 0; plane < 6; ++plane) {
        int m = 0;
        for (int i = 0; i < n; ++i) {
            const C4& cur = a[i];
            const C4& nxt = a[(i + 1) % n];
            const float dc = plane_dist(cur, plane), dn = plane_dist(nxt, plane);
            const bool ic = dc >= 0.f, in_ = dn >= 0.f;
            if (ic && m < kClipMax) b[m++] = cur;
            if (ic != in_ && m < kClipMax) {
                const C4& pi = ic ? cur : nxt;
                const C4& po = ic ? nxt : cur;
                const float di = ic ? dc : dn, dout = ic ? dn : dc;
                const float t = di / (di - dout);
                C4 q;
                q.x = pi.x + t * (po.x - pi.x);
                q.y = pi.y + t * (po.y - pi.y);
                q.z = pi.z + t * (po.z - pi.z);
                q.w = pi.w + t * (po.w - pi.w);
                b[m++] = q;
            }
        }
        n = m;
        for (int i = 0; i < n; ++i) a[i] = b[i];
        if (n < 3) return;
    }
    for (int i = 0; i < n; ++i)
        if (!(a[i].w > 0.f)) return;
    for (int k = 1; k + 1 < n; ++k) {
        const C4 t[3] = {a[0], a[k], a[k + 1]};
        Rec r;
        if (setup_tri(t, id, p, r)) out.push_back(r);
    }
}

struct DrawOrc {
    uint32_t first_index, tri_count, tri_base;
    uint64_t vtx_base;
    uint32_t object;
};

int build_records(const orc_geometry* g, const float* vp, const Params& p, std::vector<Rec>& recs,
                  orc_raster_stats* stats, int n_threads) {
    std::vector<DrawOrc> draws(g->n_draws);
    uint32_t tri_base = 0;
    for (uint32_t d = 0; d < g->n_draws; ++d) {
        const rsy_draw& s = g->draws[d];
        if (s.mesh_index >= g->n_meshes) return RSY_INVALID_ARGUMENT;
        const rsy_mesh_range& m = g->meshes[s.mesh_index];
        const uint64_t obj = s.graphic_objects_offset + s.graphics_object_index;
        if (obj >= g->n_objects) return RSY_INVALID_ARGUMENT;
        draws[d] = {m.index_offset + s.start_index, s.index_count / 3, tri_base,
                    m.vertex_buffer_offset / sizeof(rsy_position), uint32_t(obj)};
        tri_base += s.index_count / 3;
    }
    const uint32_t total = tri_base;
    for (const DrawOrc& dr : draws) {
        const rsy_mesh_range& m = g->meshes[g->draws[&dr - draws.data()].mesh_index];
        if (uint64_t(dr.first_index) + 3ull * dr.tri_count > g->n_indices) return RSY_INVALID_ARGUMENT;
        for (uint64_t i = dr.first_index; i < uint64_t(dr.first_index) + 3ull * dr.tri_count; ++i)
            if (g->indices[i] >= m.num_vertices || dr.vtx_base + g->indices[i] >= g->n_vertices)
                return RSY_INVALID_ARGUMENT;
    }
    const int nt = std::max(1, n_threads);
    std::vector<std::vector<Rec>> parts(nt);
    std::vector<uint64_t> clipped(nt, 0);
    auto work = [&](int t) {
        const uint32_t lo = uint32_t(uint64_t(total) * t / nt), hi = uint32_t(uint64_t(total) * (t + 1) / nt);
        uint32_t d = 0;
        float mvp[16];
        uint32_t cur_draw = UINT32_MAX;
        for (uint32_t tri = lo; tri < hi; ++tri) {
            while (d + 1 < draws.size() && draws[d + 1].tri_base <= tri) ++d;
            if (d != cur_draw) {
                mat_mul_fma(vp, g->objects[draws[d].object].transform, mvp);
                cur_draw = d;
            }
            const DrawOrc& dr = draws[d];
            const uint32_t local = tri - dr.tri_base;
            C4 v[3];
            for (int k = 0; k < 3; ++k) {
                const uint32_t vi = g->indices[dr.first_index + 3 * local + k];
                v[k] = transform(mvp, g->vertices[dr.vtx_base + vi].vertex);
            }
            bool reject = false;
            for (int plane = 0; plane < 6; ++plane) {
                int outside = 0;
                for (int k = 0; k < 3; ++k) outside += plane_dist(v[k], plane) < 0.f;
                if (outside == 3) reject = true;
            }
            if (reject) continue;
            if (orc::all_inside(v)) {
                Rec r;
                if (setup_tri(v, tri, p, r)) parts[t].push_back(r);
            } else {
                clipped[t]++;
                clip_and_setup(v, tri, p, parts[t]);
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
    size_t n = 0;
    for (auto& v : parts) n += v.size();
    recs.reserve(n);
    for (auto& v : parts) recs.insert(recs.end(), v.begin(), v.end());
    if (stats) {
        stats->triangles_in = total;
        stats->triangles_setup = recs.size();
        stats->triangles_clipped = 0;
        for (auto c : clipped) stats->triangles_clipped += c;
    }
    return RSY_OK;
}

Params make_params(uint32_t w, uint32_t h, int is_shadow, const rsy_raster_state* rs) {
    Params p{};
    p.width = w;
    p.height = h;
    p.shadow = is_shadow != 0;
    p.cull = rs->cull_enabled != 0;
    p.reverse = rs->reverse_winding_order_enabled != 0;
    p.bias = rs->depth_bias_enabled != 0;
    p.bias_const = rs->depth_bias_constant;
    p.bias_clamp = rs->depth_bias_clamp;
    p.bias_slope = rs->depth_bias_slope_factor;
    return p;
}

}  // namespace

extern "C" {

int orc_raster_view(const orc_geometry* g, const float view_projection[16], uint32_t width, uint32_t height,
                    int is_shadow, const rsy_raster_state* rs, uint64_t* out_vis, float* out_depth,
                    orc_raster_stats* stats, int n_threads) {
    return orc_raster_view_at(g, view_projection, width, height, is_shadow, rs, 128, 128, out_vis, out_depth, stats,
                              n_threads);
}

int orc_raster_view_at(const orc_geometry* g, const float view_projection[16], uint32_t width, uint32_t height,
                       int is_shadow, const rsy_raster_state* rs, int32_t sample_x, int32_t sample_y,
                       uint64_t* out_vis, float* out_depth, orc_raster_stats* stats, int n_threads) {
    if (!g || !rs || width == 0 || height == 0) return RSY_INVALID_ARGUMENT;
    if (is_shadow ? !out_depth : !out_vis) return RSY_INVALID_ARGUMENT;
    if (sample_x < 0 || sample_x > 255 || sample_y < 0 || sample_y > 255) return RSY_INVALID_ARGUMENT;
    Params p = make_params(width, height, is_shadow, rs);
    p.sx = sample_x;
    p.sy = sample_y;
    std::vector<Rec> recs;
    int rc = build_records(g, view_projection, p, recs, stats, n_threads);
    if (rc != RSY_OK) return rc;
    const int nt = std::max(1, n_threads);
    std::vector<uint32_t> depth_bits;
    if (is_shadow) depth_bits.assign(size_t(width) * height, 0u);
    else std::memset(out_vis, 0, sizeof(uint64_t) * size_t(width) * height);
    std::vector<uint64_t> frags(nt, 0);
    auto work = [&](int t) {
        const int32_t y0 = int32_t(uint64_t(height) * t / nt), y1 = int32_t(uint64_t(height) * (t + 1) / nt);
        for (const Rec& r : recs) {
            const int32_t ya = std::max(r.py0, y0), yb = std::min(r.py1, y1 - 1);
            if (ya > yb) continue;
            int32_t A[3], B[3];
            bool tl[3];
            for (int i = 0; i < 3; ++i) {
                const int j = (i + 1) % 3;
                A[i] = r.Y[j] - r.Y[i];
                B[i] = r.X[i] - r.X[j];
                tl[i] = A[i] > 0 || (A[i] == 0 && B[i] > 0);
            }
            for (int32_t py = ya; py <= yb; ++py) {
                const int32_t Py = py * 256 + p.sy;
                for (int32_t px = r.px0; px <= r.px1; ++px) {
                    const int32_t Px = px * 256 + p.sx;
                    bool inside = true;
                    for (int i = 0; i < 3; ++i) {
                        const int64_t e = int64_t(A[i]) * (Px - r.X[i]) + int64_t(B[i]) * (Py - r.Y[i]);
                        if (!(e > 0 || (e == 0 && tl[i]))) {
                            inside = false;
                            break;
                        }
                    }
                    if (!inside) continue;
                    float z = std::fmaf(r.dzdy, float(Py - r.Y[0]), std::fmaf(r.dzdx, float(Px - r.X[0]), r.z0));
                    z = std::fmin(std::fmax(z, 0.0f), 1.0f);
                    const uint32_t zb = fbits(z) & 0x7fffffffu;
                    const size_t o = size_t(py) * width + px;
                    ++frags[t];
                    if (is_shadow) {
                        if (zb > depth_bits[o]) depth_bits[o] = zb;
                    } else {
                        const uint64_t w = (uint64_t(zb) << 32) | uint64_t(r.id + 1u);
                        if (w > out_vis[o]) out_vis[o] = w;
                    }
                }
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
    if (is_shadow) std::memcpy(out_depth, depth_bits.data(), sizeof(float) * depth_bits.size());
    if (stats) {
        stats->fragments = 0;
        for (auto f : frags) stats->fragments += f;
    }
    return RSY_OK;
}

int orc_blend_layers(const orc_geometry* g, const float view_projection[16], uint32_t width, uint32_t height,
                     const rsy_raster_state* rs, const uint64_t* opaque_vis, uint32_t max_layers, uint64_t* layers,
                     uint32_t* count, int n_threads) {
    return orc_blend_layers_at(g, view_projection, width, height, rs, 128, 128, opaque_vis, max_layers, layers, count,
                               n_threads);
}

int orc_blend_layers_at(const orc_geometry* g, const float view_projection[16], uint32_t width, uint32_t height,
                        const rsy_raster_state* rs, int32_t sample_x, int32_t sample_y, const uint64_t* opaque_vis,
                        uint32_t max_layers, uint64_t* layers, uint32_t* count, int n_threads) {
    if (!g || !rs || !opaque_vis || !count || width == 0 || height == 0) return RSY_INVALID_ARGUMENT;
    if (max_layers && !layers) return RSY_INVALID_ARGUMENT;
    if (sample_x < 0 || sample_x > 255 || sample_y < 0 || sample_y > 255) return RSY_INVALID_ARGUMENT;
    Params p = make_params(width, height, 0, rs);
    p.sx = sample_x;
    p.sy = sample_y;
    p.cull = false;  // the blended pass draws both faces (Graphics.cpp:5675)
    std::vector<Rec> recs;
    int rc = build_records(g, view_projection, p, recs, nullptr, n_threads);
    if (rc != RSY_OK) return rc;
    const size_t np = size_t(width) * height;
    std::memset(count, 0, sizeof(uint32_t) * np);
    if (max_layers) std::memset(layers, 0, sizeof(uint64_t) * np * max_layers);
    const int nt = std::max(1, n_threads);
    std::vector<int> overflow(nt, 0);
    auto work = [&](int t) {
        const int32_t y0 = int32_t(uint64_t(height) * t / nt), y1 = int32_t(uint64_t(height) * (t + 1) / nt);
        // the fragments of this thread's rows: (key, depth bits) per pixel
        std::vector<std::vector<std::pair<uint32_t, uint32_t>>> frags(size_t(std::max(0, y1 - y0)) * width);
        for (const Rec& r : recs) {
            const int32_t ya = std::max(r.py0, y0), yb = std::min(r.py1, y1 - 1);
            if (ya > yb) continue;
            int32_t A[3], B[3];
            bool tl[3];
            for (int i = 0; i < 3; ++i) {
                const int j = (i + 1) % 3;
                A[i] = r.Y[j] - r.Y[i];
                B[i] = r.X[i] - r.X[j];
                tl[i] = A[i] > 0 || (A[i] == 0 && B[i] > 0);
            }
            for (int32_t py = ya; py <= yb; ++py) {
                const int32_t Py = py * 256 + p.sy;
                for (int32_t px = r.px0; px <= r.px1; ++px) {
                    const int32_t Px = px * 256 + p.sx;
                    bool inside = true;
                    for (int i = 0; i < 3; ++i) {
                        const int64_t e = int64_t(A[i]) * (Px - r.X[i]) + int64_t(B[i]) * (Py - r.Y[i]);
                        if (!(e > 0 || (e == 0 && tl[i]))) {
                            inside = false;
                            break;
                        }
                    }
                    if (!inside) continue;
                    float z = std::fmaf(r.dzdy, float(Py - r.Y[0]), std::fmaf(r.dzdx, float(Px - r.X[0]), r.z0));
                    z = std::fmin(std::fmax(z, 0.0f), 1.0f);
                    const uint32_t zb = fbits(z) & 0x7fffffffu;
                    const size_t o = size_t(py) * width + px;
                    if (zb < uint32_t(opaque_vis[o] >> 32)) continue;  // GEQUAL against the opaque depth
                    frags[size_t(py - y0) * width + px].push_back({r.id + 1u, zb});
                }
            }
        }
        for (int32_t py = y0; py < y1; ++py)
            for (uint32_t px = 0; px < width; ++px) {
                auto& f = frags[size_t(py - y0) * width + px];
                std::sort(f.begin(), f.end());  // primitive order
                const size_t o = size_t(py) * width + px;
                count[o] = uint32_t(f.size());
                if (f.size() > max_layers) overflow[t] = 1;
                for (size_t k = 0; k < f.size() && k < max_layers; ++k)
                    layers[k * np + o] = (uint64_t(f[k].second) << 32) | f[k].first;
            }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
    for (int o : overflow)
        if (o) return RSY_OVERFLOW;
    return RSY_OK;
}

uint64_t orc_count_tile_pairs(const orc_geometry* g, const float view_projection[16], uint32_t width,
                              uint32_t height, int is_shadow, const rsy_raster_state* rs, uint32_t tile,
                              int n_threads) {
    if (!g || !rs || tile == 0) return 0;
    const Params p = make_params(width, height, is_shadow, rs);
    std::vector<Rec> recs;
    if (build_records(g, view_projection, p, recs, nullptr, n_threads) != RSY_OK) return 0;
    uint64_t pairs = 0;
    for (const Rec& r : recs)
        pairs += uint64_t(r.px1 / int32_t(tile) - r.px0 / int32_t(tile) + 1) *
                 uint64_t(r.py1 / int32_t(tile) - r.py0 / int32_t(tile) + 1);
    return pairs;
}

int orc_tile_counts(const orc_geometry* g, const float view_projection[16], uint32_t width, uint32_t height,
                    int is_shadow, const rsy_raster_state* rs, uint32_t tile, uint32_t* counts, uint64_t* areas,
                    int n_threads) {
    if (!g || !rs || tile == 0 || !counts) return RSY_INVALID_ARGUMENT;
    const Params p = make_params(width, height, is_shadow, rs);
    std::vector<Rec> recs;
    if (build_records(g, view_projection, p, recs, nullptr, n_threads) != RSY_OK) return RSY_ERROR;
    const uint32_t tx = (width + tile - 1) / tile, ty = (height + tile - 1) / tile;
    std::fill(counts, counts + size_t(tx) * ty, 0u);
    if (areas) std::fill(areas, areas + size_t(tx) * ty + 2, uint64_t(0));
    const int32_t T = int32_t(tile);
    for (const Rec& r : recs) {
        const uint64_t nb = uint64_t(r.px1 / T - r.px0 / T + 1) * uint64_t(r.py1 / T - r.py0 / T + 1);
        if (nb > 4 && areas) areas[size_t(tx) * ty] += 1, areas[size_t(tx) * ty + 1] += nb;
        for (int32_t y = r.py0 / T; y <= r.py1 / T; ++y)
            for (int32_t x = r.px0 / T; x <= r.px1 / T; ++x) {
                ++counts[size_t(y) * tx + x];
                if (areas) {
                    const int32_t w = std::min(r.px1, x * T + T - 1) - std::max(r.px0, x * T) + 1;
                    const int32_t h = std::min(r.py1, y * T + T - 1) - std::max(r.py0, y * T) + 1;
                    areas[size_t(y) * tx + x] += uint64_t(w) * uint64_t(h);
                }
            }
    }
    return RSY_OK;
}

}  // extern "C"
