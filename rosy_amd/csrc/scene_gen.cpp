// Seeded synthetic scenes for the five raster-and-shade workloads.
//
// The reference renders Packager-produced .rsy assets; none ship with it, so the workloads are
// procedural: meshes in the rosy_asset::position layout (Asset.h:92-99), one object record per
// mesh built with the Node.cpp world-transform closed form, one surface per mesh in the
// surface_graphics_data layout (Types.h:39-48), gpu_material records (Graphics.cpp:345-360) and
// decoded seeded value-noise textures.  Camera and light matrices come from the closed forms in
// host_math.cpp.  Every random draw comes from SplitMix64 streams derived from one seed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <thread>
#include <tuple>
#include <vector>

#include "host_math.h"
#include "rosy_amd.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// SplitMix64 (Steele, Lea, Flood 2014).
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
    double range(double a, double b) { return a + (b - a) * uniform(); }
    uint32_t below(uint32_t n) { return static_cast<uint32_t>(uniform() * n) % n; }
    Rng fork(uint64_t tag) { return Rng(next() ^ (tag * 0xd1b54a32d192ed03ull)); }
};

struct V3 {
    double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 normalize(V3 a) {
    const double l = std::sqrt(dot(a, a));
    return l > 0 ? a * (1.0 / l) : V3{0, 0, 0};
}

struct MeshBuild {
    std::vector<rsy_position> v;
    std::vector<uint32_t> idx;
    uint32_t add(V3 p, V3 n, V3 t, double w, double u, double vv) {
        rsy_position q{};
        q.vertex[0] = float(p.x);
        q.vertex[1] = float(p.y);
        q.vertex[2] = float(p.z);
        q.normal[0] = float(n.x);
        q.normal[1] = float(n.y);
        q.normal[2] = float(n.z);
        q.tangents[0] = float(t.x);
        q.tangents[1] = float(t.y);
        q.tangents[2] = float(t.z);
        q.tangents[3] = float(w);
        q.color[0] = 1.f;  // position::color default (Asset.h:97)
        q.color[1] = 0.f;
        q.color[2] = 0.f;
        q.color[3] = 1.f;
        q.texture_coordinates[0] = float(u);
        q.texture_coordinates[1] = float(vv);
        v.push_back(q);
        return uint32_t(v.size() - 1);
    }
    void tri(uint32_t a, uint32_t b, uint32_t c) {
        idx.push_back(a);
        idx.push_back(b);
        idx.push_back(c);
    }
    uint32_t tris() const { return uint32_t(idx.size() / 3); }
};

// Per-vertex tangents from UV derivatives (Lengyel 2001), Gram-Schmidt against the normal, with
// handedness w = sign(dot(cross(N, T), B)).
void compute_tangents(MeshBuild& m) {
    std::vector<V3> ts(m.v.size(), V3{0, 0, 0}), bs(m.v.size(), V3{0, 0, 0});
    for (size_t t = 0; t < m.idx.size(); t += 3) {
        const uint32_t i0 = m.idx[t], i1 = m.idx[t + 1], i2 = m.idx[t + 2];
        auto P = [&](uint32_t i) { return V3{m.v[i].vertex[0], m.v[i].vertex[1], m.v[i].vertex[2]}; };
        const V3 e1 = P(i1) - P(i0), e2 = P(i2) - P(i0);
        const double du1 = m.v[i1].texture_coordinates[0] - m.v[i0].texture_coordinates[0];
        const double dv1 = m.v[i1].texture_coordinates[1] - m.v[i0].texture_coordinates[1];
        const double du2 = m.v[i2].texture_coordinates[0] - m.v[i0].texture_coordinates[0];
        const double dv2 = m.v[i2].texture_coordinates[1] - m.v[i0].texture_coordinates[1];
        const double det = du1 * dv2 - du2 * dv1;
        if (std::fabs(det) < 1e-20) continue;
        const double r = 1.0 / det;
        const V3 T = (e1 * dv2 - e2 * dv1) * r;
        const V3 B = (e2 * du1 - e1 * du2) * r;
        for (uint32_t i : {i0, i1, i2}) {
            ts[i] = ts[i] + T;
            bs[i] = bs[i] + B;
        }
    }
    for (size_t i = 0; i < m.v.size(); ++i) {
        const V3 n{m.v[i].normal[0], m.v[i].normal[1], m.v[i].normal[2]};
        V3 t = ts[i] - n * dot(n, ts[i]);
        if (dot(t, t) < 1e-24) {
            t = std::fabs(n.x) < 0.9 ? cross(n, V3{1, 0, 0}) : cross(n, V3{0, 1, 0});
        }
        t = normalize(t);
        const double w = dot(cross(n, t), bs[i]) < 0.0 ? -1.0 : 1.0;
        m.v[i].tangents[0] = float(t.x);
        m.v[i].tangents[1] = float(t.y);
        m.v[i].tangents[2] = float(t.z);
        m.v[i].tangents[3] = float(w);
    }
}

// Octahedral geodesic sphere: each octahedron face split into n x n small triangles (8 n^2 total;
// n = 16 gives 2,048).  Lattice vertices are shared across faces (keyed by integer coordinates),
// so the surface is closed.  Radius 1 + U(-jitter, jitter) per shared vertex; analytic normals;
// octahedral UV mapping.
void build_sphere(MeshBuild& m, int n, double jitter, Rng& rng) {
    std::map<std::tuple<int, int, int>, uint32_t> shared;
    auto vert = [&](int X, int Y, int Z) -> uint32_t {
        auto key = std::make_tuple(X, Y, Z);
        auto it = shared.find(key);
        if (it != shared.end()) return it->second;
        const V3 d = normalize(V3{double(X), double(Y), double(Z)});
        const double r = 1.0 + rng.range(-jitter, jitter);
        const double l1 = std::fabs(d.x) + std::fabs(d.y) + std::fabs(d.z);
        double px = d.x / l1, pz = d.z / l1;
        if (d.y < 0.0) {
            const double fx = (1.0 - std::fabs(pz)) * (px >= 0 ? 1.0 : -1.0);
            const double fz = (1.0 - std::fabs(px)) * (pz >= 0 ? 1.0 : -1.0);
            px = fx;
            pz = fz;
        }
        const uint32_t id = m.add(d * r, d, V3{1, 0, 0}, 1.0, px * 0.5 + 0.5, pz * 0.5 + 0.5);
        shared.emplace(key, id);
        return id;
    };
    for (int sx : {1, -1})
        for (int sy : {1, -1})
            for (int sz : {1, -1}) {
                int a[3] = {sx, 0, 0}, b[3] = {0, sy, 0}, c[3] = {0, 0, sz};
                if (sx * sy * sz < 0) std::swap(b, c);  // keep CCW seen from outside
                std::vector<uint32_t> grid((n + 1) * (n + 1), 0);
                for (int i = 0; i <= n; ++i)
                    for (int j = 0; i + j <= n; ++j) {
                        const int k = n - i - j;
                        grid[i * (n + 1) + j] = vert(a[0] * k + b[0] * i + c[0] * j, a[1] * k + b[1] * i + c[1] * j,
                                                     a[2] * k + b[2] * i + c[2] * j);
                    }
                for (int i = 0; i < n; ++i)
                    for (int j = 0; i + j < n; ++j) {
                        m.tri(grid[i * (n + 1) + j], grid[(i + 1) * (n + 1) + j], grid[i * (n + 1) + j + 1]);
                        if (i + j < n - 1)
                            m.tri(grid[(i + 1) * (n + 1) + j], grid[(i + 1) * (n + 1) + j + 1],
                                  grid[i * (n + 1) + j + 1]);
                    }
            }
    compute_tangents(m);
}

// One subdivided rectangle with outward normal n = u x v, `cols` quads along u and `rows` full rows
// along v plus a last row of `partial` quads (0 = none).  Triangles (p00,p10,p11), (p00,p11,p01)
// are counter-clockwise seen from the normal side.
void build_grid(MeshBuild& m, V3 origin, V3 u, V3 v, double su, double sv, int cols, int rows, int partial,
                double tile) {
    const V3 n = normalize(cross(u, v));
    const int total_rows = rows + (partial > 0 ? 1 : 0);
    const double du = su / cols, dv = sv / std::max(total_rows, 1);
    for (int r = 0; r < total_rows; ++r) {
        const int ncols = r < rows ? cols : partial;
        const uint32_t base = uint32_t(m.v.size());
        for (int j = 0; j <= 1; ++j)
            for (int i = 0; i <= ncols; ++i) {
                const double a = i * du, b = (r + j) * dv;
                const V3 p = origin + u * a + v * b;
                m.add(p, n, u, 1.0, a / tile, b / tile);
            }
        const uint32_t stride = uint32_t(ncols + 1);
        for (int i = 0; i < ncols; ++i) {
            const uint32_t p00 = base + i, p10 = base + i + 1, p01 = base + stride + i, p11 = base + stride + i + 1;
            m.tri(p00, p10, p11);
            m.tri(p00, p11, p01);
        }
    }
}

// Rows of a grid share their boundary vertices' positions exactly only when the row bases agree;
// build_grid recomputes them from the same expression, so adjacent rows are crack-free.

// Box of half extents (hx, hy, hz) resting on y = 0, each face an n x n grid: 12 n^2 triangles.
void build_box(MeshBuild& m, double hx, double hy, double hz, int n, double tile) {
    struct Face {
        V3 nrm, u, v;
        double hu, hv, hn;
    };
    const Face faces[6] = {
        {{1, 0, 0}, {0, 0, -1}, {0, 1, 0}, hz, hy, hx},  {{-1, 0, 0}, {0, 0, 1}, {0, 1, 0}, hz, hy, hx},
        {{0, 1, 0}, {1, 0, 0}, {0, 0, -1}, hx, hz, hy},  {{0, -1, 0}, {1, 0, 0}, {0, 0, 1}, hx, hz, hy},
        {{0, 0, 1}, {1, 0, 0}, {0, 1, 0}, hx, hy, hz},   {{0, 0, -1}, {-1, 0, 0}, {0, 1, 0}, hx, hy, hz},
    };
    for (const Face& f : faces) {
        const V3 origin = f.nrm * f.hn - f.u * f.hu - f.v * f.hv + V3{0, hy, 0};
        build_grid(m, origin, f.u, f.v, 2 * f.hu, 2 * f.hv, n, n, 0, tile);
    }
}

// Column: cylinder of radius r and height h on y = 0 with s segments, k rings and fan caps:
// 2 s k + 2 s triangles.  The seam column reuses the first column's positions exactly.
void build_column(MeshBuild& m, double r, double h, int s, int k, double tile) {
    std::vector<V3> ring(s);
    for (int i = 0; i < s; ++i) {
        const double th = -2.0 * kPi * i / s;
        ring[i] = V3{r * std::cos(th), 0.0, r * std::sin(th)};
    }
    const double circ = 2.0 * kPi * r;
    const uint32_t base = uint32_t(m.v.size());
    for (int j = 0; j <= k; ++j)
        for (int i = 0; i <= s; ++i) {
            const V3 q = ring[i % s];
            const V3 p{q.x, h * j / k, q.z};
            const V3 n = normalize(V3{q.x, 0, q.z});
            const V3 t{n.z, 0.0, -n.x};  // u direction: (sin th, 0, -cos th)
            m.add(p, n, t, 1.0, circ * i / s / tile, h * j / k / tile);
        }
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < s; ++i) {
            const uint32_t p00 = base + j * (s + 1) + i, p10 = p00 + 1, p01 = p00 + (s + 1), p11 = p01 + 1;
            m.tri(p00, p10, p11);
            m.tri(p00, p11, p01);
        }
    for (int cap = 0; cap < 2; ++cap) {
        const double y = cap == 0 ? h : 0.0;
        const V3 n{0, cap == 0 ? 1.0 : -1.0, 0};
        const uint32_t c = m.add(V3{0, y, 0}, n, V3{1, 0, 0}, cap == 0 ? 1.0 : -1.0, 0.0, 0.0);
        const uint32_t rb = uint32_t(m.v.size());
        for (int i = 0; i < s; ++i)
            m.add(V3{ring[i].x, y, ring[i].z}, n, V3{1, 0, 0}, cap == 0 ? 1.0 : -1.0, ring[i].x / tile,
                  ring[i].z / tile);
        for (int i = 0; i < s; ++i) {
            const uint32_t a = rb + i, b = rb + (i + 1) % s;
            if (cap == 0)
                m.tri(c, a, b);
            else
                m.tri(c, b, a);
        }
    }
}

// ---- textures --------------------------------------------------------------------------------

double srgb_encode(double l) {
    l = std::min(1.0, std::max(0.0, l));
    return l <= 0.0031308 ? 12.92 * l : 1.055 * std::pow(l, 1.0 / 2.4) - 0.055;
}
uint8_t to_byte(double x) { return uint8_t(std::lround(std::min(1.0, std::max(0.0, x)) * 255.0)); }

// Periodic value noise: random lattice values on a period x period grid, smoothstep blend.
struct ValueNoise {
    int period;
    std::vector<float> lat;
    ValueNoise(int p, Rng& rng) : period(p), lat(size_t(p) * p) {
        for (auto& x : lat) x = float(rng.uniform());
    }
    double at(double x, double y) const {  // x, y in lattice units
        const double fx = std::floor(x), fy = std::floor(y);
        const int ix = int(fx), iy = int(fy);
        double tx = x - fx, ty = y - fy;
        tx = tx * tx * (3 - 2 * tx);
        ty = ty * ty * (3 - 2 * ty);
        auto L = [&](int i, int j) {
            i = ((i % period) + period) % period;
            j = ((j % period) + period) % period;
            return double(lat[size_t(j) * period + i]);
        };
        const double a = L(ix, iy) + (L(ix + 1, iy) - L(ix, iy)) * tx;
        const double b = L(ix, iy + 1) + (L(ix + 1, iy + 1) - L(ix, iy + 1)) * tx;
        return a + (b - a) * ty;
    }
};

struct Fbm {
    std::vector<ValueNoise> oct;
    int base;
    Fbm(int base_period, int octaves, Rng& rng) : base(base_period) {
        for (int o = 0; o < octaves; ++o) oct.emplace_back(base_period << o, rng);
    }
    // u, v in [0,1) texture space; result in [0,1].
    double at(double u, double v) const {
        double s = 0, w = 0, a = 1;
        for (const auto& n : oct) {
            s += a * n.at(u * n.period, v * n.period);
            w += a;
            a *= 0.5;
        }
        return s / w;
    }
};

struct MaterialTextures {
    std::vector<uint8_t> albedo, normal, mr;
};

struct Blob {
    double u, v, r, strength;
};

// `cutout`: the albedo alpha is 0.5 + 0.5 sin(8 pi u) sin(8 pi v) (translucent materials: holes
// under an alpha mask, a varying alpha under blending); otherwise opaque (255).
// `fine`: detail per material at one of five scales, the coarsest octave at size/128 << k cycles
// (k = 0..4: up to 512 cycles in the finest octave of a 1024^2 map), stripes of 3 to 200 cycles
// (log-uniform) and blobs of radius 0.01-0.08: much of it above what a 256^2 grid holds (128 cycles).
void make_material_textures(MaterialTextures& out, int size, Rng rng, bool stripes_and_blobs, bool cutout,
                            bool fine = false) {
    const double hue = rng.uniform(), sat = rng.range(0.2, 0.8);
    const bool metal = rng.uniform() < 0.35;
    const int base = std::max(2, size / 128) << (fine ? int(rng.below(5)) : 0);
    Fbm albedo_n(base, 3, rng), tint_n(base, 2, rng), height_n(base * 2, 2, rng), metal_n(base, 2, rng),
        rough_n(base, 3, rng);
    const double stripe_angle = rng.range(0, kPi);
    const double stripe_freq = fine ? std::round(std::exp(rng.range(std::log(3.0), std::log(200.0))))
                                    : double(int(rng.range(3, 9)));
    std::vector<Blob> blobs;
    if (stripes_and_blobs) {
        const int nb = 6 + int(rng.below(6));
        const double r0 = fine ? 0.01 : 0.04, r1 = fine ? 0.08 : 0.12;
        for (int i = 0; i < nb; ++i) blobs.push_back({rng.uniform(), rng.uniform(), rng.range(r0, r1), rng.range(-0.5, 0.5)});
    }
    // base colour from hue/saturation
    auto hsv = [](double h, double s, double v, double rgb[3]) {
        const double i = std::floor(h * 6), f = h * 6 - i;
        const double p = v * (1 - s), q = v * (1 - f * s), t = v * (1 - (1 - f) * s);
        switch (int(i) % 6) {
            case 0: rgb[0] = v, rgb[1] = t, rgb[2] = p; break;
            case 1: rgb[0] = q, rgb[1] = v, rgb[2] = p; break;
            case 2: rgb[0] = p, rgb[1] = v, rgb[2] = t; break;
            case 3: rgb[0] = p, rgb[1] = q, rgb[2] = v; break;
            case 4: rgb[0] = t, rgb[1] = p, rgb[2] = v; break;
            default: rgb[0] = v, rgb[1] = p, rgb[2] = q; break;
        }
    };
    double base_rgb[3];
    hsv(hue, sat, 0.9, base_rgb);
    out.albedo.resize(size_t(size) * size * 4);
    out.normal.resize(size_t(size) * size * 2);
    out.mr.resize(size_t(size) * size * 2);
    const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](int y0, int y1) {
        const double eps = 1.0 / size;
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < size; ++x) {
                const double u = (x + 0.5) / size, v = (y + 0.5) / size;
                double feat = 0.0;
                if (stripes_and_blobs) {
                    const double s = std::sin(2 * kPi * stripe_freq * (u * std::cos(stripe_angle) + v * std::sin(stripe_angle)));
                    feat = 0.5 * s;
                    for (const Blob& b : blobs) {
                        double du = std::fabs(u - b.u), dv = std::fabs(v - b.v);
                        du = std::min(du, 1 - du);
                        dv = std::min(dv, 1 - dv);
                        feat += b.strength * std::exp(-(du * du + dv * dv) / (b.r * b.r));
                    }
                }
                const double lum = 0.45 + 0.5 * albedo_n.at(u, v) + 0.25 * feat;
                const double tint = tint_n.at(u, v) - 0.5;
                const size_t o = (size_t(y) * size + x);
                for (int c = 0; c < 3; ++c) {
                    const double lin = base_rgb[c] * lum * (1.0 + 0.4 * tint * (c == 0 ? 1 : c == 1 ? -0.5 : -1));
                    out.albedo[o * 4 + c] = to_byte(srgb_encode(lin));
                }
                out.albedo[o * 4 + 3] =
                    cutout ? to_byte(0.5 + 0.5 * std::sin(8 * kPi * u) * std::sin(8 * kPi * v)) : uint8_t(255);
                const double h0 = height_n.at(u, v), hx = height_n.at(u + eps, v), hy = height_n.at(u, v + eps);
                const double nx = std::max(-0.6, std::min(0.6, (h0 - hx) * size * 0.012));
                const double ny = std::max(-0.6, std::min(0.6, (h0 - hy) * size * 0.012));
                out.normal[o * 2 + 0] = to_byte(nx * 0.5 + 0.5);
                out.normal[o * 2 + 1] = to_byte(ny * 0.5 + 0.5);
                const double mval = metal ? 0.75 + 0.25 * metal_n.at(u, v) : 0.2 * metal_n.at(u, v);
                // roughness kept in [0.3, 1] (linear) so the GGX lobe stays well conditioned
                const double rough = std::min(1.0, std::max(0.3, 0.35 + 0.6 * rough_n.at(u, v) + 0.2 * feat));
                out.mr[o * 2 + 0] = to_byte(srgb_encode(mval));
                out.mr[o * 2 + 1] = to_byte(srgb_encode(rough));
            }
    };
    std::vector<std::thread> th;
    const int chunk = (size + int(nthreads) - 1) / int(nthreads);
    for (unsigned t = 0; t < nthreads; ++t) {
        const int y0 = int(t) * chunk, y1 = std::min(size, y0 + chunk);
        if (y0 < y1) th.emplace_back(work, y0, y1);
    }
    for (auto& t : th) t.join();
}

// ---- scene assembly ---------------------------------------------------------------------------

struct SceneStore : rsy_scene {
    std::vector<rsy_position> v_vertices;
    std::vector<uint32_t> v_indices;
    std::vector<rsy_mesh_range> v_meshes;
    std::vector<rsy_object_record> v_objects;
    std::vector<rsy_draw> v_surfaces;
    std::vector<rsy_gpu_material> v_materials;
    std::vector<MaterialTextures> tex_store;
    std::vector<rsy_texture_desc> v_textures;
    std::vector<rsy_sampler> v_samplers;
    std::vector<rsy_camera_view> v_views;
    std::vector<rsy_shadow_cascade> v_cascades;
};

struct Workload {
    uint32_t width, height, shadow_size, total_tris, n_meshes, min_tris, max_tris, n_materials, tex_size;
    bool stripes;
    uint32_t n_views;
    uint32_t n_cascades;
};

Workload workload_defaults(uint32_t w) {
    switch (w) {
        case 0: return {640, 360, 1024, 2048, 1, 2048, 2048, 1, 256, false, 1, 1};
        case 1: return {1920, 1080, 2048, 262144, 120, 64, 32768, 16, 2048, false, 1, 1};
        case 2: return {2560, 1440, 2048, 1048576, 106, 64, 65536, 16, 2048, false, 1, 4};
        case 3: return {1920, 1080, 2048, 262144, 120, 64, 32768, 64, 1024, true, 1, 1};
        default: return {3840, 2160, 2048, 2097152, 116, 64, 131072, 64, 1024, true, 4, 4};
    }
}

// Log-uniform counts on [lo, hi] mapped by a power law t = lo (u/lo)^g so that the total hits
// `total` while staying inside [lo, hi]; counts are even.
std::vector<uint32_t> triangle_budget(uint32_t n, uint32_t lo, uint32_t hi, uint32_t total, Rng& rng) {
    std::vector<double> u(n);
    for (auto& x : u) x = std::exp(rng.range(std::log(double(lo)), std::log(double(hi))));
    auto sum_for = [&](double g) {
        double s = 0;
        for (double x : u) s += lo * std::pow(x / lo, g);
        return s;
    };
    double g0 = 0.0, g1 = 1.0;
    if (sum_for(1.0) < total) {  // need expansion: allow g > 1 up to the hi clamp
        g1 = 4.0;
    }
    for (int it = 0; it < 200; ++it) {
        const double gm = 0.5 * (g0 + g1);
        if (sum_for(gm) < total) g0 = gm; else g1 = gm;
    }
    std::vector<uint32_t> t(n);
    for (uint32_t i = 0; i < n; ++i) {
        double x = lo * std::pow(u[i] / lo, 0.5 * (g0 + g1));
        x = std::min(double(hi), std::max(double(lo), x));
        t[i] = std::max(lo, uint32_t(std::lround(x / 2.0)) * 2u);
    }
    return t;
}

void fill_scene_data(rsy_scene_data& sd, const rsy_camera_view& cam, const float sun[4], const float shadow[16]) {
    std::memset(&sd, 0, sizeof(sd));
    std::memcpy(sd.view, cam.view, sizeof(sd.view));
    std::memcpy(sd.proj, cam.proj, sizeof(sd.proj));
    std::memcpy(sd.view_projection, cam.view_projection, sizeof(sd.view_projection));
    std::memcpy(sd.shadow_projection_near, shadow, sizeof(sd.shadow_projection_near));
    std::memcpy(sd.sun_position, sun, sizeof(sd.sun_position));
    std::memcpy(sd.camera_position, cam.camera_position, sizeof(sd.camera_position));
    // defaults of Level.cpp:1241-1262 packed as in Graphics.cpp:6016-6056
    const float ambient = 0.01f;
    sd.ambient_color[0] = sd.ambient_color[1] = sd.ambient_color[2] = ambient;
    sd.ambient_color[3] = 1.f;
    for (int i = 0; i < 4; ++i) sd.sunlight_color[i] = 1.f;
    sd.flip_lights[3] = 1;
    sd.draw_extent[0] = float(cam.width);
    sd.draw_extent[1] = float(cam.height);
    sd.light_intensity = 5.2f;
    sd.fragment_output = 0;
    sd.light_enabled = 1;
    sd.tangent_space_enabled = 1;
    sd.shadows_enabled = 1;
    sd.normal_maps_enabled = 1;
    sd.brdf_lighting_enabled = 1;
}

rsy_camera_view make_camera(uint32_t w, uint32_t h, const float pos[3], float pitch, float yaw) {
    rsy_camera_view c{};
    c.width = w;
    c.height = h;
    rsy_camera_projection(double(w) / double(h), 0.1, 100.0, 60.0 * kPi / 180.0, c.proj);
    rsy_camera_view_matrix(pos, pitch, yaw, c.view);
    rsy_mat4_mul(c.proj, c.view, c.view_projection);
    c.camera_position[0] = pos[0];
    c.camera_position[1] = pos[1];
    c.camera_position[2] = pos[2];
    c.camera_position[3] = 0.f;  // Camera.cpp:275-276 copies xyz only
    return c;
}

// Cascade fit (a rule of this implementation; the reference renders one cascade): the light-space
// AABB of the camera frustum slice [prev split, split] of every camera view, with the depth range of
// the whole scene so that casters outside the slice still land in the map.  Same sign convention as
// the reference's light projection (Level.cpp:505-511): x and y negated, depth 1 nearest the sun.
void fit_cascades(const std::vector<rsy_camera_view>& views, const rsy::dmat4& light_view, const double scene_lo[3],
                  const double scene_hi[3], const std::vector<double>& splits, uint32_t size,
                  std::vector<rsy_shadow_cascade>& out) {
    // scene depth range in light space
    double zlo = 1e30, zhi = -1e30;
    for (int i = 0; i < 8; ++i) {
        const double p[4] = {(i & 1) ? scene_hi[0] : scene_lo[0], (i & 2) ? scene_hi[1] : scene_lo[1],
                             (i & 4) ? scene_hi[2] : scene_lo[2], 1.0};
        double q[4];
        rsy::d_mul_vec(light_view, p, q);
        zlo = std::min(zlo, q[2]);
        zhi = std::max(zhi, q[2]);
    }
    zlo -= 1.0;
    zhi += 1.0;
    double prev = 0.1;
    for (double split : splits) {
        double xlo = 1e30, xhi = -1e30, ylo = 1e30, yhi = -1e30;
        for (const auto& v : views) {
            const double t = std::tan(30.0 * kPi / 180.0), asp = double(v.width) / v.height;
            const rsy::dmat4 inv_view = rsy::d_inverse(rsy::to_double(v.view));
            for (double d : {prev, split})
                for (int sx : {-1, 1})
                    for (int sy : {-1, 1}) {
                        const double pv[4] = {sx * d * t * asp, sy * d * t, -d, 1.0};
                        double pw[4], pl[4];
                        rsy::d_mul_vec(inv_view, pv, pw);
                        rsy::d_mul_vec(light_view, pw, pl);
                        xlo = std::min(xlo, pl[0]);
                        xhi = std::max(xhi, pl[0]);
                        ylo = std::min(ylo, pl[1]);
                        yhi = std::max(yhi, pl[1]);
                    }
        }
        // stay inside the scene box where the slice is larger than the scene
        rsy::dmat4 lp{};
        const double cx = 0.5 * (xlo + xhi), cy = 0.5 * (ylo + yhi);
        const double ex = std::max(1e-3, 0.5 * (xhi - xlo)), ey = std::max(1e-3, 0.5 * (yhi - ylo));
        rsy::at(lp, 0, 0) = -1.0 / ex;
        rsy::at(lp, 3, 0) = cx / ex;
        rsy::at(lp, 1, 1) = -1.0 / ey;
        rsy::at(lp, 3, 1) = cy / ey;
        // depth: z = zhi (farthest from the sun) -> 0, z = zlo -> 1
        rsy::at(lp, 2, 2) = -1.0 / (zhi - zlo);
        rsy::at(lp, 3, 2) = zhi / (zhi - zlo);
        rsy::at(lp, 3, 3) = 1.0;
        rsy_shadow_cascade c{};
        c.size = size;
        c.split_far = float(split);
        const rsy::fmat4 m = rsy::to_float(rsy::d_mul(lp, light_view));
        std::memcpy(c.light_view_projection, m.data(), sizeof(float) * 16);
        out.push_back(c);
        prev = split;
    }
}

}  // namespace

extern "C" {

rsy_result rsy_scene_generate(const rsy_scene_desc* desc, rsy_scene** out_scene) {
    if (!desc || !out_scene || desc->workload > 4) return RSY_INVALID_ARGUMENT;
    *out_scene = nullptr;
    SceneStore* s = new (std::nothrow) SceneStore();
    if (!s) return RSY_ALLOCATION_FAILURE;
    try {
        Workload wl = workload_defaults(desc->workload);
        if (desc->width) wl.width = desc->width;
        if (desc->height) wl.height = desc->height;
        if (desc->shadow_size) wl.shadow_size = desc->shadow_size;
        if (desc->total_triangles) wl.total_tris = desc->total_triangles;
        if (desc->n_meshes) wl.n_meshes = desc->n_meshes;
        if (desc->n_materials) wl.n_materials = desc->n_materials;
        if (desc->texture_size) wl.tex_size = desc->texture_size;
        Rng master(desc->seed * 0x9e3779b97f4a7c15ull + 0x5eed + desc->workload);
        Rng geo = master.fork(1), xf = master.fork(2), mat_rng = master.fork(3), tex_rng = master.fork(4);

        std::vector<MeshBuild> meshes;
        std::vector<rsy_object_record> objects;
        double scene_lo[3] = {-20, 0, -10}, scene_hi[3] = {20, 20, 10};
        if (desc->workload == 0) {
            MeshBuild m;
            int n = 16;
            while (8 * n * n > int(wl.total_tris) && n > 1) --n;
            build_sphere(m, n, 0.02, geo);
            meshes.push_back(std::move(m));
            rsy_object_record rec{};
            const float t[3] = {0, 0, 0};
            rsy_object_record_from_trs(t, 1.f, 0.f, &rec);
            objects.push_back(rec);
            scene_lo[0] = scene_lo[1] = scene_lo[2] = -1.1;
            scene_hi[0] = scene_hi[1] = scene_hi[2] = 1.1;
        } else {
            std::vector<uint32_t> budget = triangle_budget(wl.n_meshes, wl.min_tris, wl.max_tris, wl.total_tris, geo);
            const uint32_t ground = uint32_t(std::max_element(budget.begin(), budget.end()) - budget.begin());
            {  // rounding residue of the power-law budget goes to the ground plane
                int64_t sum = 0;
                for (uint32_t b : budget) sum += b;
                budget[ground] = uint32_t(int64_t(budget[ground]) + int64_t(wl.total_tris) - sum);
            }
            std::vector<int> kind(wl.n_meshes);  // 0 box, 1 column, 2 plane
            for (uint32_t i = 0; i < wl.n_meshes; ++i) {
                const double r = geo.uniform();
                kind[i] = i == ground ? 2 : (r < 0.4 ? 0 : r < 0.7 ? 1 : 2);
            }
            // realise boxes and columns first, then give the planes the remainder so the total is exact
            std::vector<MeshBuild> built(wl.n_meshes);
            std::vector<double> plane_su(wl.n_meshes), plane_sv(wl.n_meshes);
            int64_t deficit = 0;
            for (uint32_t i = 0; i < wl.n_meshes; ++i) {
                const uint32_t t = budget[i];
                if (kind[i] == 0) {
                    int n = std::max(1, int(std::lround(std::sqrt(t / 12.0))));
                    if (12 * n * n < int(wl.min_tris)) ++n;
                    build_box(built[i], geo.range(0.4, 2.5), geo.range(0.4, 3.0), geo.range(0.4, 2.5), n, 4.0);
                    deficit += int64_t(t) - built[i].tris();
                } else if (kind[i] == 1) {
                    const int s = std::max(6, int(std::sqrt(t / 2.0)));
                    const int k = std::max(1, (int(t / 2) + s - 1) / s - 1);
                    build_column(built[i], geo.range(0.25, 1.0), geo.range(1.5, 9.0), s, k, 4.0);
                    deficit += int64_t(t) - built[i].tris();
                }
            }
            std::vector<uint32_t> planes;
            for (uint32_t i = 0; i < wl.n_meshes; ++i)
                if (kind[i] == 2) planes.push_back(i);
            for (size_t pi = 0; pi < planes.size(); ++pi) {
                const uint32_t i = planes[pi];
                int64_t t = budget[i];
                if (i == ground) t += deficit;  // the ground plane (largest budget) absorbs the residue
                t = std::max<int64_t>(2, t);
                const int64_t quads = t / 2;
                double su, sv;
                if (i == ground) {
                    su = 40.0, sv = 20.0;
                } else {
                    su = geo.range(1.0, 6.0), sv = geo.range(1.0, 6.0);
                }
                int cols = std::max<int64_t>(1, int64_t(std::floor(std::sqrt(double(quads) * su / sv))));
                cols = int(std::min<int64_t>(cols, quads));
                const int rows = int(quads / cols), partial = int(quads - int64_t(rows) * cols);
                const V3 origin{-su / 2, 0, sv / 2};
                build_grid(built[i], origin, V3{1, 0, 0}, V3{0, 0, -1}, su, sv, cols, rows, partial, 4.0);
                plane_su[i] = su;
                plane_sv[i] = sv;
            }
            for (uint32_t i = 0; i < wl.n_meshes; ++i) {
                float t[3];
                float scale = 1.f, yaw = float(xf.range(0, 2 * kPi));
                if (i == ground) {
                    t[0] = 0, t[1] = 0, t[2] = 0, yaw = 0;
                } else if (kind[i] == 2) {
                    t[0] = float(xf.range(-17, 17)), t[1] = float(xf.range(0.5, 10)), t[2] = float(xf.range(-8, 8));
                } else {
                    t[0] = float(xf.range(-18, 18)), t[1] = 0.f, t[2] = float(xf.range(-9, 9));
                    scale = float(xf.range(0.7, 1.4));
                }
                rsy_object_record rec{};
                rsy_object_record_from_trs(t, scale, yaw, &rec);
                objects.push_back(rec);
            }
            meshes = std::move(built);
        }

        // flatten geometry
        uint64_t vbase = 0;
        uint32_t ibase = 0;
        for (size_t i = 0; i < meshes.size(); ++i) {
            rsy_mesh_range mr{};
            mr.vertex_buffer_offset = vbase * sizeof(rsy_position);
            mr.index_offset = ibase;
            mr.num_indices = uint32_t(meshes[i].idx.size());
            mr.num_vertices = uint32_t(meshes[i].v.size());
            s->v_meshes.push_back(mr);
            s->v_vertices.insert(s->v_vertices.end(), meshes[i].v.begin(), meshes[i].v.end());
            s->v_indices.insert(s->v_indices.end(), meshes[i].idx.begin(), meshes[i].idx.end());
            vbase += meshes[i].v.size();
            ibase += uint32_t(meshes[i].idx.size());
        }
        s->v_objects = objects;

        // materials + textures: three maps per material
        const uint32_t nm = wl.n_materials;
        s->tex_store.resize(nm);
        s->v_samplers.push_back(rsy_sampler{4, 4, 2, 2});  // LINEAR / REPEAT
        for (uint32_t m = 0; m < nm; ++m) {
            const bool translucent = m < desc->n_translucent;
            make_material_textures(s->tex_store[m], int(wl.tex_size), tex_rng.fork(m + 1), wl.stripes, translucent,
                                   desc->texture_detail != 0);
            const uint32_t base = uint32_t(s->v_textures.size());
            const uint32_t mips = desc->texture_mips ? RSY_MIPS_BUILD : 0u;
            s->v_textures.push_back({wl.tex_size, wl.tex_size, RSY_TEX_RGBA8_SRGB, mips, s->tex_store[m].albedo.data()});
            s->v_textures.push_back({wl.tex_size, wl.tex_size, RSY_TEX_RG8_UNORM, mips, s->tex_store[m].normal.data()});
            s->v_textures.push_back({wl.tex_size, wl.tex_size, RSY_TEX_RG8_SRGB, mips, s->tex_store[m].mr.data()});
            rsy_gpu_material g{};
            g.color[0] = g.color[1] = g.color[2] = g.color[3] = 1.f;
            g.metallic_factor = 1.f;
            g.roughness_factor = 1.f;
            g.alpha_cutoff = 0.f;
            g.alpha_mode = 0;
            if (translucent) {
                // even: BLEND with a 0.7 alpha factor; odd: MASK at cutoff 0.5 (Level.cpp's alpha modes)
                g.alpha_mode = m % 2 == 0 ? 2 : 1;
                g.color[3] = m % 2 == 0 ? 0.7f : 1.f;
                g.alpha_cutoff = m % 2 == 0 ? 0.f : 0.5f;
            }
            g.color_sampled_image_index = base;
            g.color_sampler_index = 0;
            g.normal_sampled_image_index = base + 1;
            g.normal_sampler_index = 0;
            g.metallic_sampled_image_index = base + 2;
            g.metallic_sampler_index = 0;
            g.mixmap_sampled_image_index = RSY_ABSENT;
            g.mixmap_sampler_index = RSY_ABSENT;
            s->v_materials.push_back(g);
        }
        // one surface per mesh (Level.cpp:1046-1062)
        for (uint32_t i = 0; i < uint32_t(meshes.size()); ++i) {
            rsy_draw d{};
            d.mesh_index = i;
            d.graphic_objects_offset = 0;
            d.graphics_object_index = i;
            d.material_index = desc->workload == 0 ? 0 : mat_rng.below(nm);
            d.index_count = s->v_meshes[i].num_indices;
            d.start_index = 0;
            d.blended = s->v_materials[d.material_index].alpha_mode != 0 ? 1 : 0;
            s->v_surfaces.push_back(d);
        }

        // cameras
        if (desc->workload == 0) {
            // on the sun's side of the sphere (the default sun sits at -z), looking down +z
            const float pos[3] = {0.f, 0.f, -3.f};
            s->v_views.push_back(make_camera(wl.width, wl.height, pos, 0.f, float(kPi)));
        } else {
            // the default sun sits at -z: the main camera looks from -z towards +z so lit faces face it
            const float pos[3] = {0.f, 14.f, -30.f};
            s->v_views.push_back(make_camera(wl.width, wl.height, pos, float(std::atan2(10.0, 30.0)), float(kPi)));
            if (wl.n_views > 1) {
                const uint32_t sizes[3][2] = {{1920, 1080}, {1280, 720}, {640, 360}};
                Rng cam = master.fork(5);
                for (uint32_t v = 1; v < wl.n_views && v < 4; ++v) {
                    const float yaw = float(cam.range(-0.6, 0.6));
                    const float dist = float(cam.range(22, 34));
                    const float p[3] = {float(-std::sin(double(yaw)) * dist), float(cam.range(8, 16)),
                                        float(-std::cos(double(yaw)) * dist)};
                    const float pitch = float(std::atan2(double(p[1]) - 4.0, double(dist)));
                    s->v_views.push_back(make_camera(sizes[v - 1][0], sizes[v - 1][1], p, pitch, yaw + float(kPi)));
                }
            }
        }
        // light: reference defaults (Level.cpp:1241-1245)
        float sun[4], shadow_proj[16];
        rsy_light_matrices(23.776f, 5.343f, 6.414f, 36.f, 76.838f, sun, shadow_proj);
        if (wl.n_cascades <= 1) {
            rsy_shadow_cascade c{};
            c.size = wl.shadow_size;
            c.split_far = INFINITY;
            std::memcpy(c.light_view_projection, shadow_proj, sizeof(shadow_proj));
            s->v_cascades.push_back(c);
        } else {
            fit_cascades(s->v_views, rsy::light_view_inverse(23.776, 5.343, 6.414), scene_lo, scene_hi, {5.0, 15.0, 40.0, 100.0}, wl.shadow_size,
                         s->v_cascades);
        }
        fill_scene_data(s->scene_data, s->v_views[0], sun, s->v_cascades[0].light_view_projection);

        rsy_raster_state& rs = s->raster_state;
        std::memset(&rs, 0, sizeof(rs));
        rs.cull_enabled = 1;
        rs.reverse_winding_order_enabled = 0;
        rs.depth_bias_enabled = 1;
        rs.depth_bias_constant = -21.882f;
        rs.depth_bias_clamp = -20.937f;
        rs.depth_bias_slope_factor = -3.f;
        rs.pcf_mode = desc->workload == 0 ? RSY_PCF_GRID_3X3 : RSY_PCF_REFERENCE_5TAP;
        rs.tile_size = 0;
        rs.clear_color[0] = 0.f;  // Graphics.cpp:5416-5418
        rs.clear_color[1] = 0.05f;
        rs.clear_color[2] = 0.1f;
        rs.clear_color[3] = 1.f;

        s->vertices = s->v_vertices.data();
        s->n_vertices = s->v_vertices.size();
        s->indices = s->v_indices.data();
        s->n_indices = s->v_indices.size();
        s->meshes = s->v_meshes.data();
        s->n_meshes = uint32_t(s->v_meshes.size());
        s->objects = s->v_objects.data();
        s->n_objects = uint32_t(s->v_objects.size());
        s->surfaces = s->v_surfaces.data();
        s->n_surfaces = uint32_t(s->v_surfaces.size());
        s->materials = s->v_materials.data();
        s->n_materials = uint32_t(s->v_materials.size());
        s->textures = s->v_textures.data();
        s->n_textures = uint32_t(s->v_textures.size());
        s->samplers = s->v_samplers.data();
        s->n_samplers = uint32_t(s->v_samplers.size());
        s->camera_views = s->v_views.data();
        s->n_camera_views = uint32_t(s->v_views.size());
        s->cascades = s->v_cascades.data();
        s->n_cascades = uint32_t(s->v_cascades.size());
        s->default_shadow_size = wl.shadow_size;
        s->workload = desc->workload;
    } catch (const std::bad_alloc&) {
        delete s;
        return RSY_ALLOCATION_FAILURE;
    } catch (...) {
        delete s;
        return RSY_ERROR;
    }
    *out_scene = s;
    return RSY_OK;
}

void rsy_scene_free(rsy_scene* scene) { delete static_cast<SceneStore*>(scene); }

}  // extern "C"
