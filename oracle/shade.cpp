// CPU oracle, shading half, in FP64.  TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Deferred restatement of the main pass: for each pixel of a visibility buffer, the three vertices
// of the visible triangle go through vertexMain (basic.slang:15-119), the ten varyings and the CSM
// coordinate are interpolated perspective-correctly at the pixel centre, and fragmentMain
// (basic.slang:137-344) runs with lighting.slang:5-43 and pcf.slang:1-12.  Textures are sampled
// bilinearly at mip 0 with sRGB decode before filtering (Vulkan sampler rules, Graphics.cpp:3421-3432
// and 3632-3647); the shadow compare sampler is LINEAR / CLAMP_TO_EDGE / GREATER_OR_EQUAL
// (Graphics.cpp:2697-2725).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "fixed_point.h"
#include "oracle.h"
#include "neural_oracle.h"

namespace {

struct D3 {
    double x, y, z;
};
struct D4 {
    double x, y, z, w;
};
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
D3 operator*(D3 a, D3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
D3 normalize(D3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }
double clamp01(double x) { return std::min(1.0, std::max(0.0, x)); }
D3 xyz(D4 v) { return {v.x, v.y, v.z}; }

// column-major 4x4 times vector
D4 mul4(const float* m, D4 v) {
    D4 r;
    double* o = &r.x;
    for (int row = 0; row < 4; ++row)
        o[row] = double(m[row]) * v.x + double(m[4 + row]) * v.y + double(m[8 + row]) * v.z + double(m[12 + row]) * v.w;
    return r;
}
// column-major 3x3 times vector (float3x3 normalTransform, data.slang:35)
D3 mul3(const float* m, D3 v) {
    return {double(m[0]) * v.x + double(m[3]) * v.y + double(m[6]) * v.z,
            double(m[1]) * v.x + double(m[4]) * v.y + double(m[7]) * v.z,
            double(m[2]) * v.x + double(m[5]) * v.y + double(m[8]) * v.z};
}

bool valid_index(uint32_t i) { return i < 4294967295u; }  // validDescriptorIndex (data.slang:135-137)

double srgb_eotf(double c) { return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4); }

// One bit per element; set from any thread.
struct Bits {
    std::vector<uint64_t> w;
    explicit Bits(size_t n = 0) : w((n + 63) / 64, 0) {}
    void set(size_t i) { __atomic_fetch_or(&w[i >> 6], uint64_t(1) << (i & 63), __ATOMIC_RELAXED); }
    uint64_t count() const {
        uint64_t n = 0;
        for (uint64_t x : w) n += uint64_t(__builtin_popcountll(x));
        return n;
    }
};

// Footprint recorder of orc_shade_footprint: every material-map texel (per texture, levels one
// after another), shadow-map texel, latent cell, triangle and vertex the shading reads.
struct Touch {
    std::vector<Bits> tex;                  // per texture
    std::vector<std::vector<size_t>> level_base;  // per texture: first texel of each level
    std::vector<const float*> maps;         // shadow maps, by cascade
    std::vector<Bits> map_bits;
    std::vector<Bits> latent;               // per decoder slot
    Bits tris, verts;
};
thread_local Touch* tl_touch = nullptr;

// A texture level: texels of level `level` of the description (levels stored one after another
// from level 0, as mip_levels says; 0 or 1 = level 0 only).
struct Tex {
    const rsy_texture_desc* d;
    int channels;
    bool srgb;
    int level = 0;
    int id = -1;  // index in the shading inputs' textures (footprint counting)
    int levels() const { return d->mip_levels > 1 ? int(d->mip_levels) : 1; }
    int width() const { return std::max(1, int(d->width) >> level); }
    int height() const { return std::max(1, int(d->height) >> level); }
    const uint8_t* texels() const {
        size_t o = 0;
        for (int l = 0; l < level; ++l)
            o += size_t(std::max(1, int(d->width) >> l)) * std::max(1, int(d->height) >> l) * channels;
        return static_cast<const uint8_t*>(d->pixels) + o;
    }
};

int wrap_coord(int i, int n, uint16_t mode) {
    if (mode == 0) return std::min(std::max(i, 0), n - 1);  // clamp to edge
    if (mode == 1) {                                       // mirrored repeat
        const int period = 2 * n;
        int m = ((i % period) + period) % period;
        return m < n ? m : period - 1 - m;
    }
    return ((i % n) + n) % n;  // repeat (default)
}

D4 texel(const Tex& t, int x, int y) {
    if (tl_touch && t.id >= 0)
        tl_touch->tex[size_t(t.id)].set(tl_touch->level_base[size_t(t.id)][size_t(t.level)] + size_t(y) * t.width() + x);
    const uint8_t* p = t.texels() + (size_t(y) * t.width() + x) * t.channels;
    D4 c{0, 0, 0, 1};
    double* o = &c.x;
    for (int k = 0; k < t.channels; ++k) {
        const double v = p[k] / 255.0;
        o[k] = (t.srgb && k < 3) ? srgb_eotf(v) : v;
    }
    return c;
}

D4 sample(const Tex& t, const rsy_sampler* s, double u, double v, bool minify = false) {
    const uint16_t ws = s ? s->wrap_s : 2, wt = s ? s->wrap_t : 2;
    const bool nearest = s && (minify ? s->min_filter : s->mag_filter) <= 2;
    const int W = t.width(), H = t.height();
    if (nearest) {
        const int x = wrap_coord(int(std::floor(u * W)), W, ws), y = wrap_coord(int(std::floor(v * H)), H, wt);
        return texel(t, x, y);
    }
    const double x = u * W - 0.5, y = v * H - 0.5;
    const double fx0 = std::floor(x), fy0 = std::floor(y);
    const double ax = x - fx0, ay = y - fy0;
    const int x0 = int(fx0), y0 = int(fy0);
    const int xa = wrap_coord(x0, W, ws), xb = wrap_coord(x0 + 1, W, ws);
    const int ya = wrap_coord(y0, H, wt), yb = wrap_coord(y0 + 1, H, wt);
    const D4 c00 = texel(t, xa, ya), c10 = texel(t, xb, ya), c01 = texel(t, xa, yb), c11 = texel(t, xb, yb);
    D4 r;
    const double* p00 = &c00.x;
    const double* p10 = &c10.x;
    const double* p01 = &c01.x;
    const double* p11 = &c11.x;
    double* o = &r.x;
    for (int k = 0; k < 4; ++k)
        o[k] = (1 - ax) * (1 - ay) * p00[k] + ax * (1 - ay) * p10[k] + (1 - ax) * ay * p01[k] + ax * ay * p11[k];
    return r;
}

// Vulkan's level of detail with mipmapMode LINEAR (the asset samplers, Graphics.cpp:4002-4012; no
// bias, no clamp): rho_x = |(du/dx W, dv/dx H)|, rho_y likewise, lambda = log2(max(rho_x, rho_y));
// lambda <= 0 magnifies from level 0, otherwise d = min(lambda, levels - 1) blends levels floor(d)
// and floor(d) + 1 (the last level past it) by frac(d) with the minification filter.
D4 sample_lod(const Tex& t, const rsy_sampler* s, double u, double v, double dudx, double dvdx, double dudy,
              double dvdy) {
    if (t.levels() <= 1) return sample(t, s, u, v);
    const double W = t.d->width, H = t.d->height;
    const double rho = std::max(std::hypot(dudx * W, dvdx * H), std::hypot(dudy * W, dvdy * H));
    const double lambda = std::log2(rho);
    if (!(lambda > 0.0)) return sample(t, s, u, v);
    const double d = std::min(lambda, double(t.levels() - 1));
    const int lo = int(std::floor(d)), hi = std::min(lo + 1, t.levels() - 1);
    const double f = d - std::floor(d);
    Tex a = t, b = t;
    a.level = lo;
    b.level = hi;
    const D4 ca = sample(a, s, u, v, true), cb = sample(b, s, u, v, true);
    return {ca.x + (cb.x - ca.x) * f, ca.y + (cb.y - ca.y) * f, ca.z + (cb.z - ca.z) * f, ca.w + (cb.w - ca.w) * f};
}

// One bilinear compare tap (SampleCmpLevelZero, GREATER_OR_EQUAL, CLAMP_TO_EDGE).
double cmp_tap(const float* map, int S, double u, double v, double ref, double& margin) {
    const double x = u * S - 0.5, y = v * S - 0.5;
    const double fx0 = std::floor(x), fy0 = std::floor(y);
    const double ax = x - fx0, ay = y - fy0;
    const int x0 = int(std::max(-2.0, std::min(double(S) + 1, fx0))), y0 = int(std::max(-2.0, std::min(double(S) + 1, fy0)));
    double r = 0.0;
    for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
            const int xi = std::min(std::max(x0 + i, 0), S - 1), yi = std::min(std::max(y0 + j, 0), S - 1);
            const double w = (i ? ax : 1 - ax) * (j ? ay : 1 - ay);
            const double t = map[size_t(yi) * S + xi];
            if (tl_touch)
                for (size_t k = 0; k < tl_touch->maps.size(); ++k)
                    if (tl_touch->maps[k] == map) tl_touch->map_bits[k].set(size_t(yi) * S + xi);
            if (w > 0.0) margin = std::min(margin, std::fabs(ref - t));
            r += w * (ref >= t ? 1.0 : 0.0);
        }
    return r;
}

// pcf.slang:1-12 (5 taps, offset 0.00025 UV) or a 3x3 grid of one-texel offsets.
double pcf(const float* map, int S, double x, double y, double ref, uint32_t mode, double& margin) {
    if (mode == RSY_PCF_GRID_3X3) {
        double r = 0.0;
        for (int j = -1; j <= 1; ++j)
            for (int i = -1; i <= 1; ++i) r += cmp_tap(map, S, x + double(i) / S, y + double(j) / S, ref, margin);
        return r * (1.0 / 9.0);
    }
    const double o = 0.00025;
    double r = 0.0;
    r += cmp_tap(map, S, x - o, y - o, ref, margin);
    r += cmp_tap(map, S, x + o * 2.0, y, ref, margin);
    r += cmp_tap(map, S, x, y + o * 2.0, ref, margin);
    r += cmp_tap(map, S, x + o * 2.0, y + o * 2.0, ref, margin);
    r += cmp_tap(map, S, x - o * 2.0, y - o * 2.0, ref, margin);
    return r * (1.0 / 5.0);
}

struct VOut {  // BasicVertexStageOutput (data.slang:3-22) + CSM
    D3 vertex, normal, world_normal, view_dir, light_dir, tangent, bitangent;
    D4 color;
    double u, v, sigma;
    D4 csm;
    D4 clip;
};

struct Ctx {
    const orc_geometry* g;
    const orc_shade_in* in;
    const orc_neural* nm = nullptr;
    std::vector<uint32_t> tri_base;
    std::vector<Tex> tex;
};

VOut vertex_main(const Ctx& c, const rsy_position& vin, const rsy_object_record& gd, const rsy_gpu_material* md,
                 const float* cascade_m) {
    const rsy_scene_data& sd = *c.in->scene;
    VOut o{};
    D4 nt{vin.tangents[0], vin.tangents[1], vin.tangents[2], vin.tangents[3]};
    if (sd.ignore_asset_tangent_sign == 1) nt.w = 1.0;
    if (sd.flip_tangents[0] == 1) nt.x = -nt.x;
    if (sd.flip_tangents[1] == 1) nt.y = -nt.y;
    if (sd.flip_tangents[2] == 1) nt.z = -nt.z;
    if (sd.flip_tangents[3] == 1) nt.w = -nt.w;
    bool tangent_space = sd.tangent_space_enabled == 1;
    if (md && !valid_index(md->normal_sampled_image_index)) tangent_space = false;
    const D3 vpos{vin.vertex[0], vin.vertex[1], vin.vertex[2]};
    const D3 vnorm{vin.normal[0], vin.normal[1], vin.normal[2]};
    D3 normal = vnorm, tangent = xyz(nt);
    D3 bitangent = cross(normal, tangent) * nt.w;
    if (sd.ensure_orthogonal_bitangent == 1) {
        const D3 np = bitangent * dot(normal, bitangent);
        const D3 on = normalize(normal - np);
        const D3 tp = np - bitangent * dot(tangent, bitangent);
        const D3 ot = normalize(tangent - tp);
        bitangent = cross(on, ot) * nt.w;
    }
    const D4 p4{vpos.x, vpos.y, vpos.z, 1.0};
    const D4 world = mul4(gd.transform, p4);
    const D4 cam{sd.camera_position[0], sd.camera_position[1], sd.camera_position[2], sd.camera_position[3]};
    const D4 sun{sd.sun_position[0], sd.sun_position[1], sd.sun_position[2], sd.sun_position[3]};
    D3 view_dir = xyz(cam) - xyz(world);
    D3 light_dir = xyz(sun) - xyz(world);
    if (tangent_space) {
        view_dir = xyz(mul4(gd.to_object_space_transform, cam)) - vpos;
        light_dir = xyz(mul4(gd.to_object_space_transform, sun)) - vpos;
    } else {
        normal = normalize(mul3(gd.normal_transform, vnorm));
        tangent = normalize(mul3(gd.normal_transform, xyz(nt)));
        tangent = normalize(tangent - normal * dot(normal, tangent));
    }
    if (sd.flip_lights[0] == 1) light_dir.x = -light_dir.x;
    if (sd.flip_lights[1] == 1) light_dir.y = -light_dir.y;
    if (sd.flip_lights[2] == 1) light_dir.z = -light_dir.z;
    o.vertex = xyz(world) * (1.0 / world.w);
    o.normal = normal;
    o.world_normal = mul3(gd.normal_transform, vnorm);
    o.color = {vin.color[0], vin.color[1], vin.color[2], vin.color[3]};
    o.tangent = xyz(nt);
    o.bitangent = bitangent;
    o.view_dir = view_dir;
    o.light_dir = light_dir;
    o.sigma = nt.w;
    o.u = vin.texture_coordinates[0];
    o.v = vin.texture_coordinates[1];
    if (cascade_m) {
        const D4 s = mul4(cascade_m, world);  // shadowProj * (W * p)
        o.csm = {0.5 * s.x + 0.5 * s.w, 0.5 * s.y + 0.5 * s.w, s.z, s.w};  // ndc_to_tc (basic.slang:24)
    }
    return o;
}

template <typename T>
T lerp3(const T& a, const T& b, const T& c, const double* w);
template <>
D3 lerp3(const D3& a, const D3& b, const D3& c, const double* w) {
    return a * w[0] + b * w[1] + c * w[2];
}
template <>
D4 lerp3(const D4& a, const D4& b, const D4& c, const double* w) {
    return {a.x * w[0] + b.x * w[1] + c.x * w[2], a.y * w[0] + b.y * w[1] + c.y * w[2],
            a.z * w[0] + b.z * w[1] + c.z * w[2], a.w * w[0] + b.w * w[1] + c.w * w[2]};
}

// Metallic/roughness channels: 2-channel maps carry (metallic, roughness) in (R, G); 4-channel maps
// follow the reference's glTF convention (G roughness, B metallic; basic.slang:208-220).
void mr_from(const Tex& t, const D4& s, double& rough, double& metal) {
    if (t.channels == 2) {
        metal = s.x;
        rough = s.y;
    } else {
        rough = s.y;
        metal = s.z;
    }
}

// GGX / Smith / Schlick (lighting.slang:5-43)
double distribution_ggx(D3 N, D3 H, double r) {
    const double a = r * r, a2 = a * a;
    const double ndh = std::max(dot(N, H), 0.0), ndh2 = ndh * ndh;
    double den = ndh2 * (a2 - 1.0) + 1.0;
    den = 3.14159265359 * den * den;
    return a2 / den;
}
double schlick_ggx(double ndv, double r) {
    const double rr = r + 1.0, k = rr * rr / 8.0;
    return ndv / (ndv * (1.0 - k) + k);
}
double geometry_smith(D3 N, D3 V, D3 L, double r) {
    return schlick_ggx(std::max(dot(N, L), 0.0), r) * schlick_ggx(std::max(dot(N, V), 0.0), r);
}
D3 schlick_fresnel(double vdh, D3 f0) {
    const double p = std::pow(clamp01(1.0 - vdh), 5.0);
    return f0 + (D3{1, 1, 1} - f0) * p;
}

void shade_pixel(const Ctx& c, uint32_t px, uint32_t py, uint64_t vis, double out[4], double* shadow_out,
                 double* margin_out, const float* override_shadow) {
    const orc_shade_in& in = *c.in;
    const rsy_scene_data& sd = *in.scene;
    const rsy_camera_view& view = *in.view;
    if (shadow_out) *shadow_out = 1.0;
    if (margin_out) *margin_out = std::numeric_limits<double>::infinity();
    if (vis == 0) {
        for (int k = 0; k < 4; ++k) out[k] = in.raster->clear_color[k];
        return;
    }
    const uint32_t tri = uint32_t(vis & 0xffffffffu) - 1u;
    const uint32_t d = uint32_t(std::upper_bound(c.tri_base.begin(), c.tri_base.end(), tri) - c.tri_base.begin()) - 1;
    const rsy_draw& dr = c.g->draws[d];
    const rsy_mesh_range& mesh = c.g->meshes[dr.mesh_index];
    const rsy_object_record& gd = c.g->objects[dr.graphic_objects_offset + dr.graphics_object_index];
    const rsy_gpu_material* md = dr.material_index < in.n_materials ? &in.materials[dr.material_index] : nullptr;
    const uint64_t vbase = mesh.vertex_buffer_offset / sizeof(rsy_position);
    const uint32_t local = tri - c.tri_base[d];
    const uint32_t first = mesh.index_offset + dr.start_index + 3 * local;

    // Perspective-correct barycentrics at the pixel centre.  When the triangle was rasterised
    // without clipping they come from the exact integer edge functions of the snapped triangle (the
    // one the visibility buffer holds), divided by each vertex's w; otherwise from the 2D
    // homogeneous form b ~ cross(x - u w, y - v w) over the unclipped clip positions.
    D4 clip[3];
    const rsy_position* vin[3];
    float mvp[16];
    orc::mat_mul_fma(view.view_projection, gd.transform, mvp);
    orc::C4 c32[3];
    if (tl_touch) tl_touch->tris.set(tri);
    for (int k = 0; k < 3; ++k) {
        vin[k] = &c.g->vertices[vbase + c.g->indices[first + k]];
        if (tl_touch) tl_touch->verts.set(vbase + c.g->indices[first + k]);
        const D4 p{vin[k]->vertex[0], vin[k]->vertex[1], vin[k]->vertex[2], 1.0};
        clip[k] = mul4(view.view_projection, mul4(gd.transform, p));
        c32[k] = orc::transform(mvp, vin[k]->vertex);
    }
    // With them, their derivatives along the pixel's x and y for the texture level of detail:
    // both forms are w3 = n / S with n linear in the pixel position over w, so dw3 = (dn - w3 dS) / S.
    double w3[3], dx3[3], dy3[3];
    if (orc::all_inside(c32)) {
        int32_t X[3], Y[3];
        float zs;
        for (int k = 0; k < 3; ++k) orc::snap(c32[k], view.width, view.height, X[k], Y[k], zs);
        const int64_t Px = int64_t(px) * 256 + 128, Py = int64_t(py) * 256 + 128;
        const int64_t e0 = (X[1] - Px) * (Y[2] - Py) - (X[2] - Px) * (Y[1] - Py);
        const int64_t e1 = (X[2] - Px) * (Y[0] - Py) - (X[0] - Px) * (Y[2] - Py);
        const int64_t e2 = (X[0] - Px) * (Y[1] - Py) - (X[1] - Px) * (Y[0] - Py);
        const double f0 = double(e0) / clip[0].w, f1 = double(e1) / clip[1].w, f2 = double(e2) / clip[2].w;
        const double s = f0 + f1 + f2;
        w3[0] = f0 / s;
        w3[1] = f1 / s;
        w3[2] = f2 / s;
        // e_k is linear in (Px, Py); one pixel is 256 sub-pixel units
        const double gx[3] = {256.0 * (Y[1] - Y[2]) / clip[0].w, 256.0 * (Y[2] - Y[0]) / clip[1].w,
                              256.0 * (Y[0] - Y[1]) / clip[2].w};
        const double gy[3] = {256.0 * (X[2] - X[1]) / clip[0].w, 256.0 * (X[0] - X[2]) / clip[1].w,
                              256.0 * (X[1] - X[0]) / clip[2].w};
        const double sx = gx[0] + gx[1] + gx[2], sy = gy[0] + gy[1] + gy[2];
        for (int k = 0; k < 3; ++k) {
            dx3[k] = (gx[k] - w3[k] * sx) / s;
            dy3[k] = (gy[k] - w3[k] * sy) / s;
        }
    } else {
        const double u = (px + 0.5) / view.width * 2.0 - 1.0, v = (py + 0.5) / view.height * 2.0 - 1.0;
        const D3 a{clip[0].x - u * clip[0].w, clip[1].x - u * clip[1].w, clip[2].x - u * clip[2].w};
        const D3 b{clip[0].y - v * clip[0].w, clip[1].y - v * clip[1].w, clip[2].y - v * clip[2].w};
        const D3 n = cross(a, b);
        const double s = n.x + n.y + n.z;
        // du/dx = 2 / width, dv/dy = 2 / height
        const D3 da{-clip[0].w * 2.0 / view.width, -clip[1].w * 2.0 / view.width, -clip[2].w * 2.0 / view.width};
        const D3 db{-clip[0].w * 2.0 / view.height, -clip[1].w * 2.0 / view.height, -clip[2].w * 2.0 / view.height};
        const D3 nx = cross(da, b), ny = cross(a, db);
        const double n3[3] = {n.x, n.y, n.z}, nx3[3] = {nx.x, nx.y, nx.z}, ny3[3] = {ny.x, ny.y, ny.z};
        for (int k = 0; k < 3; ++k) w3[k] = n3[k] / s;
        for (int k = 0; k < 3; ++k) {
            dx3[k] = (nx3[k] - w3[k] * (nx.x + nx.y + nx.z)) / s;
            dy3[k] = (ny3[k] - w3[k] * (ny.x + ny.y + ny.z)) / s;
        }
    }
    double duv[4] = {0, 0, 0, 0};  // du/dx, dv/dx, du/dy, dv/dy
    for (int k = 0; k < 3; ++k) {
        const double tu = vin[k]->texture_coordinates[0], tv = vin[k]->texture_coordinates[1];
        duv[0] += dx3[k] * tu;
        duv[1] += dx3[k] * tv;
        duv[2] += dy3[k] * tu;
        duv[3] += dy3[k] * tv;
    }

    const double view_depth = clip[0].w * w3[0] + clip[1].w * w3[1] + clip[2].w * w3[2];

    // cascade selection by view depth against the split distances
    int cascade = -1;
    for (uint32_t k = 0; k < in.n_cascades; ++k)
        if (view_depth <= double(in.cascades[k].split_far)) {
            cascade = int(k);
            break;
        }
    const float* cm = cascade >= 0 ? in.cascades[cascade].light_view_projection : nullptr;

    VOut vo[3];
    for (int k = 0; k < 3; ++k) vo[k] = vertex_main(c, *vin[k], gd, md, cm);
    VOut bv;
    bv.vertex = lerp3(vo[0].vertex, vo[1].vertex, vo[2].vertex, w3);
    bv.normal = lerp3(vo[0].normal, vo[1].normal, vo[2].normal, w3);
    bv.world_normal = lerp3(vo[0].world_normal, vo[1].world_normal, vo[2].world_normal, w3);
    bv.view_dir = lerp3(vo[0].view_dir, vo[1].view_dir, vo[2].view_dir, w3);
    bv.light_dir = lerp3(vo[0].light_dir, vo[1].light_dir, vo[2].light_dir, w3);
    bv.tangent = lerp3(vo[0].tangent, vo[1].tangent, vo[2].tangent, w3);
    bv.bitangent = lerp3(vo[0].bitangent, vo[1].bitangent, vo[2].bitangent, w3);
    bv.color = lerp3(vo[0].color, vo[1].color, vo[2].color, w3);
    bv.u = vo[0].u * w3[0] + vo[1].u * w3[1] + vo[2].u * w3[2];
    bv.v = vo[0].v * w3[0] + vo[1].v * w3[1] + vo[2].v * w3[2];
    bv.sigma = vo[0].sigma * w3[0] + vo[1].sigma * w3[1] + vo[2].sigma * w3[2];
    bv.csm = lerp3(vo[0].csm, vo[1].csm, vo[2].csm, w3);

    // fragmentMain (basic.slang:137-344)
    const D3 ambient{sd.ambient_color[0], sd.ambient_color[1], sd.ambient_color[2]};
    D4 out_color = bv.color;
    D4 image_color = out_color;
    double alpha = bv.color.w;
    double sampled_rough = 0.0, sampled_metal = 0.0;
    D3 light_normal = normalize(bv.normal);
    const D3 light_dir = normalize(bv.light_dir);
    const D3 bitangent = bv.bitangent;
    if (md) {
        out_color = {md->color[0], md->color[1], md->color[2], md->color[3]};
        image_color = out_color;
        alpha = std::min(double(md->color[3]), alpha);
        const bool decoded = c.nm && dr.material_index < c.nm->n_materials &&
                             c.nm->slot_of[dr.material_index] != 0xffffffffu &&
                             (!c.nm->active || c.nm->active[dr.material_index]);
        if (decoded) {
            // the decoder's outputs stand in for the three samples (two-channel normal map)
            double dec[7];
            orc::nm_decode_point(c.nm, dr.material_index, bv.u, bv.v, dec);
            if (tl_touch) {
                uint64_t cells[4];
                orc::nm_cells(c.nm, bv.u, bv.v, cells);
                for (uint64_t cell : cells) tl_touch->latent[c.nm->slot_of[dr.material_index]].set(cell);
            }
            if (sd.normal_maps_enabled == 1) {
                D3 m{dec[3] * 2 - 1, dec[4] * 2 - 1, 0.0};
                m.z = std::sqrt(std::max(0.0, 1.0 - m.x * m.x - m.y * m.y));
                const D3 n = normalize(bv.normal);
                const D3 t = normalize(bv.tangent - n * dot(bv.tangent, n));
                const D3 bb = normalize(cross(bv.normal, bv.tangent) * bv.sigma);
                light_normal = normalize(t * m.x + bb * m.y + n * m.z);
            }
            image_color = {dec[0], dec[1], dec[2], 1.0};
            alpha = std::min(1.0, alpha);
            sampled_metal = dec[5];
            sampled_rough = dec[6];
        } else {
            const uint32_t nti = md->normal_sampled_image_index;
            if (sd.normal_maps_enabled == 1 && valid_index(nti) && nti < c.tex.size()) {
                const rsy_sampler* smp = md->normal_sampler_index < in.n_samplers ? &in.samplers[md->normal_sampler_index] : nullptr;
                const D4 sm = sample_lod(c.tex[nti], smp, bv.u, bv.v, duv[0], duv[1], duv[2], duv[3]);
                D3 m{sm.x * 2 - 1, sm.y * 2 - 1, sm.z * 2 - 1};
                if (c.tex[nti].channels == 2) m.z = std::sqrt(std::max(0.0, 1.0 - m.x * m.x - m.y * m.y));
                const D3 n = normalize(bv.normal);
                const D3 t = normalize(bv.tangent - n * dot(bv.tangent, n));
                const D3 bb = normalize(cross(bv.normal, bv.tangent) * bv.sigma);
                light_normal = normalize(t * m.x + bb * m.y + n * m.z);
            }
            const uint32_t cti = md->color_sampled_image_index;
            if (valid_index(cti) && cti < c.tex.size()) {
                const rsy_sampler* smp = md->color_sampler_index < in.n_samplers ? &in.samplers[md->color_sampler_index] : nullptr;
                image_color = sample_lod(c.tex[cti], smp, bv.u, bv.v, duv[0], duv[1], duv[2], duv[3]);
                alpha = std::min(image_color.w, alpha);
            }
            const uint32_t mixti = md->mixmap_sampled_image_index, metti = md->metallic_sampled_image_index;
            if (valid_index(mixti) && mixti < c.tex.size()) {
                const rsy_sampler* smp = md->mixmap_sampler_index < in.n_samplers ? &in.samplers[md->mixmap_sampler_index] : nullptr;
                mr_from(c.tex[mixti], sample_lod(c.tex[mixti], smp, bv.u, bv.v, duv[0], duv[1], duv[2], duv[3]), sampled_rough,
                        sampled_metal);
            } else if (valid_index(metti) && metti < c.tex.size()) {
                const rsy_sampler* smp = md->metallic_sampler_index < in.n_samplers ? &in.samplers[md->metallic_sampler_index] : nullptr;
                mr_from(c.tex[metti], sample_lod(c.tex[metti], smp, bv.u, bv.v, duv[0], duv[1], duv[2], duv[3]), sampled_rough,
                        sampled_metal);
            }
        }
        // neither map present: the reference reads an invalid descriptor (basic.slang:211-213);
        // here roughness and metallic stay 0.
    }
    const D3 V = normalize(bv.view_dir);
    const D3 L = light_dir;
    const D3 N = light_normal;
    const D3 H = normalize(L + V);
    const double cos_theta = clamp01(dot(L, N)), cos_phi = clamp01(dot(H, N));
    const double l_dot_n = cos_theta, l_dot_h = clamp01(dot(L, H)), n_dot_v = clamp01(dot(N, V));
    auto set3 = [&](D3 c3) { out_color = {c3.x, c3.y, c3.z, alpha}; };
    switch (sd.fragment_output) {
        case 0: set3(xyz(image_color)); break;
        case 1: set3(normalize(bv.world_normal) * 0.5 + D3{0.5, 0.5, 0.5}); break;
        case 2: set3(normalize(bv.tangent) * 0.5 + D3{0.5, 0.5, 0.5}); break;
        case 3: set3(normalize(bv.light_dir) * 0.5 + D3{0.5, 0.5, 0.5}); break;
        case 4: set3(normalize(bv.view_dir) * 0.5 + D3{0.5, 0.5, 0.5}); break;
        case 5: set3(xyz(bv.color)); break;
        case 6: set3(light_normal * 0.5 + D3{0.5, 0.5, 0.5}); break;
        case 7: set3(bitangent * 0.5 + D3{0.5, 0.5, 0.5}); break;
        case 8: set3(D3{cos_theta, cos_theta, cos_theta} * 0.5 + D3{0.5, 0.5, 0.5}); break;
        case 9: set3(D3{cos_phi, cos_phi, cos_phi} * 0.5 + D3{0.5, 0.5, 0.5}); break;
        case 10: set3(D3{l_dot_h, l_dot_h, l_dot_h} * 0.5 + D3{0.5, 0.5, 0.5}); break;
        case 11: set3(D3{n_dot_v, n_dot_v, n_dot_v} * 0.5 + D3{0.5, 0.5, 0.5}); break;
        case 12: set3(D3{bv.u, bv.v, 0.0}); break;
        case 13: set3(D3{sampled_rough, 0.05, 0.05}); break;
        case 14: set3(D3{sampled_metal, 0.05, 0.05}); break;
        case 15: set3(D3{1 - alpha, 0.05, 0.05}); break;
        default: break;
    }
    if (sd.light_enabled > 0) {
        double shadow = 1.0;
        if (sd.shadows_enabled == 1 && cascade >= 0) {
            double margin = std::numeric_limits<double>::infinity();
            shadow = pcf(in.shadow_maps[cascade], int(in.cascades[cascade].size), bv.csm.x, bv.csm.y, bv.csm.z,
                         in.raster->pcf_mode, margin);
            if (margin_out) *margin_out = margin;
        }
        if (shadow_out) *shadow_out = shadow;
        if (override_shadow) shadow = override_shadow[size_t(py) * view.width + px];
        if (sd.brdf_lighting_enabled == 0) {
            const D3 sun_light = D3{sd.sunlight_color[0], sd.sunlight_color[1], sd.sunlight_color[2]} * std::max(cos_theta, 0.0);
            const D3 lit = sun_light * shadow;
            const D3 f{std::max(ambient.x, lit.x), std::max(ambient.y, lit.y), std::max(ambient.z, lit.z)};
            out_color = {out_color.x * f.x, out_color.y * f.y, out_color.z * f.z, out_color.w * alpha};
        } else {
            const double roughness = sampled_rough, metallic = sampled_metal;
            const D3 albedo = xyz(out_color);
            const D3 f0 = D3{0.04, 0.04, 0.04} + (albedo - D3{0.04, 0.04, 0.04}) * metallic;
            const double ndf = distribution_ggx(N, H, roughness);
            const double gg = geometry_smith(N, V, L, roughness);
            const D3 F = schlick_fresnel(std::max(dot(H, V), 0.0), f0);
            const D3 kD = (D3{1, 1, 1} - F) * (1.0 - metallic);
            const D3 numerator = F * (ndf * gg);
            const double denominator = 4.0 * n_dot_v * l_dot_n + 0.0001;
            const D3 specular = numerator * (1.0 / denominator);
            const D3 diffuse = albedo * (1.0 / 3.14159265);
            const D3 sun_light = D3{sd.sunlight_color[0], sd.sunlight_color[1], sd.sunlight_color[2]} * sd.light_intensity;
            const D3 brdf = (kD * diffuse + specular) * sun_light * l_dot_n;
            const D3 amb = ambient * diffuse;
            const D3 sh = brdf * shadow;
            out_color = {std::max(amb.x, sh.x), std::max(amb.y, sh.y), std::max(amb.z, sh.z), alpha};
        }
    }
    if (md) out_color.w = alpha;
    out[0] = out_color.x;
    out[1] = out_color.y;
    out[2] = out_color.z;
    out[3] = out_color.w;
}

}  // namespace

extern "C" {

double orc_srgb_to_linear(uint32_t code) { return srgb_eotf(double(code & 0xffu) / 255.0); }

double orc_ggx_d(double n_dot_h, double roughness) {
    return distribution_ggx(D3{0, 0, 1}, D3{std::sqrt(std::max(0.0, 1 - n_dot_h * n_dot_h)), 0, n_dot_h}, roughness);
}
double orc_schlick_ggx(double n_dot_v, double roughness) { return schlick_ggx(n_dot_v, roughness); }
double orc_schlick_fresnel(double v_dot_h, double f0) { return schlick_fresnel(v_dot_h, D3{f0, f0, f0}).x; }
double orc_pcf(const float* map, uint32_t size, double u, double v, double ref, uint32_t mode, double* margin) {
    double m = std::numeric_limits<double>::infinity();
    const double r = pcf(map, int(size), u, v, ref, mode, m);
    if (margin) *margin = m;
    return r;
}

static int shade_view(const orc_geometry* g, const orc_shade_in* in, const orc_neural* nm, const uint64_t* vis,
                      double* out_rgba, double* out_shadow, double* out_margin, const float* override_shadow,
                      int n_threads, Touch* touch = nullptr) {
    if (!g || !in || !vis || (!out_rgba && !touch) || !in->scene || !in->raster || !in->view) return RSY_INVALID_ARGUMENT;
    Ctx c;
    c.g = g;
    c.in = in;
    c.nm = nm;
    uint32_t base = 0;
    for (uint32_t d = 0; d < g->n_draws; ++d) {
        c.tri_base.push_back(base);
        base += g->draws[d].index_count / 3;
    }
    for (uint32_t t = 0; t < in->n_textures; ++t) {
        // the oracle samples stored texels: BC7 images and built chains come decoded / built, with
        // their levels given
        const rsy_texture_desc& d = in->textures[t];
        if (d.format > RSY_TEX_RG8_SRGB || d.mip_levels == RSY_MIPS_BUILD) return RSY_INVALID_ARGUMENT;
        Tex x{&d, (d.format == RSY_TEX_RG8_UNORM || d.format == RSY_TEX_RG8_SRGB) ? 2 : 4,
              d.format == RSY_TEX_RGBA8_SRGB || d.format == RSY_TEX_RG8_SRGB};
        x.id = int(t);
        c.tex.push_back(x);
    }
    const uint32_t W = in->view->width, H = in->view->height;
    const int nt = std::max(1, n_threads);
    auto work = [&](int t) {
        tl_touch = touch;
        double scratch[4];
        for (uint32_t y = uint32_t(uint64_t(H) * t / nt); y < uint32_t(uint64_t(H) * (t + 1) / nt); ++y)
            for (uint32_t x = 0; x < W; ++x) {
                const size_t o = size_t(y) * W + x;
                shade_pixel(c, x, y, vis[o], out_rgba ? out_rgba + 4 * o : scratch, out_shadow ? out_shadow + o : nullptr,
                            out_margin ? out_margin + o : nullptr, override_shadow);
            }
        tl_touch = nullptr;
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
    return RSY_OK;
}

int orc_shade_view(const orc_geometry* g, const orc_shade_in* in, const uint64_t* vis, double* out_rgba,
                   double* out_shadow, double* out_margin, const float* override_shadow, int n_threads) {
    return shade_view(g, in, nullptr, vis, out_rgba, out_shadow, out_margin, override_shadow, n_threads);
}

int orc_shade_view_neural(const orc_geometry* g, const orc_shade_in* in, const orc_neural* nm, const uint64_t* vis,
                          double* out_rgba, double* out_shadow, double* out_margin, const float* override_shadow,
                          int n_threads) {
    if (!nm || !nm->slot_of || !nm->latents || !nm->mlp) return RSY_INVALID_ARGUMENT;
    return shade_view(g, in, nm, vis, out_rgba, out_shadow, out_margin, override_shadow, n_threads);
}

int orc_shade_footprint(const orc_geometry* g, const orc_shade_in* in, const orc_neural* nm, const uint64_t* vis,
                        orc_footprint* out, int n_threads) {
    if (!g || !in || !vis || !out || !in->view) return RSY_INVALID_ARGUMENT;
    if (nm && (!nm->slot_of || !nm->latents || !nm->mlp)) return RSY_INVALID_ARGUMENT;
    Touch touch;
    uint64_t n_tris = 0;
    for (uint32_t d = 0; d < g->n_draws; ++d) n_tris += g->draws[d].index_count / 3;
    touch.tris = Bits(n_tris);
    touch.verts = Bits(g->n_vertices);
    for (uint32_t t = 0; t < in->n_textures; ++t) {
        const rsy_texture_desc& d = in->textures[t];
        std::vector<size_t> base;
        size_t n = 0;
        const int levels = d.mip_levels > 1 && d.mip_levels != RSY_MIPS_BUILD ? int(d.mip_levels) : 1;
        for (int l = 0; l < levels; ++l) {
            base.push_back(n);
            n += size_t(std::max(1u, d.width >> l)) * std::max(1u, d.height >> l);
        }
        touch.level_base.push_back(base);
        touch.tex.emplace_back(n);
    }
    for (uint32_t k = 0; k < in->n_cascades; ++k) {
        touch.maps.push_back(in->shadow_maps[k]);
        touch.map_bits.emplace_back(size_t(in->cascades[k].size) * in->cascades[k].size);
    }
    if (nm) {
        uint32_t slots = 0;
        for (uint32_t m = 0; m < nm->n_materials; ++m)
            if (nm->slot_of[m] != 0xffffffffu) slots = std::max(slots, nm->slot_of[m] + 1);
        for (uint32_t k = 0; k < slots; ++k) touch.latent.emplace_back(size_t(nm->latent_size) * nm->latent_size);
    }
    const int r = shade_view(g, in, nm, vis, nullptr, nullptr, nullptr, nullptr, n_threads, &touch);
    if (r != RSY_OK) return r;
    orc_footprint f{};
    const size_t npx = size_t(in->view->width) * in->view->height;
    for (size_t i = 0; i < npx; ++i) f.covered_pixels += vis[i] != 0;
    f.visible_triangles = touch.tris.count();
    f.visible_vertices = touch.verts.count();
    for (uint32_t t = 0; t < in->n_textures; ++t) {
        const rsy_texture_desc& d = in->textures[t];
        const uint64_t n = touch.tex[t].count();
        f.texels += n;
        f.texel_bytes += n * ((d.format == RSY_TEX_RG8_UNORM || d.format == RSY_TEX_RG8_SRGB) ? 2 : 4);
    }
    for (const Bits& b : touch.map_bits) f.shadow_texels += b.count();
    for (const Bits& b : touch.latent) f.latent_cells += b.count();
    *out = f;
    return RSY_OK;
}

int orc_sample_texture(const rsy_texture_desc* tex, const rsy_sampler* sampler, double u, double v, double out[4]) {
    if (!tex || !out || !tex->pixels) return RSY_INVALID_ARGUMENT;
    const Tex t{tex, (tex->format == RSY_TEX_RG8_UNORM || tex->format == RSY_TEX_RG8_SRGB) ? 2 : 4,
                tex->format == RSY_TEX_RGBA8_SRGB || tex->format == RSY_TEX_RG8_SRGB};
    const D4 r = sample(t, sampler, u, v);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
    return RSY_OK;
}

}  // extern "C"
