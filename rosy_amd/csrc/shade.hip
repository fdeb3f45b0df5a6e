// Deferred shading kernel (FP32) over the visibility buffers.
//
// For each pixel the visible triangle's three vertices go through vertexMain (basic.slang:15-119),
// the varyings are interpolated perspective-correctly at the pixel centre (clip-space barycentrics
// from the transformed vertices of the raster pass), and fragmentMain (basic.slang:137-344) runs with
// the GGX/Smith/Schlick terms of lighting.slang:5-43 and the bilinear-compare PCF of pcf.slang:1-12
// (compare sampler Graphics.cpp:2697-2725).  Textures are sampled at mip 0 with sRGB decode before
// filtering through a 256-entry table in LDS.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "bin_raster.h"
#include "fixed_point.h"
#include "frame_types.h"
#include "kernels.h"
#include "neural.h"
#include "rosy_amd.h"
#include "texture.h"

namespace rsy {
namespace {

struct f3 {
    float x, y, z;
};
__device__ inline f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// Unit vector by the hardware reciprocal square root (v_rsq_f32, 1 ulp): lighting directions only --
// the barycentrics, texture coordinates and shadow-map coordinates use correctly rounded arithmetic.
// (The normal maps' third component likewise uses the hardware square root, v_sqrt_f32.)
__device__ inline f3 normalize(f3 a) { return a * __builtin_amdgcn_rsqf(dot(a, a)); }
__device__ inline float clamp01(float x) { return fminf(1.0f, fmaxf(0.0f, x)); }

template <class MP>
__device__ inline float4 mul4(MP m, float4 v) {
    float4 r;
    r.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
    r.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
    r.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
    r.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
    return r;
}
template <class MP>
__device__ inline f3 mul3(MP m, f3 v) {
    return {m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
            m[2] * v.x + m[5] * v.y + m[8] * v.z};
}
__device__ inline f3 xyz(float4 v) { return {v.x, v.y, v.z}; }

__device__ inline bool valid_index(uint32_t i) { return i < 4294967295u; }

__device__ inline float cmp_tap(GlobalPtr<const float> map, int S, float u, float v, float ref) {
    const float x = u * float(S) - 0.5f, y = v * float(S) - 0.5f;
    const float fx0 = floorf(fminf(fmaxf(x, -2.0f), float(S) + 1.0f));
    const float fy0 = floorf(fminf(fmaxf(y, -2.0f), float(S) + 1.0f));
    const float ax = x - fx0, ay = y - fy0;
    const int x0 = int(fx0), y0 = int(fy0);
    const int xa = min(max(x0, 0), S - 1), xb = min(max(x0 + 1, 0), S - 1);
    const int ya = min(max(y0, 0), S - 1), yb = min(max(y0 + 1, 0), S - 1);
    const float t00 = map[size_t(ya) * S + xa], t10 = map[size_t(ya) * S + xb];
    const float t01 = map[size_t(yb) * S + xa], t11 = map[size_t(yb) * S + xb];
    return (1 - ax) * (1 - ay) * (ref >= t00 ? 1.f : 0.f) + ax * (1 - ay) * (ref >= t10 ? 1.f : 0.f) +
           (1 - ax) * ay * (ref >= t01 ? 1.f : 0.f) + ax * ay * (ref >= t11 ? 1.f : 0.f);
}

// Frame constants of a shading launch.  The scene data is read through the constant address space
// (scalar loads, shared by the waves of a CU).  The sRGB table and the cascades (selected per
// pixel) are copied into LDS at the start of each workgroup: shared_request issues the loads (one
// word or two per thread, beside the workgroup's first visibility loads) and shared_store writes
// them; the caller synchronises before the first use.
using SceneP = const __attribute__((address_space(4))) rsy_scene_data*;
struct ShadeShared {
    float lut[256];
    ShadeCascade casc[kMaxShadeCascades];
};
constexpr uint32_t kCascadeWords = uint32_t(sizeof(ShadeCascade) / 4);
static_assert(kMaxShadeCascades * kCascadeWords <= 512, "one or two words per thread");
struct SharedWords {
    float lut;
    uint32_t c0, c1;
};
__device__ inline SharedWords shared_request(const ShadeLaunch& L) {
    const uint32_t t = threadIdx.x;
    const uint32_t cw = min(L.n_cascades, kMaxShadeCascades) * kCascadeWords;
    const uint32_t* cs = reinterpret_cast<const uint32_t*>(L.cascades);
    SharedWords w;
    w.lut = L.srgb_lut[t];
    w.c0 = t < cw ? cs[t] : 0u;
    w.c1 = t + 256 < cw ? cs[t + 256] : 0u;
    return w;
}
__device__ inline void shared_store(const ShadeLaunch& L, ShadeShared& S, const SharedWords& w) {
    const uint32_t t = threadIdx.x;
    const uint32_t cw = min(L.n_cascades, kMaxShadeCascades) * kCascadeWords;
    S.lut[t] = w.lut;
    if (t < cw) reinterpret_cast<uint32_t*>(S.casc)[t] = w.c0;
    if (t + 256 < cw) reinterpret_cast<uint32_t*>(S.casc)[t + 256] = w.c1;
}

// Varyings of basicVertex (data.slang:10-22) plus csm.near, interpolated at the pixel.
struct Varyings {
    f3 normal, world_normal, view_dir, light_dir, tangent, bitangent;
    float4 color;
    float u, v, sigma;
    float su, sv;  // whole periods removed from u, v (see below)
    float4 csm;
};

// vertexMain (basic.slang:15-119) evaluated at the pixel.  Every output is affine in the vertex
// inputs except the bitangent (a cross product per vertex) and, with tangent space off, the
// normalised world normal, so those two are formed per vertex and interpolated, and the rest is
// formed once from the interpolated position, normal, tangent, colour and uv: the same values as
// running vertexMain per vertex and interpolating its outputs.
template <class GP>
__device__ Varyings pixel_varyings(SceneP sd, const rsy_position* const v[3], GP gd,
                                   bool has_material, uint32_t normal_index, const float* cascade_m, float w0,
                                   float w1, float w2, bool shift_uv) {
    Varyings o;
    const float w[3] = {w0, w1, w2};
    bool tangent_space = sd->tangent_space_enabled == 1;
    if (has_material && !valid_index(normal_index)) tangent_space = false;
    const GP W = gd;
    const GP toObj = gd + 16;
    const GP nmat = gd + 32;
    f3 P = mk(0.f, 0.f, 0.f), N = P, Tn = P, B = P, Nw = P;
    float sigma = 0.f;
    float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
    // the three vertex records (64 bytes each) are requested together before any of them is used
    float4 raw[3][4];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float4* p = reinterpret_cast<const float4*>(v[k]);
#pragma unroll
        for (int q = 0; q < 4; ++q) raw[k][q] = p[q];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        // vertex xyz, normal xyz, tangent xyzw, colour rgba, uv (the rsy_position layout)
        const float4 r0 = raw[k][0], r1 = raw[k][1], r2 = raw[k][2], r3 = raw[k][3];
        float4 nt = make_float4(r1.z, r1.w, r2.x, r2.y);
        if (sd->ignore_asset_tangent_sign == 1) nt.w = 1.f;
        if (sd->flip_tangents[0] == 1) nt.x = -nt.x;
        if (sd->flip_tangents[1] == 1) nt.y = -nt.y;
        if (sd->flip_tangents[2] == 1) nt.z = -nt.z;
        if (sd->flip_tangents[3] == 1) nt.w = -nt.w;
        const f3 n = mk(r0.w, r1.x, r1.y);
        const f3 t = mk(nt.x, nt.y, nt.z);
        f3 bt = cross(n, t) * nt.w;
        if (sd->ensure_orthogonal_bitangent == 1) {
            const f3 np = bt * dot(n, bt);
            const f3 on = normalize(n - np);
            const f3 tp = np - bt * dot(t, bt);
            const f3 ot = normalize(t - tp);
            bt = cross(on, ot) * nt.w;
        }
        P = P + mk(r0.x, r0.y, r0.z) * w[k];
        N = N + n * w[k];
        Tn = Tn + t * w[k];
        B = B + bt * w[k];
        sigma += nt.w * w[k];
        col.x += r2.z * w[k];
        col.y += r2.w * w[k];
        col.z += r3.x * w[k];
        col.w += r3.y * w[k];
        if (!tangent_space) Nw = Nw + normalize(mul3(nmat, n)) * w[k];
    }
    const float4 world = mul4(W, make_float4(P.x, P.y, P.z, 1.f));
    const float4 cam = make_float4(sd->camera_position[0], sd->camera_position[1], sd->camera_position[2], sd->camera_position[3]);
    const float4 sun = make_float4(sd->sun_position[0], sd->sun_position[1], sd->sun_position[2], sd->sun_position[3]);
    if (tangent_space) {
        o.normal = N;
        o.view_dir = xyz(mul4(toObj, cam)) - P;
        o.light_dir = xyz(mul4(toObj, sun)) - P;
    } else {
        o.normal = Nw;
        o.view_dir = xyz(cam) - xyz(world);
        o.light_dir = xyz(sun) - xyz(world);
    }
    if (sd->flip_lights[0] == 1) o.light_dir.x = -o.light_dir.x;
    if (sd->flip_lights[1] == 1) o.light_dir.y = -o.light_dir.y;
    if (sd->flip_lights[2] == 1) o.light_dir.z = -o.light_dir.z;
    o.world_normal = mul3(nmat, N);
    o.tangent = Tn;
    o.bitangent = B;
    o.color = col;
    o.sigma = sigma;
    // Texture coordinates relative to vertex 0 with a whole even number of periods removed: exact
    // for REPEAT and MIRRORED_REPEAT and keeps |uv| small, so the FP32 texel address keeps its
    // sub-texel precision at large uv.  Clamp-to-edge samplers keep the raw value.
    const float u0 = raw[0][3].z, v0c = raw[0][3].w;
    o.su = shift_uv ? 2.0f * floorf(u0 * 0.5f) : 0.f;
    o.sv = shift_uv ? 2.0f * floorf(v0c * 0.5f) : 0.f;
    o.u = (u0 - o.su) + w1 * (raw[1][3].z - u0) + w2 * (raw[2][3].z - u0);
    o.v = (v0c - o.sv) + w1 * (raw[1][3].w - v0c) + w2 * (raw[2][3].w - v0c);
    o.csm = make_float4(0.f, 0.f, 0.f, 0.f);
    if (cascade_m) {
        const float4 s4 = mul4(cascade_m, world);
        o.csm = make_float4(0.5f * s4.x + 0.5f * s4.w, 0.5f * s4.y + 0.5f * s4.w, s4.z, s4.w);  // ndc_to_tc
    }
    return o;
}

__device__ inline void mr_from(int channels, float4 s, float& rough, float& metal) {
    if (channels == 2) {
        metal = s.x;
        rough = s.y;
    } else {
        rough = s.y;
        metal = s.z;
    }
}

// GGX D = a^2 / (pi (NdotH^2 (a^2 - 1) + 1)^2).  The denominator is evaluated as
// (1 - NdotH^2) + NdotH^2 a^2 with 1 - NdotH^2 = |N x H|^2 (unit N and H): near the lobe's peak at
// low roughness, 1 - NdotH^2 from an FP32 dot product loses most of its digits (N and H are unit
// only to an ulp, which the cross product's length does not amplify), and D there runs to 10^4.
__device__ inline float distribution_ggx(f3 N, f3 H, float r) {
    const float a = r * r, a2 = a * a;
    const float ndh = fmaxf(dot(N, H), 0.f), ndh2 = ndh * ndh;
    const f3 c = cross(N, H);
    const float s2 = ndh > 0.f ? dot(c, c) : 1.f;  // NdotH clamped to 0 leaves the denominator 1
    float den = s2 + ndh2 * a2;
    den = 3.14159265359f * den * den;
    return a2 * __builtin_amdgcn_rcpf(den);  // hardware reciprocal (1 ulp), as in the other lighting terms
}
__device__ inline float schlick_ggx(float ndv, float r) {
    const float rr = r + 1.f, k = rr * rr * 0.125f;
    return ndv * __builtin_amdgcn_rcpf(ndv * (1.f - k) + k);
}

// Visible triangle of a covered pixel: its draw, vertex indices, barycentrics and cascade.
struct PixelGeom {
    uint32_t draw;        // index in the view's draw list
    uint32_t material;    // the draw's material index
    uint32_t v0, v1, v2;  // indices into the global vertex buffer
    f3 lam;               // perspective-correct barycentrics at the pixel centre
    f3 dlx, dly;          // with kLod: their derivatives along the pixel's x and y
    int cascade;
};

// kLod: also the barycentrics' screen derivatives (for the texture level of detail), exact for the
// same rational form: lam = n / S with n linear over w, so d(lam) = (dn - lam dS) / S.
template <bool kLod = false>
__device__ inline PixelGeom pixel_geometry(const ShadeLaunch& L, const ShadeShared& S, const ShadeView& V, uint32_t x,
                                           uint32_t y, unsigned long long vis) {
    PixelGeom pg;
    const uint32_t tri = uint32_t(vis & 0xffffffffull) - 1u;
    // the triangle's shading record (vertices, draw, clip w) and its raster record, loaded together
    // (a clipped triangle's raster slot holds stale data, computed with and then replaced)
    const GlobalPtr<const uint4> sr = gmem(reinterpret_cast<const uint4*>(V.shade_recs + tri));
    const uint4 sa = sr[0], sb = sr[1];
    const uint4* rq = reinterpret_cast<const uint4*>(V.recs + tri);
    const uint4 q0 = gmem(rq)[0];  // X0, Y0, X1, Y1
    const uint2 q1 = *gmem(reinterpret_cast<const uint2*>(rq + 1));  // X2, Y2
    const uint32_t id = gmem(reinterpret_cast<const uint32_t*>(rq + 2))[1];
    pg.draw = sa.w & ~kShadeNoRecord;
    pg.material = sb.w;
    pg.v0 = sa.x;
    pg.v1 = sa.y;
    pg.v2 = sa.z;
    const float w0 = __uint_as_float(sb.x), w1 = __uint_as_float(sb.y), w2 = __uint_as_float(sb.z);

    // Perspective-correct barycentrics at the pixel centre.  A triangle that needed no clipping has
    // a record in the triangle's slot (its box says so): the exact integer edge functions of the
    // snapped vertices the raster pass covered, each over its vertex's w (as e_k * w_i * w_j,
    // normalised below).  The record may hold vertices 1 and 2 exchanged, which negates all three
    // edge functions and swaps the last two.  A clipped triangle's slot is empty; its barycentrics
    // come from b ~ cross(x - u w, y - v w) over the clip positions.  The edge form is evaluated for
    // every pixel (on a clipped triangle's stale slot too, then replaced), so that the raster
    // record is requested beside the shading record rather than after it.
    f3 lam;
    {
        const int32_t X0 = int32_t(q0.x), Y0 = int32_t(q0.y), X1 = int32_t(q0.z), Y1 = int32_t(q0.w);
        const int32_t X2 = int32_t(q1.x), Y2 = int32_t(q1.y);
        const int32_t Px = int32_t(x) * 256 + 128, Py = int32_t(y) * 256 + 128;
        const float e0 = float(int64_t(X1 - Px) * (Y2 - Py) - int64_t(X2 - Px) * (Y1 - Py));
        const float e1 = float(int64_t(X2 - Px) * (Y0 - Py) - int64_t(X0 - Px) * (Y2 - Py));
        const float e2 = float(int64_t(X0 - Px) * (Y1 - Py) - int64_t(X1 - Px) * (Y0 - Py));
        const bool swapped = (id & kRecSwapped) != 0u;
        const float b1 = swapped ? e2 : e1, b2 = swapped ? e1 : e2;
        lam = mk(e0 * (w1 * w2), b1 * (w0 * w2), b2 * (w0 * w1));
        if constexpr (kLod) {
            // d e_k per pixel step (256 sub-pixel units), vertices in the record's order
            const float ex0 = 256.0f * float(Y1 - Y2), ey0 = 256.0f * float(X2 - X1);
            const float ex1 = 256.0f * float(Y2 - Y0), ey1 = 256.0f * float(X0 - X2);
            const float ex2 = 256.0f * float(Y0 - Y1), ey2 = 256.0f * float(X1 - X0);
            pg.dlx = mk(ex0 * (w1 * w2), (swapped ? ex2 : ex1) * (w0 * w2), (swapped ? ex1 : ex2) * (w0 * w1));
            pg.dly = mk(ey0 * (w1 * w2), (swapped ? ey2 : ey1) * (w0 * w2), (swapped ? ey1 : ey2) * (w0 * w1));
        }
    }
    if ((sa.w & kShadeNoRecord) != 0u) {
        const DrawGPU dr = L.draws[pg.draw];
        const float4* xv = L.xv + V.xv_base + dr.xv_off;
        const float4 c0 = xv[pg.v0 - dr.vtx_base], c1 = xv[pg.v1 - dr.vtx_base], c2 = xv[pg.v2 - dr.vtx_base];
        const float u = (float(x) + 0.5f) * (2.0f / float(V.width)) - 1.0f;
        const float v = (float(y) + 0.5f) * (2.0f / float(V.height)) - 1.0f;
        const f3 a = mk(c0.x - u * c0.w, c1.x - u * c1.w, c2.x - u * c2.w);
        const f3 b = mk(c0.y - v * c0.w, c1.y - v * c1.w, c2.y - v * c2.w);
        lam = cross(a, b);
        if constexpr (kLod) {
            const float sx = -2.0f / float(V.width), sy = -2.0f / float(V.height);
            pg.dlx = cross(mk(c0.w * sx, c1.w * sx, c2.w * sx), b);
            pg.dly = cross(a, mk(c0.w * sy, c1.w * sy, c2.w * sy));
        }
    }
    const float inv_s = 1.0f / (lam.x + lam.y + lam.z);
    lam = lam * inv_s;
    if constexpr (kLod) {
        const float dsx = pg.dlx.x + pg.dlx.y + pg.dlx.z, dsy = pg.dly.x + pg.dly.y + pg.dly.z;
        pg.dlx = (pg.dlx - lam * dsx) * inv_s;
        pg.dly = (pg.dly - lam * dsy) * inv_s;
    }
    const float view_depth = w0 * lam.x + w1 * lam.y + w2 * lam.z;
    int cascade = -1;
    for (uint32_t k = 0; k < L.n_cascades; ++k)
        if (view_depth <= S.casc[k].split_far) {
            cascade = int(k);
            break;
        }
    pg.lam = lam;
    pg.cascade = cascade;
    return pg;
}

// Access to the per-draw records (draw, object, material, sampler, texture) of a shading pass:
// plain global pointers, or, when every live lane of the wave shades the same draw, pointers into
// the constant address space -- the records are read-only while the kernel runs -- so that the
// wave-uniform loads become scalar loads.
struct LaneRecords {
    template <class T> using P = const T*;
    template <class T> __device__ static P<T> ptr(const T* p) { return p; }
    template <class T> __device__ static T load(P<T> p) { return *p; }
};
struct WaveRecords {
    template <class T> using P = const __attribute__((address_space(4))) T*;
    template <class T> __device__ static P<T> ptr(const T* p) { return (P<T>)(p); }
    // a record copied word by word (a struct copy would need a generic-address-space source)
    template <class T> __device__ static T load(P<T> p) {
        static_assert(sizeof(T) % 4 == 0, "whole words");
        T v;
        uint32_t* d = reinterpret_cast<uint32_t*>(&v);
        const P<uint32_t> w = (P<uint32_t>)(p);
#pragma unroll
        for (uint32_t k = 0; k < sizeof(T) / 4; ++k) d[k] = w[k];
        return v;
    }
};

// Result of fragmentMain for one pixel: the colour, the PCF factor, and whether the fragment is
// discarded (alpha mask below its cutoff, basic.slang:335-337).
struct Fragment {
    float4 color;
    float shadow;
    bool discard;
};

// fragmentMain (basic.slang:137-344) for one covered pixel.  When `use_dec` is set, the material
// decoder's outputs `dec` replace the albedo, normal and metallic-roughness samples of the plain
// textures.  kUniformDraw: every live lane of the wave shades the same draw.
template <bool kUniformDraw, bool kLod = false>
__device__ inline Fragment shade_pixel(const ShadeLaunch& L, const ShadeShared& S, const PixelGeom& pg, bool use_dec,
                                       const float (&dec)[kNmOut]) {
    using R = std::conditional_t<kUniformDraw, WaveRecords, LaneRecords>;
    const SceneP sd = (SceneP)(L.scene);
    const float* lut = S.lut;
    const uint32_t draw = kUniformDraw ? uint32_t(__builtin_amdgcn_readfirstlane(int(pg.draw))) : pg.draw;
    const DrawGPU dr = R::load(R::ptr(L.draws) + draw);
    const auto gd = R::ptr(L.objects) + size_t(dr.object) * 41u;
    const bool has_material = dr.material < L.n_materials;
    // dereferenced only when has_material
    const auto md = R::ptr(static_cast<const rsy_gpu_material*>(L.materials)) + (has_material ? dr.material : 0u);
    const rsy_position* verts = static_cast<const rsy_position*>(L.vertices);
    const auto samplers = R::ptr(static_cast<const rsy_sampler*>(L.samplers));
    const auto textures = R::ptr(L.textures);
    const f3 lam = pg.lam;
    const int cascade = pg.cascade;
    const float* cm = cascade >= 0 ? S.casc[cascade].m : nullptr;
    const uint32_t nti = has_material ? md->normal_sampled_image_index : RSY_ABSENT;
    const rsy_position* const vtx[3] = {&verts[pg.v0], &verts[pg.v1], &verts[pg.v2]};
    const bool shift_uv = has_material && !(dr.flags & kDrawClampSampler);
    const Varyings bv = pixel_varyings(sd, vtx, gd, has_material, nti, cm, lam.x, lam.y, lam.z, shift_uv);
    const f3 bn = bv.normal, bwn = bv.world_normal, bview = bv.view_dir, blight = bv.light_dir;
    const f3 btan = bv.tangent, bbit = bv.bitangent;
    const float4 bcol = bv.color;
    const float tu = bv.u, tv = bv.v, su = bv.su, sv = bv.sv, sigma = bv.sigma;
    // kLod: texture-coordinate derivatives along the pixel's x and y
    float dudx = 0.f, dvdx = 0.f, dudy = 0.f, dvdy = 0.f;
    if constexpr (kLod) {
        const float* t0 = vtx[0]->texture_coordinates;
        const float* t1 = vtx[1]->texture_coordinates;
        const float* t2 = vtx[2]->texture_coordinates;
        dudx = pg.dlx.x * t0[0] + pg.dlx.y * t1[0] + pg.dlx.z * t2[0];
        dvdx = pg.dlx.x * t0[1] + pg.dlx.y * t1[1] + pg.dlx.z * t2[1];
        dudy = pg.dly.x * t0[0] + pg.dly.y * t1[0] + pg.dly.z * t2[0];
        dvdy = pg.dly.x * t0[1] + pg.dly.y * t1[1] + pg.dly.z * t2[1];
    }
    auto sample = [&](const TexGPU& tg, const SamplerState& smp) {
        const TexRef t{L.texels + tg.offset, int(tg.width), int(tg.height), int(tg.channels), int(tg.srgb),
                       int(tg.levels)};
        if constexpr (kLod) return sample_tex_lod(t, smp, tu, tv, dudx, dvdx, dudy, dvdy, lut);
        else return sample_tex(t, smp, tu, tv, lut);
    };
    const float4 csm = bv.csm;

    const f3 ambient = mk(sd->ambient_color[0], sd->ambient_color[1], sd->ambient_color[2]);
    float4 out = bcol;
    float4 image = out;
    float alpha = bcol.w;
    float srough = 0.f, smetal = 0.f;
    f3 light_normal = normalize(bn);
    const f3 light_dir = normalize(blight);
    if (has_material) {
        out = make_float4(md->color[0], md->color[1], md->color[2], md->color[3]);
        image = out;
        alpha = fminf(md->color[3], alpha);
        if (use_dec) {
            if (sd->normal_maps_enabled == 1) {
                f3 m = mk(dec[3] * 2.f - 1.f, dec[4] * 2.f - 1.f, 0.f);
                m.z = __builtin_amdgcn_sqrtf(fmaxf(0.f, 1.f - m.x * m.x - m.y * m.y));  // two-channel normal map
                const f3 n = normalize(bn);
                const f3 tt = normalize(btan - n * dot(btan, n));
                const f3 bb = normalize(cross(bn, btan) * sigma);
                light_normal = normalize(tt * m.x + bb * m.y + n * m.z);
            }
            image = make_float4(dec[0], dec[1], dec[2], 1.f);
            alpha = fminf(1.f, alpha);
            smetal = dec[5];
            srough = dec[6];
        } else {
            if (sd->normal_maps_enabled == 1 && valid_index(nti) && nti < L.n_textures) {
                const TexGPU tg = R::load(textures + nti);
                const uint32_t si = md->normal_sampler_index;
                const SamplerState smp = si < L.n_samplers ? sampler_state(R::load(samplers + si)) : SamplerState{};
                const float4 s = sample(tg, smp);
                f3 m = mk(s.x * 2.f - 1.f, s.y * 2.f - 1.f, s.z * 2.f - 1.f);
                if (tg.channels == 2) m.z = __builtin_amdgcn_sqrtf(fmaxf(0.f, 1.f - m.x * m.x - m.y * m.y));
                const f3 n = normalize(bn);
                const f3 tt = normalize(btan - n * dot(btan, n));
                const f3 bb = normalize(cross(bn, btan) * sigma);
                light_normal = normalize(tt * m.x + bb * m.y + n * m.z);
            }
            const uint32_t cti = md->color_sampled_image_index;
            if (valid_index(cti) && cti < L.n_textures) {
                const TexGPU tg = R::load(textures + cti);
                const uint32_t si = md->color_sampler_index;
                const SamplerState smp = si < L.n_samplers ? sampler_state(R::load(samplers + si)) : SamplerState{};
                image = sample(tg, smp);
                alpha = fminf(image.w, alpha);
            }
            const uint32_t mixti = md->mixmap_sampled_image_index, metti = md->metallic_sampled_image_index;
            const bool use_mix = valid_index(mixti) && mixti < L.n_textures;
            const uint32_t mti = use_mix ? mixti : metti;
            const uint32_t msi = use_mix ? md->mixmap_sampler_index : md->metallic_sampler_index;
            if (valid_index(mti) && mti < L.n_textures) {
                const TexGPU tg = R::load(textures + mti);
                const SamplerState smp = msi < L.n_samplers ? sampler_state(R::load(samplers + msi)) : SamplerState{};
                mr_from(int(tg.channels), sample(tg, smp), srough, smetal);
            }
        }
    }
    const f3 Vd = normalize(bview);
    const f3 Ld = light_dir;
    const f3 N = light_normal;
    const f3 H = normalize(Ld + Vd);
    const float cos_theta = clamp01(dot(Ld, N)), cos_phi = clamp01(dot(H, N));
    const float l_dot_n = cos_theta, l_dot_h = clamp01(dot(Ld, H)), n_dot_v = clamp01(dot(N, Vd));
    const f3 half = mk(0.5f, 0.5f, 0.5f);
    f3 c3;
    bool set = true;
    switch (sd->fragment_output) {
        case 0: c3 = mk(image.x, image.y, image.z); break;
        case 1: c3 = normalize(bwn) * 0.5f + half; break;
        case 2: c3 = normalize(btan) * 0.5f + half; break;
        case 3: c3 = normalize(blight) * 0.5f + half; break;
        case 4: c3 = normalize(bview) * 0.5f + half; break;
        case 5: c3 = mk(bcol.x, bcol.y, bcol.z); break;
        case 6: c3 = light_normal * 0.5f + half; break;
        case 7: c3 = bbit * 0.5f + half; break;
        case 8: c3 = mk(cos_theta, cos_theta, cos_theta) * 0.5f + half; break;
        case 9: c3 = mk(cos_phi, cos_phi, cos_phi) * 0.5f + half; break;
        case 10: c3 = mk(l_dot_h, l_dot_h, l_dot_h) * 0.5f + half; break;
        case 11: c3 = mk(n_dot_v, n_dot_v, n_dot_v) * 0.5f + half; break;
        case 12: c3 = mk(tu + su, tv + sv, 0.f); break;
        case 13: c3 = mk(srough, 0.05f, 0.05f); break;
        case 14: c3 = mk(smetal, 0.05f, 0.05f); break;
        case 15: c3 = mk(1.f - alpha, 0.05f, 0.05f); break;
        default: set = false; c3 = mk(0.f, 0.f, 0.f); break;
    }
    if (set) out = make_float4(c3.x, c3.y, c3.z, alpha);
    float shadow = 1.f;
    if (sd->light_enabled > 0) {
        if (sd->shadows_enabled == 1 && cascade >= 0) {
            const ShadeCascade& C = S.casc[cascade];
            const GlobalPtr<const float> map = gmem(C.map);
            const int Sz = int(C.size);
            if (L.pcf_mode == RSY_PCF_GRID_3X3) {
                float r = 0.f;
                const float inv = 1.0f / float(Sz);
                for (int j = -1; j <= 1; ++j)
                    for (int i = -1; i <= 1; ++i) r += cmp_tap(map, Sz, csm.x + float(i) * inv, csm.y + float(j) * inv, csm.z);
                shadow = r * (1.0f / 9.0f);
            } else {
                const float so = 0.00025f;
                float r = 0.f;
                r += cmp_tap(map, Sz, csm.x - so, csm.y - so, csm.z);
                r += cmp_tap(map, Sz, csm.x + so * 2.f, csm.y, csm.z);
                r += cmp_tap(map, Sz, csm.x, csm.y + so * 2.f, csm.z);
                r += cmp_tap(map, Sz, csm.x + so * 2.f, csm.y + so * 2.f, csm.z);
                r += cmp_tap(map, Sz, csm.x - so * 2.f, csm.y - so * 2.f, csm.z);
                shadow = r * (1.0f / 5.0f);
            }
        }
        const f3 sun_col = mk(sd->sunlight_color[0], sd->sunlight_color[1], sd->sunlight_color[2]);
        if (sd->brdf_lighting_enabled == 0) {
            const f3 lit = sun_col * fmaxf(cos_theta, 0.f) * shadow;
            out = make_float4(out.x * fmaxf(ambient.x, lit.x), out.y * fmaxf(ambient.y, lit.y),
                              out.z * fmaxf(ambient.z, lit.z), out.w * alpha);
        } else {
            const float roughness = srough, metallic = smetal;
            const f3 albedo = mk(out.x, out.y, out.z);
            const f3 f0 = mk(0.04f, 0.04f, 0.04f) + (albedo - mk(0.04f, 0.04f, 0.04f)) * metallic;
            const float ndf = distribution_ggx(N, H, roughness);
            const float g = schlick_ggx(fmaxf(dot(N, Ld), 0.f), roughness) * schlick_ggx(fmaxf(dot(N, Vd), 0.f), roughness);
            const float q = clamp01(1.f - fmaxf(dot(H, Vd), 0.f));
            const float q5 = q * q * q * q * q;
            const f3 F = f0 + (mk(1.f, 1.f, 1.f) - f0) * q5;
            const f3 kD = (mk(1.f, 1.f, 1.f) - F) * (1.f - metallic);
            const f3 spec = F * (ndf * g * __builtin_amdgcn_rcpf(4.f * n_dot_v * l_dot_n + 0.0001f));
            const f3 diffuse = albedo * (1.0f / 3.14159265f);
            const f3 sun_light = sun_col * sd->light_intensity;
            const f3 brdf = (kD * diffuse + spec) * sun_light * l_dot_n;
            const f3 amb = ambient * diffuse;
            const f3 sh = brdf * shadow;
            out = make_float4(fmaxf(amb.x, sh.x), fmaxf(amb.y, sh.y), fmaxf(amb.z, sh.z), alpha);
        }
    }
    bool discard = false;
    if (has_material) {
        discard = md->alpha_mode == 1u && alpha < md->alpha_cutoff;
        out.w = alpha;
    }
    return Fragment{out, shadow, discard};
}

__device__ inline void store_fragment(const ShadeView& V, size_t o, const Fragment& f) {
    gmem(V.color)[o] = f.color;
    if (V.shadow_factor) gmem(V.shadow_factor)[o] = f.shadow;
}


}  // namespace

// True when every live lane of the wave holds the same value.
__device__ inline bool wave_uniform(uint32_t v) {
    const uint32_t v0 = uint32_t(__builtin_amdgcn_readfirstlane(int(v)));
    return __builtin_amdgcn_ballot_w64(v != v0) == 0ull;
}

template <bool kNeural, bool kLod>
__global__ __launch_bounds__(256) void k_shade(ShadeLaunch L) {
    __shared__ ShadeShared S;
    const SharedWords sw = shared_request(L);

    uint32_t vi = 0;
    while (vi + 1 < L.n_views && L.views[vi + 1].block_base <= blockIdx.x) ++vi;
    const ShadeView& V = L.views[vi];
    const uint32_t lb = blockIdx.x - V.block_base;
    const uint32_t x = (lb % V.blocks_x) * 16 + (threadIdx.x & 15);
    const uint32_t y = (lb / V.blocks_x) * 16 + (threadIdx.x >> 4);
    const bool inside = x < V.width && y < V.height;
    const size_t o = size_t(y) * V.width + x;
    // the pixel's visibility word is requested beside the workgroup's constants
    const unsigned long long vis_word = inside ? gmem(V.vis)[o] : 0ull;
    shared_store(L, S, sw);
    __syncthreads();
    if constexpr (!kNeural) {
        if (!inside) return;
        const unsigned long long vis = vis_word;
        if (vis == 0ull) {
            gmem(V.color)[o] = make_float4(L.clear_color[0], L.clear_color[1], L.clear_color[2], L.clear_color[3]);
            if (V.shadow_factor) gmem(V.shadow_factor)[o] = 1.f;
            return;
        }
        const float none[kNmOut] = {};
        const PixelGeom pg = pixel_geometry<kLod>(L, S, V, x, y, vis);
        store_fragment(V, o, wave_uniform(pg.draw) ? shade_pixel<true, kLod>(L, S, pg, false, none)
                                                   : shade_pixel<false, kLod>(L, S, pg, false, none));
    } else {
        // Every lane stays live up to the wave-wide decoder MLP; lanes without a decoded material
        // feed zero features and ignore the result.
        const unsigned long long vis = vis_word;
        PixelGeom pg{};
        bool decode = false;
        float f[kNmFeat] = {};
        if (vis != 0ull) {
            pg = pixel_geometry<kLod>(L, S, V, x, y, vis);
            // the material's decoder flag and latent slot load beside the texture coordinates
            const uint32_t mat = pg.material;
            const bool known = mat < L.n_materials;
            const uint32_t active = known ? L.nm_active[mat] : 0u, slot = known ? L.nm_slot[mat] : 0u;
            const rsy_position* verts = static_cast<const rsy_position*>(L.vertices);
            const float* t0 = verts[pg.v0].texture_coordinates;
            const float* t1 = verts[pg.v1].texture_coordinates;
            const float* t2 = verts[pg.v2].texture_coordinates;
            const float u0 = t0[0], v0 = t0[1], u1 = t1[0], v1 = t1[1], u2 = t2[0], v2 = t2[1];
            if (active) {
                decode = true;
                // relative to vertex 0 with whole even periods removed (the grid repeats with period 1)
                const float u = (u0 - 2.0f * floorf(u0 * 0.5f)) + pg.lam.y * (u1 - u0) + pg.lam.z * (u2 - u0);
                const float v = (v0 - 2.0f * floorf(v0 * 0.5f)) + pg.lam.y * (v1 - v0) + pg.lam.z * (v2 - v0);
                nm_features(L.nm_lat16 + size_t(slot) * L.nm_L * L.nm_L, L.nm_L, u, v, f);
            }
        }
        float dec[kNmOut] = {};
        if (__builtin_amdgcn_ballot_w64(decode) != 0ull) nm_mlp_wave(static_cast<const NmWeights*>(L.nm_weights), f, dec);
        if (!inside) return;
        if (vis == 0ull) {
            gmem(V.color)[o] = make_float4(L.clear_color[0], L.clear_color[1], L.clear_color[2], L.clear_color[3]);
            if (V.shadow_factor) gmem(V.shadow_factor)[o] = 1.f;
            return;
        }
        store_fragment(V, o, wave_uniform(pg.draw) ? shade_pixel<true, kLod>(L, S, pg, decode, dec)
                                                   : shade_pixel<false, kLod>(L, S, pg, decode, dec));
    }
}

// The 4-sample main pass (Graphics.cpp:5566-5575: a 4-sample colour attachment resolved with
// VK_RESOLVE_MODE_AVERAGE_BIT).  Without sample shading the fragment shader runs once per pixel and
// primitive, at the pixel centre, and its colour goes to every sample the primitive won; samples
// no primitive covers keep the clear colour.  The block's (pixel, distinct primitive) pairs are
// numbered by a prefix sum and shaded as one list by all 256 threads (so a wave is not held back
// by its edge pixels), their colours kept in LDS; each pixel then writes the average of its four
// samples (and, with a blended pass to follow, the sample colours).  The debug shadow factor is
// that of sample 0's primitive.
template <bool kNeural, bool kLod>
__global__ __launch_bounds__(256) void k_shade_ms(ShadeLaunch L) {
    __shared__ ShadeShared S;
    __shared__ uint16_t jobs[1024];       // pixel (thread) << 2 | first sample of the primitive
    __shared__ float4 job_color[1024];
    __shared__ float shadow0[256];
    __shared__ uint32_t keys[4][256];     // the pixel's sample keys, kept for the resolve
    __shared__ uint32_t scan_lds[4];
    const SharedWords sw = shared_request(L);

    uint32_t vi = 0;
    while (vi + 1 < L.n_views && L.views[vi + 1].block_base <= blockIdx.x) ++vi;
    const ShadeView& V = L.views[vi];
    const uint32_t lb = blockIdx.x - V.block_base;
    const uint32_t bx = (lb % V.blocks_x) * 16, by = (lb / V.blocks_x) * 16;
    const uint32_t x = bx + (threadIdx.x & 15), y = by + (threadIdx.x >> 4);
    const bool inside = x < V.width && y < V.height;
    const size_t o = size_t(y) * V.width + x, plane = size_t(V.width) * V.height;
    unsigned long long v[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) v[s] = inside ? gmem(V.vis)[s * plane + o] : 0ull;
    shared_store(L, S, sw);
    const uint32_t k0 = uint32_t(v[0]), k1 = uint32_t(v[1]), k2 = uint32_t(v[2]), k3 = uint32_t(v[3]);
    // bit s: sample s is the first to hold its primitive (or the background)
    const uint32_t todo = inside ? 1u | (k1 != k0 ? 2u : 0u) | (k2 != k0 && k2 != k1 ? 4u : 0u) |
                                       (k3 != k0 && k3 != k1 && k3 != k2 ? 8u : 0u)
                                 : 0u;
    keys[0][threadIdx.x] = k0;
    keys[1][threadIdx.x] = k1;
    keys[2][threadIdx.x] = k2;
    keys[3][threadIdx.x] = k3;
    uint32_t total;
    const uint32_t first = block_exclusive_scan(uint32_t(__builtin_popcount(todo)), scan_lds, total);
    {
        uint32_t m = todo, at = first;
        while (m) {
            jobs[at++] = uint16_t((threadIdx.x << 2) | uint32_t(__builtin_ctz(m)));
            m &= m - 1u;
        }
    }
    __syncthreads();
    const float4 clear = make_float4(L.clear_color[0], L.clear_color[1], L.clear_color[2], L.clear_color[3]);
    const uint32_t rounds = (total + 255u) / 256u;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t j = r * 256u + threadIdx.x;
        const bool live = j < total;
        const uint32_t job = live ? jobs[j] : 0u;
        const uint32_t t = job >> 2, smp = job & 3u;
        const uint32_t jx = bx + (t & 15u), jy = by + (t >> 4);
        const size_t jo = size_t(jy) * V.width + jx;
        const unsigned long long w = live ? gmem(V.vis)[smp * plane + jo] : 0ull;
        const bool covered = live && uint32_t(w) != 0u;
        PixelGeom pg{};
        if (covered) pg = pixel_geometry<kLod>(L, S, V, jx, jy, w);
        float dec[kNmOut] = {};
        bool decode = false;
        if constexpr (kNeural) {
            float f[kNmFeat] = {};
            if (covered) {
                const uint32_t mat = pg.material;
                const bool known = mat < L.n_materials;
                const uint32_t active = known ? L.nm_active[mat] : 0u, slot = known ? L.nm_slot[mat] : 0u;
                const rsy_position* verts = static_cast<const rsy_position*>(L.vertices);
                const float* t0 = verts[pg.v0].texture_coordinates;
                const float* t1 = verts[pg.v1].texture_coordinates;
                const float* t2 = verts[pg.v2].texture_coordinates;
                const float u0 = t0[0], v0 = t0[1], u1 = t1[0], v1 = t1[1], u2 = t2[0], v2 = t2[1];
                if (active) {
                    decode = true;
                    const float u = (u0 - 2.0f * floorf(u0 * 0.5f)) + pg.lam.y * (u1 - u0) + pg.lam.z * (u2 - u0);
                    const float vv = (v0 - 2.0f * floorf(v0 * 0.5f)) + pg.lam.y * (v1 - v0) + pg.lam.z * (v2 - v0);
                    nm_features(L.nm_lat16 + size_t(slot) * L.nm_L * L.nm_L, L.nm_L, u, vv, f);
                }
            }
            if (__builtin_amdgcn_ballot_w64(decode) != 0ull) nm_mlp_wave(static_cast<const NmWeights*>(L.nm_weights), f, dec);
        }
        if (!live) continue;
        float4 c = clear;
        float sh = 1.f;
        if (covered) {
            const Fragment fr = wave_uniform(pg.draw) ? shade_pixel<true, kLod>(L, S, pg, decode, dec)
                                                      : shade_pixel<false, kLod>(L, S, pg, decode, dec);
            c = fr.color;
            sh = fr.shadow;
        }
        job_color[j] = c;
        if (smp == 0u) shadow0[t] = sh;
    }
    __syncthreads();
    if (!inside) return;
    // the average of the four samples: each primitive's colour times the samples it holds
    const uint32_t q0 = keys[0][threadIdx.x], q1 = keys[1][threadIdx.x], q2 = keys[2][threadIdx.x],
                   q3 = keys[3][threadIdx.x];
    uint32_t m = 1u | (q1 != q0 ? 2u : 0u) | (q2 != q0 && q2 != q1 ? 4u : 0u) | (q3 != q0 && q3 != q1 && q3 != q2 ? 8u : 0u);
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t at = first;
    while (m) {
        const int s = __builtin_ctz(m);
        m &= m - 1u;
        const uint32_t key = s == 0 ? q0 : s == 1 ? q1 : s == 2 ? q2 : q3;
        const uint32_t same = (q0 == key ? 1u : 0u) | (q1 == key ? 2u : 0u) | (q2 == key ? 4u : 0u) | (q3 == key ? 8u : 0u);
        const float4 c = job_color[at++];
        const float n = float(__builtin_popcount(same));
        acc = make_float4(acc.x + c.x * n, acc.y + c.y * n, acc.z + c.z * n, acc.w + c.w * n);
        if (V.sample_color)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (same & (1u << q)) gmem(V.sample_color)[q * plane + o] = c;
    }
    gmem(V.color)[o] = make_float4(acc.x * 0.25f, acc.y * 0.25f, acc.z * 0.25f, acc.w * 0.25f);
    if (V.shadow_factor) gmem(V.shadow_factor)[o] = shadow0[threadIdx.x];
}

// The blended pass (Graphics.cpp:5655-5691): the blended draw list in primitive order over the
// opaque frame, depth test GEQUAL against the opaque depth without depth writes, no culling, each
// fragment shaded with the plain material maps (fragmentMain, basic.slang:137-344), discarded under
// the alpha mask (basic.slang:335-337) and otherwise blended with the reference's blend state
// (Graphics.cpp:5657-5664): colour = src.rgb * src.a + dst.rgb * (1 - src.a), alpha = src.a.
//
// Three launches.  k_blend_collect builds every pixel's list of passing fragments, sorted by key
// (triangle id + 1 in the blended list, so primitive order); k_blend_shade shades all fragments of
// the frame in parallel; k_blend_resolve blends each pixel's list in order over its colour.

// Shaded colour of a fragment the alpha mask discards: all bits of w set (a NaN that shading does
// not produce).
constexpr uint32_t kDiscarded = 0xffffffffu;

// One workgroup per 16x16 bin (T = 16), quarter (T = 32) or sixteenth (T = 64) of a bin of a blend view, so that
// a crowded bin spreads over several workgroups.  Two raster sweeps over the bin's blended records:
// the first counts each pixel's passing fragments in LDS, a prefix sum turns the counts into list
// offsets (the workgroup takes its range of the frame's fragment array with one device atomic),
// and the second writes each fragment's key into its pixel's list.  The lists stay in LDS when
// they fit (global memory otherwise), are sorted per pixel (insertion sort; lists are short), and
// go out as fragment records plus per-pixel (first fragment, count) entries.  A workgroup whose
// fragments do not fit the frame's fragment array writes empty lists and flags the overflow (the
// frame is re-run with a larger array).
constexpr int kBlendSub = 16;  // sub-tile edge of the fragment collection

template <int T, int N>
__global__ __launch_bounds__(256) void k_blend_collect(const BlendLaunch B) {
    constexpr int U = kBlendSub, S = U + 1, Q = (T / U) * (T / U);
    static_assert(U * U == 256, "one pixel per thread");
    constexpr uint32_t kLdsFrags = 8192u;
    // list entries: the key, and with 4 samples the passing-sample mask in bits 28-31
    constexpr uint32_t kKeyMask = N == 4 ? 0x0fffffffu : 0xffffffffu;
    __shared__ uint32_t opaque[N * U * S], count[U * S], cursor[U * S];
    __shared__ uint32_t lds_keys[kLdsFrags];
    __shared__ TriRec chunk_recs[kChunkRecs];
    __shared__ uint8_t owner[kChunkRecs * U];
    __shared__ uint32_t row_start[kChunkRecs];
    __shared__ uint32_t scan_lds[4];
    __shared__ uint32_t base_lds;

    const uint32_t bin = B.views[B.first_blend].tile_base + blockIdx.x / Q;
    const uint32_t off = B.bin_offsets[bin];
    const uint32_t end = uint32_t(min(uint64_t(off) + B.bin_counts[bin], uint64_t(B.list_cap)));
    if (end <= off) return;  // no blended triangle touches this bin
    uint32_t view = B.first_blend;
    while (view + 1 < B.n_views && B.views[view + 1].tile_base <= bin) ++view;
    const ViewGPU& V = B.views[view];
    const uint32_t local_bin = bin - V.tile_base, q = blockIdx.x % Q;
    const int ox = int(local_bin % V.tiles_x) * T + int(q % (T / U)) * U;
    const int oy = int(local_bin / V.tiles_x) * T + int(q / (T / U)) * U;
    const int width = int(V.width), height = int(V.height);
    if (ox >= width || oy >= height) return;  // a sub-tile past the view's edge
    const unsigned long long* vis = reinterpret_cast<const unsigned long long*>(B.views[V.source].out);
    uint2* pixel_lists = reinterpret_cast<uint2*>(V.out);
    const uint32_t sview = view - B.first_blend;

    // thread t owns pixel t of the sub-tile (row-major)
    const int px = ox + int(threadIdx.x) % U, py = oy + int(threadIdx.x) / U;
    const int pi = (int(threadIdx.x) / U) * S + int(threadIdx.x) % U;
    const bool inside = px < width && py < height;
    const size_t plane = size_t(width) * height;
#pragma unroll
    for (int s = 0; s < N; ++s)
        opaque[s * U * S + pi] = inside ? uint32_t(vis[s * plane + size_t(py) * width + px] >> 32) : 0xffffffffu;
    count[pi] = 0u;
    __syncthreads();
    auto sweep = [&](const auto& op) {
        for (uint32_t base = off; base < end; base += kItemRecs) {
            const uint32_t n = min(end - base, kItemRecs);
            const uint32_t ri = threadIdx.x < n ? B.lists[base + threadIdx.x] : 0xffffffffu;
            const bool have = ri < B.rec_cap_total;
            const uint4* src = reinterpret_cast<const uint4*>(B.recs + (have ? ri : 0u));
            const uint4 w0 = src[0], w1 = src[1], w2 = src[2];
            raster_item<U>(op, chunk_recs, owner, row_start, scan_lds, have, w0, w1, w2, n, ox, oy);
        }
    };
    if constexpr (N == 4) sweep(CountFragMS<U>{count, opaque});  // ends on a barrier
    else sweep(CountFrag<U>{count, opaque});

    const uint32_t n = count[pi];
    uint32_t total;
    const uint32_t first = block_exclusive_scan(n, scan_lds, total);
    cursor[pi] = first;
    if (threadIdx.x == 0) base_lds = atomicAdd(&B.counters[B.first_blend * kCtrPerView + kCtrFrags], total);
    __syncthreads();
    const uint32_t base = base_lds;
    if (uint64_t(base) + total > B.frag_cap) {
        // out of fragment space: empty lists for this sub-tile, empty records for its reserved range
        if (threadIdx.x == 0) atomicOr(&B.counters[view * kCtrPerView + kCtrOverflow], kOverflowFrags);
        for (uint32_t j = base + threadIdx.x; j < B.frag_cap; j += 256) B.frags[j] = make_uint4(0u, 0u, 0u, 0u);
        if (inside) pixel_lists[size_t(py) * width + px] = make_uint2(0u, 0u);
        return;
    }
    uint32_t* keys = total <= kLdsFrags ? lds_keys : B.frag_keys + base;
    if constexpr (N == 4) sweep(FillFragMS<U>{cursor, opaque, keys});  // ends on a barrier
    else sweep(FillFrag<U>{cursor, opaque, keys});

    uint32_t* l = keys + first;
    for (uint32_t a = 1; a < n; ++a) {
        const uint32_t v = l[a];
        uint32_t b = a;
        for (; b > 0 && (l[b - 1] & kKeyMask) > (v & kKeyMask); --b) l[b] = l[b - 1];
        l[b] = v;
    }
    if constexpr (N == 4)
        for (uint32_t a = 0; a < n; ++a) B.frag_mask[base + first + a] = uint8_t(l[a] >> 28);
    if (inside) pixel_lists[size_t(py) * width + px] = make_uint2(base + first, n);
    // fragment records layer by layer (every pixel's first fragment, then every second one, ...),
    // so that a wave of the shading kernel holds neighbouring pixels of one layer; each record
    // carries the index of its colour in the per-pixel lists
    const uint32_t xy = uint32_t(px) | (uint32_t(py) << 16);
    uint32_t layer_base = base;
    for (uint32_t a = 0;; ++a) {
        uint32_t in_layer;
        const uint32_t rank = block_exclusive_scan(a < n ? 1u : 0u, scan_lds, in_layer);
        if (in_layer == 0u) break;
        if (a < n) B.frags[layer_base + rank] = make_uint4(l[a] & kKeyMask, xy, sview, base + first + a);
        layer_base += in_layer;
    }
}

// Every fragment of the frame shaded at its pixel centre (a grid-stride loop over the count the
// collection left), its colour stored at the index its record names.  Records of sub-tiles that ran
// out of space have key 0 and are skipped.
template <bool kLod>
__global__ __launch_bounds__(256) void k_blend_shade(const BlendLaunch B, const ShadeLaunch L) {
    __shared__ ShadeShared S;
    shared_store(L, S, shared_request(L));
    __syncthreads();
    const uint32_t n = min(B.counters[B.first_blend * kCtrPerView + kCtrFrags], B.frag_cap);
    const float none[kNmOut] = {};
    for (uint32_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
        const uint4 f = B.frags[j];
        if (f.x == 0u) continue;
        const ShadeView& SV = L.views[f.z];
        const uint32_t x = f.y & 0xffffu, y = f.y >> 16;
        const size_t o = size_t(y) * SV.width + x;
        const PixelGeom pg = pixel_geometry<kLod>(L, S, SV, x, y, (unsigned long long)f.x);
        const Fragment fr = wave_uniform(pg.draw) ? shade_pixel<true, kLod>(L, S, pg, false, none)
                                                  : shade_pixel<false, kLod>(L, S, pg, false, none);
        B.frag_color[f.w] = fr.discard ? make_float4(0.f, 0.f, 0.f, __uint_as_float(kDiscarded)) : fr.color;
    }
}

// Each pixel of a blend view whose bin holds blended records blends its sorted fragments in order;
// with 4 samples each fragment goes into the samples it passed at, and the pixel's colour becomes
// the average of its samples again.
__device__ inline float4 blend_over(float4 dst, float4 f) {
    const float a = f.w;
    return make_float4(f.x * a + dst.x * (1.f - a), f.y * a + dst.y * (1.f - a), f.z * a + dst.z * (1.f - a), a);
}

template <int T, int N>
__global__ __launch_bounds__(256) void k_blend_resolve(const BlendLaunch B, const ShadeLaunch L) {
    uint32_t vi = 0;
    while (vi + 1 < L.n_views && L.views[vi + 1].block_base <= blockIdx.x) ++vi;
    const ShadeView& SV = L.views[vi];
    const ViewGPU& V = B.views[B.first_blend + vi];
    const uint32_t lb = blockIdx.x - SV.block_base;
    const uint32_t x = (lb % SV.blocks_x) * 16 + (threadIdx.x & 15);
    const uint32_t y = (lb / SV.blocks_x) * 16 + (threadIdx.x >> 4);
    if (x >= SV.width || y >= SV.height) return;
    // the collection wrote this pixel's list only when its bin has records inside the bin lists'
    // capacity (a frame that overflowed them is re-run); the same test here
    const uint32_t bin = V.tile_base + (y / T) * V.tiles_x + x / T;
    const uint32_t off = B.bin_offsets[bin];
    if (uint64_t(off) + B.bin_counts[bin] <= off || off >= B.list_cap) return;
    const size_t o = size_t(y) * SV.width + x;
    const uint2 l = reinterpret_cast<const uint2*>(V.out)[o];
    if (l.y == 0u || uint64_t(l.x) + l.y > B.frag_cap) return;
    if constexpr (N == 4) {
        const size_t plane = size_t(SV.width) * SV.height;
        float4 c[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) c[s] = SV.sample_color[s * plane + o];
        for (uint32_t k = 0; k < l.y; ++k) {
            const float4 f = B.frag_color[l.x + k];
            if (__float_as_uint(f.w) == kDiscarded) continue;
            const uint32_t m = B.frag_mask[l.x + k];
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (m & (1u << s)) c[s] = blend_over(c[s], f);
        }
        SV.color[o] = make_float4((c[0].x + c[1].x + c[2].x + c[3].x) * 0.25f, (c[0].y + c[1].y + c[2].y + c[3].y) * 0.25f,
                                  (c[0].z + c[1].z + c[2].z + c[3].z) * 0.25f, (c[0].w + c[1].w + c[2].w + c[3].w) * 0.25f);
    } else {
        float4 c = SV.color[o];
        for (uint32_t k = 0; k < l.y; ++k) {
            const float4 f = B.frag_color[l.x + k];
            if (__float_as_uint(f.w) == kDiscarded) continue;
            c = blend_over(c, f);
        }
        SV.color[o] = c;
    }
}

hipError_t launch_blend_fragments(const BlendLaunch& b, const ShadeLaunch& l, hipStream_t s) {
    if (b.blend_bins == 0 || l.total_blocks == 0) return hipSuccess;
    const uint32_t shade_blocks = max(1u, min(2048u, (b.frag_cap + 255) / 256));
    if (b.tile == 16 && b.samples == 1)
        hipLaunchKernelGGL((k_blend_collect<16, 1>), dim3(b.blend_bins), dim3(256), 0, s, b);
    else if (b.tile == 32 && b.samples == 1)
        hipLaunchKernelGGL((k_blend_collect<32, 1>), dim3(b.blend_bins * 4), dim3(256), 0, s, b);
    else if (b.tile == 64 && b.samples == 1)
        hipLaunchKernelGGL((k_blend_collect<64, 1>), dim3(b.blend_bins * 16), dim3(256), 0, s, b);
    else if (b.tile == 32 && b.samples == 4)
        hipLaunchKernelGGL((k_blend_collect<32, 4>), dim3(b.blend_bins * 4), dim3(256), 0, s, b);
    else
        return hipErrorInvalidValue;
    if (l.lod)
        hipLaunchKernelGGL(k_blend_shade<true>, dim3(shade_blocks), dim3(256), 0, s, b, l);
    else
        hipLaunchKernelGGL(k_blend_shade<false>, dim3(shade_blocks), dim3(256), 0, s, b, l);
    return hipGetLastError();
}

hipError_t launch_blend_resolve(const BlendLaunch& b, const ShadeLaunch& l, hipStream_t s) {
    if (b.blend_bins == 0 || l.total_blocks == 0) return hipSuccess;
    if (b.tile == 16 && b.samples == 1)
        hipLaunchKernelGGL((k_blend_resolve<16, 1>), dim3(l.total_blocks), dim3(256), 0, s, b, l);
    else if (b.tile == 32 && b.samples == 1)
        hipLaunchKernelGGL((k_blend_resolve<32, 1>), dim3(l.total_blocks), dim3(256), 0, s, b, l);
    else if (b.tile == 64 && b.samples == 1)
        hipLaunchKernelGGL((k_blend_resolve<64, 1>), dim3(l.total_blocks), dim3(256), 0, s, b, l);
    else if (b.tile == 32 && b.samples == 4)
        hipLaunchKernelGGL((k_blend_resolve<32, 4>), dim3(l.total_blocks), dim3(256), 0, s, b, l);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_shade(const ShadeLaunch& l, hipStream_t s) {
    if (l.total_blocks == 0) return hipSuccess;
    // kernels by samples per pixel, decoder, and whether some texture has mip levels
#define RSY_SHADE(K, N, LOD) hipLaunchKernelGGL((K<N, LOD>), dim3(l.total_blocks), dim3(256), 0, s, l)
    const bool lod = l.lod != 0, neural = l.neural != 0;
    if (l.samples == 4) {
        if (neural) { if (lod) RSY_SHADE(k_shade_ms, true, true); else RSY_SHADE(k_shade_ms, true, false); }
        else { if (lod) RSY_SHADE(k_shade_ms, false, true); else RSY_SHADE(k_shade_ms, false, false); }
    } else {
        if (neural) { if (lod) RSY_SHADE(k_shade, true, true); else RSY_SHADE(k_shade, true, false); }
        else { if (lod) RSY_SHADE(k_shade, false, true); else RSY_SHADE(k_shade, false, false); }
    }
#undef RSY_SHADE
    return hipGetLastError();
}

}  // namespace rsy
