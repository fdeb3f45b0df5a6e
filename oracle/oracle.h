/*
 * oracle.h -- CPU oracle of the raster-and-shade frame path.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and
 * only as the checker or the timed CPU baseline.  The product (rosy_amd) never links or calls it.
 *
 * What it restates (citations relative to the reference tree):
 *   - shadow.slang:6-20 and basic.slang:15-119: clip = (VP * W) * p for every vertex;
 *   - the Vulkan fixed-function state of Graphics.cpp:4988-5045, 5195-5292, 5553-5653: cull FRONT
 *     with front = CW, depth test GREATER_OR_EQUAL with later draws winning ties, depth clear 0,
 *     depth bias in the shadow pass only; rasterised here by the fixed-point rules below;
 *   - basic.slang:137-344, lighting.slang:5-43, pcf.slang:1-12 and the compare sampler of
 *     Graphics.cpp:2697-2725: deferred Cook-Torrance + PCF shading, in FP64.
 *
 * Fixed-point rules (identical by definition in the HIP kernels; Vulkan leaves them to the driver):
 *   clip row r    = fma(M[2][r], z, fma(M[1][r], y, fma(M[0][r], x, M[3][r]))), M = VP * W by an
 *                   fmaf chain in k order;
 *   clip planes   = w-z >= 0 (near, reverse-Z), z >= 0 (far), |x|,|y| <= 4w guard band;
 *                   Sutherland-Hodgman in that order, intersections from the inside vertex, fan;
 *   snap          = X = rint(256 * fma(x/w, W/2, W/2)), Y likewise; 8 sub-pixel bits;
 *   cull          = area2 = (X1-X0)(Y2-Y0)-(X2-X0)(Y1-Y0); front iff area2 > 0 (front = CW);
 *   coverage      = pixel centres, int64 edge functions, top-left rule (y down);
 *   depth         = plane from the snapped vertices in FP32 (no contraction), clamped to [0,1];
 *   depth bias    = o = m*slope + r*constant, m = max|dz/dx|,|dz/dy| per pixel, r = 2^(e-23);
 *   visibility    = max over fragments of (depth bits << 32 | tri + 1).
 */
#ifndef ROSY_ORACLE_H
#define ROSY_ORACLE_H

#include <stdint.h>

#include "rosy_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_geometry {
    const rsy_position* vertices;
    uint64_t n_vertices;
    const uint32_t* indices;
    uint64_t n_indices;
    const rsy_mesh_range* meshes;
    uint32_t n_meshes;
    uint32_t n_objects;
    const rsy_object_record* objects;
    const rsy_draw* draws; /* the draw list of the view, in order */
    uint32_t n_draws;
    uint32_t reserved;
} orc_geometry;

typedef struct orc_raster_stats {
    uint64_t triangles_in;
    uint64_t triangles_setup; /* records after clip/cull/bbox */
    uint64_t triangles_clipped;
    uint64_t fragments;       /* covered pixel samples */
} orc_raster_stats;

/* Rasterise one view.  is_shadow = 0: camera view, writes out_vis (w*h u64, zero = background).
 * is_shadow = 1: depth-only shadow view with depth bias, writes out_depth (w*h float). */
int orc_raster_view(const orc_geometry* g, const float view_projection[16], uint32_t width, uint32_t height,
                    int is_shadow, const rsy_raster_state* rs, uint64_t* out_vis, float* out_depth,
                    orc_raster_stats* stats, int n_threads);

/* The same with coverage and depth taken at the point (sample_x, sample_y) / 256 inside each pixel
 * instead of its centre (128, 128): one sample plane of a multisampled view. */
int orc_raster_view_at(const orc_geometry* g, const float view_projection[16], uint32_t width, uint32_t height,
                       int is_shadow, const rsy_raster_state* rs, int32_t sample_x, int32_t sample_y,
                       uint64_t* out_vis, float* out_depth, orc_raster_stats* stats, int n_threads);

/* The blended pass's fragments at every pixel of a camera view (Graphics.cpp:5655-5691: no culling,
 * no depth write, GEQUAL against the opaque depth of opaque_vis), in primitive order: layer k of
 * pixel o is layers[k * w*h + o] = (depth bits << 32) | (triangle id + 1) in the view's draw list,
 * 0 past the pixel's count.  RSY_OVERFLOW when a pixel holds more than max_layers fragments. */
int orc_blend_layers(const orc_geometry* g, const float view_projection[16], uint32_t width, uint32_t height,
                     const rsy_raster_state* rs, const uint64_t* opaque_vis, uint32_t max_layers, uint64_t* layers,
                     uint32_t* count, int n_threads);

/* The same at the sample point (sample_x, sample_y) / 256 of each pixel, against that sample
 * plane's opaque visibility. */
int orc_blend_layers_at(const orc_geometry* g, const float view_projection[16], uint32_t width, uint32_t height,
                        const rsy_raster_state* rs, int32_t sample_x, int32_t sample_y, const uint64_t* opaque_vis,
                        uint32_t max_layers, uint64_t* layers, uint32_t* count, int n_threads);

/* Number of (record, tile) pairs when records are binned by their pixel bounding boxes into
 * tile x tile bins (the bin traffic term of the roofline). */
uint64_t orc_count_tile_pairs(const orc_geometry* g, const float view_projection[16], uint32_t width,
                              uint32_t height, int is_shadow, const rsy_raster_state* rs, uint32_t tile,
                              int n_threads);

/* Records per tile x tile bin (row-major, ceil(w/tile) x ceil(h/tile)) for the same binning, and
 * (nullable) the pixels of the records' boxes inside each bin, followed by two totals: records
 * touching more than 4 bins and their (record, bin) pairs. */
int orc_tile_counts(const orc_geometry* g, const float view_projection[16], uint32_t width, uint32_t height,
                    int is_shadow, const rsy_raster_state* rs, uint32_t tile, uint32_t* counts, uint64_t* areas,
                    int n_threads);

typedef struct orc_shade_in {
    const rsy_gpu_material* materials;
    uint32_t n_materials;
    uint32_t n_textures;
    const rsy_texture_desc* textures;
    const rsy_sampler* samplers;
    uint32_t n_samplers;
    uint32_t n_cascades;
    const rsy_shadow_cascade* cascades;
    const float* const* shadow_maps; /* n_cascades maps, size*size floats each */
    const rsy_scene_data* scene;
    const rsy_raster_state* raster;
    const rsy_camera_view* view;
} orc_shade_in;

/* FP64 deferred shading of one camera view from its visibility buffer.
 * out_rgba: w*h*4 doubles.  out_shadow (nullable): per-pixel PCF factor (1 where not evaluated).
 * out_margin (nullable): per-pixel min |ref - texel| over compares with non-zero weight.
 * override_shadow (nullable): use these per-pixel shadow factors instead of the PCF result. */
int orc_shade_view(const orc_geometry* g, const orc_shade_in* in, const uint64_t* vis, double* out_rgba,
                   double* out_shadow, double* out_margin, const float* override_shadow, int n_threads);

/* Neural material decoder parameters (see neural.cpp for the definition). */
typedef struct orc_neural {
    uint32_t latent_size;
    uint32_t n_materials;
    const uint32_t* slot_of; /* per material: latent grid slot, 0xFFFFFFFF = none */
    const uint32_t* active;  /* per material: 1 = decoded, 0 = plain maps (nullable: all slots decoded) */
    const float* latents;    /* [slots][L][L][8] */
    const float* mlp;        /* [583]: W1 [32][10], b1 [32], W2 [7][32], b2 [7] */
} orc_neural;

/* orc_shade_view with the maps of every active material replaced by the decoder's outputs. */
int orc_shade_view_neural(const orc_geometry* g, const orc_shade_in* in, const orc_neural* nm, const uint64_t* vis,
                          double* out_rgba, double* out_shadow, double* out_margin, const float* override_shadow,
                          int n_threads);
/* What shading one view reads, counted by running the FP64 shading of every pixel with each read
 * recorded once (the algorithmic bytes of the shading kernel): unique texels of the material maps
 * (every level the level of detail reaches), unique shadow-map texels under the PCF taps, unique
 * latent cells the decoder fetches (nm non-null) and the visible triangles and their vertices. */
typedef struct orc_footprint {
    uint64_t covered_pixels;
    uint64_t visible_triangles;
    uint64_t visible_vertices;
    uint64_t texels;        /* unique material-map texels read */
    uint64_t texel_bytes;   /* their bytes as sampled: 4 per RGBA8 (and decoded BC7) texel, 2 per RG8 */
    uint64_t shadow_texels; /* unique shadow-map texels read (4 bytes each) */
    uint64_t latent_cells;  /* unique latent cells read (8 channels each) */
} orc_footprint;
int orc_shade_footprint(const orc_geometry* g, const orc_shade_in* in, const orc_neural* nm, const uint64_t* vis,
                        orc_footprint* out, int n_threads);

/* Decoder outputs at material[n], uv[n][2] -> out[n][7]. */
int orc_nm_decode(const orc_neural* nm, const uint32_t* material, const float* uv, uint32_t n, double* out);
/* Loss and gradients of the training objective (mean squared error over samples and channels, no
 * clamp) for samples material[n], uv[n][2], target[n][7].  grad_mlp [583]; grad_latents
 * [n_slots][L][L][8]; min_abs_preact (nullable) [n]: min_j |z_j| of each sample's hidden layer. */
int orc_nm_gradients(const orc_neural* nm, uint32_t n_slots, const uint32_t* material, const float* uv,
                     const float* target, uint32_t n, double* loss, double* grad_mlp, double* grad_latents,
                     double* min_abs_preact);
/* The plain sampler (bilinear or nearest, wrap modes, sRGB decode before filtering) at (u, v). */
int orc_sample_texture(const rsy_texture_desc* tex, const rsy_sampler* sampler, double u, double v, double out[4]);

/* Exact sRGB EOTF (basic.slang:121-128 form) of an 8-bit code. */
double orc_srgb_to_linear(uint32_t code);

/* Terms of the oracle's BRDF and shadow filter, exported so tests can pin them on their own. */
double orc_ggx_d(double n_dot_h, double roughness);          /* lighting.slang:10-22 */
double orc_schlick_ggx(double n_dot_v, double roughness);    /* lighting.slang:24-33 */
double orc_schlick_fresnel(double v_dot_h, double f0);       /* lighting.slang:5-8 (one channel) */
/* PCF shadow factor at (u, v) with reference depth ref over a size x size map (pcf.slang:1-12 for
 * RSY_PCF_REFERENCE_5TAP, 3x3 one-texel grid for RSY_PCF_GRID_3X3). */
double orc_pcf(const float* map, uint32_t size, double u, double v, double ref, uint32_t mode, double* margin);

/* BC7 blocks of a width x height image (ceil(w/4) x ceil(h/4) blocks of 16 bytes, row-major)
 * decoded to RGBA8 (R first); texels past the image's edge are dropped. */
int orc_bc7_decode(const uint8_t* blocks, uint32_t width, uint32_t height, uint8_t* rgba);
/* The partition tables as subset per texel (64 x 16 two-subset, 64 x 16 three-subset) and the
 * anchor texels (64 of subset 1 in two-subset partitions, then 64 of subset 1 and 64 of subset 2
 * in three-subset ones). */
int orc_bc7_tables(uint8_t* two, uint8_t* three, uint8_t* anchors);

#ifdef __cplusplus
}
#endif

#endif
