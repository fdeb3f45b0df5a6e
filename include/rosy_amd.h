/*
 * rosy_amd.h -- C ABI of the MI355X raster-and-shade frame path.
 *
 * This header is the drop-in boundary that replaces the renderer-facing surface of Rosy's Vulkan
 * renderer (src/Engine/Graphics.h:11-21 and graphics_device in src/Engine/Graphics.cpp).  Every
 * record type keeps the reference's byte layout (packed, little-endian, static_assert'ed sizes),
 * every entry point names the reference function it stands in for, and every call returns the
 * reference's numeric rosy::result code (src/Engine/Types.h:13-31).  No C++ types or exceptions
 * cross this boundary; all pointers are borrowed for the duration of the call only.
 */
#ifndef ROSY_AMD_H
#define ROSY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSY_ABI_VERSION 1u
#define RSY_ABSENT 0xFFFFFFFFu /* UINT32_MAX == not present (src/Asset/Asset.h:38-45) */

/* rosy::result, same numeric values (src/Engine/Types.h:13-31). */
typedef enum rsy_result {
    RSY_OK = 0,
    RSY_STATE_CHANGED = 1,
    RSY_ERROR = 2,
    RSY_INVALID_ARGUMENT = 3,
    RSY_ALLOCATION_FAILURE = 4,
    RSY_GRAPHICS_INIT_FAILURE = 5,
    RSY_GRAPHICS_FRAME_FAILURE = 6,
    RSY_GRAPHICS_SWAPCHAIN_FAILURE = 7,
    RSY_OPEN_FAILED = 8,
    RSY_WRITE_FAILED = 9,
    RSY_COPY_FAILED = 10,
    RSY_READ_FAILED = 11,
    RSY_CREATE_FAILED = 12,
    RSY_UPDATE_FAILED = 13,
    RSY_OVERFLOW = 14,
    RSY_INVALID_STATE = 15
} rsy_result;

#pragma pack(push, 1)

/* Vertex: rosy_asset::position (src/Asset/Asset.h:92-99) == BasicInputVertex
 * (src/shaders/data.slang:24-30).  64 bytes, AoS. */
typedef struct rsy_position {
    float vertex[3];
    float normal[3];
    float tangents[4]; /* w = handedness */
    float color[4];
    float texture_coordinates[2];
} rsy_position;

/* gpu_material (src/Engine/Graphics.cpp:345-360) == BasicMaterialData (src/shaders/data.slang:77-91).
 * Image indices refer to the texture table given to rsy_set_materials, sampler indices to the
 * sampler table.  RSY_ABSENT marks a missing map.  64 bytes. */
typedef struct rsy_gpu_material {
    float color[4];
    float metallic_factor;
    float roughness_factor;
    float alpha_cutoff;
    uint32_t alpha_mode;
    uint32_t color_sampled_image_index;
    uint32_t color_sampler_index;
    uint32_t normal_sampled_image_index;
    uint32_t normal_sampler_index;
    uint32_t metallic_sampled_image_index;
    uint32_t metallic_sampler_index;
    uint32_t mixmap_sampled_image_index;
    uint32_t mixmap_sampler_index;
} rsy_gpu_material;

/* graphic_object_data (src/Engine/Graphics.cpp:422-427) == BasicGraphicsData
 * (src/shaders/data.slang:32-36).  Column-major matrices.  164 bytes. */
typedef struct rsy_object_record {
    float transform[16];
    float to_object_space_transform[16];
    float normal_transform[9];
} rsy_object_record;

/* gpu_scene_data (src/Engine/Graphics.cpp:281-306) == SceneData (src/shaders/data.slang:45-70).
 * 408 bytes.  csm_index_sampler / csm_index_near are kept for layout; the shadow map used is the
 * one the frame renders. */
typedef struct rsy_scene_data {
    float view[16];
    float proj[16];
    float view_projection[16];
    float shadow_projection_near[16];
    float sun_position[4];
    float camera_position[4];
    float ambient_color[4];
    float sunlight_color[4];
    uint32_t flip_lights[4];
    uint32_t flip_tangents[4];
    float draw_extent[2];
    float light_intensity;
    uint32_t csm_index_sampler;
    uint32_t csm_index_near;
    uint32_t fragment_output;
    uint32_t light_enabled;
    uint32_t tangent_space_enabled;
    uint32_t shadows_enabled;
    uint32_t normal_maps_enabled;
    uint32_t inverse_bnt;
    uint32_t ignore_asset_tangent_sign;
    uint32_t ensure_orthogonal_bitangent;
    uint32_t brdf_lighting_enabled;
} rsy_scene_data;

/* surface_graphics_data (src/Engine/Types.h:39-48) with fixed-width fields.  48 bytes. */
typedef struct rsy_draw {
    uint64_t mesh_index;
    uint64_t graphic_objects_offset;
    uint64_t graphics_object_index;
    uint64_t material_index; /* RSY_ABSENT (as u64 0xFFFFFFFF) == no material */
    uint32_t index_count;
    uint32_t start_index;
    uint32_t blended;
    uint32_t reserved;
} rsy_draw;

/* gpu_mesh_buffers (src/Engine/Graphics.cpp:376-381) plus the vertex count.  The vertex offset
 * is in bytes into the global vertex buffer, as in Graphics.cpp:4322; indices are mesh-local.
 * 24 bytes. */
typedef struct rsy_mesh_range {
    uint64_t vertex_buffer_offset;
    uint32_t index_offset;
    uint32_t num_indices;
    uint32_t num_vertices;
    uint32_t reserved;
} rsy_mesh_range;

/* rosy_asset::sampler (src/Asset/Asset.h:48-54): filter codes 0-2 nearest, 3-5 / absent linear
 * (filter_to_val Graphics.cpp:250-264); wrap 0 clamp-to-edge, 1 mirrored repeat, 2 repeat
 * (wrap_to_val Graphics.cpp:266-278).  8 bytes. */
typedef struct rsy_sampler {
    uint16_t min_filter;
    uint16_t mag_filter;
    uint16_t wrap_s;
    uint16_t wrap_t;
} rsy_sampler;

#pragma pack(pop)

/* Texture formats.  Colour, metallic-roughness and mixmap images are sRGB-encoded, normal maps
 * UNORM (Graphics.cpp:3632-3647).  The reference's images are BC7 (VK_FORMAT_BC7_SRGB_BLOCK /
 * VK_FORMAT_BC7_UNORM_BLOCK): BC7 data is decoded on the GPU when it is uploaded and kept as RGBA8
 * texels. */
typedef enum rsy_texture_format {
    RSY_TEX_RGBA8_SRGB = 0,
    RSY_TEX_RGBA8_UNORM = 1,
    RSY_TEX_RG8_UNORM = 2,
    RSY_TEX_RG8_SRGB = 3,
    RSY_TEX_BC7_SRGB = 4,  /* 16-byte blocks of 4x4 texels, ceil(w/4) x ceil(h/4) per level */
    RSY_TEX_BC7_UNORM = 5
} rsy_texture_format;

/* mip_levels: 0 or 1 = level 0 only; n = n levels in `pixels`, level 0 first, level l being
 * max(1, width >> l) x max(1, height >> l) (the DDS layout of the reference's images, whose chains
 * run down to 1 x 1); RSY_MIPS_BUILD = the library builds the full chain from level 0 with the
 * reference's packager filter (Packager.cpp:216-308: 2x2 box on linear, alpha-premultiplied colour
 * for sRGB images; on the stored values otherwise).  Sampled with the reference's samplers
 * (VK_SAMPLER_MIPMAP_MODE_LINEAR, no LOD clamp: Graphics.cpp:4002-4012): trilinear filtering at
 * the level of detail of the pixel's texture-coordinate derivatives. */
#define RSY_MIPS_BUILD 0xFFFFFFFFu
typedef struct rsy_texture_desc {
    uint32_t width;
    uint32_t height;
    uint32_t format;     /* rsy_texture_format */
    uint32_t mip_levels;
    const void* pixels;  /* level 0: width*height*bytes_per_texel (or BC7 blocks), row-major, row 0 = v 0 */
} rsy_texture_desc;

/* Fixed-function state the reference sets with dynamic state (Graphics.cpp:4988-5045, 5226-5231)
 * from draw_config_state / light_read_write_state (Types.h:131-172). */
typedef enum rsy_pcf_mode {
    RSY_PCF_REFERENCE_5TAP = 0, /* pcf.slang:1-12: 5 bilinear-compare taps, offset 0.00025 UV */
    RSY_PCF_GRID_3X3 = 1        /* 9 bilinear-compare taps at one-texel spacing */
} rsy_pcf_mode;

typedef struct rsy_raster_state {
    uint32_t cull_enabled;                  /* cull FRONT when set (Graphics.cpp:5010-5015) */
    uint32_t reverse_winding_order_enabled; /* front = CCW when set, else CW */
    uint32_t depth_bias_enabled;            /* shadow pass only (Graphics.cpp:5224-5231) */
    float depth_bias_constant;
    float depth_bias_clamp;
    float depth_bias_slope_factor;
    uint32_t pcf_mode;  /* rsy_pcf_mode */
    uint32_t tile_size; /* raster bin size in pixels: 8, 16, 32 or 64 (16 or more with a blended pass,
                           32 with 4 samples); 0 = autotuned (32 with 4 samples) */
    float clear_color[4];
    /* Samples per pixel of the camera views: 1 (0 = 1) or 4.  The reference's main pass renders
     * into a 4-sample colour attachment resolved by averaging (Graphics.cpp:1794-1808, 5566-5575);
     * with 4 samples, coverage and depth are taken at the standard sample locations and the
     * fragment shader runs once per pixel and primitive.  Shadow views are single-sampled. */
    uint32_t msaa_samples;
    uint32_t reserved;
} rsy_raster_state;

/* A camera view rendered into a visibility buffer and shaded. */
typedef struct rsy_camera_view {
    uint32_t width;
    uint32_t height;
    uint32_t reserved[2];
    float view[16];
    float proj[16];
    float view_projection[16];
    float camera_position[4]; /* w kept as given (the reference leaves it 0) */
} rsy_camera_view;

/* A shadow-map view (depth only).  split_far is the view-space distance up to which the cascade
 * is selected by the shading kernel; the single reference cascade uses +inf. */
typedef struct rsy_shadow_cascade {
    uint32_t size; /* square map, size x size texels */
    float split_far;
    uint32_t reserved[2];
    float light_view_projection[16];
} rsy_shadow_cascade;

typedef enum rsy_shade_mode {
    RSY_SHADE_NONE = 0,   /* visibility buffers and shadow maps only */
    RSY_SHADE_PLAIN = 1,  /* Cook-Torrance + PCF over the plain textures (basic.slang:137-344) */
    RSY_SHADE_NEURAL = 2  /* same BRDF; material maps decoded by the neural material decoder */
} rsy_shade_mode;

/* Views of one frame: its camera views, its cascades, and one blend view per camera view when the
 * blended list is not empty.  A frame with more is refused with RSY_INVALID_ARGUMENT. */
#define RSY_MAX_VIEWS 16u

typedef struct rsy_frame_desc {
    uint32_t n_camera_views; /* 0 = one view from the scene data (view_projection, draw_extent) */
    uint32_t n_cascades;     /* 0 = one cascade from scene data shadow_projection_near */
    const rsy_camera_view* camera_views;
    const rsy_shadow_cascade* cascades;
    uint32_t shade_mode;      /* rsy_shade_mode */
    uint32_t default_shadow_size; /* used when n_cascades == 0 */
} rsy_frame_desc;

typedef struct rsy_config {
    uint32_t abi_version;      /* RSY_ABI_VERSION */
    int32_t device;            /* HIP device ordinal */
    uint32_t max_window_width; /* rosy::config (Types.h:33-37) */
    uint32_t max_window_height;
} rsy_config;

typedef struct rsy_frame_stats {
    /* graphics_stats (src/Engine/DebugUI.h:7-13), counted as in Graphics.cpp:5285-5286 */
    uint64_t triangle_count;
    uint64_t draw_call_count;
    /* per-frame raster counters */
    uint64_t triangles_setup;   /* survived clip, cull and bbox rejection (all views) */
    uint64_t triangles_clipped; /* needed near/far/guard-band clipping */
    uint64_t tile_pairs;        /* (triangle, tile) bin entries */
    uint32_t tile_size;
    uint32_t overflowed;        /* times the last frame was re-run after a record, clip-queue, bin or
                                   fragment buffer overflowed and grew (0: it fit the first time) */
    float gpu_ms;               /* device time of the last frame (HIP events) */
    float stage_ms[8];          /* with RSY_DEBUG_STAGE_TIMING (HIP events on the context stream):
                                   [0] geometry (mvp, transform, setup, clip), [1] binning (scan,
                                   scatter), [2] tile raster kernel, [3] shading kernel, [4] blended
                                   pass, [5] the last rsy_present (draw image + blit) */
} rsy_frame_stats;

typedef struct rsy_ctx rsy_ctx;

/* ---------------------------------------------------------------------------------------------
 * Context lifetime.  Replaces graphics::init / graphics_device::init (Graphics.cpp:544-738,
 * 6119-6170) and graphics::deinit.  One context per thread; calls on a context are serialised.
 * ------------------------------------------------------------------------------------------- */
rsy_result rsy_create(const rsy_config* cfg, rsy_ctx** out_ctx);
void rsy_destroy(rsy_ctx* ctx);
const char* rsy_last_error(const rsy_ctx* ctx);
uint32_t rsy_abi_version(void);

/* ---------------------------------------------------------------------------------------------
 * Data entry points (graphics_device::set_asset, Graphics.cpp:3525-4689).
 * ------------------------------------------------------------------------------------------- */

/* Global vertex buffer (AoS 64-byte vertices), global index buffer (mesh-local u32 indices) and
 * the mesh table.  Replaces the vertex/index upload in set_asset (Graphics.cpp:4310-4560). */
rsy_result rsy_set_geometry(rsy_ctx* ctx, const rsy_position* vertices, uint64_t n_vertices,
                            const uint32_t* indices, uint64_t n_indices,
                            const rsy_mesh_range* meshes, uint32_t n_meshes);

/* Materials, samplers and decoded texture images.  Replaces the image, sampler and material
 * upload in set_asset (Graphics.cpp:3590-4160). */
rsy_result rsy_set_materials(rsy_ctx* ctx, const rsy_gpu_material* materials, uint32_t n_materials,
                             const rsy_texture_desc* textures, uint32_t n_textures,
                             const rsy_sampler* samplers, uint32_t n_samplers);

/* Object records and surfaces; builds the shadow (all), opaque (!blended) and blended lists in
 * the order given.  Replaces graphics_device::set_graphic_objects (Graphics.cpp:4691-4921). */
rsy_result rsy_set_graphic_objects(rsy_ctx* ctx, const rsy_object_record* objects, uint32_t n_objects,
                                   const rsy_draw* surfaces, uint32_t n_surfaces);

/* Partial object-record update.  Replaces update_graphic_objects (Graphics.cpp:4923-4927). */
rsy_result rsy_update_graphic_objects(rsy_ctx* ctx, uint64_t offset, const rsy_object_record* objects,
                                      uint32_t n_objects);

/* Per-frame camera + light uniforms and flags.  Replaces graphics_device::update
 * (Graphics.cpp:6016-6061); the record is copied, not retained. */
rsy_result rsy_update(rsy_ctx* ctx, const rsy_scene_data* scene_data, const rsy_raster_state* state);

/* Render one frame: every shadow cascade (shadow list), then every camera view (opaque list) into
 * visibility buffers, then deferred shading.  Replaces graphics_device::render
 * (Graphics.cpp:4934-6014) up to the MSAA resolve.  Work is enqueued on the context's stream;
 * outputs stay valid until the next render call. */
rsy_result rsy_render(rsy_ctx* ctx, const rsy_frame_desc* frame);

/* Capture the frame described by `frame` as a HIP graph and replay it `n_replays` times. */
rsy_result rsy_render_graph(rsy_ctx* ctx, const rsy_frame_desc* frame, uint32_t n_replays);

/* Wait for the context's stream. */
rsy_result rsy_synchronize(rsy_ctx* ctx);

/* ---------------------------------------------------------------------------------------------
 * Outputs (host readback; synchronises the stream).
 * ------------------------------------------------------------------------------------------- */

/* Visibility word per pixel: (float depth bits << 32) | (triangle id + 1); 0 = background.
 * Triangle ids count triangles over the opaque list in order.  With 4 samples the buffer holds
 * one plane per sample (n_pixels = 4 * width * height; sample s of pixel o at s * width * height
 * + o). */
rsy_result rsy_read_visibility(rsy_ctx* ctx, uint32_t view, uint64_t* out, uint64_t n_pixels);
/* Linear RGBA32F colour per pixel (the reference's draw image before the blit; with 4 samples the
 * resolved average of the samples). */
rsy_result rsy_read_color(rsy_ctx* ctx, uint32_t view, float* out_rgba, uint64_t n_pixels);
/* Shadow-map depth per texel (float, 0 = clear). */
rsy_result rsy_read_shadow_map(rsy_ctx* ctx, uint32_t cascade, float* out, uint64_t n_texels);
/* Texels of one mip level of texture `texture` as the sampler reads them (BC7 decoded to RGBA8,
 * built levels included): max(1, w >> level) * max(1, h >> level) * channels bytes (4 for RGBA8 and
 * BC7, 2 for RG8).  Counterpart of inspecting the uploaded VkImage; for validation. */
rsy_result rsy_read_texture(rsy_ctx* ctx, uint32_t texture, uint32_t level, uint8_t* out, uint64_t n_bytes);
/* Per-pixel shadow factor written when RSY_DEBUG_SHADOW_FACTOR is enabled (else zeros). */
rsy_result rsy_read_shadow_factor(rsy_ctx* ctx, uint32_t view, float* out, uint64_t n_pixels);
rsy_result rsy_get_frame_stats(rsy_ctx* ctx, rsy_frame_stats* out);

/* Device pointers of the last frame's outputs, for consumers on the same device. */
rsy_result rsy_get_device_outputs(rsy_ctx* ctx, uint32_t view, const void** vis, const void** color);

/* Debug and tuning switches. */
#define RSY_DEBUG_SHADOW_FACTOR 1u /* also write the per-pixel PCF shadow factor */
#define RSY_DEBUG_STAGE_TIMING 2u  /* record HIP events between stages */
#define RSY_DEBUG_MIN_CAPACITY 4u  /* start the frame's record, clip-queue, bin-pair, bin-list and
                                      blended-fragment capacities small, so that frames overflow and
                                      take the regrowth path (takes effect at the next new layout),
                                      and give a new layout's first frame 64 raster workgroups, too
                                      few for its work items, so that it is re-run with a grid that
                                      fits */
rsy_result rsy_set_debug_flags(rsy_ctx* ctx, uint32_t flags);

/* The output stage of the last shaded frame for one camera view (Graphics.cpp:1908, 5835-5880):
 * its colour becomes the RGBA16F draw image (round to nearest even), which is blitted with a
 * linear filter (vkCmdBlitImage2 mapping: destination texel centre (x + 0.5) * src / dst,
 * clamp-to-edge) to a width x height B8G8R8A8_SRGB image (colour through the sRGB encode, alpha
 * linear, both rounded to 8 bits).  Enqueued on the context stream; the readbacks synchronise. */
rsy_result rsy_present(rsy_ctx* ctx, uint32_t view, uint32_t width, uint32_t height);
rsy_result rsy_read_draw_image(rsy_ctx* ctx, uint16_t* out_rgba16f, uint64_t n_pixels);
rsy_result rsy_read_present(rsy_ctx* ctx, uint8_t* out_bgra8, uint64_t n_pixels);

/* Autotuner: replays the given frame n_iters times per candidate bin size (8, 16, 32, 64; only 32
 * with 4 samples, 16 and up with a blended pass) and keeps the fastest for later frames with
 * tile_size == 0.  Writes the chosen size. */
rsy_result rsy_autotune(rsy_ctx* ctx, const rsy_frame_desc* frame, uint32_t n_iters, uint32_t* chosen_tile);

/* ---------------------------------------------------------------------------------------------
 * Allocation-free frame.  No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * rsy_render_noalloc renders the frame rsy_render would, with the same kernels, but calls no
 * hipMalloc, hipFree or allocating routine: the frame's working memory (transformed vertices,
 * triangle records, bins, bin lists, raster work items, counters and view tables) is carved from
 * one static device array of RSY_STATIC_SCRATCH_BYTES per GPU, and the outputs are the caller's
 * device buffers.  A frame whose working memory exceeds the array is rejected with RSY_OVERFLOW
 * before anything runs (rsy_last_error states the size); one whose bin lists overflow at run time
 * returns RSY_OVERFLOW afterwards (there is no regrowth; the outputs are then undefined).
 * Limits: single-sampled frames (msaa_samples 1) without a blended pass (the blended list empty, or
 * RSY_SHADE_NONE); the per-pixel shadow factor and stage timings are not written.
 * The call is synchronous: it returns when the frame is complete.  Every call on a GPU shares that
 * GPU's array and its contents are those of the last call: calls that overlap on one GPU (two
 * contexts on two threads, say) race.  The scene set-up calls (rsy_set_geometry, rsy_set_materials,
 * rsy_set_graphic_objects, rsy_update) still allocate, as they do for rsy_render.  The draw lists
 * must be built: after rsy_set_geometry, call rsy_set_graphic_objects before this frame
 * (RSY_INVALID_STATE otherwise); rsy_set_materials keeps them built. */
#define RSY_STATIC_SCRATCH_BYTES (128ull << 20)
typedef struct rsy_frame_targets {
    uint32_t n_camera_views;     /* the frame's camera views */
    uint32_t n_cascades;         /* the frame's shadow cascades (the default one when the frame names none) */
    uint64_t* const* visibility; /* per camera view: width*height words as rsy_read_visibility gives them */
    float* const* color;         /* per camera view: width*height RGBA32F (may be null with RSY_SHADE_NONE) */
    float* const* shadow_maps;   /* per cascade: size*size f32 depths */
} rsy_frame_targets;           /* each pointer is device memory of the context's device */
rsy_result rsy_render_noalloc(rsy_ctx* ctx, const rsy_frame_desc* frame, const rsy_frame_targets* targets);

/* ---------------------------------------------------------------------------------------------
 * Neural material decoder (no counterpart in the reference: an alternative to sampling the three
 * material maps of basic.slang:172-221, judged against the plain sampler).
 *
 * Each eligible material gets an 8-channel latent grid of latent_size^2 cells; one MLP shared by
 * all materials maps [bilinear latent (8), 1-|2ax-1|, 1-|2ay-1|] -> 32 ReLU -> 7 outputs clamped to
 * [0,1]: linear albedo rgb, normal-map r, g, metallic, roughness -- the plain sampler's values for
 * those maps.  Eligible: RGBA8 opaque albedo, two-channel normal and metallic-roughness maps,
 * bilinear REPEAT samplers.  Training (Adam on the GPU) regresses the decoder onto plain-sampler
 * values at random points; evaluation decodes every albedo texel centre, computes the PSNR over
 * the 7 channels (peak 1) and activates the materials at or above psnr_threshold.  Shade mode
 * RSY_SHADE_NEURAL decodes active materials (FP16 latents, FP16 MFMA with FP32 accumulation) and
 * samples the plain maps for every other material.
 * ------------------------------------------------------------------------------------------- */
typedef enum rsy_neural_init_mode {
    RSY_NEURAL_INIT_FROM_MAPS = 0, /* latent channels 0..6 = plain values at cell centres, MLP = identity */
    RSY_NEURAL_INIT_RANDOM = 1     /* uniform latents in [-0.1, 0.1], random MLP */
} rsy_neural_init_mode;

typedef enum rsy_neural_weights {
    RSY_NEURAL_WEIGHTS_FP16 = 0, /* FP16 weights */
    RSY_NEURAL_WEIGHTS_FP8 = 1   /* OCP FP8 E4M3 weights with one FP32 scale per layer (b2 stays FP32) */
} rsy_neural_weights;

typedef struct rsy_neural_config {
    uint32_t latent_size;    /* cells per side: a power of two in [4, 4096] */
    uint32_t batch;          /* samples per step, a multiple of 256 */
    uint32_t init;           /* rsy_neural_init_mode */
    uint32_t region;         /* training region side in cells: a power of two in [2, 32], <= latent_size */
    float lr_mlp;            /* Adam learning rates */
    float lr_latent;
    float beta1, beta2, eps;
    float psnr_threshold;    /* dB */
    uint32_t decay_steps;    /* cosine learning-rate decay to 0 over this many steps (0: constant) */
    uint32_t weights;        /* rsy_neural_weights: storage of the inference MLP weights */
    uint64_t seed;
} rsy_neural_config;

typedef struct rsy_neural_stats {
    uint32_t n_materials;    /* materials of the context */
    uint32_t n_eligible;     /* materials with a latent grid */
    uint32_t n_active;       /* materials decoded when rendering (PSNR >= threshold) */
    uint32_t steps;          /* Adam steps run so far */
    float train_ms;          /* device time of the last rsy_neural_train call */
    float eval_ms;           /* device time of the last rsy_neural_evaluate call */
    float last_loss;         /* mean squared error of the last step */
    float reserved;
} rsy_neural_stats;

/* Fills the defaults: latent 256, batch 65536, init from maps, region 16, lr 1e-3 / 1e-3,
 * betas 0.9 / 0.999, eps 1e-8, threshold 40 dB, no decay, seed 1. */
void rsy_neural_default_config(rsy_neural_config* cfg);
/* Allocates and initialises the decoder for the context's current materials (rsy_set_materials
 * first).  Re-initialising discards any trained state. */
rsy_result rsy_neural_init(rsy_ctx* ctx, const rsy_neural_config* cfg);
/* Runs n_steps Adam steps on the GPU.  loss_history (nullable) receives each step's mean squared
 * error.  Invalidates the last evaluation. */
rsy_result rsy_neural_train(rsy_ctx* ctx, uint32_t n_steps, float* loss_history);
/* Exports the FP16 inference parameters, measures every eligible material's PSNR and sets the
 * active set.  psnr (nullable, n_materials entries) receives the PSNR, NaN for ineligible ones. */
rsy_result rsy_neural_evaluate(rsy_ctx* ctx, float* psnr, uint32_t n_materials);
rsy_result rsy_neural_get_stats(rsy_ctx* ctx, rsy_neural_stats* out);
/* Overrides the active set after an evaluation (active[m] = 0 forces the plain maps; 1 is only
 * accepted for eligible materials). */
rsy_result rsy_neural_set_active(rsy_ctx* ctx, const uint32_t* active, uint32_t n_materials);
/* FP32 master parameters (checkpoint / restore): latents [eligible][L][L][8] in material order,
 * MLP [583] = W1 [32][10], b1 [32], W2 [7][32], b2 [7]. */
rsy_result rsy_neural_get_params(rsy_ctx* ctx, float* latents, uint64_t n_latents, float* mlp, uint32_t n_mlp);
rsy_result rsy_neural_set_params(rsy_ctx* ctx, const float* latents, uint64_t n_latents, const float* mlp,
                                 uint32_t n_mlp);
/* Exact loss and gradients of the training objective on a given batch (the training kernels run
 * over the whole grid with the caller's samples; no update): material[n], uv[n][2], target[n][7].
 * Requires latent_size <= 32.  grad_mlp [583], grad_latents as in rsy_neural_get_params. */
rsy_result rsy_neural_gradients(rsy_ctx* ctx, const uint32_t* material, const float* uv, const float* target,
                                uint32_t n, float* loss, float* grad_mlp, float* grad_latents, uint64_t n_latents);
/* Switches the inference weight storage (rsy_neural_weights); takes effect at the next
 * rsy_neural_evaluate (neural frames are refused until then). */
rsy_result rsy_neural_set_weights(rsy_ctx* ctx, uint32_t weights);
/* The MLP the inference path evaluates, as FP32 values in the rsy_neural_get_params layout: FP16- or
 * FP8-rounded (times the layer scale) W1, b1, W2 and FP32 b2.  After rsy_neural_evaluate. */
rsy_result rsy_neural_get_inference_mlp(rsy_ctx* ctx, float* mlp, uint32_t n_mlp);
/* Decoder outputs (inference path, after rsy_neural_evaluate) or, with targets != 0, the plain
 * sampler's values that training regresses onto, at material[n], uv[n][2] -> out[n][7]. */
rsy_result rsy_neural_decode(rsy_ctx* ctx, const uint32_t* material, const float* uv, uint32_t n, uint32_t targets,
                             float* out);

/* ---------------------------------------------------------------------------------------------
 * Input producers: closed-form restatements of the reference's camera, light and node math.
 * All matrices are column-major float[16] (glm value_ptr order).
 * ------------------------------------------------------------------------------------------- */

/* synthetic_camera::get_projection (src/Engine/Camera.cpp:62-83): reverse-Z, Vulkan y flip.
 * fov in radians. */
void rsy_camera_projection(double aspect, double near_plane, double far_plane, double fov, float out[16]);
/* synthetic_camera::get_view_matrix / get_rotation_matrix (Camera.cpp:85-98). */
void rsy_camera_view_matrix(const float position[3], float pitch, float yaw, float out[16]);
/* Level.cpp:464-528: sun position and shadowProjNear = lp * inverse(light_line_rot * T(0,0,d)). */
void rsy_light_matrices(float sun_distance, float sun_pitch, float sun_yaw, float cascade_level,
                        float orthographic_depth, float out_sun_position[4], float out_shadow_projection[16]);
/* node_state::get_world_space_transform (Node.cpp:96-103) with identity coordinate system and
 * parent, plus to-object = inverse(W) and normal = transpose(inverse(mat3(W))) (Node.cpp:225-247). */
void rsy_object_record_from_trs(const float translate[3], float scale, float yaw, rsy_object_record* out);
/* Matrix product c = a * b (column-major, FP32 fmaf chain in k order). */
void rsy_mat4_mul(const float a[16], const float b[16], float c[16]);

/* ---------------------------------------------------------------------------------------------
 * .rsy asset files: the Packager's binary format (src/Asset/Asset.h:9-149, Asset.cpp:6-848),
 * read and written on the host, and packed into the inputs of rsy_set_geometry and
 * rsy_set_materials the way set_asset packs them (Graphics.cpp:3525-4689).
 *
 * Layout (little endian, the reference's structs with natural alignment): an 80-byte header
 * (magic "RSYF" = 0x52535946, version 1, endianness 1, coordinate system float[16], root scene);
 * six u64 counts (materials, samplers, scenes, nodes, images, meshes); the 68-byte materials; the
 * 8-byte samplers; per scene a u64 node count and its u32 node indices; per node seven u64 sizes
 * (1, 1, 1, 1, 1, children, name bytes) then translate float[3], scale, yaw, transform float[16],
 * mesh id, u32 children and the name; per image two u64 sizes (1, name bytes), the u32 image type
 * and the name; per mesh three u64 counts (positions, indices, surfaces) then the 64-byte
 * positions, u32 indices and 36-byte surfaces.  A node's coordinate system and is_world_node are
 * not part of the file (Asset.h:64-70).
 *
 * Errors follow asset::read / asset::write: RSY_OPEN_FAILED when a file cannot be opened,
 * RSY_WRITE_FAILED when it cannot be written, RSY_READ_FAILED for a short file, a magic, version
 * or endianness mismatch, node or image size fields other than 1, or counts that exceed the bytes
 * left (nothing is allocated for them), RSY_INVALID_ARGUMENT for null pointers or a capacity too
 * small.  Image pixels live outside the file (the engine loads them by name): the caller decodes
 * them and passes them to rsy_set_materials in the asset's image order.
 * ------------------------------------------------------------------------------------------- */
typedef struct rsy_asset rsy_asset;

/* rosy_asset::material (Asset.h:26-46): 68 bytes with the compiler's padding made explicit. */
typedef struct rsy_asset_material {
    uint8_t double_sided;
    uint8_t pad0[3];
    float base_color_factor[4];
    float metallic_factor;
    float roughness_factor;
    uint8_t alpha_mode;
    uint8_t pad1[3];
    float alpha_cutoff;
    uint32_t color_image_index; /* RSY_ABSENT == not present, likewise below */
    uint32_t color_sampler_index;
    uint32_t normal_image_index;
    uint32_t normal_sampler_index;
    uint32_t metallic_image_index;
    uint32_t metallic_sampler_index;
    uint32_t mixmap_image_index;
    uint32_t mixmap_sampler_index;
} rsy_asset_material;

/* rosy_asset::surface (Asset.h:56-63): 36 bytes. */
typedef struct rsy_asset_surface {
    uint32_t start_index;
    uint32_t count;
    uint32_t material;
    float min_bounds[3];
    float max_bounds[3];
} rsy_asset_surface;

/* rosy_asset::node (Asset.h:64-84) as stored. */
typedef struct rsy_asset_node {
    float world_translate[3];
    float world_scale;
    float world_yaw;
    float transform[16];
    uint32_t mesh_id; /* RSY_ABSENT == no mesh */
    const uint32_t* child_nodes;
    uint64_t n_child_nodes;
    const char* name; /* not NUL-terminated */
    uint64_t name_len;
} rsy_asset_node;

typedef struct rsy_asset_scene { /* rosy_asset::scene (Asset.h:86-89) */
    const uint32_t* nodes;
    uint64_t n_nodes;
} rsy_asset_scene;

typedef struct rsy_asset_image { /* rosy_asset::image (Asset.h:107-117) */
    uint32_t image_type;         /* 0 color, 1 normal map, 2 metallic-roughness, 3 mixmap */
    const char* name;
    uint64_t name_len;
} rsy_asset_image;

typedef struct rsy_asset_mesh { /* rosy_asset::mesh (Asset.h:100-105) */
    const rsy_position* positions;
    uint64_t n_positions;
    const uint32_t* indices; /* mesh-local */
    uint64_t n_indices;
    const rsy_asset_surface* surfaces;
    uint64_t n_surfaces;
} rsy_asset_mesh;

/* An asset in memory (rosy_asset::asset, Asset.h:125-148), as borrowed arrays. */
typedef struct rsy_asset_desc {
    float coordinate_system[16];
    uint32_t root_scene;
    const rsy_asset_material* materials;
    uint64_t n_materials;
    const rsy_sampler* samplers;
    uint64_t n_samplers;
    const rsy_asset_scene* scenes;
    uint64_t n_scenes;
    const rsy_asset_node* nodes;
    uint64_t n_nodes;
    const rsy_asset_image* images;
    uint64_t n_images;
    const rsy_asset_mesh* meshes;
    uint64_t n_meshes;
} rsy_asset_desc;

/* asset::write (Asset.cpp:39-403) into memory: with out == NULL only *size is written (the file
 * size); else the file bytes go to out[0, capacity). */
rsy_result rsy_asset_encode(const rsy_asset_desc* desc, void* out, uint64_t capacity, uint64_t* size);
/* asset::write to a file. */
rsy_result rsy_asset_write(const rsy_asset_desc* desc, const char* path);
/* asset::read (Asset.cpp:405-848) from memory or a file; the asset owns copies of everything. */
rsy_result rsy_asset_decode(const void* data, uint64_t size, rsy_asset** out_asset);
rsy_result rsy_asset_read(const char* path, rsy_asset** out_asset);
/* The asset's contents as a description whose arrays stay valid until rsy_asset_destroy. */
rsy_result rsy_asset_describe(const rsy_asset* asset, rsy_asset_desc* out);
void rsy_asset_destroy(rsy_asset* asset);
/* Sizes of the packed geometry: vertices and indices over all meshes, and the mesh count. */
rsy_result rsy_asset_geometry_size(const rsy_asset* asset, uint64_t* n_vertices, uint64_t* n_indices,
                                   uint32_t* n_meshes);
/* The global vertex buffer (meshes' positions back to back), index buffer (mesh-local indices back
 * to back) and mesh table as set_asset builds them (Graphics.cpp:4310-4333): byte offset of each
 * mesh's first vertex, its first index and index count. */
rsy_result rsy_asset_pack_geometry(const rsy_asset* asset, rsy_position* vertices, uint64_t n_vertices,
                                   uint32_t* indices, uint64_t n_indices, rsy_mesh_range* meshes, uint32_t n_meshes);
/* gpu_material records as set_asset fills them (Graphics.cpp:4066-4152), with texture indices into
 * the asset's image list and sampler indices into its sampler list: a map whose image index is
 * out of range is absent, and a present map whose sampler index is out of range uses the default
 * sampler (RSY_ABSENT here: REPEAT, linear). */
rsy_result rsy_asset_pack_materials(const rsy_asset* asset, rsy_gpu_material* materials, uint32_t n_materials);

/* ---------------------------------------------------------------------------------------------
 * Synthetic scene generator for the five workloads (seeded; host-only).
 * ------------------------------------------------------------------------------------------- */
typedef struct rsy_scene_desc {
    uint32_t workload;        /* 0 sphere, 1 scene-scale, 2 batched views, 3 learned decoder, 4 combined */
    uint32_t texture_mips;    /* 1: every texture gets its full mip chain (RSY_MIPS_BUILD); 0: level 0 only */
    uint64_t seed;
    /* overrides, 0 = the workload's default */
    uint32_t width;
    uint32_t height;
    uint32_t shadow_size;
    uint32_t total_triangles;
    uint32_t n_meshes;
    uint32_t n_materials;
    uint32_t texture_size;
    uint32_t n_translucent;   /* materials [0, n) translucent (default 0): even BLEND, odd MASK, with a
                                 sin(8 pi u) sin(8 pi v) albedo alpha; their surfaces go to the blended list */
    uint32_t texture_detail;  /* 1: each material's maps carry detail at one of five scales up to 512
                                 cycles across the map, stripes up to 200 cycles and small blobs (content
                                 above a 256^2 latent grid's Nyquist rate); 0 (default): up to 32 cycles */
} rsy_scene_desc;

typedef struct rsy_scene {
    const rsy_position* vertices;
    uint64_t n_vertices;
    const uint32_t* indices;
    uint64_t n_indices;
    const rsy_mesh_range* meshes;
    uint32_t n_meshes;
    uint32_t n_objects;
    const rsy_object_record* objects;
    const rsy_draw* surfaces;
    uint32_t n_surfaces;
    uint32_t n_materials;
    const rsy_gpu_material* materials;
    const rsy_texture_desc* textures;
    uint32_t n_textures;
    uint32_t n_samplers;
    const rsy_sampler* samplers;
    rsy_scene_data scene_data;
    rsy_raster_state raster_state;
    uint32_t n_camera_views;
    uint32_t n_cascades;
    const rsy_camera_view* camera_views;
    const rsy_shadow_cascade* cascades;
    uint32_t default_shadow_size;
    uint32_t workload;
} rsy_scene;

rsy_result rsy_scene_generate(const rsy_scene_desc* desc, rsy_scene** out_scene);
void rsy_scene_free(rsy_scene* scene);

#ifdef __cplusplus
} /* extern "C" */
#endif

#ifdef __cplusplus
static_assert(sizeof(rsy_position) == 64, "rsy_position must be 64 bytes (Asset.h:92-99)");
static_assert(sizeof(rsy_gpu_material) == 64, "rsy_gpu_material must be 64 bytes (Graphics.cpp:345-360)");
static_assert(sizeof(rsy_object_record) == 164, "rsy_object_record must be 164 bytes (Graphics.cpp:422-427)");
static_assert(sizeof(rsy_scene_data) == 408, "rsy_scene_data must be 408 bytes (Graphics.cpp:281-306)");
static_assert(sizeof(rsy_draw) == 48, "rsy_draw must be 48 bytes");
static_assert(sizeof(rsy_mesh_range) == 24, "rsy_mesh_range must be 24 bytes");
static_assert(sizeof(rsy_sampler) == 8, "rsy_sampler must be 8 bytes (Asset.h:48-54)");
static_assert(sizeof(rsy_asset_material) == 68, "rsy_asset_material must be 68 bytes (Asset.h:26-46)");
static_assert(sizeof(rsy_asset_surface) == 36, "rsy_asset_surface must be 36 bytes (Asset.h:56-63)");
#endif

#endif /* ROSY_AMD_H */
