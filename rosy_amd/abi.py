"""ctypes mirror of include/rosy_amd.h (the C ABI of the raster-and-shade frame path).

The record layouts are the reference's packed structs (rosy_asset::position Asset.h:92-99,
gpu_material Graphics.cpp:345-360, graphic_object_data Graphics.cpp:422-427, gpu_scene_data
Graphics.cpp:281-306, surface_graphics_data Types.h:39-48); sizes are checked at import.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librosy_amd.so")

# rosy::result (Types.h:13-31)
RESULT_NAMES = {
    0: "ok", 1: "state_changed", 2: "error", 3: "invalid_argument", 4: "allocation_failure",
    5: "graphics_init_failure", 6: "graphics_frame_failure", 7: "graphics_swapchain_failure",
    8: "open_failed", 9: "write_failed", 10: "copy_failed", 11: "read_failed", 12: "create_failed",
    13: "update_failed", 14: "overflow", 15: "invalid_state",
}
RSY_OK, RSY_ERROR, RSY_INVALID_ARGUMENT, RSY_ALLOCATION_FAILURE = 0, 2, 3, 4
RSY_GRAPHICS_INIT_FAILURE, RSY_READ_FAILED, RSY_OVERFLOW, RSY_INVALID_STATE = 5, 11, 14, 15
RSY_ABSENT = 0xFFFFFFFF
RSY_ABI_VERSION = 1
SHADE_NONE, SHADE_PLAIN, SHADE_NEURAL = 0, 1, 2
PCF_REFERENCE_5TAP, PCF_GRID_3X3 = 0, 1
DEBUG_SHADOW_FACTOR, DEBUG_STAGE_TIMING, DEBUG_MIN_CAPACITY = 1, 2, 4
TEX_RGBA8_SRGB, TEX_RGBA8_UNORM, TEX_RG8_UNORM, TEX_RG8_SRGB, TEX_BC7_SRGB, TEX_BC7_UNORM = 0, 1, 2, 3, 4, 5
MIPS_BUILD = 0xFFFFFFFF

F16 = C.c_float * 16
F4 = C.c_float * 4


class Position(C.Structure):
    _pack_ = 1
    _fields_ = [("vertex", C.c_float * 3), ("normal", C.c_float * 3), ("tangents", C.c_float * 4),
                ("color", C.c_float * 4), ("texture_coordinates", C.c_float * 2)]


class GpuMaterial(C.Structure):
    _pack_ = 1
    _fields_ = [("color", F4), ("metallic_factor", C.c_float), ("roughness_factor", C.c_float),
                ("alpha_cutoff", C.c_float), ("alpha_mode", C.c_uint32),
                ("color_sampled_image_index", C.c_uint32), ("color_sampler_index", C.c_uint32),
                ("normal_sampled_image_index", C.c_uint32), ("normal_sampler_index", C.c_uint32),
                ("metallic_sampled_image_index", C.c_uint32), ("metallic_sampler_index", C.c_uint32),
                ("mixmap_sampled_image_index", C.c_uint32), ("mixmap_sampler_index", C.c_uint32)]


class ObjectRecord(C.Structure):
    _pack_ = 1
    _fields_ = [("transform", F16), ("to_object_space_transform", F16), ("normal_transform", C.c_float * 9)]


class SceneData(C.Structure):
    _pack_ = 1
    _fields_ = [("view", F16), ("proj", F16), ("view_projection", F16), ("shadow_projection_near", F16),
                ("sun_position", F4), ("camera_position", F4), ("ambient_color", F4), ("sunlight_color", F4),
                ("flip_lights", C.c_uint32 * 4), ("flip_tangents", C.c_uint32 * 4), ("draw_extent", C.c_float * 2),
                ("light_intensity", C.c_float), ("csm_index_sampler", C.c_uint32), ("csm_index_near", C.c_uint32),
                ("fragment_output", C.c_uint32), ("light_enabled", C.c_uint32),
                ("tangent_space_enabled", C.c_uint32), ("shadows_enabled", C.c_uint32),
                ("normal_maps_enabled", C.c_uint32), ("inverse_bnt", C.c_uint32),
                ("ignore_asset_tangent_sign", C.c_uint32), ("ensure_orthogonal_bitangent", C.c_uint32),
                ("brdf_lighting_enabled", C.c_uint32)]


class Draw(C.Structure):
    _pack_ = 1
    _fields_ = [("mesh_index", C.c_uint64), ("graphic_objects_offset", C.c_uint64),
                ("graphics_object_index", C.c_uint64), ("material_index", C.c_uint64),
                ("index_count", C.c_uint32), ("start_index", C.c_uint32), ("blended", C.c_uint32),
                ("reserved", C.c_uint32)]


class MeshRange(C.Structure):
    _pack_ = 1
    _fields_ = [("vertex_buffer_offset", C.c_uint64), ("index_offset", C.c_uint32), ("num_indices", C.c_uint32),
                ("num_vertices", C.c_uint32), ("reserved", C.c_uint32)]


class Sampler(C.Structure):
    _pack_ = 1
    _fields_ = [("min_filter", C.c_uint16), ("mag_filter", C.c_uint16), ("wrap_s", C.c_uint16),
                ("wrap_t", C.c_uint16)]


class TextureDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32), ("mip_levels", C.c_uint32),
                ("pixels", C.c_void_p)]


class RasterState(C.Structure):
    _fields_ = [("cull_enabled", C.c_uint32), ("reverse_winding_order_enabled", C.c_uint32),
                ("depth_bias_enabled", C.c_uint32), ("depth_bias_constant", C.c_float),
                ("depth_bias_clamp", C.c_float), ("depth_bias_slope_factor", C.c_float),
                ("pcf_mode", C.c_uint32), ("tile_size", C.c_uint32), ("clear_color", F4),
                ("msaa_samples", C.c_uint32), ("reserved", C.c_uint32)]


class CameraView(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("reserved", C.c_uint32 * 2), ("view", F16),
                ("proj", F16), ("view_projection", F16), ("camera_position", F4)]


class ShadowCascade(C.Structure):
    _fields_ = [("size", C.c_uint32), ("split_far", C.c_float), ("reserved", C.c_uint32 * 2),
                ("light_view_projection", F16)]


class FrameDesc(C.Structure):
    _fields_ = [("n_camera_views", C.c_uint32), ("n_cascades", C.c_uint32),
                ("camera_views", C.POINTER(CameraView)), ("cascades", C.POINTER(ShadowCascade)),
                ("shade_mode", C.c_uint32), ("default_shadow_size", C.c_uint32)]


class Config(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("device", C.c_int32), ("max_window_width", C.c_uint32),
                ("max_window_height", C.c_uint32)]


class FrameTargets(C.Structure):
    _fields_ = [("n_camera_views", C.c_uint32), ("n_cascades", C.c_uint32),
                ("visibility", C.POINTER(C.c_void_p)), ("color", C.POINTER(C.c_void_p)),
                ("shadow_maps", C.POINTER(C.c_void_p))]


STATIC_SCRATCH_BYTES = 128 << 20
MAX_VIEWS = 16


class FrameStats(C.Structure):
    _fields_ = [("triangle_count", C.c_uint64), ("draw_call_count", C.c_uint64), ("triangles_setup", C.c_uint64),
                ("triangles_clipped", C.c_uint64), ("tile_pairs", C.c_uint64), ("tile_size", C.c_uint32),
                ("overflowed", C.c_uint32), ("gpu_ms", C.c_float), ("stage_ms", C.c_float * 8)]


class SceneDesc(C.Structure):
    _fields_ = [("workload", C.c_uint32), ("texture_mips", C.c_uint32), ("seed", C.c_uint64), ("width", C.c_uint32),
                ("height", C.c_uint32), ("shadow_size", C.c_uint32), ("total_triangles", C.c_uint32),
                ("n_meshes", C.c_uint32), ("n_materials", C.c_uint32), ("texture_size", C.c_uint32),
                ("n_translucent", C.c_uint32), ("texture_detail", C.c_uint32)]


class Scene(C.Structure):
    _fields_ = [("vertices", C.POINTER(Position)), ("n_vertices", C.c_uint64),
                ("indices", C.POINTER(C.c_uint32)), ("n_indices", C.c_uint64),
                ("meshes", C.POINTER(MeshRange)), ("n_meshes", C.c_uint32), ("n_objects", C.c_uint32),
                ("objects", C.POINTER(ObjectRecord)), ("surfaces", C.POINTER(Draw)), ("n_surfaces", C.c_uint32),
                ("n_materials", C.c_uint32), ("materials", C.POINTER(GpuMaterial)),
                ("textures", C.POINTER(TextureDesc)), ("n_textures", C.c_uint32), ("n_samplers", C.c_uint32),
                ("samplers", C.POINTER(Sampler)), ("scene_data", SceneData), ("raster_state", RasterState),
                ("n_camera_views", C.c_uint32), ("n_cascades", C.c_uint32),
                ("camera_views", C.POINTER(CameraView)), ("cascades", C.POINTER(ShadowCascade)),
                ("default_shadow_size", C.c_uint32), ("workload", C.c_uint32)]


class NeuralConfig(C.Structure):
    _fields_ = [("latent_size", C.c_uint32), ("batch", C.c_uint32), ("init", C.c_uint32), ("region", C.c_uint32),
                ("lr_mlp", C.c_float), ("lr_latent", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float), ("psnr_threshold", C.c_float), ("decay_steps", C.c_uint32),
                ("weights", C.c_uint32), ("seed", C.c_uint64)]


class NeuralStats(C.Structure):
    _fields_ = [("n_materials", C.c_uint32), ("n_eligible", C.c_uint32), ("n_active", C.c_uint32),
                ("steps", C.c_uint32), ("train_ms", C.c_float), ("eval_ms", C.c_float), ("last_loss", C.c_float),
                ("reserved", C.c_float)]


class AssetMaterial(C.Structure):
    """rosy_asset::material (Asset.h:26-46), natural alignment: 68 bytes."""
    _fields_ = [("double_sided", C.c_uint8), ("pad0", C.c_uint8 * 3), ("base_color_factor", F4),
                ("metallic_factor", C.c_float), ("roughness_factor", C.c_float), ("alpha_mode", C.c_uint8),
                ("pad1", C.c_uint8 * 3), ("alpha_cutoff", C.c_float),
                ("color_image_index", C.c_uint32), ("color_sampler_index", C.c_uint32),
                ("normal_image_index", C.c_uint32), ("normal_sampler_index", C.c_uint32),
                ("metallic_image_index", C.c_uint32), ("metallic_sampler_index", C.c_uint32),
                ("mixmap_image_index", C.c_uint32), ("mixmap_sampler_index", C.c_uint32)]


class AssetSurface(C.Structure):
    """rosy_asset::surface (Asset.h:56-63): 36 bytes."""
    _fields_ = [("start_index", C.c_uint32), ("count", C.c_uint32), ("material", C.c_uint32),
                ("min_bounds", C.c_float * 3), ("max_bounds", C.c_float * 3)]


class AssetNode(C.Structure):
    _fields_ = [("world_translate", C.c_float * 3), ("world_scale", C.c_float), ("world_yaw", C.c_float),
                ("transform", F16), ("mesh_id", C.c_uint32), ("child_nodes", C.POINTER(C.c_uint32)),
                ("n_child_nodes", C.c_uint64), ("name", C.c_void_p), ("name_len", C.c_uint64)]


class AssetScene(C.Structure):
    _fields_ = [("nodes", C.POINTER(C.c_uint32)), ("n_nodes", C.c_uint64)]


class AssetImage(C.Structure):
    _fields_ = [("image_type", C.c_uint32), ("name", C.c_void_p), ("name_len", C.c_uint64)]


class AssetMesh(C.Structure):
    _fields_ = [("positions", C.POINTER(Position)), ("n_positions", C.c_uint64),
                ("indices", C.POINTER(C.c_uint32)), ("n_indices", C.c_uint64),
                ("surfaces", C.POINTER(AssetSurface)), ("n_surfaces", C.c_uint64)]


class AssetDesc(C.Structure):
    _fields_ = [("coordinate_system", F16), ("root_scene", C.c_uint32),
                ("materials", C.POINTER(AssetMaterial)), ("n_materials", C.c_uint64),
                ("samplers", C.POINTER(Sampler)), ("n_samplers", C.c_uint64),
                ("scenes", C.POINTER(AssetScene)), ("n_scenes", C.c_uint64),
                ("nodes", C.POINTER(AssetNode)), ("n_nodes", C.c_uint64),
                ("images", C.POINTER(AssetImage)), ("n_images", C.c_uint64),
                ("meshes", C.POINTER(AssetMesh)), ("n_meshes", C.c_uint64)]


RSY_OPEN_FAILED, RSY_WRITE_FAILED = 8, 9

NEURAL_WEIGHTS_FP16 = 0
NEURAL_WEIGHTS_FP8 = 1
NEURAL_INIT_FROM_MAPS = 0
NEURAL_INIT_RANDOM = 1
NEURAL_PARAMS = 583   # W1 [32][10], b1 [32], W2 [7][32], b2 [7]
NEURAL_CHANNELS = 8
NEURAL_OUTPUTS = 7

EXPECTED_SIZES = {Position: 64, GpuMaterial: 64, ObjectRecord: 164, SceneData: 408, Draw: 48, MeshRange: 24,
                  Sampler: 8, AssetMaterial: 68, AssetSurface: 36}
for _t, _n in EXPECTED_SIZES.items():
    assert C.sizeof(_t) == _n, (_t.__name__, C.sizeof(_t), _n)

# exported symbols and their prototypes: (name, restype, argtypes)
PROTOTYPES = [
    ("rsy_abi_version", C.c_uint32, []),
    ("rsy_create", C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    ("rsy_destroy", None, [C.c_void_p]),
    ("rsy_last_error", C.c_char_p, [C.c_void_p]),
    ("rsy_set_geometry", C.c_int, [C.c_void_p, C.POINTER(Position), C.c_uint64, C.POINTER(C.c_uint32), C.c_uint64,
                                   C.POINTER(MeshRange), C.c_uint32]),
    ("rsy_set_materials", C.c_int, [C.c_void_p, C.POINTER(GpuMaterial), C.c_uint32, C.POINTER(TextureDesc),
                                    C.c_uint32, C.POINTER(Sampler), C.c_uint32]),
    ("rsy_set_graphic_objects", C.c_int, [C.c_void_p, C.POINTER(ObjectRecord), C.c_uint32, C.POINTER(Draw),
                                          C.c_uint32]),
    ("rsy_update_graphic_objects", C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(ObjectRecord), C.c_uint32]),
    ("rsy_update", C.c_int, [C.c_void_p, C.POINTER(SceneData), C.POINTER(RasterState)]),
    ("rsy_render", C.c_int, [C.c_void_p, C.POINTER(FrameDesc)]),
    ("rsy_render_graph", C.c_int, [C.c_void_p, C.POINTER(FrameDesc), C.c_uint32]),
    ("rsy_synchronize", C.c_int, [C.c_void_p]),
    ("rsy_read_visibility", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64]),
    ("rsy_read_color", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64]),
    ("rsy_read_shadow_map", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64]),
    ("rsy_read_texture", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint64]),
    ("rsy_read_shadow_factor", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64]),
    ("rsy_get_frame_stats", C.c_int, [C.c_void_p, C.POINTER(FrameStats)]),
    ("rsy_get_device_outputs", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    ("rsy_set_debug_flags", C.c_int, [C.c_void_p, C.c_uint32]),
    ("rsy_present", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("rsy_read_draw_image", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("rsy_read_present", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("rsy_autotune", C.c_int, [C.c_void_p, C.POINTER(FrameDesc), C.c_uint32, C.POINTER(C.c_uint32)]),
    ("rsy_render_noalloc", C.c_int, [C.c_void_p, C.POINTER(FrameDesc), C.POINTER(FrameTargets)]),
    ("rsy_neural_default_config", None, [C.POINTER(NeuralConfig)]),
    ("rsy_neural_init", C.c_int, [C.c_void_p, C.POINTER(NeuralConfig)]),
    ("rsy_neural_train", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    ("rsy_neural_evaluate", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    ("rsy_neural_get_stats", C.c_int, [C.c_void_p, C.POINTER(NeuralStats)]),
    ("rsy_neural_set_active", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    ("rsy_neural_get_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32]),
    ("rsy_neural_set_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32]),
    ("rsy_neural_gradients", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                       C.POINTER(C.c_float), C.c_void_p, C.c_void_p, C.c_uint64]),
    ("rsy_neural_set_weights", C.c_int, [C.c_void_p, C.c_uint32]),
    ("rsy_neural_get_inference_mlp", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    ("rsy_neural_decode", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    ("rsy_camera_projection", None, [C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_float)]),
    ("rsy_camera_view_matrix", None, [C.POINTER(C.c_float), C.c_float, C.c_float, C.POINTER(C.c_float)]),
    ("rsy_light_matrices", None, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float),
                                  C.POINTER(C.c_float)]),
    ("rsy_object_record_from_trs", None, [C.POINTER(C.c_float), C.c_float, C.c_float, C.POINTER(ObjectRecord)]),
    ("rsy_mat4_mul", None, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("rsy_asset_encode", C.c_int, [C.POINTER(AssetDesc), C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("rsy_asset_write", C.c_int, [C.POINTER(AssetDesc), C.c_char_p]),
    ("rsy_asset_decode", C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("rsy_asset_read", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("rsy_asset_describe", C.c_int, [C.c_void_p, C.POINTER(AssetDesc)]),
    ("rsy_asset_destroy", None, [C.c_void_p]),
    ("rsy_asset_geometry_size", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_uint32)]),
    ("rsy_asset_pack_geometry", C.c_int, [C.c_void_p, C.POINTER(Position), C.c_uint64, C.POINTER(C.c_uint32),
                                          C.c_uint64, C.POINTER(MeshRange), C.c_uint32]),
    ("rsy_asset_pack_materials", C.c_int, [C.c_void_p, C.POINTER(GpuMaterial), C.c_uint32]),
    ("rsy_scene_generate", C.c_int, [C.POINTER(SceneDesc), C.POINTER(C.POINTER(Scene))]),
    ("rsy_scene_free", None, [C.POINTER(Scene)]),
]


def bind(lib):
    for name, res, args in PROTOTYPES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
