"""ctypes binding of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, and
only as the checker or the timed CPU baseline; the product (rosy_amd) never uses it.
"""
import ctypes as C
import os

import numpy as np

from rosy_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class Geometry(C.Structure):
    _fields_ = [("vertices", C.POINTER(abi.Position)), ("n_vertices", C.c_uint64),
                ("indices", C.POINTER(C.c_uint32)), ("n_indices", C.c_uint64),
                ("meshes", C.POINTER(abi.MeshRange)), ("n_meshes", C.c_uint32), ("n_objects", C.c_uint32),
                ("objects", C.POINTER(abi.ObjectRecord)), ("draws", C.POINTER(abi.Draw)), ("n_draws", C.c_uint32),
                ("reserved", C.c_uint32)]


class RasterStats(C.Structure):
    _fields_ = [("triangles_in", C.c_uint64), ("triangles_setup", C.c_uint64), ("triangles_clipped", C.c_uint64),
                ("fragments", C.c_uint64)]


class ShadeIn(C.Structure):
    _fields_ = [("materials", C.POINTER(abi.GpuMaterial)), ("n_materials", C.c_uint32), ("n_textures", C.c_uint32),
                ("textures", C.POINTER(abi.TextureDesc)), ("samplers", C.POINTER(abi.Sampler)),
                ("n_samplers", C.c_uint32), ("n_cascades", C.c_uint32),
                ("cascades", C.POINTER(abi.ShadowCascade)), ("shadow_maps", C.POINTER(C.POINTER(C.c_float))),
                ("scene", C.POINTER(abi.SceneData)), ("raster", C.POINTER(abi.RasterState)),
                ("view", C.POINTER(abi.CameraView))]


class Footprint(C.Structure):
    _fields_ = [("covered_pixels", C.c_uint64), ("visible_triangles", C.c_uint64), ("visible_vertices", C.c_uint64),
                ("texels", C.c_uint64), ("texel_bytes", C.c_uint64), ("shadow_texels", C.c_uint64),
                ("latent_cells", C.c_uint64)]


class Neural(C.Structure):
    _fields_ = [("latent_size", C.c_uint32), ("n_materials", C.c_uint32), ("slot_of", C.c_void_p),
                ("active", C.c_void_p), ("latents", C.c_void_p), ("mlp", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.orc_raster_view.restype = C.c_int
        L.orc_raster_view.argtypes = [C.POINTER(Geometry), C.POINTER(C.c_float), C.c_uint32, C.c_uint32, C.c_int,
                                      C.POINTER(abi.RasterState), C.c_void_p, C.c_void_p, C.POINTER(RasterStats),
                                      C.c_int]
        L.orc_raster_view_at.restype = C.c_int
        L.orc_raster_view_at.argtypes = [C.POINTER(Geometry), C.POINTER(C.c_float), C.c_uint32, C.c_uint32, C.c_int,
                                         C.POINTER(abi.RasterState), C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                         C.POINTER(RasterStats), C.c_int]
        L.orc_blend_layers_at.restype = C.c_int
        L.orc_blend_layers_at.argtypes = [C.POINTER(Geometry), C.POINTER(C.c_float), C.c_uint32, C.c_uint32,
                                          C.POINTER(abi.RasterState), C.c_int32, C.c_int32, C.c_void_p, C.c_uint32,
                                          C.c_void_p, C.c_void_p, C.c_int]
        L.orc_bc7_decode.restype = C.c_int
        L.orc_bc7_decode.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        L.orc_bc7_tables.restype = C.c_int
        L.orc_bc7_tables.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_count_tile_pairs.restype = C.c_uint64
        L.orc_count_tile_pairs.argtypes = [C.POINTER(Geometry), C.POINTER(C.c_float), C.c_uint32, C.c_uint32, C.c_int,
                                           C.POINTER(abi.RasterState), C.c_uint32, C.c_int]
        L.orc_blend_layers.restype = C.c_int
        L.orc_blend_layers.argtypes = [C.POINTER(Geometry), C.POINTER(C.c_float), C.c_uint32, C.c_uint32,
                                       C.POINTER(abi.RasterState), C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                       C.c_int]
        L.orc_tile_counts.restype = C.c_int
        L.orc_tile_counts.argtypes = [C.POINTER(Geometry), C.POINTER(C.c_float), C.c_uint32, C.c_uint32, C.c_int,
                                      C.POINTER(abi.RasterState), C.c_uint32, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_shade_view.restype = C.c_int
        L.orc_shade_view.argtypes = [C.POINTER(Geometry), C.POINTER(ShadeIn), C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_int]
        L.orc_shade_view_neural.restype = C.c_int
        L.orc_shade_view_neural.argtypes = [C.POINTER(Geometry), C.POINTER(ShadeIn), C.POINTER(Neural), C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_shade_footprint.restype = C.c_int
        L.orc_shade_footprint.argtypes = [C.POINTER(Geometry), C.POINTER(ShadeIn), C.POINTER(Neural), C.c_void_p,
                                          C.POINTER(Footprint), C.c_int]
        L.orc_nm_decode.restype = C.c_int
        L.orc_nm_decode.argtypes = [C.POINTER(Neural), C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
        L.orc_nm_gradients.restype = C.c_int
        L.orc_nm_gradients.argtypes = [C.POINTER(Neural), C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                       C.POINTER(C.c_double), C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_sample_texture.restype = C.c_int
        L.orc_sample_texture.argtypes = [C.POINTER(abi.TextureDesc), C.POINTER(abi.Sampler), C.c_double, C.c_double,
                                         C.POINTER(C.c_double)]
        L.orc_srgb_to_linear.restype = C.c_double
        L.orc_srgb_to_linear.argtypes = [C.c_uint32]
        for name in ("orc_ggx_d", "orc_schlick_ggx", "orc_schlick_fresnel"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double, C.c_double]
        L.orc_pcf.restype = C.c_double
        L.orc_pcf.argtypes = [C.POINTER(C.c_float), C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint32,
                              C.POINTER(C.c_double)]
        _lib = L
    return _lib


def threads():
    return max(1, min(os.cpu_count() or 1, 16))


SHADOW_LIST, OPAQUE_LIST, BLENDED_LIST = 0, 1, 2


def draw_list(scene, which):
    """Shadow list = every surface; opaque list = surfaces with blended == 0; blended list = the
    others, in surface order (Graphics.cpp:4716-4735).  `which` is a list number, or a bool for
    the shadow (True) / opaque (False) list."""
    which = (SHADOW_LIST if which else OPAQUE_LIST) if isinstance(which, bool) else which
    surf = scene.surfaces()
    keep = [d for d in surf if which == SHADOW_LIST or bool(d.blended) == (which == BLENDED_LIST)]
    arr = (abi.Draw * max(1, len(keep)))()
    for i, d in enumerate(keep):
        arr[i] = d
    return arr, len(keep)


def geometry(scene, shadow):
    s = scene.s
    draws, n = draw_list(scene, shadow)
    g = Geometry(s.vertices, s.n_vertices, s.indices, s.n_indices, s.meshes, s.n_meshes, s.n_objects, s.objects,
                 C.cast(draws, C.POINTER(abi.Draw)), n, 0)
    g._keep = draws
    return g


# The standard sample locations of a 4-sample image (Vulkan: (0.375, 0.125), (0.875, 0.375),
# (0.125, 0.625), (0.625, 0.875) of the pixel), in 1/256 pixel.
MSAA4_SAMPLES = ((96, 32), (224, 96), (32, 160), (160, 224))


def raster_view(scene, view_projection, width, height, shadow, raster_state, n_threads=None, sample=None):
    """One view's visibility (camera) or depth (shadow); coverage and depth at the pixel centre, or at
    `sample` = (x, y) in 1/256 pixel."""
    g = geometry(scene, shadow)
    vp = (C.c_float * 16)(*view_projection)
    st = RasterStats()
    sx, sy = sample if sample is not None else (128, 128)
    if shadow:
        out = np.zeros((height, width), dtype=np.float32)
        rc = lib().orc_raster_view_at(C.byref(g), vp, width, height, 1, C.byref(raster_state), sx, sy, None,
                                      out.ctypes.data, C.byref(st), n_threads or threads())
    else:
        out = np.zeros((height, width), dtype=np.uint64)
        rc = lib().orc_raster_view_at(C.byref(g), vp, width, height, 0, C.byref(raster_state), sx, sy,
                                      out.ctypes.data, None, C.byref(st), n_threads or threads())
    if rc != 0:
        raise RuntimeError(f"orc_raster_view failed: {rc}")
    return out, st


def raster_view_msaa(scene, view_projection, width, height, raster_state, n_threads=None):
    """The four sample planes [4, h, w] of a 4-sample camera view: each the single-sample raster at
    that sample location."""
    return np.stack([raster_view(scene, view_projection, width, height, False, raster_state, n_threads, smp)[0]
                     for smp in MSAA4_SAMPLES])


def tile_pairs(scene, view_projection, width, height, shadow, raster_state, tile, n_threads=None):
    g = geometry(scene, shadow)
    vp = (C.c_float * 16)(*view_projection)
    return int(lib().orc_count_tile_pairs(C.byref(g), vp, width, height, 1 if shadow else 0, C.byref(raster_state),
                                          tile, n_threads or threads()))


class NeuralParams:
    """Decoder parameters for the oracle: slot_of / active per material, latents [slots, L, L, 8] and
    mlp [583] as the float values to evaluate (e.g. FP16-rounded for the inference path)."""

    def __init__(self, slot_of, latents, mlp, active=None):
        self.slot_of = np.ascontiguousarray(slot_of, dtype=np.uint32)
        self.latents = np.ascontiguousarray(latents, dtype=np.float32)
        self.mlp = np.ascontiguousarray(mlp, dtype=np.float32)
        self.active = None if active is None else np.ascontiguousarray(active, dtype=np.uint32)
        self.struct = Neural(self.latents.shape[1], self.slot_of.size, self.slot_of.ctypes.data,
                             self.active.ctypes.data if self.active is not None else None,
                             self.latents.ctypes.data, self.mlp.ctypes.data)


def nm_decode(params, material, uv):
    m = np.ascontiguousarray(material, dtype=np.uint32)
    uv = np.ascontiguousarray(uv, dtype=np.float32)
    out = np.zeros((m.size, 7), dtype=np.float64)
    rc = lib().orc_nm_decode(C.byref(params.struct), m.ctypes.data, uv.ctypes.data, m.size, out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"orc_nm_decode failed: {rc}")
    return out


def nm_gradients(params, material, uv, target):
    """(loss, grad_mlp [583], grad_latents like params.latents, min |pre-activation| per sample)."""
    m = np.ascontiguousarray(material, dtype=np.uint32)
    uv = np.ascontiguousarray(uv, dtype=np.float32)
    t = np.ascontiguousarray(target, dtype=np.float32)
    g_mlp = np.zeros(583, dtype=np.float64)
    g_lat = np.zeros(params.latents.shape, dtype=np.float64)
    margin = np.zeros(m.size, dtype=np.float64)
    loss = C.c_double()
    rc = lib().orc_nm_gradients(C.byref(params.struct), params.latents.shape[0], m.ctypes.data, uv.ctypes.data,
                                t.ctypes.data, m.size, C.byref(loss), g_mlp.ctypes.data, g_lat.ctypes.data,
                                margin.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"orc_nm_gradients failed: {rc}")
    return loss.value, g_mlp, g_lat, margin


def sample_texture(desc, sampler, u, v):
    out = (C.c_double * 4)()
    rc = lib().orc_sample_texture(C.byref(desc), C.byref(sampler) if sampler is not None else None, u, v, out)
    if rc != 0:
        raise RuntimeError(f"orc_sample_texture failed: {rc}")
    return np.array(out[:])


def tile_counts(scene, view_projection, width, height, shadow, raster_state, tile, n_threads=None):
    """(records per bin, box pixels per bin), each shaped (ceil(h / tile), ceil(w / tile))."""
    g = geometry(scene, shadow)
    vp = (C.c_float * 16)(*view_projection)
    shape = ((height + tile - 1) // tile, (width + tile - 1) // tile)
    out = np.zeros(shape, dtype=np.uint32)
    area = np.zeros(shape[0] * shape[1] + 2, dtype=np.uint64)
    rc = lib().orc_tile_counts(C.byref(g), vp, width, height, 1 if shadow else 0, C.byref(raster_state), tile,
                               out.ctypes.data, area.ctypes.data, n_threads or threads())
    if rc != 0:
        raise RuntimeError(f"orc_tile_counts failed: {rc}")
    tile_counts.wide = (int(area[-2]), int(area[-1]))  # records over 4 bins, their pairs
    return out, area[:-2].reshape(shape)


def shade_view(scene, view, cascades, shadow_maps, vis, scene_data=None, raster_state=None, override_shadow=None,
               n_threads=None, neural=None, which=OPAQUE_LIST):
    """FP64 shading of one camera view.  Returns (rgba float64 [h,w,4], shadow factor [h,w], margin [h,w]).
    With `neural` (NeuralParams) the maps of the active materials are replaced by the decoder.  The
    visibility words name triangles of draw list `which`."""
    s = scene.s
    g = geometry(scene, which)
    maps = [np.ascontiguousarray(m, dtype=np.float32) for m in shadow_maps]
    ptrs = (C.POINTER(C.c_float) * max(1, len(maps)))(*[m.ctypes.data_as(C.POINTER(C.c_float)) for m in maps])
    casc = (abi.ShadowCascade * max(1, len(cascades)))(*cascades)
    sd = scene_data if scene_data is not None else s.scene_data
    rs = raster_state if raster_state is not None else s.raster_state
    cv = abi.CameraView()
    C.memmove(C.byref(cv), C.byref(view), C.sizeof(abi.CameraView))
    si = ShadeIn(s.materials, s.n_materials, s.n_textures, stored_textures(scene), s.samplers, s.n_samplers,
                 len(cascades),
                 C.cast(casc, C.POINTER(abi.ShadowCascade)), C.cast(ptrs, C.POINTER(C.POINTER(C.c_float))),
                 C.pointer(sd), C.pointer(rs), C.pointer(cv))
    h, w = vis.shape
    vis = np.ascontiguousarray(vis, dtype=np.uint64)
    rgba = np.zeros((h, w, 4), dtype=np.float64)
    shadow = np.zeros((h, w), dtype=np.float64)
    margin = np.zeros((h, w), dtype=np.float64)
    ov = None
    if override_shadow is not None:
        ov = np.ascontiguousarray(override_shadow, dtype=np.float32)
    if neural is not None:
        rc = lib().orc_shade_view_neural(C.byref(g), C.byref(si), C.byref(neural.struct), vis.ctypes.data,
                                         rgba.ctypes.data, shadow.ctypes.data, margin.ctypes.data,
                                         ov.ctypes.data if ov is not None else None, n_threads or threads())
    else:
        rc = lib().orc_shade_view(C.byref(g), C.byref(si), vis.ctypes.data, rgba.ctypes.data, shadow.ctypes.data,
                                  margin.ctypes.data, ov.ctypes.data if ov is not None else None,
                                  n_threads or threads())
    if rc != 0:
        raise RuntimeError(f"orc_shade_view failed: {rc}")
    return rgba, shadow, margin


def shade_footprint(scene, view, cascades, shadow_maps, vis, scene_data=None, raster_state=None, neural=None,
                    n_threads=None):
    """What shading the view reads (unique texels and their bytes, shadow-map texels, latent cells,
    visible triangles and vertices): a dict of the orc_footprint fields."""
    s = scene.s
    g = geometry(scene, OPAQUE_LIST)
    maps = [np.ascontiguousarray(m, dtype=np.float32) for m in shadow_maps]
    ptrs = (C.POINTER(C.c_float) * max(1, len(maps)))(*[m.ctypes.data_as(C.POINTER(C.c_float)) for m in maps])
    casc = (abi.ShadowCascade * max(1, len(cascades)))(*cascades)
    sd = scene_data if scene_data is not None else s.scene_data
    rs = raster_state if raster_state is not None else s.raster_state
    cv = abi.CameraView()
    C.memmove(C.byref(cv), C.byref(view), C.sizeof(abi.CameraView))
    si = ShadeIn(s.materials, s.n_materials, s.n_textures, stored_textures(scene), s.samplers, s.n_samplers,
                 len(cascades), C.cast(casc, C.POINTER(abi.ShadowCascade)),
                 C.cast(ptrs, C.POINTER(C.POINTER(C.c_float))), C.pointer(sd), C.pointer(rs), C.pointer(cv))
    vis = np.ascontiguousarray(vis, dtype=np.uint64)
    out = Footprint()
    rc = lib().orc_shade_footprint(C.byref(g), C.byref(si), C.byref(neural.struct) if neural is not None else None,
                                   vis.ctypes.data, C.byref(out), n_threads or threads())
    if rc != 0:
        raise RuntimeError(f"orc_shade_footprint failed: {rc}")
    return {k: int(getattr(out, k)) for k, _ in Footprint._fields_}


def blend_layers(scene, view, opaque_vis, max_layers=16, raster_state=None, n_threads=None, sample=None):
    """The blended list's fragments per pixel in primitive order: (layers [K, h, w] u64 visibility
    words, count [h, w]); at the pixel centre, or at `sample` = (x, y) in 1/256 pixel against that
    sample plane's opaque visibility."""
    g = geometry(scene, BLENDED_LIST)
    h, w = opaque_vis.shape
    vp = (C.c_float * 16)(*view.view_projection[:])
    rs = raster_state if raster_state is not None else scene.s.raster_state
    layers = np.zeros((max_layers, h, w), dtype=np.uint64)
    count = np.zeros((h, w), dtype=np.uint32)
    ov = np.ascontiguousarray(opaque_vis, dtype=np.uint64)
    sx, sy = sample if sample is not None else (128, 128)
    rc = lib().orc_blend_layers_at(C.byref(g), vp, w, h, C.byref(rs), sx, sy, ov.ctypes.data, max_layers,
                                   layers.ctypes.data, count.ctypes.data, n_threads or threads())
    if rc != 0:
        raise RuntimeError(f"orc_blend_layers failed: {rc}")
    return layers, count


def blend_view(scene, view, cascades, shadow_maps, opaque_vis, opaque_rgba, max_layers=16, scene_data=None,
               raster_state=None, n_threads=None, sample=None):
    """The blended pass over an opaque frame, FP64: each pixel's blended fragments in primitive
    order are shaded (fragmentMain, at the pixel centre), discarded when alpha mode 1 puts alpha
    below the cutoff (basic.slang:335-337), and blended as colour = src.rgb * src.a + dst.rgb *
    (1 - src.a), alpha = src.a (Graphics.cpp:5657-5664).  With `sample`, the fragments are those
    covering that sample location and passing the depth test there (one sample plane)."""
    layers, count = blend_layers(scene, view, opaque_vis, max_layers, raster_state, n_threads, sample)
    out = np.array(opaque_rgba, dtype=np.float64, copy=True)
    draws, n = draw_list(scene, BLENDED_LIST)
    tri_base = np.cumsum([0] + [draws[i].index_count // 3 for i in range(n)])
    nm = scene.s.n_materials
    modes = np.array([scene.s.materials[i].alpha_mode for i in range(nm)] + [0], dtype=np.int64)
    cuts = np.array([scene.s.materials[i].alpha_cutoff for i in range(nm)] + [0.0], dtype=np.float64)
    mat_of_draw = np.array([min(int(draws[i].material_index), nm) for i in range(n)] + [nm], dtype=np.int64)
    for k in range(int(count.max()) if count.size else 0):
        has = count > k
        rgba, _, _ = shade_view(scene, view, cascades, shadow_maps, np.where(has, layers[k], 0), scene_data,
                                raster_state, n_threads=n_threads, which=BLENDED_LIST)
        tri = (layers[k] & 0xFFFFFFFF).astype(np.int64) - 1
        d = np.where(has, np.searchsorted(tri_base, tri, side="right") - 1, n)
        mat = mat_of_draw[np.clip(d, 0, n)]  # index nm: no material (never discarded)
        mode, cut = modes[mat], cuts[mat]
        a = rgba[..., 3]
        keep = has & ~((mode == 1) & (a < cut))
        src = rgba[..., :3] * a[..., None] + out[..., :3] * (1.0 - a[..., None])
        out[..., :3] = np.where(keep[..., None], src, out[..., :3])
        out[..., 3] = np.where(keep, a, out[..., 3])
    return out, count


def msaa_view(scene, view, cascades, shadow_maps, vis_planes, scene_data=None, raster_state=None, n_threads=None,
              neural=None, blended=False, max_layers=16):
    """A 4-sample camera view shaded and resolved, FP64 (the main pass of Graphics.cpp:5566-5575:
    a 4-sample colour attachment resolved with VK_RESOLVE_MODE_AVERAGE_BIT).  The fragment shader
    runs once per pixel and primitive at the pixel centre (no sample shading), so sample s holds the
    shading of its plane's primitive at the centre (the clear colour where uncovered); with
    `blended` the blended pass then runs per sample plane.  Returns (resolved rgba [h, w, 4],
    sample colours [4, h, w, 4], blended layer counts per plane [4, h, w] or None, the smallest
    shadow-compare margin of the opaque shading over the planes [h, w])."""
    rs = raster_state if raster_state is not None else scene.s.raster_state
    samples, counts, margins = [], [], []
    for s, smp in enumerate(MSAA4_SAMPLES):
        rgba, _, margin = shade_view(scene, view, cascades, shadow_maps, vis_planes[s], scene_data, rs,
                                     n_threads=n_threads, neural=neural)
        margins.append(margin)
        if blended:
            rgba, count = blend_view(scene, view, cascades, shadow_maps, vis_planes[s], rgba, max_layers, scene_data,
                                     rs, n_threads, sample=smp)
            counts.append(count)
        samples.append(rgba)
    samples = np.stack(samples)
    return samples.mean(axis=0), samples, (np.stack(counts) if blended else None), np.min(margins, axis=0)


def bc7_decode(blocks, width, height):
    """BC7 blocks (bytes, ceil(w/4) * ceil(h/4) * 16) decoded to RGBA8 [h, w, 4]."""
    b = np.frombuffer(bytes(blocks), np.uint8)
    assert b.size == ((width + 3) // 4) * ((height + 3) // 4) * 16
    out = np.zeros((height, width, 4), np.uint8)
    rc = lib().orc_bc7_decode(b.ctypes.data, width, height, out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"orc_bc7_decode failed: {rc}")
    return out


def bc7_tables():
    """(two-subset [64, 16], three-subset [64, 16], anchors [3, 64]) as the decoder uses them."""
    two, three, anchors = np.zeros((64, 16), np.uint8), np.zeros((64, 16), np.uint8), np.zeros((3, 64), np.uint8)
    lib().orc_bc7_tables(two.ctypes.data, three.ctypes.data, anchors.ctypes.data)
    return two, three, anchors


def stored_textures(scene):
    """The scene's textures as the sampler reads them, for the oracle's shading: BC7 images decoded
    by the oracle's decoder and chains to build (mip_levels = MIPS_BUILD) built by the oracle's
    packager filter (oracle/textures.py), as RGBA8 / RG8 with every level given.  Cached on the
    scene."""
    cached = getattr(scene, "_oracle_textures", None)
    if cached is not None:
        return cached[0]
    from . import textures as tx
    s = scene.s
    n = s.n_textures
    out = (abi.TextureDesc * max(1, n))()
    keep = []
    for i in range(n):
        d = s.textures[i]
        bc7 = d.format in (abi.TEX_BC7_SRGB, abi.TEX_BC7_UNORM)
        build = d.mip_levels == abi.MIPS_BUILD
        if not bc7 and not build:
            out[i] = d
            continue
        ch = 2 if d.format in (abi.TEX_RG8_UNORM, abi.TEX_RG8_SRGB) else 4
        srgb = d.format in (abi.TEX_RGBA8_SRGB, abi.TEX_RG8_SRGB, abi.TEX_BC7_SRGB)
        given = 1 if build else max(1, d.mip_levels)
        levels, at = [], 0
        for l in range(given):
            w, h = max(1, d.width >> l), max(1, d.height >> l)
            nbytes = ((w + 3) // 4) * ((h + 3) // 4) * 16 if bc7 else w * h * ch
            raw = C.string_at(d.pixels + at, nbytes)
            at += nbytes
            levels.append(bc7_decode(raw, w, h) if bc7 else np.frombuffer(raw, np.uint8).reshape(h, w, ch))
        if build:
            levels = tx.mip_chain(levels[0], srgb)
        data = np.concatenate([np.ascontiguousarray(v).reshape(-1) for v in levels])
        keep.append(data)
        fmt = d.format if not bc7 else (abi.TEX_RGBA8_SRGB if srgb else abi.TEX_RGBA8_UNORM)
        out[i] = abi.TextureDesc(d.width, d.height, fmt, len(levels), data.ctypes.data)
    scene._oracle_textures = (out, keep)
    return out

