"""Hand-built scenes for the raster and shading tests (small, exact inputs).

A CustomScene exposes the same surface as rosy_amd.Scene (``.s`` is an rsy_scene record, plus
surfaces(), camera_views(), cascades() and frame_desc()), so it can be given both to the oracle and
to a Renderer.
"""
import ctypes as C

import numpy as np

from rosy_amd import abi

IDENTITY = [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]


def ndc_from_subpixel(X, extent):
    """NDC coordinate whose snap lands exactly on sub-pixel X for a view of `extent` pixels:
    snap = rint(256 * (x * extent/2 + extent/2)), so x = X / (128 * extent) - 1 (exact in FP32
    for the small integers used here)."""
    return np.float32(X / (128.0 * extent) - 1.0)


class CustomScene:
    def __init__(self, positions, triangles, width, height, draws=None, view_projection=None,
                 shadow_projection=None, shadow_size=64, normals=None, uvs=None, tangents=None,
                 textures=None, material=True, objects=None, object_of_draw=None, blended=None,
                 materials=None, material_of_draw=None, samplers=None, colors=None):
        pos = np.asarray(positions, dtype=np.float32).reshape(-1, 3)
        tri = np.asarray(triangles, dtype=np.uint32).reshape(-1, 3)
        nv = len(pos)
        self._verts = (abi.Position * max(1, nv))()
        for i, p in enumerate(pos):
            v = self._verts[i]
            v.vertex[:] = [float(x) for x in p]
            n = normals[i] if normals is not None else (0.0, 0.0, -1.0)
            v.normal[:] = [float(x) for x in n]
            t = tangents[i] if tangents is not None else (1.0, 0.0, 0.0, 1.0)
            v.tangents[:] = [float(x) for x in t]
            v.color[:] = [float(x) for x in colors[i]] if colors is not None else [1.0, 0.0, 0.0, 1.0]
            uv = uvs[i] if uvs is not None else (0.5 * (p[0] + 1.0), 0.5 * (p[1] + 1.0))
            v.texture_coordinates[:] = [float(x) for x in uv]
        idx = tri.reshape(-1)
        self._idx = (C.c_uint32 * max(1, len(idx)))(*[int(x) for x in idx])
        self._mesh = (abi.MeshRange * 1)()
        self._mesh[0] = abi.MeshRange(0, 0, len(idx), nv, 0)
        if draws is None:
            draws = [(0, len(tri))]
        objs = objects if objects is not None else [IDENTITY]
        self._objs = (abi.ObjectRecord * len(objs))()
        for i, m in enumerate(objs):
            self._objs[i].transform[:] = [float(x) for x in m]
            inv = np.linalg.inv(np.array(m, dtype=np.float64).reshape(4, 4).T)
            self._objs[i].to_object_space_transform[:] = [float(x) for x in inv.T.reshape(-1)]
            nm = inv[:3, :3].T
            self._objs[i].normal_transform[:] = [float(x) for x in nm.T.reshape(-1)]
        self._draws = (abi.Draw * len(draws))()
        for i, (first_tri, count) in enumerate(draws):
            d = self._draws[i]
            d.mesh_index = 0
            d.graphic_objects_offset = 0
            d.graphics_object_index = object_of_draw[i] if object_of_draw else 0
            d.material_index = (material_of_draw[i] if material_of_draw else 0) if material else 0xFFFFFFFF
            d.index_count = 3 * count
            d.start_index = 3 * first_tri
            d.blended = 1 if blended and blended[i] else 0
        # material: 1x1 textures unless given
        self._tex_pixels = []
        texs = textures if textures is not None else [
            (1, 1, abi.TEX_RGBA8_SRGB, bytes([200, 150, 100, 255])),
            (1, 1, abi.TEX_RG8_UNORM, bytes([128, 128])),
            (1, 1, abi.TEX_RG8_SRGB, bytes([0, 230])),
        ]
        self._tex = (abi.TextureDesc * len(texs))()
        for i, tex in enumerate(texs):
            w, h, fmt, data = tex[:4]
            levels = tex[4] if len(tex) > 4 else 0  # mip_levels: 0 / 1 one level, n given, MIPS_BUILD
            buf = C.create_string_buffer(bytes(data), len(data))
            self._tex_pixels.append(buf)
            self._tex[i] = abi.TextureDesc(w, h, fmt, levels, C.cast(buf, C.c_void_p))
        # materials: each a dict of overrides of the default (white, textures 0 / 1 / 2, opaque)
        mats = materials if materials is not None else [{}]
        self._mat = (abi.GpuMaterial * len(mats))()
        for k, over in enumerate(mats):
            m = self._mat[k]
            m.color[:] = [1.0, 1.0, 1.0, 1.0]
            m.metallic_factor = 1.0
            m.roughness_factor = 1.0
            m.color_sampled_image_index, m.color_sampler_index = 0, 0
            m.normal_sampled_image_index, m.normal_sampler_index = 1, 0
            m.metallic_sampled_image_index, m.metallic_sampler_index = 2, 0
            m.mixmap_sampled_image_index = m.mixmap_sampler_index = abi.RSY_ABSENT
            for key, val in over.items():
                if key == "color":
                    m.color[:] = [float(x) for x in val]
                else:
                    setattr(m, key, val)
        # samplers: (min_filter, mag_filter, wrap_s, wrap_t); default one LINEAR / REPEAT sampler
        smps = samplers if samplers is not None else [(4, 4, 2, 2)]
        self._smp = (abi.Sampler * len(smps))()
        for i, sm in enumerate(smps):
            self._smp[i] = abi.Sampler(*sm)
        vp = view_projection if view_projection is not None else IDENTITY
        sp = shadow_projection if shadow_projection is not None else IDENTITY
        self._views = (abi.CameraView * 1)()
        cv = self._views[0]
        cv.width, cv.height = width, height
        cv.view[:] = IDENTITY
        cv.proj[:] = IDENTITY
        cv.view_projection[:] = [float(x) for x in vp]
        cv.camera_position[:] = [0.0, 0.0, -10.0, 0.0]
        self._casc = (abi.ShadowCascade * 1)()
        self._casc[0].size = shadow_size
        self._casc[0].split_far = float("inf")
        self._casc[0].light_view_projection[:] = [float(x) for x in sp]

        s = abi.Scene()
        s.vertices = C.cast(self._verts, C.POINTER(abi.Position))
        s.n_vertices = nv
        s.indices = C.cast(self._idx, C.POINTER(C.c_uint32))
        s.n_indices = len(idx)
        s.meshes = C.cast(self._mesh, C.POINTER(abi.MeshRange))
        s.n_meshes = 1
        s.objects = C.cast(self._objs, C.POINTER(abi.ObjectRecord))
        s.n_objects = len(objs)
        s.surfaces = C.cast(self._draws, C.POINTER(abi.Draw))
        s.n_surfaces = len(draws)
        s.materials = C.cast(self._mat, C.POINTER(abi.GpuMaterial))
        s.n_materials = len(mats)
        s.textures = C.cast(self._tex, C.POINTER(abi.TextureDesc))
        s.n_textures = len(texs)
        s.samplers = C.cast(self._smp, C.POINTER(abi.Sampler))
        s.n_samplers = len(smps)
        sd = s.scene_data
        sd.view[:] = IDENTITY
        sd.proj[:] = IDENTITY
        sd.view_projection[:] = [float(x) for x in vp]
        sd.shadow_projection_near[:] = [float(x) for x in sp]
        sd.sun_position[:] = [0.0, 0.0, -20.0, 1.0]
        sd.camera_position[:] = [0.0, 0.0, -10.0, 0.0]
        sd.ambient_color[:] = [0.01, 0.01, 0.01, 1.0]
        sd.sunlight_color[:] = [1.0, 1.0, 1.0, 1.0]
        sd.flip_lights[:] = [0, 0, 0, 1]
        sd.draw_extent[:] = [float(width), float(height)]
        sd.light_intensity = 5.2
        sd.light_enabled = 1
        sd.tangent_space_enabled = 1
        sd.shadows_enabled = 1
        sd.normal_maps_enabled = 1
        sd.brdf_lighting_enabled = 1
        rs = s.raster_state
        rs.cull_enabled = 0
        rs.depth_bias_enabled = 0
        rs.depth_bias_constant = -21.882
        rs.depth_bias_clamp = -20.937
        rs.depth_bias_slope_factor = -3.0
        rs.pcf_mode = abi.PCF_REFERENCE_5TAP
        rs.tile_size = 0
        rs.clear_color[:] = [0.0, 0.05, 0.1, 1.0]
        s.camera_views = C.cast(self._views, C.POINTER(abi.CameraView))
        s.n_camera_views = 1
        s.cascades = C.cast(self._casc, C.POINTER(abi.ShadowCascade))
        s.n_cascades = 1
        s.default_shadow_size = shadow_size
        self.s = s
        self.n_triangles = len(tri)

    def surfaces(self):
        return [self.s.surfaces[i] for i in range(self.s.n_surfaces)]

    def camera_views(self):
        return [self.s.camera_views[i] for i in range(self.s.n_camera_views)]

    def cascades(self):
        return [self.s.cascades[i] for i in range(self.s.n_cascades)]

    def frame_desc(self, shade_mode=abi.SHADE_PLAIN, explicit=True):
        f = abi.FrameDesc()
        f.shade_mode = shade_mode
        f.default_shadow_size = self.s.default_shadow_size
        if explicit:
            f.n_camera_views = 1
            f.camera_views = self.s.camera_views
            f.n_cascades = 1
            f.cascades = self.s.cascades
        return f

    def close(self):
        pass
