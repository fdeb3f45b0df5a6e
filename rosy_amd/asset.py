"""ctypes harness for the .rsy asset reader/writer (the rsy_asset_* entry points of
include/rosy_amd.h; format of src/Asset/Asset.h:9-149 and Asset.cpp:6-848).

An asset is held here as plain Python and numpy data (`Asset`); `encode`/`write` go through
rsy_asset_encode / rsy_asset_write and `decode`/`read` through rsy_asset_decode / rsy_asset_read
+ rsy_asset_describe.  `Handle` keeps a decoded asset in the library to pack it into the frame
path's geometry and material inputs (rsy_asset_pack_geometry / rsy_asset_pack_materials).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi, lib, RosyError

MATERIAL_DTYPE = np.dtype([
    ("double_sided", "u1"), ("pad0", "u1", 3), ("base_color_factor", "<f4", 4), ("metallic_factor", "<f4"),
    ("roughness_factor", "<f4"), ("alpha_mode", "u1"), ("pad1", "u1", 3), ("alpha_cutoff", "<f4"),
    ("color_image_index", "<u4"), ("color_sampler_index", "<u4"), ("normal_image_index", "<u4"),
    ("normal_sampler_index", "<u4"), ("metallic_image_index", "<u4"), ("metallic_sampler_index", "<u4"),
    ("mixmap_image_index", "<u4"), ("mixmap_sampler_index", "<u4")])
SAMPLER_DTYPE = np.dtype([("min_filter", "<u2"), ("mag_filter", "<u2"), ("wrap_s", "<u2"), ("wrap_t", "<u2")])
SURFACE_DTYPE = np.dtype([("start_index", "<u4"), ("count", "<u4"), ("material", "<u4"),
                          ("min_bounds", "<f4", 3), ("max_bounds", "<f4", 3)])
assert MATERIAL_DTYPE.itemsize == 68 and SAMPLER_DTYPE.itemsize == 8 and SURFACE_DTYPE.itemsize == 36

IDENTITY = tuple(float(i % 5 == 0) for i in range(16))


@dataclass
class Node:
    world_translate: tuple = (0.0, 0.0, 0.0)
    world_scale: float = 1.0
    world_yaw: float = 0.0
    transform: tuple = IDENTITY
    mesh_id: int = abi.RSY_ABSENT
    children: list = field(default_factory=list)
    name: bytes = b""


@dataclass
class Image:
    image_type: int = 0
    name: bytes = b""


@dataclass
class Mesh:
    positions: np.ndarray = field(default_factory=lambda: np.zeros((0, 16), np.float32))  # (n, 16) float32
    indices: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    surfaces: np.ndarray = field(default_factory=lambda: np.zeros(0, SURFACE_DTYPE))


@dataclass
class Asset:
    coordinate_system: tuple = IDENTITY
    root_scene: int = 0
    materials: np.ndarray = field(default_factory=lambda: np.zeros(0, MATERIAL_DTYPE))
    samplers: np.ndarray = field(default_factory=lambda: np.zeros(0, SAMPLER_DTYPE))
    scenes: list = field(default_factory=list)  # lists of node indices
    nodes: list = field(default_factory=list)
    images: list = field(default_factory=list)
    meshes: list = field(default_factory=list)


def _ptr(a, ctype):
    return C.cast(a.ctypes.data, C.POINTER(ctype)) if a.size else C.POINTER(ctype)()


def _desc(asset):
    """An AssetDesc over `asset` plus the objects that must stay alive while it is used."""
    keep = []
    d = abi.AssetDesc()
    d.coordinate_system[:] = [float(x) for x in asset.coordinate_system]
    d.root_scene = asset.root_scene
    mats = np.ascontiguousarray(asset.materials, MATERIAL_DTYPE)
    smps = np.ascontiguousarray(asset.samplers, SAMPLER_DTYPE)
    keep += [mats, smps]
    d.materials, d.n_materials = _ptr(mats, abi.AssetMaterial), len(mats)
    d.samplers, d.n_samplers = _ptr(smps, abi.Sampler), len(smps)
    scenes = (abi.AssetScene * max(1, len(asset.scenes)))()
    for i, s in enumerate(asset.scenes):
        a = np.ascontiguousarray(s, np.uint32)
        keep.append(a)
        scenes[i].nodes, scenes[i].n_nodes = _ptr(a, C.c_uint32), len(a)
    d.scenes, d.n_scenes = scenes, len(asset.scenes)
    nodes = (abi.AssetNode * max(1, len(asset.nodes)))()
    for i, n in enumerate(asset.nodes):
        ch = np.ascontiguousarray(n.children, np.uint32)
        name = C.create_string_buffer(bytes(n.name), max(1, len(n.name)))
        keep += [ch, name]
        nd = nodes[i]
        nd.world_translate[:] = [float(x) for x in n.world_translate]
        nd.world_scale, nd.world_yaw = n.world_scale, n.world_yaw
        nd.transform[:] = [float(x) for x in n.transform]
        nd.mesh_id = n.mesh_id
        nd.child_nodes, nd.n_child_nodes = _ptr(ch, C.c_uint32), len(ch)
        nd.name, nd.name_len = C.cast(name, C.c_void_p), len(n.name)
    d.nodes, d.n_nodes = nodes, len(asset.nodes)
    images = (abi.AssetImage * max(1, len(asset.images)))()
    for i, im in enumerate(asset.images):
        name = C.create_string_buffer(bytes(im.name), max(1, len(im.name)))
        keep.append(name)
        images[i].image_type, images[i].name, images[i].name_len = im.image_type, C.cast(name, C.c_void_p), len(im.name)
    d.images, d.n_images = images, len(asset.images)
    meshes = (abi.AssetMesh * max(1, len(asset.meshes)))()
    for i, m in enumerate(asset.meshes):
        pos = np.ascontiguousarray(m.positions, np.float32).reshape(-1, 16)
        idx = np.ascontiguousarray(m.indices, np.uint32)
        srf = np.ascontiguousarray(m.surfaces, SURFACE_DTYPE)
        keep += [pos, idx, srf]
        meshes[i].positions, meshes[i].n_positions = _ptr(pos, abi.Position), len(pos)
        meshes[i].indices, meshes[i].n_indices = _ptr(idx, C.c_uint32), len(idx)
        meshes[i].surfaces, meshes[i].n_surfaces = _ptr(srf, abi.AssetSurface), len(srf)
    d.meshes, d.n_meshes = meshes, len(asset.meshes)
    keep += [scenes, nodes, images, meshes]
    return d, keep


def _check(code, what):
    if code != abi.RSY_OK:
        raise RosyError(code, what)


def encode(asset):
    """The .rsy file bytes of `asset` (rsy_asset_encode)."""
    d, keep = _desc(asset)
    n = C.c_uint64()
    _check(lib().rsy_asset_encode(C.byref(d), None, 0, C.byref(n)), "rsy_asset_encode")
    out = np.zeros(n.value, np.uint8)
    _check(lib().rsy_asset_encode(C.byref(d), out.ctypes.data, n.value, C.byref(n)), "rsy_asset_encode")
    return out.tobytes()


def write(asset, path):
    d, keep = _desc(asset)
    _check(lib().rsy_asset_write(C.byref(d), str(path).encode()), "rsy_asset_write")


def _arr(ptr, n, dtype, shape=None):
    if n == 0:
        return np.zeros(shape if shape else 0, dtype)
    raw = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint8)), shape=(n * np.dtype(dtype).itemsize,))
    a = raw.view(dtype).copy()
    return a.reshape(shape) if shape else a


class Handle:
    """A decoded asset owned by the library."""

    def __init__(self, ptr):
        self.ptr = ptr

    def close(self):
        if self.ptr:
            lib().rsy_asset_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def describe(self):
        d = abi.AssetDesc()
        _check(lib().rsy_asset_describe(self.ptr, C.byref(d)), "rsy_asset_describe")
        a = Asset(coordinate_system=tuple(d.coordinate_system), root_scene=d.root_scene,
                  materials=_arr(d.materials, d.n_materials, MATERIAL_DTYPE),
                  samplers=_arr(d.samplers, d.n_samplers, SAMPLER_DTYPE))
        for i in range(d.n_scenes):
            a.scenes.append([int(x) for x in _arr(d.scenes[i].nodes, d.scenes[i].n_nodes, np.uint32)])
        for i in range(d.n_nodes):
            n = d.nodes[i]
            a.nodes.append(Node(tuple(n.world_translate), n.world_scale, n.world_yaw, tuple(n.transform), n.mesh_id,
                                [int(x) for x in _arr(n.child_nodes, n.n_child_nodes, np.uint32)],
                                C.string_at(n.name, n.name_len) if n.name_len else b""))
        for i in range(d.n_images):
            im = d.images[i]
            a.images.append(Image(im.image_type, C.string_at(im.name, im.name_len) if im.name_len else b""))
        for i in range(d.n_meshes):
            m = d.meshes[i]
            a.meshes.append(Mesh(_arr(m.positions, m.n_positions * 16, np.float32, (m.n_positions, 16)),
                                 _arr(m.indices, m.n_indices, np.uint32),
                                 _arr(m.surfaces, m.n_surfaces, SURFACE_DTYPE)))
        return a

    def pack_geometry(self):
        """(vertices (n, 16) float32, indices uint32, mesh table as MeshRange array) as set_asset packs them."""
        nv, ni, nm = C.c_uint64(), C.c_uint64(), C.c_uint32()
        _check(lib().rsy_asset_geometry_size(self.ptr, C.byref(nv), C.byref(ni), C.byref(nm)), "rsy_asset_geometry_size")
        v = np.zeros((nv.value, 16), np.float32)
        i = np.zeros(ni.value, np.uint32)
        m = (abi.MeshRange * max(1, nm.value))()
        _check(lib().rsy_asset_pack_geometry(self.ptr, _ptr(v, abi.Position), nv.value, _ptr(i, C.c_uint32), ni.value,
                                             m, nm.value), "rsy_asset_pack_geometry")
        return v, i, m[:nm.value]

    def pack_materials(self, n):
        out = (abi.GpuMaterial * max(1, n))()
        _check(lib().rsy_asset_pack_materials(self.ptr, out, n), "rsy_asset_pack_materials")
        return out[:n]


def from_scene(scene):
    """A generated workload (rosy_amd.Scene) as an asset: one mesh per scene mesh with the surfaces
    that draw it, the materials, samplers and images (named texture_<i>.dds, typed by format), and one
    node per graphic object in a single scene."""

    s = scene.s
    verts = np.ctypeslib.as_array(scene.vertices.view(np.float32)).reshape(-1, 16)
    idx = scene.indices
    draws = scene.surfaces()
    meshes = []
    for k in range(s.n_meshes):
        m = s.meshes[k]
        v0 = m.vertex_buffer_offset // 64
        own = sorted({(d.start_index, d.index_count, int(d.material_index)) for d in draws if d.mesh_index == k})
        srf = np.zeros(len(own), SURFACE_DTYPE)
        for i, (start, count, mat) in enumerate(own):
            srf[i]["start_index"], srf[i]["count"], srf[i]["material"] = start, count, mat & 0xFFFFFFFF
        meshes.append(Mesh(verts[v0:v0 + m.num_vertices].copy(),
                                 idx[m.index_offset:m.index_offset + m.num_indices].copy(), srf))
    mats = np.zeros(s.n_materials, MATERIAL_DTYPE)
    for i in range(s.n_materials):
        g = s.materials[i]
        mats[i]["base_color_factor"] = list(g.color)
        mats[i]["metallic_factor"], mats[i]["roughness_factor"] = g.metallic_factor, g.roughness_factor
        mats[i]["alpha_mode"], mats[i]["alpha_cutoff"] = g.alpha_mode, g.alpha_cutoff
        for a, b in (("color_image_index", "color_sampled_image_index"), ("color_sampler_index", "color_sampler_index"),
                     ("normal_image_index", "normal_sampled_image_index"),
                     ("normal_sampler_index", "normal_sampler_index"),
                     ("metallic_image_index", "metallic_sampled_image_index"),
                     ("metallic_sampler_index", "metallic_sampler_index"),
                     ("mixmap_image_index", "mixmap_sampled_image_index"),
                     ("mixmap_sampler_index", "mixmap_sampler_index")):
            mats[i][a] = getattr(g, b)
    smps = np.array([(s.samplers[i].min_filter, s.samplers[i].mag_filter, s.samplers[i].wrap_s, s.samplers[i].wrap_t)
                     for i in range(s.n_samplers)], SAMPLER_DTYPE)
    types = {abi.TEX_RGBA8_SRGB: 0, abi.TEX_RGBA8_UNORM: 0, abi.TEX_RG8_UNORM: 1, abi.TEX_RG8_SRGB: 2}
    images = [Image(types[s.textures[i].format], f"texture_{i}.dds".encode()) for i in range(s.n_textures)]
    nodes = []
    for o in range(s.n_objects):
        mesh = next((int(d.mesh_index) for d in draws if d.graphics_object_index == o), abi.RSY_ABSENT)
        nodes.append(Node(transform=tuple(s.objects[o].transform), mesh_id=mesh, name=f"object_{o}".encode()))
    return Asset(materials=mats, samplers=smps, scenes=[list(range(len(nodes)))], nodes=nodes, images=images,
                       meshes=meshes)


def open_bytes(data):
    buf = np.frombuffer(bytes(data), np.uint8)
    p = C.c_void_p()
    _check(lib().rsy_asset_decode(buf.ctypes.data if buf.size else None, buf.size, C.byref(p)), "rsy_asset_decode")
    return Handle(p)


def open_file(path):
    p = C.c_void_p()
    _check(lib().rsy_asset_read(str(path).encode(), C.byref(p)), "rsy_asset_read")
    return Handle(p)


def decode(data):
    with open_bytes(data) as h:
        return h.describe()


def read(path):
    with open_file(path) as h:
        return h.describe()
