"""The seeded workload generator: exact triangle totals, closed meshes, outward winding,
determinism and the layouts the C ABI expects."""
import ctypes as C

import numpy as np
import pytest

import rosy_amd
from rosy_amd import abi


def digest(sc):
    v = sc.vertices
    return hash((v.tobytes(), sc.indices.tobytes()))


def test_sphere_has_2048_triangles_and_is_closed():
    sc = rosy_amd.Scene(0, seed=1)
    assert sc.n_triangles == 2048
    tri = sc.indices.reshape(-1, 3)
    edges = {}
    for a, b, c in tri:
        for u, v in ((a, b), (b, c), (c, a)):
            edges[(u, v)] = edges.get((u, v), 0) + 1
    # closed and consistently oriented: every directed edge appears once, its reverse once
    assert all(n == 1 for n in edges.values())
    assert all((v, u) in edges for (u, v) in edges)
    pos = sc.vertices["vertex"]
    r = np.linalg.norm(pos, axis=1)
    assert r.min() >= 0.98 - 1e-6 and r.max() <= 1.02 + 1e-6
    # outward winding: face normal points away from the centre
    p = pos[tri]
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert (np.einsum("ij,ij->i", n, p.mean(axis=1)) > 0).all()


@pytest.mark.parametrize("workload,total", [(1, 262144), (3, 262144)])
def test_scene_scale_totals_are_exact(workload, total):
    sc = rosy_amd.Scene(workload, seed=1, texture_size=64)
    assert sc.n_triangles == total
    assert sc.s.n_surfaces == 120 and sc.s.n_objects == 120
    counts = [d.index_count // 3 for d in sc.surfaces()]
    assert min(counts) >= 64 - 64 and sum(counts) == total
    idx = sc.indices
    for m in range(sc.s.n_meshes):
        mr = sc.s.meshes[m]
        seg = idx[mr.index_offset: mr.index_offset + mr.num_indices]
        assert seg.max() < mr.num_vertices


def test_larger_workload_totals_and_views():
    sc = rosy_amd.Scene(2, seed=1, texture_size=64)
    assert sc.n_triangles == 1048576 and sc.s.n_cascades == 4
    assert [c.split_far for c in sc.cascades()] == [5.0, 15.0, 40.0, 100.0]
    sc4 = rosy_amd.Scene(4, seed=1, texture_size=64, total_triangles=200000)
    assert sc4.s.n_camera_views == 4
    assert [(v.width, v.height) for v in sc4.camera_views()] == [(3840, 2160), (1920, 1080), (1280, 720), (640, 360)]


def test_generator_is_deterministic_per_seed():
    a = rosy_amd.Scene(1, seed=7, texture_size=64)
    b = rosy_amd.Scene(1, seed=7, texture_size=64)
    c = rosy_amd.Scene(1, seed=8, texture_size=64)
    assert digest(a) == digest(b)
    assert digest(a) != digest(c)
    ta = np.ctypeslib.as_array(abi.C.cast(a.s.textures[0].pixels, abi.C.POINTER(abi.C.c_uint8)), shape=(64 * 64 * 4,))
    tb = np.ctypeslib.as_array(abi.C.cast(b.s.textures[0].pixels, abi.C.POINTER(abi.C.c_uint8)), shape=(64 * 64 * 4,))
    assert np.array_equal(ta, tb)


def test_scene_data_defaults_follow_level_cpp():
    sc = rosy_amd.Scene(1, seed=1, texture_size=64)
    sd = sc.s.scene_data
    assert abs(sd.light_intensity - 5.2) < 1e-6 and abs(sd.ambient_color[0] - 0.01) < 1e-7
    assert sd.light_enabled == sd.tangent_space_enabled == sd.shadows_enabled == sd.normal_maps_enabled == 1
    assert sd.brdf_lighting_enabled == 1 and sd.fragment_output == 0
    assert list(sd.flip_lights) == [0, 0, 0, 1]
    assert sd.camera_position[3] == 0.0  # Camera.cpp:275-276 copies xyz only
    rs = sc.s.raster_state
    assert rs.cull_enabled == 1 and rs.depth_bias_enabled == 1
    assert abs(rs.depth_bias_constant + 21.882) < 1e-5 and abs(rs.depth_bias_slope_factor + 3) < 1e-6


def test_translucent_materials_and_their_alpha():
    kw = dict(width=160, height=90, shadow_size=256, total_triangles=8192, n_meshes=24, texture_size=32)
    sc = rosy_amd.Scene(3, seed=4, n_translucent=6, **kw)
    opaque = rosy_amd.Scene(3, seed=4, **kw)
    mats = [sc.s.materials[i] for i in range(sc.s.n_materials)]
    assert [m.alpha_mode for m in mats[:8]] == [2, 1, 2, 1, 2, 1, 0, 0]
    assert [m.color[3] for m in mats[:3]] == pytest.approx([0.7, 1.0, 0.7])
    assert [m.alpha_cutoff for m in mats[:3]] == pytest.approx([0.0, 0.5, 0.0])
    # surfaces of translucent materials go to the blended list; geometry is unchanged
    assert all(d.blended == (d.material_index < 6) for d in sc.surfaces())
    assert np.array_equal(sc.indices, opaque.indices) and sc.vertices.tobytes() == opaque.vertices.tobytes()
    def albedo(scene, m):
        t = scene.s.textures[mats[m].color_sampled_image_index]
        assert (t.width, t.height, t.format) == (32, 32, abi.TEX_RGBA8_SRGB)
        return np.ctypeslib.as_array(C.cast(t.pixels, C.POINTER(C.c_uint8)), shape=(32, 32, 4)).copy()

    for m in (0, 1, 6):
        a = albedo(sc, m)
        assert np.array_equal(a[..., :3], albedo(opaque, m)[..., :3])
        if m == 6:
            assert (a[..., 3] == 255).all()
        else:
            u = (np.arange(32) + 0.5) / 32
            want = np.round(np.clip(0.5 + 0.5 * np.outer(np.sin(8 * np.pi * u), np.sin(8 * np.pi * u)), 0, 1) * 255)
            assert np.array_equal(a[..., 3], want.astype(np.uint8))
