"""A workload read from a .rsy file feeds the frame path: the asset's packed geometry and
materials, given to rsy_set_geometry / rsy_set_materials, render the same visibility buffer,
shadow map and colour as the workload loaded directly (bit for bit), and the visibility buffer
matches the CPU oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rosy_amd
from rosy_amd import abi, asset
from rosy_amd.asset import from_scene as scene_asset

pytestmark = pytest.mark.gpu


def test_rendering_from_an_asset_file_matches_the_direct_load(tmp_path):
    scene = rosy_amd.Scene(1, seed=1)
    path = tmp_path / "scene.rsy"
    asset.write(scene_asset(scene), path)
    frame = scene.frame_desc(abi.SHADE_PLAIN)
    direct = rosy_amd.Renderer(device=0)
    direct.load(scene)
    direct.render(frame)
    want = (direct.visibility(0), direct.shadow_map(0), direct.color(0))
    direct.close()

    s = scene.s
    with asset.open_file(path) as h:
        v, i, m = h.pack_geometry()
        g = h.pack_materials(s.n_materials)
    meshes = (abi.MeshRange * len(m))(*m)
    mats = (abi.GpuMaterial * len(g))(*g)
    r = rosy_amd.Renderer(device=0)
    L = rosy_amd.lib()
    r._c(L.rsy_set_geometry(r.ctx, C.cast(v.ctypes.data, C.POINTER(abi.Position)), len(v),
                            C.cast(i.ctypes.data, C.POINTER(C.c_uint32)), len(i), meshes, len(m)), "rsy_set_geometry")
    r._c(L.rsy_set_materials(r.ctx, mats, len(g), s.textures, s.n_textures, s.samplers, s.n_samplers),
         "rsy_set_materials")
    r._c(L.rsy_set_graphic_objects(r.ctx, s.objects, s.n_objects, s.surfaces, s.n_surfaces), "rsy_set_graphic_objects")
    r.update(s.scene_data, s.raster_state)
    r.scene = scene
    r.render(frame)
    got = (r.visibility(0), r.shadow_map(0), r.color(0))
    r.close()
    assert np.array_equal(got[0], want[0])
    assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32))
    assert np.array_equal(got[2].view(np.uint32), want[2].view(np.uint32))
    cam = scene.camera_views()[0]
    o, _ = oracle.raster_view(scene, cam.view_projection[:], cam.width, cam.height, False, s.raster_state)
    assert np.array_equal(got[0], o)
    scene.close()
