"""A generated workload written as a .rsy asset and read back packs into exactly the geometry and
material inputs the workload itself feeds rsy_set_geometry / rsy_set_materials (the packing of
set_asset, Graphics.cpp:4066-4152, 4310-4333).  Host only."""
import numpy as np
import pytest

import rosy_amd
from rosy_amd import asset
from rosy_amd.asset import from_scene as scene_asset


@pytest.mark.parametrize("workload", [0, 1])
def test_scene_round_trips_through_an_asset_file(workload, tmp_path):
    scene = rosy_amd.Scene(workload, seed=1)
    a = scene_asset(scene)
    path = tmp_path / "scene.rsy"
    asset.write(a, path)
    with asset.open_file(path) as h:
        v, i, m = h.pack_geometry()
        g = h.pack_materials(scene.s.n_materials)
    s = scene.s
    assert np.array_equal(v.view(np.uint32), np.ctypeslib.as_array(scene.vertices.view(np.float32)).reshape(-1, 16)
                          .view(np.uint32))
    assert np.array_equal(i, scene.indices)
    assert len(m) == s.n_meshes
    for k in range(s.n_meshes):
        assert (m[k].vertex_buffer_offset, m[k].index_offset, m[k].num_indices, m[k].num_vertices) == \
            (s.meshes[k].vertex_buffer_offset, s.meshes[k].index_offset, s.meshes[k].num_indices,
             s.meshes[k].num_vertices)
    for k in range(s.n_materials):
        assert bytes(g[k]) == bytes(s.materials[k]), k
    scene.close()
