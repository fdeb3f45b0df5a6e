"""A generated workload restated as a .rsy asset (meshes with their surfaces, materials, images,
samplers, one node per graphic object), for the asset tests."""
import numpy as np

from rosy_amd import abi, asset


def scene_asset(scene):
    s = scene.s
    verts = np.ctypeslib.as_array(scene.vertices.view(np.float32)).reshape(-1, 16)
    idx = scene.indices
    draws = scene.surfaces()
    meshes = []
    for k in range(s.n_meshes):
        m = s.meshes[k]
        v0 = m.vertex_buffer_offset // 64
        own = sorted({(d.start_index, d.index_count, int(d.material_index)) for d in draws if d.mesh_index == k})
        srf = np.zeros(len(own), asset.SURFACE_DTYPE)
        for i, (start, count, mat) in enumerate(own):
            srf[i]["start_index"], srf[i]["count"], srf[i]["material"] = start, count, mat & 0xFFFFFFFF
        meshes.append(asset.Mesh(verts[v0:v0 + m.num_vertices].copy(),
                                 idx[m.index_offset:m.index_offset + m.num_indices].copy(), srf))
    mats = np.zeros(s.n_materials, asset.MATERIAL_DTYPE)
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
                     for i in range(s.n_samplers)], asset.SAMPLER_DTYPE)
    types = {abi.TEX_RGBA8_SRGB: 0, abi.TEX_RGBA8_UNORM: 0, abi.TEX_RG8_UNORM: 1, abi.TEX_RG8_SRGB: 2}
    images = [asset.Image(types[s.textures[i].format], f"texture_{i}.dds".encode()) for i in range(s.n_textures)]
    nodes = []
    for o in range(s.n_objects):
        mesh = next((int(d.mesh_index) for d in draws if d.graphics_object_index == o), abi.RSY_ABSENT)
        nodes.append(asset.Node(transform=tuple(s.objects[o].transform), mesh_id=mesh, name=f"object_{o}".encode()))
    return asset.Asset(materials=mats, samplers=smps, scenes=[list(range(len(nodes)))], nodes=nodes, images=images,
                       meshes=meshes)
