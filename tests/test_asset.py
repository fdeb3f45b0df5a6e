""".rsy asset reader/writer (rsy_asset_*) against the restatement of the format in
oracle/rsy_format.py (Asset.h:9-149, Asset.cpp:39-848): byte-exact encoding, decoding of the
restatement's bytes, file round trips, and the reader's failure cases (short files at every field
boundary, header mismatches, size fields, counts larger than the file).  The reference ships no
.rsy fixtures, so the pins are the reference's own struct definitions and write order, restated
independently of the library.  Host only."""
import os
import struct

import numpy as np
import pytest

from oracle import rsy_format as fmt
from rosy_amd import abi, asset, RosyError


def random_asset(seed, n_materials=3, n_samplers=2, n_scenes=2, n_nodes=4, n_images=3, n_meshes=3):
    rng = np.random.default_rng(seed)
    f32 = lambda n: [float(x) for x in rng.standard_normal(n).astype(np.float32)]
    idx = lambda hi: int(rng.integers(0, hi)) if rng.random() < 0.8 else 0xFFFFFFFF
    mats = []
    for _ in range(n_materials):
        mats.append((int(rng.integers(0, 2)), *f32(4), *f32(2), int(rng.integers(0, 3)), *f32(1),
                     *[idx(max(1, n_images)) if k % 2 == 0 else idx(max(1, n_samplers)) for k in range(8)]))
    samplers = [tuple(int(x) for x in rng.integers(0, 6, 4)) for _ in range(n_samplers)]
    scenes = [[int(x) for x in rng.integers(0, max(1, n_nodes), rng.integers(0, 4))] for _ in range(n_scenes)]
    nodes = []
    for k in range(n_nodes):
        nodes.append({"world_translate": f32(3), "world_scale": f32(1)[0], "world_yaw": f32(1)[0],
                      "transform": f32(16), "mesh_id": idx(max(1, n_meshes)),
                      "children": [int(x) for x in rng.integers(0, max(1, n_nodes), rng.integers(0, 3))],
                      "name": f"node_{k}".encode() * int(rng.integers(0, 3))})
    images = [(int(rng.integers(0, 4)), f"img{k}.dds".encode()) for k in range(n_images)]
    meshes = []
    for _ in range(n_meshes):
        nv = int(rng.integers(0, 9))
        positions = [tuple(f32(16)) for _ in range(nv)]
        indices = [int(x) for x in rng.integers(0, max(1, nv), 3 * int(rng.integers(0, 5)))]
        surfaces = [(int(rng.integers(0, 10)), int(rng.integers(0, 10)), idx(max(1, n_materials)), *f32(6))
                    for _ in range(int(rng.integers(0, 3)))]
        meshes.append((positions, indices, surfaces))
    return {"coordinate_system": f32(16), "root_scene": int(rng.integers(0, 3)), "materials": mats,
            "samplers": samplers, "scenes": scenes, "nodes": nodes, "images": images, "meshes": meshes}


def to_asset(d):
    """The restatement's dict as the harness's Asset."""
    mats = np.zeros(len(d["materials"]), asset.MATERIAL_DTYPE)
    for i, m in enumerate(d["materials"]):
        mats[i]["double_sided"] = m[0]
        mats[i]["base_color_factor"] = m[1:5]
        mats[i]["metallic_factor"], mats[i]["roughness_factor"] = m[5], m[6]
        mats[i]["alpha_mode"], mats[i]["alpha_cutoff"] = m[7], m[8]
        for k, name in enumerate(asset.MATERIAL_DTYPE.names[8:]):
            mats[i][name] = m[9 + k]
    smps = np.array([tuple(s) for s in d["samplers"]], asset.SAMPLER_DTYPE)
    nodes = [asset.Node(tuple(n["world_translate"]), n["world_scale"], n["world_yaw"], tuple(n["transform"]),
                        n["mesh_id"], list(n["children"]), bytes(n["name"])) for n in d["nodes"]]
    images = [asset.Image(t, bytes(name)) for t, name in d["images"]]
    meshes = []
    for positions, indices, surfaces in d["meshes"]:
        srf = np.zeros(len(surfaces), asset.SURFACE_DTYPE)
        for i, s in enumerate(surfaces):
            srf[i] = (s[0], s[1], s[2], s[3:6], s[6:9])
        meshes.append(asset.Mesh(np.array(positions, np.float32).reshape(-1, 16), np.array(indices, np.uint32), srf))
    return asset.Asset(tuple(d["coordinate_system"]), d["root_scene"], mats, smps, [list(s) for s in d["scenes"]],
                       nodes, images, meshes)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_encoding_matches_the_restatement_byte_for_byte(seed):
    d = random_asset(seed)
    assert asset.encode(to_asset(d)) == fmt.encode(d)


@pytest.mark.parametrize("seed", [4, 5])
def test_decoding_the_restatements_bytes(seed):
    d = random_asset(seed)
    got = asset.decode(fmt.encode(d))
    # what the library read, written again by the restatement, is the same file
    assert asset.encode(got) == fmt.encode(d)
    assert fmt.decode(asset.encode(got)) == fmt.decode(fmt.encode(d))
    assert [n.name for n in got.nodes] == [bytes(n["name"]) for n in d["nodes"]]
    assert [list(s) for s in got.scenes] == d["scenes"]


def test_empty_asset_and_file_round_trip(tmp_path):
    empty = {"coordinate_system": list(asset.IDENTITY), "root_scene": 0, "materials": [], "samplers": [],
             "scenes": [], "nodes": [], "images": [], "meshes": []}
    b = asset.encode(to_asset(empty))
    assert b == fmt.encode(empty) and len(b) == 80 + 48
    d = random_asset(7, n_meshes=5)
    p = tmp_path / "level.rsy"
    asset.write(to_asset(d), p)
    assert p.read_bytes() == fmt.encode(d)
    assert asset.encode(asset.read(p)) == fmt.encode(d)


def test_a_short_file_fails_at_every_field_boundary():
    d = random_asset(8, n_materials=2, n_samplers=1, n_scenes=1, n_nodes=2, n_images=2, n_meshes=2)
    b = fmt.encode(d)
    cuts = set(fmt.boundaries(d)) | {o + 1 for o in fmt.boundaries(d)} | {o - 1 for o in fmt.boundaries(d)}
    cuts = sorted(c for c in cuts if 0 <= c < len(b))
    assert len(cuts) > 20
    for c in cuts:
        with pytest.raises(RosyError) as e:
            asset.decode(b[:c])
        assert e.value.code == abi.RSY_READ_FAILED, c
        with pytest.raises(fmt.FormatError):
            fmt.decode(b[:c])
    assert asset.encode(asset.decode(b)) == b


@pytest.mark.parametrize("field,value", [(0, 0x46595352), (1, 2), (2, 0)])
def test_header_mismatches_fail(field, value):
    b = bytearray(fmt.encode(random_asset(9)))
    struct.pack_into("<I", b, 4 * field, value)
    with pytest.raises(RosyError) as e:
        asset.decode(bytes(b))
    assert e.value.code == abi.RSY_READ_FAILED


def test_node_and_image_size_fields_must_be_one():
    d = random_asset(10, n_materials=0, n_samplers=0, n_scenes=0, n_nodes=1, n_images=1, n_meshes=0)
    b = bytearray(fmt.encode(d))
    node_sizes = 80 + 48
    for k in range(5):
        c = bytearray(b)
        struct.pack_into("<Q", c, node_sizes + 8 * k, 2)
        with pytest.raises(RosyError) as e:
            asset.decode(bytes(c))
        assert e.value.code == abi.RSY_READ_FAILED
    image_sizes = node_sizes + 56 + 88 + 4 * len(d["nodes"][0]["children"]) + len(d["nodes"][0]["name"])
    assert struct.unpack_from("<Q", b, image_sizes)[0] == 1
    c = bytearray(b)
    struct.pack_into("<Q", c, image_sizes, 0)
    with pytest.raises(RosyError) as e:
        asset.decode(bytes(c))
    assert e.value.code == abi.RSY_READ_FAILED


@pytest.mark.parametrize("which", range(6))
def test_counts_larger_than_the_file_fail_without_allocating(which):
    b = bytearray(fmt.encode(random_asset(11)))
    struct.pack_into("<Q", b, 80 + 8 * which, 1 << 61)
    with pytest.raises(RosyError) as e:
        asset.decode(bytes(b))
    assert e.value.code == abi.RSY_READ_FAILED


def test_mesh_counts_larger_than_the_file_fail():
    d = random_asset(12, n_materials=0, n_samplers=0, n_scenes=0, n_nodes=0, n_images=0, n_meshes=1)
    for k in range(3):
        b = bytearray(fmt.encode(d))
        struct.pack_into("<Q", b, 80 + 48 + 8 * k, (1 << 58) + 3)
        with pytest.raises(RosyError) as e:
            asset.decode(bytes(b))
        assert e.value.code == abi.RSY_READ_FAILED


def test_open_and_write_failures(tmp_path):
    with pytest.raises(RosyError) as e:
        asset.read(tmp_path / "missing.rsy")
    assert e.value.code == abi.RSY_OPEN_FAILED
    with pytest.raises(RosyError) as e:
        asset.write(to_asset(random_asset(13)), tmp_path / "no_such_dir" / "x.rsy")
    assert e.value.code == abi.RSY_OPEN_FAILED


def test_pack_geometry_concatenates_meshes_as_set_asset_does():
    d = random_asset(14, n_meshes=4)
    with asset.open_bytes(fmt.encode(d)) as h:
        v, i, m = h.pack_geometry()
    pos = [p for ms in d["meshes"] for p in ms[0]]
    idx = [x for ms in d["meshes"] for x in ms[1]]
    assert np.array_equal(v, np.array(pos, np.float32).reshape(-1, 16))
    assert np.array_equal(i, np.array(idx, np.uint32))
    vo = io = 0
    for k, (positions, indices, _) in enumerate(d["meshes"]):
        assert (m[k].vertex_buffer_offset, m[k].index_offset, m[k].num_indices, m[k].num_vertices) == \
            (64 * vo, io, len(indices), len(positions))
        vo += len(positions)
        io += len(indices)


def test_pack_materials_maps_indices_as_set_asset_does():
    d = random_asset(15, n_materials=8, n_samplers=2, n_images=3)
    # out-of-range image and sampler indices on purpose
    d["materials"][0] = d["materials"][0][:9] + (7, 0, 1, 5, 0xFFFFFFFF, 1, 2, 0)
    with asset.open_bytes(fmt.encode(d)) as h:
        g = h.pack_materials(len(d["materials"]))
    n_images, n_samplers = len(d["images"]), len(d["samplers"])
    for m, out in zip(d["materials"], g):
        assert list(out.color) == pytest.approx(list(m[1:5]), abs=0)
        assert (out.metallic_factor, out.roughness_factor, out.alpha_mode) == (pytest.approx(m[5], abs=0),
                                                                                pytest.approx(m[6], abs=0), m[7])
        assert out.alpha_cutoff == pytest.approx(m[8], abs=0)
        fields = [("color_sampled_image_index", "color_sampler_index"),
                  ("normal_sampled_image_index", "normal_sampler_index"),
                  ("metallic_sampled_image_index", "metallic_sampler_index"),
                  ("mixmap_sampled_image_index", "mixmap_sampler_index")]
        for k, (fi, fs) in enumerate(fields):
            image, sampler = m[9 + 2 * k], m[10 + 2 * k]
            want_image = image if image < n_images else abi.RSY_ABSENT
            want_sampler = sampler if (image < n_images and sampler < n_samplers) else abi.RSY_ABSENT
            assert (getattr(out, fi), getattr(out, fs)) == (want_image, want_sampler)
    with pytest.raises(RosyError) as e:
        with asset.open_bytes(fmt.encode(d)) as h:
            h.pack_materials(len(d["materials"]) + 1)
    assert e.value.code == abi.RSY_INVALID_ARGUMENT


def test_library_exports_the_asset_entry_points():
    import rosy_amd
    names = [n for n, _, _ in abi.PROTOTYPES if n.startswith("rsy_asset_")]
    assert len(names) == 9
    for n in names:
        assert hasattr(rosy_amd.lib(), n)
