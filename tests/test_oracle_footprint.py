"""Pins for the oracle's shading footprint (orc_shade_footprint): the unique texels, shadow-map
texels, latent cells, triangles and vertices the shading of a view reads, which give the shading
kernel's algorithmic bytes.  Each case has a count fixed by its construction: 1x1 maps read one
texel each; a quad covering the view with texture coordinates 0..1 on 8x8 maps reads every texel;
a full-view quad under an identity shadow projection reads every texel of a shadow map of the view's
size; a 4x4 latent grid under the same quad is read whole."""
import numpy as np

import oracle
from rosy_amd import abi
from lit_scene import GEOMETRIES, lit_scene
from scenes import CustomScene


def footprint(sc, shadow_map, neural=None):
    cv = sc.camera_views()[0]
    vis, _ = oracle.raster_view(sc, cv.view_projection[:], cv.width, cv.height, False, sc.s.raster_state)
    return vis, oracle.shade_footprint(sc, cv, sc.cascades(), [shadow_map], vis, neural=neural)


def full_quad(size=64, textures=None, shadows=True):
    # corners at -1 / +1 (the view's edges), uv 0..1, normal towards the camera
    pos = [(-1.0, -1.0, 0.5), (1.0, -1.0, 0.5), (1.0, 1.0, 0.5), (-1.0, 1.0, 0.5)]
    uvs = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
    sc = CustomScene(pos, [0, 1, 2, 0, 2, 3], size, size, uvs=uvs, textures=textures, shadow_size=size)
    sc.s.scene_data.shadows_enabled = 1 if shadows else 0
    return sc


def test_one_texel_maps_and_a_quad():
    sc = lit_scene(0, 128, GEOMETRIES["normal"][0])  # no normal map, 1x1 albedo and metallic maps
    vis, f = footprint(sc, np.zeros((64, 64), np.float32))
    assert f["covered_pixels"] == 26 * 26 == (vis != 0).sum()
    assert f["visible_triangles"] == 2 and f["visible_vertices"] == 4
    assert f["texels"] == 2 and f["texel_bytes"] == 4 + 2
    assert f["shadow_texels"] == 0 and f["latent_cells"] == 0  # shadows off, no decoder


def test_full_view_quad_reads_every_texel_and_shadow_texel():
    rng = np.random.default_rng(3)
    tex = [(8, 8, abi.TEX_RGBA8_SRGB, rng.integers(0, 256, 256, dtype=np.uint8).tobytes()),
           (8, 8, abi.TEX_RG8_UNORM, rng.integers(0, 256, 128, dtype=np.uint8).tobytes()),
           (8, 8, abi.TEX_RG8_SRGB, rng.integers(0, 256, 128, dtype=np.uint8).tobytes())]
    sc = full_quad(64, tex)
    vis, f = footprint(sc, np.zeros((64, 64), np.float32))
    assert f["covered_pixels"] == 64 * 64
    assert f["texels"] == 3 * 64 and f["texel_bytes"] == 64 * (4 + 2 + 2)
    assert f["shadow_texels"] == 64 * 64
    # without shadows the map is not read
    sc = full_quad(64, tex, shadows=False)
    _, f = footprint(sc, np.zeros((64, 64), np.float32))
    assert f["shadow_texels"] == 0 and f["texels"] == 3 * 64


def test_decoder_reads_its_latent_cells_instead_of_the_maps():
    rng = np.random.default_rng(4)
    tex = [(8, 8, abi.TEX_RGBA8_SRGB, bytes([200, 150, 100, 255] * 64)),
           (8, 8, abi.TEX_RG8_UNORM, rng.integers(0, 256, 128, dtype=np.uint8).tobytes()),
           (8, 8, abi.TEX_RG8_SRGB, rng.integers(0, 256, 128, dtype=np.uint8).tobytes())]
    sc = full_quad(64, tex, shadows=False)
    L = 4
    params = oracle.NeuralParams([0], rng.uniform(-0.5, 0.5, (1, L, L, 8)).astype(np.float32),
                                 rng.uniform(-0.5, 0.5, 583).astype(np.float32))
    _, f = footprint(sc, np.zeros((64, 64), np.float32), neural=params)
    assert f["latent_cells"] == L * L and f["texels"] == 0
    # the fallback (material inactive) reads the maps again
    params = oracle.NeuralParams([0], params.latents, params.mlp, active=[0])
    _, f = footprint(sc, np.zeros((64, 64), np.float32), neural=params)
    assert f["latent_cells"] == 0 and f["texels"] == 3 * 64
