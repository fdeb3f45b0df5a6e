"""Pins of the blended-pass oracle (orc_blend_layers, oracle.blend_view) on a hand-built scene:
which fragments each pixel holds (coverage by the quads, the GEQUAL test against the opaque quad,
both faces drawn), their primitive order, and the blend arithmetic with lighting off, where every
fragment's colour is its albedo sample: colour = src * a + dst * (1 - a), alpha = a, and an alpha
mask below its cutoff leaves the pixel unchanged.  Host only."""
import numpy as np

import oracle
from blend_scene import QUADS, blend_scene, depth_at


def test_fragments_per_pixel_follow_the_quads():
    s = blend_scene(64)
    cam = s.camera_views()[0]
    vis, _ = oracle.raster_view(s, cam.view_projection[:], 64, 64, False, s.s.raster_state)
    layers, count = oracle.blend_layers(s, cam, vis, max_layers=8)
    blended = [k for k, q in enumerate(QUADS) if q[6]]
    opaque_z = {k: q[4] for k, q in enumerate(QUADS) if not q[6]}
    for py in range(0, 64, 3):
        for px in range(0, 64, 3):
            quads = depth_at(px, py, 64)
            front = max([QUADS[k][4] for k in quads if k in opaque_z], default=0.0)
            want = [k for k in quads if k in blended and QUADS[k][4] >= front]
            assert count[py, px] == len(want), (px, py)
            keys = [int(layers[i, py, px] & 0xFFFFFFFF) for i in range(count[py, px])]
            assert keys == sorted(keys)  # primitive order
            # key = triangle id + 1 in the blended list: two triangles per quad, in draw order
            assert [blended.index(k) for k in want] == [(kk - 1) // 2 for kk in keys]
    assert count.max() >= 5


def test_blend_arithmetic_with_lighting_off():
    s = blend_scene(64, light=False)
    cam = s.camera_views()[0]
    vis, _ = oracle.raster_view(s, cam.view_projection[:], 64, 64, False, s.s.raster_state)
    base, _, _ = oracle.shade_view(s, cam, s.cascades(), [np.zeros((64, 64), np.float32)], vis)
    out, count = oracle.blend_view(s, cam, s.cascades(), [np.zeros((64, 64), np.float32)], vis, base)
    lin = lambda c: c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4
    albedo1 = np.array([lin(200 / 255), lin(150 / 255), lin(100 / 255)])
    albedo3 = np.array([lin(30 / 255), lin(90 / 255), lin(240 / 255)])
    # pixel in quads 0 (opaque), 1 (a 0.5) and 5 (a 0.7), in that order
    px, py = 20, 39
    assert depth_at(px, py, 64) == [0, 1, 5]
    a1, a3 = 0.5, float(np.float32(0.7))  # the materials' colour alpha (FP32 records)
    want = base[py, px, :3]
    want = albedo1 * a1 + want * (1 - a1)
    want = albedo3 * a3 + want * (1 - a3)
    assert np.allclose(out[py, px, :3], want, atol=1e-12) and out[py, px, 3] == a3
    # outside the opaque quad, quad 3 (behind it elsewhere) lies over the clear colour
    px, py = 61, 32
    assert depth_at(px, py, 64) == [3] and count[py, px] == 1
    assert np.allclose(out[py, px, :3], albedo3 * a3 + base[py, px, :3] * (1 - a3), atol=1e-12)
    # the mask quad alone: drawn with its texture alpha at or above the cutoff, else discarded
    for px, py in ((50, 4), (30, 3), (58, 4), (24, 5)):
        assert depth_at(px, py, 64) == [2] and count[py, px] == 1
        a = oracle.sample_texture(s.s.textures[3], s.s.samplers[0], 0.5 * ((px + 0.5) / 32.0), 0.5 * ((py + 0.5) / 32.0))[3]
        assert abs(a - 0.6) > 1e-3  # the pixel's uv here is the interpolated one up to ~1e-5
        if a < 0.6:
            assert np.array_equal(out[py, px], base[py, px])
        else:
            assert np.isclose(out[py, px, 3], a, atol=1e-4)
    # a pixel with no blended fragment keeps the opaque colour
    assert np.array_equal(out[5, 5], base[5, 5])
    # both outcomes of the mask occur over the quad
    region = count[:8, 20:62] == 1
    changed = np.any(out[:8, 20:62] != base[:8, 20:62], axis=-1)
    assert changed[region].any() and (~changed[region]).any()
