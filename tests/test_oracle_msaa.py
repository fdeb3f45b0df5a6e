"""Pins for the oracle's 4-sample main pass (Graphics.cpp:1794-1808, 5566-5575: a 4-sample colour
attachment resolved by averaging; no sample shading).

A sample plane is the single-sample raster at that sample location, so each plane's coverage is
checked against the exact rational top-left rule at the sample points, its depth against the plane
equation there, and the sample locations against the standard 4-sample pattern (a vertical edge
swept across one pixel uncovers the samples in the order of their x positions).  The resolve is
checked on closed forms: a pixel inside one triangle resolves to its centre shading, a pixel whose
samples split between a triangle and the background resolves to the coverage-weighted mean."""
import numpy as np
import pytest

import oracle
from rosy_amd import abi
from scenes import CustomScene
from test_oracle_raster import (W, H, brute_force_mask, positions_from_subpixels, random_triangles, state)


def test_sample_locations_are_the_standard_four():
    # Vulkan's standard locations for 4 samples, in pixel units
    want = [(0.375, 0.125), (0.875, 0.375), (0.125, 0.625), (0.625, 0.875)]
    assert [(x / 256, y / 256) for x, y in oracle.MSAA4_SAMPLES] == want


@pytest.mark.parametrize("seed", [4, 5])
def test_each_plane_matches_the_exact_rule_at_its_sample(seed):
    for pts in random_triangles(30, seed):
        sc = CustomScene(positions_from_subpixels(pts), [0, 1, 2], W, H)
        vp = sc.camera_views()[0].view_projection[:]
        planes = oracle.raster_view_msaa(sc, vp, W, H, state())
        for s, smp in enumerate(oracle.MSAA4_SAMPLES):
            want = brute_force_mask(pts, smp)
            assert np.array_equal(planes[s] != 0, want), (pts, s)


def test_a_vertical_edge_uncovers_samples_in_x_order():
    # a rectangle from sub-pixel x = e to the right edge, one pixel column (4) holding the edge
    for e, covered in ((4 * 256 + 20, {0, 1, 2, 3}), (4 * 256 + 32, {0, 1, 2, 3}), (4 * 256 + 33, {0, 1, 3}),
                       (4 * 256 + 100, {1, 3}), (4 * 256 + 161, {1}), (4 * 256 + 224, {1}),
                       (4 * 256 + 225, set())):
        pts = [(e, 0), (W * 256, 0), (W * 256, H * 256), (e, H * 256)]
        sc = CustomScene(positions_from_subpixels(pts), [0, 1, 2, 0, 2, 3], W, H)
        planes = oracle.raster_view_msaa(sc, sc.camera_views()[0].view_projection[:], W, H, state())
        got = {s for s in range(4) if planes[s][8, 4] != 0}
        assert got == covered, (e, got)  # the left edge is inclusive (top-left rule)
        assert all((planes[s][8, 5] != 0) for s in range(4)) and not any(planes[s][8, 3] for s in range(4))


def test_depth_is_the_plane_at_each_sample():
    pts = [(128, 128), (15 * 256 + 128, 128), (128, 15 * 256 + 128)]
    zs = [0.25, 0.75, 0.5]
    pos = [(p[0], p[1], z) for p, z in zip(positions_from_subpixels(pts), zs)]
    sc = CustomScene(pos, [0, 1, 2], W, H)
    planes = oracle.raster_view_msaa(sc, sc.camera_views()[0].view_projection[:], W, H, state())
    for s, (sx, sy) in enumerate(oracle.MSAA4_SAMPLES):
        for py, px in np.argwhere(planes[s] != 0):
            Px, Py = px * 256 + sx, py * 256 + sy
            exact = zs[0] + (zs[1] - zs[0]) * (Px - 128) / (15 * 256) + (zs[2] - zs[0]) * (Py - 128) / (15 * 256)
            got = np.uint32(int(planes[s][py, px]) >> 32).view(np.float32)
            assert abs(float(got) - exact) <= 4 * np.spacing(np.float32(exact))


def test_resolve_is_the_coverage_weighted_mean():
    # a triangle over the background; unlit so shading is the material colour (closed form)
    pts = [(3 * 256 + 40, 2 * 256 + 10), (13 * 256 + 200, 5 * 256 + 77), (6 * 256 + 150, 14 * 256 + 230)]
    sc = CustomScene(positions_from_subpixels(pts), [0, 1, 2], W, H)
    sd = sc.s.scene_data
    sd.light_enabled = 0
    cam = sc.camera_views()[0]
    rs = state()
    planes = oracle.raster_view_msaa(sc, cam.view_projection[:], W, H, rs)
    resolved, samples, _, _ = oracle.msaa_view(sc, cam, sc.cascades(), [np.zeros((64, 64), np.float32)], planes,
                                            raster_state=rs)
    centre, _ = oracle.raster_view(sc, cam.view_projection[:], W, H, False, rs)
    shaded, _, _ = oracle.shade_view(sc, cam, sc.cascades(), [np.zeros((64, 64), np.float32)], centre, raster_state=rs)
    n = (planes != 0).sum(axis=0)
    inside = (n == 4) & (centre != 0)
    assert inside.sum() > 20
    assert np.allclose(resolved[inside], shaded[inside], rtol=0, atol=1e-12)
    clear = np.array(rs.clear_color[:], np.float64)
    edge = (n > 0) & (n < 4)
    assert edge.sum() > 10
    colour = shaded[inside][0]  # one flat material colour over the triangle
    want = (n[edge, None] * colour + (4 - n[edge, None]) * clear) / 4
    assert np.allclose(resolved[edge], want, rtol=0, atol=1e-12)
    assert np.allclose(resolved[n == 0], clear, rtol=0, atol=0)
