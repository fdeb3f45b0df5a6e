"""Pins for the oracle's texture level of detail (Vulkan's LOD and mipmapping rules with the asset
samplers' VK_SAMPLER_MIPMAP_MODE_LINEAR, Graphics.cpp:4002-4012) on closed forms: a texture whose
level l is a constant colour turns the sampled colour into the LOD itself, and quads whose texture
coordinates are known functions of the pixel give the footprint exactly -- affine scales (whole and
fractional levels, magnification, the clamp at the last level, anisotropic footprints take the
larger axis) and a receding plane under a perspective projection (per-row derivatives of
v = s (1 + y) / (2 (2 - y)))."""
import numpy as np
import pytest

import oracle
from lod_scene import (LEVEL_RED, affine_quad, expected_level_red, level_coded_texture, receding_footprint,
                       receding_quad)


def shaded_red(scene, size):
    cam = scene.camera_views()[0]
    vis, _ = oracle.raster_view(scene, cam.view_projection[:], size, size, False, scene.s.raster_state)
    assert (vis != 0).all()
    rgba, _, _ = oracle.shade_view(scene, cam, scene.cascades(), [np.zeros((64, 64), np.float32)], vis)
    return rgba[..., 0] * 255.0


@pytest.mark.parametrize("su,sv", [(2.0, 2.0), (3.0, 3.0), (0.25, 0.25), (40.0, 40.0), (2.0, 8.0), (5.0, 1.0)])
def test_affine_footprints(su, sv):
    size = 32
    red = shaded_red(affine_quad(size, su, sv, level_coded_texture()), size)
    rho = max(64 * su / size, 64 * sv / size)
    assert np.allclose(red, expected_level_red(rho), atol=1e-9)


def test_whole_levels_are_sampled_alone():
    assert expected_level_red(4.0) == LEVEL_RED[2] and expected_level_red(0.5) == LEVEL_RED[0]
    assert expected_level_red(1e6) == LEVEL_RED[6]


def test_receding_plane_follows_the_perspective_derivative():
    size, s = 32, 1.0
    red = shaded_red(receding_quad(size, s, level_coded_texture()), size)
    for py in range(size):
        for px in range(size):
            if px == py:
                continue  # on the diagonal the top-left rule picks the triangle
            want = expected_level_red(receding_footprint(size, s, px, py))
            assert abs(red[py, px] - want) <= 1e-6, (px, py, red[py, px], want)
    assert red.min() < LEVEL_RED[1] + 1 and red.max() > LEVEL_RED[2] + 10  # the pixels span several levels
