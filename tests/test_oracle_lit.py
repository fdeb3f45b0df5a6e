"""Closed-form pins for the oracle's lit Cook-Torrance assembly (basic.slang:301-331 with
lighting.slang:5-43).

At one pixel of a flat quad the geometry is chosen so that every cosine is known exactly (normal
incidence: all 1; oblique: NdotL = 4/5, NdotH = VdotH = 3/sqrt(10)).  The expected colour is the
published formula evaluated at those numbers, over metallic 0 / 1 and three roughness values.
Whole-formula special cases reduce further: at normal incidence with roughness 1, D = 1/pi and
G = 1, so a metal's colour is f0 / (pi (4 + 1e-4)) * sun * intensity.  A dropped (1 - metallic)
in kD, a missing factor 4 or 1e-4 in the specular denominator, a diffuse pi of the wrong precision
or a wrong Smith k each move a value here by more than the 1e-12 tolerance."""
import math

import numpy as np
import pytest

import oracle
from lit_scene import GEOMETRIES, METALS, PIXEL, ROUGHS, ALBEDO_SRGB, AMBIENT, INTENSITY, SUN_COLOR, \
    closed_form, lit_scene, srgb_eotf
from test_oracle_shading import scene_data


def shade(sc, sd=None):
    cv = sc.camera_views()[0]
    vis, _ = oracle.raster_view(sc, cv.view_projection[:], cv.width, cv.height, False, sc.s.raster_state)
    sm = np.zeros((64, 64), np.float32)
    rgba, _, _ = oracle.shade_view(sc, cv, sc.cascades(), [sm], vis, scene_data=sd)
    return vis, rgba


@pytest.mark.parametrize("geometry", list(GEOMETRIES))
@pytest.mark.parametrize("metal", METALS)
@pytest.mark.parametrize("rough", ROUGHS)
def test_lit_pixel_matches_the_closed_form(geometry, metal, rough):
    sun, cosines = GEOMETRIES[geometry]
    sc = lit_scene(metal, rough, sun)
    vis, rgba = shade(sc)
    x, y = PIXEL
    assert vis[y, x] != 0
    want = closed_form(metal / 255.0, rough / 255.0, **cosines)
    got = rgba[y, x, :3]
    assert np.allclose(got, want, rtol=1e-12, atol=1e-14), (got, want)
    assert rgba[y, x, 3] == 1.0


def test_pixel_geometry_is_the_stated_one():
    # the debug outputs at the pixel: light direction, view direction and n.l (basic.slang:236-288)
    sun, _ = GEOMETRIES["oblique"]
    sc = lit_scene(0, 255, sun)
    x, y = PIXEL
    sd = scene_data(sc, fragment_output=3, light_enabled=0)
    _, rgba = shade(sc, sd)
    assert np.allclose(rgba[y, x, :3] * 2 - 1, [0.6, 0.0, -0.8], atol=1e-13)
    _, rgba = shade(sc, scene_data(sc, fragment_output=4, light_enabled=0))
    assert np.allclose(rgba[y, x, :3] * 2 - 1, [0.0, 0.0, -1.0], atol=1e-13)
    _, rgba = shade(sc, scene_data(sc, fragment_output=8, light_enabled=0))
    assert abs(rgba[y, x, 0] * 2 - 1 - 0.8) < 1e-13


def test_rough_metal_at_normal_incidence_reduces_to_f0_over_pi():
    # roughness 1: a = 1, D = 1/pi; k = 1/2, G1(1) = 1; metallic 1: kD = 0, f0 = albedo, F(1) = f0
    sc = lit_scene(255, 255, GEOMETRIES["normal"][0])
    _, rgba = shade(sc)
    x, y = PIXEL
    for ch in range(3):
        f0 = srgb_eotf(ALBEDO_SRGB[ch])
        want = f0 / 3.14159265359 / (4.0 + 1e-4) * SUN_COLOR[ch] * INTENSITY
        assert abs(rgba[y, x, ch] - want) < 1e-14


def test_rough_dielectric_at_normal_incidence():
    # metallic 0, roughness 1, normal incidence: f0 = F = 0.04, kD = 0.96, D = 1/pi, G = 1
    sc = lit_scene(0, 255, GEOMETRIES["normal"][0])
    _, rgba = shade(sc)
    x, y = PIXEL
    for ch in range(3):
        alb = srgb_eotf(ALBEDO_SRGB[ch])
        want = (0.96 * alb / 3.14159265 + 0.04 / 3.14159265359 / (4.0 + 1e-4)) * SUN_COLOR[ch] * INTENSITY
        assert abs(rgba[y, x, ch] - want) < 1e-14


def test_weak_sun_leaves_the_ambient_term():
    # max(ambient * diffuse, brdf): with a sun 1e-9 as bright the ambient side wins
    sc = lit_scene(0, 128, GEOMETRIES["oblique"][0])
    sd = scene_data(sc, light_intensity=1e-9)
    _, rgba = shade(sc, sd)
    x, y = PIXEL
    want = [AMBIENT[ch] * srgb_eotf(ALBEDO_SRGB[ch]) / 3.14159265 for ch in range(3)]
    assert np.allclose(rgba[y, x, :3], want, rtol=1e-12, atol=0)


def test_closed_form_is_not_the_unlit_value():
    # guard on the pins themselves: the lit values differ from albedo and from ambient * diffuse
    for geometry, (sun, cos) in GEOMETRIES.items():
        v = closed_form(0.0, 0.5, **cos)
        for ch in range(3):
            alb = srgb_eotf(ALBEDO_SRGB[ch])
            assert v[ch] > 5 * AMBIENT[ch] * alb / 3.14159265 and abs(v[ch] - alb) > 1e-3
    # and the oblique geometry has the stated cosines
    h = np.array([0.6, 0.0, -0.8]) + np.array([0.0, 0.0, -1.0])
    h /= np.linalg.norm(h)
    assert abs(h @ np.array([0, 0, -1.0]) - 3 / math.sqrt(10)) < 1e-15
