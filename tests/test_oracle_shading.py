"""Pins for the oracle's FP64 shading half (basic.slang, lighting.slang, pcf.slang).

Each term is pinned by what the mathematics fixes: the GGX lobe integrates to one over the
hemisphere, Smith/Schlick and Fresnel take known values at their end points, the sRGB EOTF has
published values, PCF over constant and single-step maps has closed-form results, and whole-frame
special cases (lighting off, light behind the surface, full shadow) reduce to albedo or ambient."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
import rosy_amd
from rosy_amd import abi
from scenes import CustomScene

L = oracle.lib


def test_srgb_eotf_published_values():
    f = L().orc_srgb_to_linear
    assert f(0) == 0.0 and f(255) == 1.0
    # linear segment below 0.04045 (code 10 = 0.0392), power segment above (code 11 = 0.0431)
    assert abs(f(10) - (10 / 255) / 12.92) < 1e-15
    assert abs(f(11) - ((11 / 255 + 0.055) / 1.055) ** 2.4) < 1e-15
    assert abs(f(128) - 0.21586) < 1e-5  # sRGB 50% grey
    vals = [f(i) for i in range(256)]
    assert all(b > a for a, b in zip(vals, vals[1:]))


@pytest.mark.parametrize("r", [0.3, 0.5, 0.8, 1.0])
def test_ggx_lobe_is_normalised(r):
    # integral over the hemisphere of D(h) (n.h) dw = 1 (Trowbridge-Reitz / GGX normalisation)
    n = 20000
    th = (np.arange(n) + 0.5) * (math.pi / 2) / n
    d = np.array([L().orc_ggx_d(math.cos(t), r) for t in th])
    integral = 2 * math.pi * np.sum(d * np.cos(th) * np.sin(th)) * (math.pi / 2) / n
    assert abs(integral - 1.0) < 2e-3


def test_ggx_special_values():
    for ndh in (0.0, 0.3, 0.9, 1.0):
        assert abs(L().orc_ggx_d(ndh, 1.0) - 1 / math.pi) < 1e-12  # a = 1: flat lobe
    # peak value 1 / (pi a^2) at n.h = 1, a = r^2
    assert abs(L().orc_ggx_d(1.0, 0.5) - 1 / (math.pi * 0.25 ** 2)) < 1e-9


def test_smith_schlick_and_fresnel_end_points():
    g = L().orc_schlick_ggx
    for r in (0.0, 0.3, 1.0):
        assert abs(g(1.0, r) - 1.0) < 1e-15 and g(0.0, r) == 0.0
    assert abs(g(0.5, 1.0) - 2 / 3) < 1e-15  # k = (1+1)^2/8 = 1/2
    F = L().orc_schlick_fresnel
    assert abs(F(1.0, 0.04) - 0.04) < 1e-15 and abs(F(0.0, 0.04) - 1.0) < 1e-15
    assert abs(F(0.5, 0.04) - (0.04 + 0.96 / 32)) < 1e-15


def pcf(m, u, v, ref, mode):
    m = np.ascontiguousarray(m, dtype=np.float32)
    margin = C.c_double()
    return L().orc_pcf(m.ctypes.data_as(C.POINTER(C.c_float)), m.shape[0], u, v, ref, mode, C.byref(margin)), margin.value


@pytest.mark.parametrize("mode", [abi.PCF_REFERENCE_5TAP, abi.PCF_GRID_3X3])
def test_pcf_constant_maps(mode):
    S = 64
    assert pcf(np.zeros((S, S)), 0.37, 0.61, 0.3, mode)[0] == 1.0    # empty map: lit
    assert pcf(np.full((S, S), 0.9), 0.37, 0.61, 0.3, mode)[0] == 0.0  # occluder nearer the sun: shadow
    assert pcf(np.full((S, S), 0.375), 0.37, 0.61, 0.375, mode)[0] == 1.0  # equal depth passes GREATER_OR_EQUAL


def test_pcf_single_step_edge_closed_forms():
    S = 64
    m = np.zeros((S, S))
    m[:, :32] = 0.9  # occluders left of the texel boundary at u = 0.5
    # 3x3 one-texel grid centred on the boundary: columns give 0, 1/2 and 1 -> 1/2
    assert abs(pcf(m, 0.5, 0.5, 0.3, abi.PCF_GRID_3X3)[0] - 0.5) < 1e-12
    # reference 5 taps: mean lit fraction 1/2 + S*o*(-1 + 2 + 0 + 2 - 2)/5 with o = 0.00025
    want = 0.5 + S * 0.00025 * 1 / 5
    assert abs(pcf(m, 0.5, 0.5, 0.3, abi.PCF_REFERENCE_5TAP)[0] - want) < 1e-12
    # clamp to edge: sampling far outside the map reads the border texels
    assert pcf(m, -3.0, 0.5, 0.3, abi.PCF_GRID_3X3)[0] == 0.0
    assert pcf(m, 4.0, 0.5, 0.3, abi.PCF_GRID_3X3)[0] == 1.0


def quad_scene(normal_z=-1.0, sun_z=-20.0, textures=None, material=True):
    # a camera-facing square at z = 0.5 (identity view-projection), normal towards -z
    pos = [(-0.8, -0.8, 0.5), (0.8, -0.8, 0.5), (0.8, 0.8, 0.5), (-0.8, 0.8, 0.5)]
    nrm = [(0.0, 0.0, normal_z)] * 4
    sc = CustomScene(pos, [0, 1, 2, 0, 2, 3], 32, 32, normals=nrm, textures=textures, material=material)
    sc.s.scene_data.sun_position[:] = [0.0, 0.0, sun_z, 1.0]
    return sc


def shade(sc, sd=None, shadow_map=None):
    cv = sc.camera_views()[0]
    vis, _ = oracle.raster_view(sc, cv.view_projection[:], cv.width, cv.height, False, sc.s.raster_state)
    sm = shadow_map if shadow_map is not None else np.zeros((64, 64), np.float32)
    rgba, sh, _ = oracle.shade_view(sc, cv, sc.cascades(), [sm], vis, scene_data=sd)
    return vis, rgba, sh


def scene_data(sc, **kw):
    sd = abi.SceneData()
    C.memmove(C.byref(sd), C.byref(sc.s.scene_data), C.sizeof(sd))
    for k, v in kw.items():
        setattr(sd, k, v)
    return sd


def test_lighting_off_outputs_albedo_and_clear_colour():
    sc = quad_scene()
    vis, rgba, _ = shade(sc, scene_data(sc, light_enabled=0))
    eotf = L().orc_srgb_to_linear
    want = [eotf(200), eotf(150), eotf(100), 1.0]  # 1x1 RGBA8 sRGB albedo (200,150,100,255)
    covered = vis != 0
    assert covered.sum() > 300
    assert np.allclose(rgba[covered], want, atol=1e-15)
    assert np.allclose(rgba[~covered], [0.0, 0.05, 0.1, 1.0], atol=1e-7)  # Graphics.cpp:5416-5418


def test_light_behind_surface_leaves_ambient_times_diffuse():
    sc = quad_scene(sun_z=+20.0)  # sun behind the quad: n.l <= 0 everywhere
    vis, rgba, _ = shade(sc)
    eotf = L().orc_srgb_to_linear
    albedo = np.array([eotf(200), eotf(150), eotf(100)])
    covered = vis != 0
    assert np.allclose(rgba[covered][:, :3], 0.01 * albedo / 3.14159265, atol=1e-12)


def test_full_shadow_leaves_ambient_times_diffuse():
    sc = quad_scene()
    vis, rgba, sh = shade(sc, shadow_map=np.full((64, 64), 1.0, np.float32))
    covered = vis != 0
    assert (sh[covered] == 0.0).all()
    eotf = L().orc_srgb_to_linear
    albedo = np.array([eotf(200), eotf(150), eotf(100)])
    assert np.allclose(rgba[covered][:, :3], 0.01 * albedo / 3.14159265, atol=1e-12)


def test_debug_outputs_vertex_colour_and_uv():
    sc = rosy_amd.Scene(0, seed=1)
    cv = sc.camera_views()[0]
    vis, _ = oracle.raster_view(sc, cv.view_projection[:], cv.width, cv.height, False, sc.s.raster_state)
    sm = np.zeros((sc.cascades()[0].size,) * 2, np.float32)
    covered = vis != 0
    rgba, _, _ = oracle.shade_view(sc, cv, sc.cascades(), [sm], vis, scene_data=scene_data(sc, light_enabled=0, fragment_output=5))
    assert np.allclose(rgba[covered], [1.0, 0.0, 0.0, 1.0], atol=1e-12)  # position::color default (1,0,0,1)
    rgba, _, _ = oracle.shade_view(sc, cv, sc.cascades(), [sm], vis, scene_data=scene_data(sc, light_enabled=0, fragment_output=12))
    uv = rgba[covered][:, :2]
    assert uv.min() >= -1e-12 and uv.max() <= 1 + 1e-12  # octahedral sphere UVs lie in [0,1]
    rgba, _, _ = oracle.shade_view(sc, cv, sc.cascades(), [sm], vis, scene_data=scene_data(sc, light_enabled=0, fragment_output=1))
    n = rgba[covered][:, :3] * 2 - 1
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-9)
    # sphere normals point to the camera (which looks at the sphere from z = -3)
    assert (n[:, 2] < 0.2).mean() > 0.95


def test_lit_sphere_is_brighter_towards_the_sun():
    sc = rosy_amd.Scene(0, seed=1)
    cv = sc.camera_views()[0]
    vis, _ = oracle.raster_view(sc, cv.view_projection[:], cv.width, cv.height, False, sc.s.raster_state)
    sm = np.zeros((sc.cascades()[0].size,) * 2, np.float32)
    rgba, sh, _ = oracle.shade_view(sc, cv, sc.cascades(), [sm], vis)
    covered = vis != 0
    assert np.allclose(sh[covered], 1.0, atol=1e-12)  # empty shadow map: fully lit
    lum = rgba[..., :3].sum(axis=2)
    ys, xs = np.nonzero(covered)
    med = np.median(ys)
    # the default sun is above the horizon: the upper half (small y) of the sphere is brighter
    top = lum[ys[ys < med], xs[ys < med]].mean()
    bottom = lum[ys[ys > med], xs[ys > med]].mean()
    assert top > 2 * bottom
