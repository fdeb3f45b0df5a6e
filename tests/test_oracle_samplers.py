"""Pins for the oracle's samplers and material-map branches.

Samplers follow the Vulkan specification's texel addressing and filtering (the reference's samplers,
Graphics.cpp:250-278, 3421-3432): nearest i = floor(u w); linear u' = u w - 1/2 with weights
frac(u'); clamp-to-edge clamp(i, 0, w-1); repeat i mod w; mirrored repeat (w-1) - mirror((i mod 2w) - w)
with mirror(a) = a for a >= 0 and -(1 + a) otherwise; sRGB texels decoded before filtering.  The
material branches (basic.slang:172-221) are pinned on the sampler scene through the debug outputs:
roughness, metallic and alpha at a pixel equal the named map's sample there."""
import math

import numpy as np
import pytest

import oracle
from rosy_amd import abi
from sampler_scene import MATERIALS, NEAREST, SAMPLERS, sampler_scene, textures
from test_oracle_shading import scene_data


def srgb(c):
    v = c / 255.0
    return np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4)


def wrap(i, n, mode):
    if mode == 0:
        return min(max(i, 0), n - 1)
    if mode == 2:
        return i % n
    t = i % (2 * n)
    a = t - n
    return (n - 1) - (a if a >= 0 else -(1 + a))


def spec_sample(tex, w, h, ch, is_srgb, smp, u, v):
    """Vulkan's filtering of a w x h texture of `ch` channels at (u, v)."""
    t = np.frombuffer(tex, np.uint8).reshape(h, w, ch).astype(np.float64)
    t = srgb(t) if is_srgb else t / 255.0
    if is_srgb and ch == 4:
        t[..., 3] = np.frombuffer(tex, np.uint8).reshape(h, w, ch)[..., 3] / 255.0  # alpha stays linear
    nearest = smp[1] <= 2
    ws, wt = smp[2], smp[3]
    if nearest:
        return t[wrap(math.floor(v * h), h, wt), wrap(math.floor(u * w), w, ws)]
    x, y = u * w - 0.5, v * h - 0.5
    x0, y0 = math.floor(x), math.floor(y)
    a, b = x - x0, y - y0
    xa, xb = wrap(x0, w, ws), wrap(x0 + 1, w, ws)
    ya, yb = wrap(y0, h, wt), wrap(y0 + 1, h, wt)
    return (1 - a) * (1 - b) * t[ya, xa] + a * (1 - b) * t[ya, xb] + (1 - a) * b * t[yb, xa] + a * b * t[yb, xb]


def test_wrap_functions_match_the_specification_examples():
    # mirrored repeat of a 4-texel row from i = -6: 2 3 3 2 1 0 | 0 1 2 3 | 3 2 1 0 | 0 1
    assert [wrap(i, 4, 1) for i in range(-6, 10)] == [2, 3, 3, 2, 1, 0, 0, 1, 2, 3, 3, 2, 1, 0, 0, 1]
    assert [wrap(i, 3, 0) for i in (-5, -1, 0, 2, 3, 9)] == [0, 0, 0, 2, 2, 2]
    assert [wrap(i, 3, 2) for i in (-4, -1, 0, 3, 5)] == [2, 2, 0, 0, 2]


@pytest.mark.parametrize("si", range(len(SAMPLERS)))
@pytest.mark.parametrize("ti", [0, 1, 5, 7])
def test_sampler_addressing_and_filtering(si, ti):
    w, h, fmt, data = textures()[ti]
    ch = 2 if fmt in (abi.TEX_RG8_UNORM, abi.TEX_RG8_SRGB) else 4
    is_srgb = fmt in (abi.TEX_RGBA8_SRGB, abi.TEX_RG8_SRGB)
    buf = np.frombuffer(data, np.uint8).copy()
    desc = abi.TextureDesc(w, h, fmt, 1, buf.ctypes.data)
    smp = abi.Sampler(*SAMPLERS[si])
    rng = np.random.default_rng(si * 31 + ti)
    for u, v in rng.uniform(-2.7, 3.4, size=(64, 2)):
        # keep nearest lookups away from texel edges (FP rounding there is not what is pinned)
        if SAMPLERS[si][1] <= 2 and (abs(u * w - round(u * w)) < 1e-9 or abs(v * h - round(v * h)) < 1e-9):
            continue
        got = oracle.sample_texture(desc, smp, float(u), float(v))
        want = spec_sample(data, w, h, ch, is_srgb, SAMPLERS[si], float(u), float(v))
        assert np.allclose(got[:ch], want, atol=1e-12), (u, v, got, want)


def shade_debug(sc, mode):
    cv = sc.camera_views()[0]
    vis, _ = oracle.raster_view(sc, cv.view_projection[:], cv.width, cv.height, False, sc.s.raster_state)
    sm = np.zeros((64, 64), np.float32)
    rgba, _, _ = oracle.shade_view(sc, cv, sc.cascades(), [sm], vis,
                                   scene_data=scene_data(sc, fragment_output=mode, light_enabled=0))
    return vis, rgba


@pytest.fixture(scope="module")
def debug_planes():
    sc = sampler_scene()
    vis, uv = shade_debug(sc, 12)
    _, rough = shade_debug(sc, 13)
    _, metal = shade_debug(sc, 14)
    _, alpha = shade_debug(sc, 15)
    _, col = shade_debug(sc, 5)
    _, img = shade_debug(sc, 0)
    return sc, vis, uv, rough[..., 0], metal[..., 0], 1 - alpha[..., 0], col, img


def _sample(sc, tex_index, smp_index, u, v):
    w, h, fmt, data = textures()[tex_index]
    buf = np.frombuffer(data, np.uint8).copy()
    return oracle.sample_texture(abi.TextureDesc(w, h, fmt, 1, buf.ctypes.data), abi.Sampler(*SAMPLERS[smp_index]), u, v)


def test_material_map_branches(debug_planes):
    sc, vis, uv, rough, metal, alpha, col, img = debug_planes
    covered = vis != 0
    tri = (vis.astype(np.int64) & 0xffffffff) - 1
    rng = np.random.default_rng(2)
    ys, xs = np.nonzero(covered)
    seen = set()
    for i in rng.choice(len(ys), size=600, replace=False):
        y, x = ys[i], xs[i]
        m = int(tri[y, x]) // 2
        seen.add(m)
        md = MATERIALS[m]
        u, v = float(uv[y, x, 0]), float(uv[y, x, 1])
        if md["mixmap_sampled_image_index"] != abi.RSY_ABSENT:
            s = _sample(sc, md["mixmap_sampled_image_index"], md["mixmap_sampler_index"], u, v)
            want_r, want_m = s[1], s[2]          # four-channel: G roughness, B metallic
        elif md["metallic_sampled_image_index"] != abi.RSY_ABSENT:
            t = md["metallic_sampled_image_index"]
            s = _sample(sc, t, md["metallic_sampler_index"], u, v)
            want_r, want_m = (s[1], s[0]) if textures()[t][2] in (abi.TEX_RG8_SRGB, abi.TEX_RG8_UNORM) else (s[1], s[2])
        else:
            want_r = want_m = 0.0                # neither map: 0 (the reference reads an absent descriptor)
        assert abs(rough[y, x] - want_r) < 1e-12 and abs(metal[y, x] - want_m) < 1e-12, (m, x, y)
        # alpha = min(material alpha, vertex alpha, albedo alpha): at most the material's and the map's
        a = md.get("color", (1, 1, 1, 1))[3]
        if md["color_sampled_image_index"] != abi.RSY_ABSENT:
            s = _sample(sc, md["color_sampled_image_index"], md["color_sampler_index"], u, v)
            assert np.allclose(img[y, x, :3], s[:3], atol=1e-12)
            a = min(a, s[3])
        else:
            assert np.allclose(img[y, x, :3], md["color"][:3], atol=1e-7)
        assert alpha[y, x] <= a + 1e-7
    assert seen == set(range(9))


def test_alpha_is_the_minimum_of_its_sources(debug_planes):
    sc, vis, uv, rough, metal, alpha, col, img = debug_planes
    tri = (vis.astype(np.int64) & 0xffffffff) - 1
    # material 5 (no maps, colour alpha 0.7): min(0.7, vertex alpha) -- 0.7 where the vertex alpha
    # is larger, below it where smaller (vertex alphas run from 0.25 to 1 over the quad)
    sel = (vis != 0) & (tri // 2 == 5)
    assert sel.sum() > 500
    a = alpha[sel]
    assert (a <= np.float32(0.7) + 1e-7).all()
    assert (np.abs(a - np.float32(0.7)) < 1e-7).sum() > 20 and a.min() < 0.6
    # material 7 (colour alpha 0.5 and an albedo alpha of at least 40/255): never above 0.5
    sel = (vis != 0) & (tri // 2 == 7)
    assert (alpha[sel] <= 0.5 + 1e-12).all() and alpha[sel].min() < 0.45


def test_nearest_lookups_stay_clear_of_texel_edges(debug_planes):
    # precondition of the GPU comparison: FP32 and FP64 texture coordinates pick the same texel
    sc, vis, uv, *_ = debug_planes
    tri = (vis.astype(np.int64) & 0xffffffff) - 1
    for m, md in enumerate(MATERIALS):
        sel = (vis != 0) & (tri // 2 == m)
        for key in ("color", "normal", "metallic", "mixmap"):
            t, s = md[f"{key}_sampled_image_index"], md[f"{key}_sampler_index"]
            if t == abi.RSY_ABSENT or s not in NEAREST:
                continue
            w, h = textures()[t][:2]
            for c, n in ((0, w), (1, h)):
                f = uv[sel][:, c] * n
                assert np.abs(f - np.round(f)).min() > 1e-4, (m, key)
