"""A scene that reaches every sampler and material variant of the reference's shading path.

Nine quads (one material each) in a 3 x 3 grid, texture coordinates running from -1.3 to 2.2 so
that addressing modes matter:
  - samplers: nearest and linear filtering x clamp-to-edge, mirrored repeat and repeat, plus one
    with different modes per axis (Graphics.cpp:250-278, 3421-3432);
  - normal maps: two-channel RG8 (z reconstructed), four-channel RGBA8 with z from the texture
    (basic.slang:177) and a BC7 one (decoded at upload, Graphics.cpp:3632-3647);
  - metallic maps: two-channel (R metallic, G roughness) and four-channel (G roughness, B metallic,
    basic.slang:211-220); mixmaps, which take precedence over the metallic map (basic.slang:200-209);
  - a material with no maps at all (colour from the material, basic.slang:144-168) and alpha from
    the minimum of vertex, material and texture alpha.
Texture sizes include non-powers of two (5 x 7, 6 x 5).
"""
import numpy as np

from rosy_amd import abi
from scenes import CustomScene

A = abi.RSY_ABSENT
U0, U1 = -1.3, 2.2

# (min_filter, mag_filter, wrap_s, wrap_t); filter codes 0-2 nearest, 3-5 linear; wrap 0 clamp,
# 1 mirrored repeat, 2 repeat
SAMPLERS = [
    (4, 4, 2, 2),  # 0 linear repeat
    (0, 0, 0, 0),  # 1 nearest clamp
    (4, 4, 0, 0),  # 2 linear clamp
    (1, 1, 1, 1),  # 3 nearest mirrored
    (5, 5, 1, 1),  # 4 linear mirrored
    (2, 2, 2, 2),  # 5 nearest repeat
    (3, 3, 0, 1),  # 6 linear, clamp in s, mirrored in t
]
NEAREST = {1, 3, 5}


def _bc7_blocks(rng, n):
    """n random BC7 blocks of modes 0-7 (the mode is the position of the lowest set bit)."""
    out = bytearray(rng.integers(0, 256, size=16 * n, dtype=np.uint8).tobytes())
    for b in range(n):
        mode = int(rng.integers(0, 8))
        out[16 * b] = ((out[16 * b] >> (mode + 1)) << (mode + 1)) | (1 << mode)
    return bytes(out)


def textures(seed=5):
    rng = np.random.default_rng(seed)

    def img(w, h, ch, lo=0, rough=None):
        a = rng.integers(lo, 256, size=(w * h, ch), dtype=np.uint8)
        if rough is not None:  # roughness channel from code 96 (0.117 after sRGB decoding)
            a[:, rough] = rng.integers(96, 256, size=w * h, dtype=np.uint8)
        return a.tobytes()

    # Roughness starts at 0.117: below it the GGX peak turns FP32 rounding of the normal and half
    # vector into more than 1e-3 of an HDR highlight (1.5e-3 at roughness 0.047 and value 8.6).
    return [
        (8, 8, abi.TEX_RGBA8_SRGB, img(8, 8, 4, 0)),     # 0 albedo (alpha random too)
        (8, 8, abi.TEX_RG8_UNORM, img(8, 8, 2)),         # 1 two-channel normal map
        (8, 8, abi.TEX_RG8_SRGB, img(8, 8, 2, rough=1)),  # 2 two-channel metallic-roughness
        (8, 8, abi.TEX_RGBA8_UNORM, img(8, 8, 4)),       # 3 four-channel normal map
        (8, 8, abi.TEX_RGBA8_SRGB, img(8, 8, 4, rough=1)),  # 4 four-channel metallic (G rough, B metal)
        (6, 5, abi.TEX_RGBA8_SRGB, img(6, 5, 4, rough=1)),  # 5 mixmap
        (8, 8, abi.TEX_BC7_UNORM, _bc7_blocks(rng, 4)),  # 6 BC7 normal map
        (5, 7, abi.TEX_RGBA8_SRGB, img(5, 7, 4, 40)),    # 7 albedo with alpha
    ]


def _mat(color=None, albedo=(A, A), normal=(A, A), metal=(A, A), mix=(A, A)):
    m = {
        "color_sampled_image_index": albedo[0], "color_sampler_index": albedo[1],
        "normal_sampled_image_index": normal[0], "normal_sampler_index": normal[1],
        "metallic_sampled_image_index": metal[0], "metallic_sampler_index": metal[1],
        "mixmap_sampled_image_index": mix[0], "mixmap_sampler_index": mix[1],
    }
    if color is not None:
        m["color"] = color
    return m


MATERIALS = [
    _mat(albedo=(0, 1), normal=(1, 1), metal=(2, 1)),                  # nearest clamp throughout
    _mat(albedo=(0, 2), normal=(1, 4), metal=(2, 3)),                  # linear clamp / mirrored
    _mat(albedo=(0, 3), normal=(3, 5), metal=(2, 4)),                  # four-channel normal map
    _mat(albedo=(0, 4), normal=(1, 0), metal=(4, 6)),                  # four-channel metallic map
    _mat(albedo=(7, 5), normal=(1, 2), metal=(2, 0), mix=(5, 3)),      # mixmap over metallic
    _mat(color=(0.3, 0.6, 0.9, 0.7)),                                  # no maps at all
    _mat(albedo=(0, 6), normal=(6, 2), metal=(4, 5)),                  # BC7 normal map
    _mat(color=(1.0, 1.0, 1.0, 0.5), albedo=(7, 0), normal=(1, 0), metal=(2, 0)),  # alpha sources
    _mat(albedo=(0, 5), mix=(5, 2)),                                   # mixmap, no normal map
]

WIDTH, HEIGHT = 150, 102


def sampler_scene(seed=5):
    rng = np.random.default_rng(seed + 1)
    pos, uvs, cols, tris, draws = [], [], [], [], []
    for k in range(9):
        cx, cy = k % 3, k // 3
        # cell of 2/3 NDC, quad inset by a non-round margin
        x0, x1 = -1 + cx * 2 / 3 + 0.031, -1 + (cx + 1) * 2 / 3 - 0.027
        y0, y1 = -1 + cy * 2 / 3 + 0.029, -1 + (cy + 1) * 2 / 3 - 0.033
        base = len(pos)
        pos += [(x0, y0, 0.5), (x1, y0, 0.5), (x1, y1, 0.5), (x0, y1, 0.5)]
        uvs += [(U0, U0), (U1, U0), (U1, U1), (U0, U1)]
        for _ in range(4):
            cols.append((float(rng.uniform(0.2, 1)), float(rng.uniform(0.2, 1)), float(rng.uniform(0.2, 1)),
                         float(rng.uniform(0.25, 1))))
        draws.append((len(tris) // 3, 2))
        tris += [base, base + 1, base + 2, base, base + 2, base + 3]
    sc = CustomScene(pos, tris, WIDTH, HEIGHT, draws=draws, uvs=uvs, colors=cols, textures=textures(seed),
                     materials=MATERIALS, material_of_draw=list(range(9)), samplers=SAMPLERS)
    sd = sc.s.scene_data
    sd.sun_position[:] = [6.0, -4.0, -20.0, 1.0]
    sd.shadows_enabled = 0
    return sc
