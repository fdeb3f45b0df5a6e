"""A hand-built scene for the blended pass: an opaque quad and six blended quads (two BLEND
materials, one alpha MASK material with a texture whose alpha varies) at different depths, some
behind the opaque quad, overlapping up to six deep in draw order."""
import numpy as np

from rosy_amd import abi
from scenes import CustomScene

# (x0, x1, y0, y1, depth, material, blended); identity view-projection: depth is the reverse-Z value
QUADS = [
    (-0.8, 0.8, -0.8, 0.8, 0.30, 0, False),   # opaque
    (-0.9, 0.2, -0.5, 0.9, 0.50, 1, True),    # BLEND, alpha 0.5
    (-0.3, 0.9, -0.9, 0.4, 0.60, 2, True),    # MASK, cutoff 0.6, alpha from the texture
    (-0.95, 0.95, -0.2, 0.2, 0.10, 3, True),  # BLEND behind the opaque quad: only outside it
    (-0.2, 0.5, -0.7, 0.7, 0.70, 1, True),
    (-0.5, 0.35, -0.35, 0.3, 0.40, 3, True),
    (-0.1, 0.25, -0.25, 0.25, 0.55, 1, True),
]


def blend_scene(size=64, light=True, quads=QUADS):
    pos, tri, draws, mats_of, blended = [], [], [], [], []
    for k, (x0, x1, y0, y1, z, m, b) in enumerate(quads):
        base = len(pos)
        pos += [(x0, y0, z), (x1, y0, z), (x1, y1, z), (x0, y1, z)]
        tri += [(base, base + 1, base + 2), (base, base + 2, base + 3)]
        draws.append((2 * k, 2))
        mats_of.append(m)
        blended.append(b)
    textures = [
        (1, 1, abi.TEX_RGBA8_SRGB, bytes([200, 150, 100, 255])),
        (1, 1, abi.TEX_RG8_UNORM, bytes([128, 128])),
        (1, 1, abi.TEX_RG8_SRGB, bytes([0, 230])),
        # 2x2 albedo whose alpha runs 0.3 .. 0.9 (bilinear, REPEAT): the mask cuts part of its quad
        (2, 2, abi.TEX_RGBA8_SRGB, bytes([40, 200, 90, 76, 40, 200, 90, 230, 220, 60, 60, 230, 220, 60, 60, 76])),
        (1, 1, abi.TEX_RGBA8_SRGB, bytes([30, 90, 240, 255])),
    ]
    materials = [
        {},
        {"color": (1.0, 1.0, 1.0, 0.5), "alpha_mode": 2},
        {"alpha_mode": 1, "alpha_cutoff": 0.6, "color_sampled_image_index": 3},
        {"color": (1.0, 1.0, 1.0, 0.7), "alpha_mode": 2, "color_sampled_image_index": 4},
    ]
    s = CustomScene(pos, tri, size, size, draws=draws, textures=textures, materials=materials,
                    material_of_draw=mats_of, blended=blended)
    sd = s.s.scene_data
    sd.shadows_enabled = 0
    sd.light_enabled = 1 if light else 0
    return s


def depth_at(px, py, size, quads=QUADS):
    """The draws (in order) whose quads cover pixel centre (px, py) of a size x size view."""
    x = (px + 0.5) * 2.0 / size - 1.0
    y = (py + 0.5) * 2.0 / size - 1.0
    return [k for k, (x0, x1, y0, y1, z, m, b) in enumerate(quads) if x0 < x < x1 and y0 < y < y1]
