"""Scenes for the texture level of detail: an unlit full-view quad textured with a mip-mapped albedo,
so the shaded colour is the sampled albedo, with closed forms for its texture-coordinate
derivatives."""
import numpy as np

from rosy_amd import abi
from scenes import CustomScene

LEVEL_RED = [0, 40, 80, 120, 160, 200, 240]  # red of each level of the level-coded 64 x 64 texture

# clip = (x, y, 0.1 z, z): a camera looking down +z (w = z, constant reverse-Z depth 0.1)
PERSPECTIVE = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0.1, 1, 0, 0, 0, 0]


def level_coded_texture():
    """64 x 64 UNORM RGBA8 with all 7 levels given, level l the constant (LEVEL_RED[l], 0, 0, 255)."""
    data = b""
    for l, red in enumerate(LEVEL_RED):
        n = max(1, 64 >> l)
        data += np.tile(np.array([red, 0, 0, 255], np.uint8), n * n).tobytes()
    return (64, 64, abi.TEX_RGBA8_UNORM, data, 7)


def _scene(pos, uvs, size, texture, vp=None):
    s = CustomScene(pos, [0, 1, 2, 0, 2, 3], size, size, uvs=uvs, view_projection=vp,
                    textures=[texture, (1, 1, abi.TEX_RG8_UNORM, bytes([128, 128])),
                              (1, 1, abi.TEX_RG8_SRGB, bytes([0, 230]))])
    sd = s.s.scene_data
    sd.light_enabled = 0
    sd.shadows_enabled = 0
    return s


def affine_quad(size, su, sv, texture):
    """The view covered by a quad (w = 1) with uv from (0, 0) to (su, sv): du/dx = su / size and
    dv/dy = sv / size on every pixel."""
    pos = [(-1.0, -1.0, 0.5), (1.0, -1.0, 0.5), (1.0, 1.0, 0.5), (-1.0, 1.0, 0.5)]
    return _scene(pos, [(0.0, 0.0), (su, 0.0), (su, sv), (0.0, sv)], size, texture)


def receding_quad(size, s, texture):
    """A plane from depth 1 (bottom of the view) to depth 3 (top) under PERSPECTIVE, filling the view,
    uv from (0, 0) to (s, s) at its corners (a trapezoid in object space, two triangles)."""
    pos = [(-1.0, -1.0, 1.0), (1.0, -1.0, 1.0), (3.0, 3.0, 3.0), (-3.0, 3.0, 3.0)]
    return _scene(pos, [(0.0, 0.0), (s, 0.0), (s, s), (0.0, s)], size, texture, vp=PERSPECTIVE)


def receding_footprint(size, s, px, py):
    """Level-0 texels per pixel (64 x 64 texture) of receding_quad at pixel (px, py), in closed form.
    At NDC (x, y) the plane's depth is z = 3 / (2 - y) and its object point (x z, y z); uv is affine in
    the object point on each triangle: u = s/2 X - s/4 Y + s/4 below the diagonal x = y (vertices 0,
    1, 2), u = s/6 X + s/12 Y + s/4 above it (vertices 0, 2, 3), and v = s/4 (Y + 1) on both."""
    x = (px + 0.5) * 2.0 / size - 1.0
    y = (py + 0.5) * 2.0 / size - 1.0
    z = 3.0 / (2.0 - y)
    dz = 3.0 / (2.0 - y) ** 2  # dz/dy
    a, b = (s / 2, -s / 4) if x > y else (s / 6, s / 12)
    du_dx, du_dy = a * z, (a * x + b * y) * dz + b * z
    dv_dx, dv_dy = 0.0, 1.5 * s / (2.0 - y) ** 2
    k = 2.0 / size * 64  # NDC per pixel, times level-0 texels per uv unit
    return max(np.hypot(du_dx * k, dv_dx * k), np.hypot(du_dy * k, dv_dy * k))


def expected_level_red(rho):
    """Red of the level-coded texture sampled at footprint rho (level-0 texels per pixel)."""
    lam = np.log2(rho)
    if not lam > 0:
        return LEVEL_RED[0]
    d = min(lam, len(LEVEL_RED) - 1)
    lo = int(np.floor(d))
    hi = min(lo + 1, len(LEVEL_RED) - 1)
    f = d - lo
    return LEVEL_RED[lo] + (LEVEL_RED[hi] - LEVEL_RED[lo]) * f
