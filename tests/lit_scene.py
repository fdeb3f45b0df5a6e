"""A flat lit quad whose shading at one pixel has a closed form (basic.slang:301-331,
lighting.slang:5-43), shared by the oracle pins and the GPU test.

The quad lies at z = 0.5 under an identity view-projection in a 32x32 view, with vertices on the
sub-pixel grid (x = X / 4096 - 1 is exact in FP32), so the barycentrics at a pixel centre are exact
and the interpolated world position there is the pixel centre itself.  Pixel (15, 15) sits at
P0 = (-1/32, -1/32, 0.5).  The camera and the sun are placed relative to P0 so that V and L have
chosen directions there; the normal is the vertex normal (0, 0, -1) (no normal map).
"""
import math

from rosy_amd import abi
from scenes import CustomScene

PIXEL = (15, 15)                 # (x, y)
P0 = (-1.0 / 32.0, -1.0 / 32.0, 0.5)
EDGE = 13.0 / 16.0               # 0.8125: sub-pixel 768 / 7424 in a 32-pixel view
ALBEDO_SRGB = (200, 150, 100)
# exact in FP32, the scene data's storage type
SUN_COLOR = (1.0, 0.875, 0.75)
INTENSITY = 2.5
AMBIENT = (0.0078125, 0.015625, 0.03125)


def srgb_eotf(code):
    """IEC 61966-2-1 sRGB decoding of an 8-bit code."""
    v = code / 255.0
    return v / 12.92 if v <= 0.04045 else ((v + 0.055) / 1.055) ** 2.4


def lit_scene(metal_code, rough_code, sun_offset, camera_offset=(0.0, 0.0, -10.0)):
    """The quad with a 1x1 sRGB albedo and a 1x1 UNORM metallic-roughness map (R = metallic,
    G = roughness); the sun at P0 + sun_offset, the camera at P0 + camera_offset; no shadows."""
    e = EDGE
    pos = [(-e, -e, 0.5), (e, -e, 0.5), (e, e, 0.5), (-e, e, 0.5)]
    textures = [
        (1, 1, abi.TEX_RGBA8_SRGB, bytes([*ALBEDO_SRGB, 255])),
        (1, 1, abi.TEX_RG8_UNORM, bytes([128, 128])),
        (1, 1, abi.TEX_RG8_UNORM, bytes([metal_code, rough_code])),
    ]
    sc = CustomScene(pos, [0, 1, 2, 0, 2, 3], 32, 32, textures=textures,
                     materials=[{"normal_sampled_image_index": abi.RSY_ABSENT}])
    sd = sc.s.scene_data
    sd.sun_position[:] = [P0[0] + sun_offset[0], P0[1] + sun_offset[1], P0[2] + sun_offset[2], 1.0]
    sd.camera_position[:] = [P0[0] + camera_offset[0], P0[1] + camera_offset[1], P0[2] + camera_offset[2], 0.0]
    sd.sunlight_color[:] = [*SUN_COLOR, 1.0]
    sd.light_intensity = INTENSITY
    sd.ambient_color[:] = [*AMBIENT, 1.0]
    sd.shadows_enabled = 0
    return sc


def closed_form(metal, rough, n_dot_l, n_dot_v, n_dot_h, v_dot_h, intensity=INTENSITY):
    """fragmentMain's BRDF branch at given cosines: GGX D = a^2 / (pi (NdotH^2 (a^2 - 1) + 1)^2) with
    a = r^2 and pi = 3.14159265359; Smith G = G1(NdotL) G1(NdotV), G1(x) = x / (x (1 - k) + k),
    k = (r + 1)^2 / 8; Schlick F = f0 + (1 - f0)(1 - VdotH)^5, f0 = lerp(0.04, albedo, metal);
    colour = max(ambient * diffuse, ((1 - F)(1 - metal) diffuse + D G F / (4 NdotV NdotL + 1e-4))
    * sun * intensity * NdotL) with diffuse = albedo / 3.14159265."""
    albedo = [srgb_eotf(c) for c in ALBEDO_SRGB]
    a = rough * rough
    a2 = a * a
    d = a2 / (3.14159265359 * (n_dot_h * n_dot_h * (a2 - 1.0) + 1.0) ** 2)
    k = (rough + 1.0) ** 2 / 8.0
    g = (n_dot_l / (n_dot_l * (1 - k) + k)) * (n_dot_v / (n_dot_v * (1 - k) + k))
    out = []
    for ch in range(3):
        f0 = 0.04 + (albedo[ch] - 0.04) * metal
        f = f0 + (1.0 - f0) * (1.0 - v_dot_h) ** 5
        kd = (1.0 - f) * (1.0 - metal)
        spec = d * g * f / (4.0 * n_dot_v * n_dot_l + 0.0001)
        diffuse = albedo[ch] / 3.14159265
        brdf = (kd * diffuse + spec) * SUN_COLOR[ch] * intensity * n_dot_l
        out.append(max(AMBIENT[ch] * diffuse, brdf))
    return out


# The two geometries at pixel (15, 15), with N = (0, 0, -1) and the camera straight in front:
#   normal incidence: L = V = N, so H = N and every cosine is 1;
#   oblique: L = (3, 0, -4) / 5 (sun offset (12, 0, -16)), so NdotL = 4/5 and H = (1, 0, -3) / sqrt(10):
#   NdotH = VdotH = 3 / sqrt(10).
GEOMETRIES = {
    "normal": ((0.0, 0.0, -20.0), dict(n_dot_l=1.0, n_dot_v=1.0, n_dot_h=1.0, v_dot_h=1.0)),
    "oblique": ((12.0, 0.0, -16.0), dict(n_dot_l=0.8, n_dot_v=1.0, n_dot_h=3 / math.sqrt(10), v_dot_h=3 / math.sqrt(10))),
}
METALS = [0, 255]               # metallic 0 and 1
ROUGHS = [77, 128, 255]         # roughness 77/255 (~0.3), 128/255 (~0.5), 1
