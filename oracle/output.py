"""CPU restatement of the output stage (test infrastructure: the oracle of rsy_present; nothing in
the product imports it).

The reference renders into an RGBA16F draw image (Graphics.cpp:1908) and blits it to the
B8G8R8A8_SRGB swapchain image with a linear filter (vkCmdBlitImage2, Graphics.cpp:5835-5880).
Restated here in FP64 from the Vulkan specification: the blit maps each destination texel centre
to the source coordinate (x + 0.5) * src / dst, filters bilinearly over texel centres at i + 0.5
with clamp-to-edge, and the destination format encodes colour with the sRGB OETF (alpha linear),
rounding to 8 bits.
"""
import numpy as np


def draw_image(color):
    """RGBA32F colour -> RGBA16F draw image, round to nearest even."""
    return np.asarray(color, np.float32).astype(np.float16)


def srgb_encode(c):
    """Linear -> sRGB (the sRGB OETF), then 8-bit rounding (round half to even)."""
    c = np.clip(np.asarray(c, np.float64), 0.0, 1.0)
    s = np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(c, 1.0 / 2.4) - 0.055)
    return np.rint(s * 255.0).astype(np.uint8)


def blit(draw, width, height):
    """Linear-filtered blit of an (h, w, 4) image to (height, width, 4), FP64."""
    src = np.asarray(draw, np.float64)
    sh, sw = src.shape[:2]
    u = (np.arange(width) + 0.5) * (sw / width) - 0.5
    v = (np.arange(height) + 0.5) * (sh / height) - 0.5
    fu, fv = np.floor(u), np.floor(v)
    au, av = (u - fu)[None, :, None], (v - fv)[:, None, None]
    x0 = np.clip(fu.astype(np.int64), 0, sw - 1)
    x1 = np.clip(fu.astype(np.int64) + 1, 0, sw - 1)
    y0 = np.clip(fv.astype(np.int64), 0, sh - 1)
    y1 = np.clip(fv.astype(np.int64) + 1, 0, sh - 1)
    c00, c10 = src[y0][:, x0], src[y0][:, x1]
    c01, c11 = src[y1][:, x0], src[y1][:, x1]
    return (1 - au) * (1 - av) * c00 + au * (1 - av) * c10 + (1 - au) * av * c01 + au * av * c11


def present(color, width, height):
    """The B8G8R8A8_SRGB image (height, width, 4) of a colour buffer: blue, green, red, alpha."""
    rgba = blit(draw_image(color), width, height)
    out = np.empty((height, width, 4), np.uint8)
    out[..., 0] = srgb_encode(rgba[..., 2])
    out[..., 1] = srgb_encode(rgba[..., 1])
    out[..., 2] = srgb_encode(rgba[..., 0])
    out[..., 3] = np.rint(np.clip(rgba[..., 3], 0.0, 1.0) * 255.0).astype(np.uint8)
    return out
