"""CPU restatement of the texture pipeline's mip chains -- test infrastructure: the oracle the
library's mip builder (launch_build_mip, rosy_amd/csrc/bc7.hip) is checked against; nothing in the
product imports it.

The reference's packager builds each image's chain with NVTT (Packager.cpp:216-308): for sRGB
images, toLinearFromSrgb, premultiplyAlpha, buildNextMipmap(MipmapFilter_Box), demultiplyAlpha,
toSrgb; for normal maps buildNextMipmap(MipmapFilter_Box) on the stored values.  NVTT is closed and
is not used: the box filter is restated as the mean of the 2 x 2 texels above each texel (the edge
texel repeated on an axis one texel long, an odd level's last row or column unread), in FP64, with
the sRGB transfer functions of the Vulkan data-format specification and rounding half up to 8 bits.
"""
import numpy as np


def srgb_decode(v):
    v = np.asarray(v, np.float64)
    return np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4)


def srgb_encode(x):
    x = np.asarray(x, np.float64)
    return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(np.maximum(x, 0.0), 1.0 / 2.4) - 0.055)


def to8(x):
    return np.floor(np.clip(x, 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8)


def next_level(img, srgb):
    """Level l + 1 of an [h, w, ch] uint8 level l."""
    h, w, ch = img.shape
    dh, dw = max(1, h >> 1), max(1, w >> 1)
    ys = np.minimum(np.stack([2 * np.arange(dh), 2 * np.arange(dh) + 1]), h - 1)
    xs = np.minimum(np.stack([2 * np.arange(dw), 2 * np.arange(dw) + 1]), w - 1)
    v = img.astype(np.float64) / 255.0
    quad = [v[ys[j]][:, xs[k]] for j in range(2) for k in range(2)]  # four [dh, dw, ch]
    if srgb and ch == 4:
        a = [q[..., 3] for q in quad]
        col = sum(srgb_decode(q[..., :3]) * qa[..., None] for q, qa in zip(quad, a)) * 0.25
        alpha = sum(a) * 0.25
        col = np.where(alpha[..., None] > 0.0, col / np.where(alpha > 0.0, alpha, 1.0)[..., None], col)
        return np.concatenate([to8(srgb_encode(col)), to8(alpha)[..., None]], axis=-1)
    mean = sum(srgb_decode(q) if srgb else q for q in quad) * 0.25
    return to8(srgb_encode(mean)) if srgb else to8(mean)


def mip_chain(img, srgb):
    """All levels down to 1 x 1, level 0 first."""
    levels = [np.asarray(img, np.uint8)]
    while levels[-1].shape[0] > 1 or levels[-1].shape[1] > 1:
        levels.append(next_level(levels[-1], srgb))
    return levels
