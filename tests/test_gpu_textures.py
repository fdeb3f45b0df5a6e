"""The texture pipeline on the GPU through the C ABI: BC7 blocks decoded at upload bit-exact against
the oracle's decoder (every mode, the reserved mode, images whose sides are not multiples of 4,
given mip levels), and built mip chains bit-exact against the oracle's packager filter (sRGB with
premultiplied alpha, UNORM, two-channel images)."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rosy_amd
from oracle import textures as tx
from rosy_amd import abi
from blend_scene import blend_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    s = blend_scene(16)
    r = rosy_amd.Renderer(device=0)
    r.load(s)
    yield r
    r.close()


def upload(r, textures):
    """rsy_set_materials with no materials and the given (width, height, format, mip_levels, bytes)."""
    descs = (abi.TextureDesc * len(textures))()
    keep = []
    for i, (w, h, fmt, levels, data) in enumerate(textures):
        buf = C.create_string_buffer(bytes(data), len(data))
        keep.append(buf)
        descs[i] = abi.TextureDesc(w, h, fmt, levels, C.cast(buf, C.c_void_p))
    return rosy_amd.lib().rsy_set_materials(r.ctx, None, 0, descs, len(textures), None, 0)


def read(r, t, level, w, h, ch):
    lw, lh = max(1, w >> level), max(1, h >> level)
    out = np.zeros((lh, lw, ch), np.uint8)
    assert rosy_amd.lib().rsy_read_texture(r.ctx, t, level, out.ctypes.data, out.size) == abi.RSY_OK
    return out


def random_blocks(rng, n):
    b = rng.integers(0, 256, (n, 16), dtype=np.uint8)
    modes = rng.integers(0, 9, n)  # 8: leave the first byte random (mode 0 bit pattern or reserved)
    for i, m in enumerate(modes):
        if m < 8:
            b[i, 0] = (b[i, 0] & ~np.uint8((1 << (m + 1)) - 1)) | np.uint8(1 << m)
    b[0, 0] = 0  # one reserved block
    return b


def test_bc7_decode_matches_the_oracle_for_every_mode(renderer):
    rng = np.random.default_rng(5)
    w, h = 52, 35  # 13 x 9 blocks, the last row cut to 3 texels
    blocks = random_blocks(rng, 13 * 9).tobytes()
    assert upload(renderer, [(w, h, abi.TEX_BC7_SRGB, 1, blocks), (w, h, abi.TEX_BC7_UNORM, 1, blocks)]) == abi.RSY_OK
    want = oracle.bc7_decode(blocks, w, h)
    for t in range(2):
        got = read(renderer, t, 0, w, h, 4)
        assert np.array_equal(got, want), np.argwhere(got != want)[:4]


def test_bc7_given_levels_and_built_levels(renderer):
    rng = np.random.default_rng(6)
    w, h = 24, 17
    sizes = [(max(1, w >> l), max(1, h >> l)) for l in range(3)]
    levels = [random_blocks(rng, ((lw + 3) // 4) * ((lh + 3) // 4)).tobytes() for lw, lh in sizes]
    assert upload(renderer, [(w, h, abi.TEX_BC7_SRGB, 3, b"".join(levels)),
                             (w, h, abi.TEX_BC7_SRGB, abi.MIPS_BUILD, levels[0])]) == abi.RSY_OK
    for l, (lw, lh) in enumerate(sizes):
        assert np.array_equal(read(renderer, 0, l, w, h, 4), oracle.bc7_decode(levels[l], lw, lh))
    chain = tx.mip_chain(oracle.bc7_decode(levels[0], w, h), srgb=True)
    assert len(chain) == 5
    for l, want in enumerate(chain):
        assert np.array_equal(read(renderer, 1, l, w, h, 4), want), l


@pytest.mark.parametrize("fmt,ch,srgb", [(abi.TEX_RGBA8_SRGB, 4, True), (abi.TEX_RGBA8_UNORM, 4, False),
                                         (abi.TEX_RG8_SRGB, 2, True), (abi.TEX_RG8_UNORM, 2, False)])
def test_built_chains_match_the_packager_filter(renderer, fmt, ch, srgb):
    rng = np.random.default_rng(7 + fmt)
    w, h = 37, 20
    img = rng.integers(0, 256, (h, w, ch), dtype=np.uint8)
    if ch == 4:
        img[rng.random((h, w)) < 0.2, 3] = 0  # some fully transparent texels
    assert upload(renderer, [(w, h, fmt, abi.MIPS_BUILD, img.tobytes())]) == abi.RSY_OK
    chain = tx.mip_chain(img, srgb)
    for l, want in enumerate(chain):
        got = read(renderer, 0, l, w, h, ch)
        assert np.array_equal(got, want), (l, int((got != want).sum()))


def test_level_counts_are_checked(renderer):
    img = bytes(8 * 8 * 4)
    assert upload(renderer, [(8, 8, abi.TEX_RGBA8_UNORM, 5, img * 2)]) == abi.RSY_INVALID_ARGUMENT  # 8x8 has 4
    assert upload(renderer, [(8, 8, abi.TEX_RGBA8_UNORM, 4, img * 2)]) == abi.RSY_OK
    out = np.zeros(4, np.uint8)
    assert rosy_amd.lib().rsy_read_texture(renderer.ctx, 0, 4, out.ctypes.data, 4) == abi.RSY_INVALID_ARGUMENT
    assert rosy_amd.lib().rsy_read_texture(renderer.ctx, 0, 3, out.ctypes.data, 4) == abi.RSY_OK
