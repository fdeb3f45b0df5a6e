"""Pins for the oracle's mip chains (oracle/textures.py: the packager's filter of
Packager.cpp:216-308 restated): closed forms of the 2 x 2 box on UNORM and sRGB images, the effect
of premultiplied alpha, chain lengths down to 1 x 1, and constant images staying constant."""
import numpy as np

from oracle import textures as tx


def test_chain_lengths_and_extents():
    img = np.zeros((20, 37, 4), np.uint8)
    chain = tx.mip_chain(img, srgb=False)
    assert [c.shape[:2] for c in chain] == [(20, 37), (10, 18), (5, 9), (2, 4), (1, 2), (1, 1)]
    assert len(tx.mip_chain(np.zeros((1, 1, 2), np.uint8), True)) == 1


def test_unorm_box_is_the_rounded_mean():
    img = np.array([[[10, 0], [20, 255]], [[30, 0], [41, 255]]], np.uint8)
    got = tx.next_level(img, srgb=False)
    assert got.shape == (1, 1, 2)
    assert list(got[0, 0]) == [int(np.floor((10 + 20 + 30 + 41) / 4 + 0.5)), 128]  # 25.25 -> 25, 127.5 -> 128


def test_srgb_box_averages_linear_light():
    # black and white average to linear 0.5, which is sRGB 188 (not 128)
    img = np.array([[[0, 0, 0, 255], [255, 255, 255, 255]], [[0, 0, 0, 255], [255, 255, 255, 255]]], np.uint8)
    got = tx.next_level(img, srgb=True)
    assert list(got[0, 0]) == [188, 188, 188, 255]
    assert abs(1.055 * 0.5 ** (1 / 2.4) - 0.055 - 187.5 / 255) < 0.5 / 255


def test_premultiplied_alpha_ignores_transparent_colour():
    # a fully transparent texel contributes nothing to the colour, only to the coverage
    img = np.array([[[255, 0, 0, 255], [0, 255, 0, 0]], [[255, 0, 0, 255], [0, 0, 255, 0]]], np.uint8)
    got = tx.next_level(img, srgb=True)
    assert list(got[0, 0]) == [255, 0, 0, 128]


def test_constant_images_stay_constant_and_srgb_round_trips():
    for srgb in (False, True):
        for ch in (2, 4):
            img = np.full((16, 8, ch), 77, np.uint8)
            if ch == 4:
                img[..., 3] = 255
            for level in tx.mip_chain(img, srgb)[1:]:
                assert (level == img[0, 0]).all()
    v = np.arange(256) / 255.0
    assert np.array_equal(tx.to8(tx.srgb_encode(tx.srgb_decode(v))), np.arange(256))
