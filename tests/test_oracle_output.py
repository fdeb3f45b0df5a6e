"""Pins of the output-stage oracle (oracle/output.py): the sRGB encode at known values, the blit's
closed forms (identity, 2x2 box at half size, constant images, clamp-to-edge) and the FP16
rounding of the draw image.  Host only."""
import numpy as np

from oracle import output


def test_srgb_encode_known_values():
    lin = np.array([0.0, 1.0, 0.5, 0.18, 0.0031308, 0.001, 2.0, -1.0])
    # 0.5 linear is sRGB 187.5 -> 188; 18 % grey is sRGB 118; below the knee the slope is 12.92
    assert list(output.srgb_encode(lin)) == [0, 255, 188, 118, 10, 3, 255, 0]
    # monotone over the whole 16-bit range of inputs
    ramp = output.srgb_encode(np.linspace(0, 1, 65537))
    assert np.all(np.diff(ramp.astype(int)) >= 0) and ramp[0] == 0 and ramp[-1] == 255


def test_blit_closed_forms():
    rng = np.random.default_rng(1)
    img = rng.random((6, 8, 4))
    assert np.array_equal(output.blit(img, 8, 6), img)  # same size: texel centres map to texel centres
    half = output.blit(img, 4, 3)
    box = img.reshape(3, 2, 4, 2, 4).mean(axis=(1, 3))
    assert np.allclose(half, box, atol=1e-15)
    const = np.full((5, 7, 4), 0.3)
    assert np.allclose(output.blit(const, 13, 2), 0.3)
    one = rng.random((1, 1, 4))
    assert np.allclose(output.blit(one, 9, 4), one)  # clamp-to-edge everywhere
    # a horizontal ramp upsampled 2x: interior samples sit a quarter texel from the centres
    ramp = np.tile(np.arange(4, dtype=np.float64)[None, :, None], (1, 1, 4))
    up = output.blit(ramp, 8, 1)[0, :, 0]
    assert np.allclose(up, [0.0, 0.25, 0.75, 1.25, 1.75, 2.25, 2.75, 3.0])


def test_draw_image_rounds_to_nearest_even():
    x = np.array([1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 65504.0, 1e-8, 0.1], np.float32)
    h = output.draw_image(x)
    assert list(h.astype(np.float64)) == [1.0, 1.0 + 2 * 2.0 ** -10, 65504.0, 0.0, np.float64(np.float16(0.1))]


def test_present_channel_order():
    c = np.zeros((1, 1, 4), np.float32)
    c[0, 0] = (1.0, 0.5, 0.0, 0.5)
    p = output.present(c, 1, 1)
    assert list(p[0, 0]) == [0, 188, 255, 128]  # B, G, R, A
