"""The output stage on the GPU (rsy_present): the RGBA16F draw image equals the shaded colour
rounded to FP16 (bit for bit), and the B8G8R8A8_SRGB blit matches the FP64 oracle within one
8-bit step (the FP32 sRGB encode can land on the other side of a rounding boundary), at the
frame's own size, half size and an odd upscale."""
import numpy as np
import pytest

import rosy_amd
from oracle import output
from rosy_amd import abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def shaded():
    scene = rosy_amd.Scene(0, seed=1)
    r = rosy_amd.Renderer(device=0)
    r.load(scene)
    r.render(scene.frame_desc(abi.SHADE_PLAIN))
    yield r, r.color(0)
    r.close()
    scene.close()


@pytest.mark.parametrize("size", [None, (320, 180), (997, 541)])
def test_present_matches_the_oracle(shaded, size):
    r, color = shaded
    h, w = color.shape[:2]
    dw, dh = size if size else (w, h)
    draw, img = r.present(0, dw, dh)
    assert np.array_equal(draw.view(np.uint16), output.draw_image(color).view(np.uint16))
    want = output.present(color, dw, dh)
    diff = np.abs(img.astype(int) - want.astype(int))
    assert diff.max() <= 1, diff.max()
    assert (diff == 0).mean() > 0.999


def test_present_errors(shaded):
    r, _ = shaded
    assert rosy_amd.lib().rsy_present(r.ctx, 0, 0, 10) == abi.RSY_INVALID_ARGUMENT  # empty extent
    assert rosy_amd.lib().rsy_present(r.ctx, 0, 20000, 10) == abi.RSY_INVALID_ARGUMENT
    assert rosy_amd.lib().rsy_present(r.ctx, 5, 10, 10) == abi.RSY_INVALID_ARGUMENT  # no such view
