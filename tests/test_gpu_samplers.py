"""GPU parity for the reference's sampler and material variants, and lit pixels against their closed
form.

The sampler scene reaches nearest / linear filtering with clamp-to-edge, mirrored repeat and repeat
addressing, two- and four-channel and BC7 normal maps, two- and four-channel metallic maps, mixmaps,
a material with no maps and every alpha source (Graphics.cpp:250-278, 3632-3647; basic.slang:144-221).
Visibility is bit-exact and colour within 1e-3 of the FP64 oracle, lit and in the debug outputs that
show the sampled roughness, metallic and alpha directly.  The lit quad's pixel is compared with the
closed-form Cook-Torrance value of basic.slang:301-331 at 1e-3."""
import numpy as np
import pytest

import oracle
import rosy_amd
from rosy_amd import abi
from lit_scene import GEOMETRIES, METALS, PIXEL, ROUGHS, closed_form, lit_scene
from sampler_scene import sampler_scene
from test_gpu_parity import check_frame
from test_oracle_shading import scene_data

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    r = rosy_amd.Renderer(device=0)
    yield r
    r.close()


def test_sampler_and_material_variants(renderer):
    sc = sampler_scene()
    renderer.load(sc)
    errs = check_frame(renderer, sc)
    print("sampler scene lit colour max abs err", errs)
    # the sampled values themselves: albedo (0), roughness (13), metallic (14), 1 - alpha (15)
    for mode in (0, 13, 14, 15, 6):
        sd = scene_data(sc, fragment_output=mode, light_enabled=0)
        renderer.update(sd, sc.s.raster_state)
        e = check_frame(renderer, sc, sd=sd)
        print(f"fragment_output {mode}: max abs err {e}")


@pytest.mark.parametrize("geometry", list(GEOMETRIES))
def test_lit_pixels_match_the_closed_form(renderer, geometry):
    sun, cosines = GEOMETRIES[geometry]
    worst = 0.0
    for metal in METALS:
        for rough in ROUGHS:
            sc = lit_scene(metal, rough, sun)
            renderer.load(sc)
            renderer.render(sc.frame_desc(abi.SHADE_PLAIN))
            x, y = PIXEL
            assert renderer.visibility(0)[y, x] != 0
            got = renderer.color(0)[y, x, :3].astype(np.float64)
            want = np.array(closed_form(metal / 255.0, rough / 255.0, **cosines))
            err = float(np.abs(got - want).max())
            assert err <= 1e-3, (metal, rough, got, want)
            assert np.allclose(got, want, rtol=2e-5, atol=0), (metal, rough, got, want)
            worst = max(worst, err)
            sc.close()
    print(f"{geometry}: worst abs err vs closed form {worst:.2e}")
