"""4-sample frames on the GPU through the C ABI, against the oracle (the main pass of
Graphics.cpp:1794-1808, 5566-5575: a 4-sample colour attachment resolved by averaging).

- visibility: every sample plane bit-exact against the oracle's single-sample raster at that
  sample location; shadow maps (single-sampled) bit-exact;
- colour: the resolved colour within 1e-3 (max abs) of the FP64 oracle (each pixel's primitives
  shaded once at the pixel centre, the four samples averaged), plain and decoder shading, and with a
  blended pass (per-sample depth test and blending, then the average).  Pixels whose opaque shading
  sits on a shadow-compare tie (|ref - texel| < 1e-5) in some sample plane are left out of the bar:
  a depth compare is a step function, so FP32 and FP64 can land on either side of it;
- graph replays give the eager frame; invalid sample counts are rejected."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rosy_amd
from rosy_amd import abi
from blend_scene import blend_scene

pytestmark = pytest.mark.gpu

TOL = 1e-3
TIE = 1e-5


def msaa_state(scene, samples=4, tile=None):
    rs = abi.RasterState()
    C.memmove(C.byref(rs), C.byref(scene.s.raster_state), C.sizeof(rs))
    rs.msaa_samples = samples
    if tile is not None:
        rs.tile_size = tile
    return rs


def render(scene, shade=abi.SHADE_PLAIN, setup=None):
    r = rosy_amd.Renderer(device=0)
    r.load(scene)
    rs = msaa_state(scene)
    r.update(scene.s.scene_data, rs)
    if setup:
        setup(r)
    frame = scene.frame_desc(shade)
    r.render(frame)
    return r, rs, frame


def check_colour(r, scene, rs, neural=None, blended=False):
    cam, cas = scene.camera_views()[0], scene.cascades()
    vis = r.visibility(0)
    assert vis.shape == (4, cam.height, cam.width)
    want = oracle.raster_view_msaa(scene, cam.view_projection[:], cam.width, cam.height, rs)
    assert np.array_equal(vis, want), f"{int((vis != want).sum())} sample words differ"
    maps = []
    for ci, cs in enumerate(cas):
        g = r.shadow_map(ci)
        o, _ = oracle.raster_view(scene, cs.light_view_projection[:], cs.size, cs.size, True, rs)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
        maps.append(g)
    resolved, _, counts, margin = oracle.msaa_view(scene, cam, cas, maps, vis, raster_state=rs, neural=neural,
                                                   blended=blended, max_layers=32)
    tie = margin < TIE
    err = float(np.abs(r.color(0).astype(np.float64) - resolved)[~tie].max())
    print(f"4-sample colour max abs error {err:.3e} ({tie.mean():.4f} of pixels on a shadow tie)")
    assert tie.mean() < 0.01 and err <= TOL
    keys = vis & np.uint64(0xFFFFFFFF)
    mixed = (keys != keys[0]).any(axis=0)
    assert mixed.sum() > 0  # edge pixels hold more than one primitive (or the background)
    return vis, counts


def test_generated_scene_plain():
    s = rosy_amd.Scene(1, seed=3, width=256, height=160, shadow_size=512, total_triangles=20000, n_meshes=24,
                       texture_size=64)
    r, rs, frame = render(s)
    vis, _ = check_colour(r, s, rs)
    # the resolve is not the centre sample: the single-sample frame differs at edge pixels only
    col = r.color(0)
    r.update(s.s.scene_data, msaa_state(s, 1))
    r.render(frame)
    one = r.color(0)
    keys = vis & np.uint64(0xFFFFFFFF)
    uniform = (keys == keys[0]).all(axis=0)
    single = r.visibility(0) & np.uint64(0xFFFFFFFF)
    same = uniform & (single == keys[0])
    assert np.abs(col - one)[same].max() <= 1e-5 and np.abs(col - one)[~uniform].max() > 1e-3
    r.close()


def test_generated_scene_decoder():
    s = rosy_amd.Scene(3, seed=2, total_triangles=20000, n_meshes=24, width=320, height=192, shadow_size=512,
                       texture_size=64, n_materials=3)

    def train(r):
        r.neural_init(latent_size=32, psnr_threshold=0.0)
        r.neural_train(60)
        r.neural_evaluate()

    r, rs, _ = render(s, abi.SHADE_NEURAL, train)
    lat, _ = r.neural_params()
    f16 = lat.astype(np.float16).astype(np.float32)
    params = oracle.NeuralParams(np.arange(3), f16, r.neural_inference_mlp())
    check_colour(r, s, rs, neural=params)
    r.close()


@pytest.mark.parametrize("size", [64, 120])
def test_blended_pass_per_sample(size):
    s = blend_scene(size)
    r, rs, frame = render(s)
    _, counts = check_colour(r, s, rs, blended=True)
    assert counts.max() >= 4
    col = r.color(0)
    r.render_graph(frame, 3)
    assert np.array_equal(r.color(0), col)
    r.close()


def test_generated_scene_with_translucent_materials():
    s = rosy_amd.Scene(3, seed=3, width=320, height=180, shadow_size=512, total_triangles=16384, n_meshes=30,
                       texture_size=64, n_translucent=16)
    r, rs, frame = render(s)
    check_colour(r, s, rs, blended=True)
    col = r.color(0)
    r.render_graph(frame, 3)
    assert np.array_equal(r.color(0), col)
    r.close()


def test_sample_count_and_bin_size_are_checked():
    s = blend_scene(32)
    r = rosy_amd.Renderer(device=0)
    r.load(s)
    for bad in (msaa_state(s, 2), msaa_state(s, 8), msaa_state(s, 4, tile=64)):
        assert lib_update(r, s, bad) == abi.RSY_INVALID_ARGUMENT
    r.update(s.s.scene_data, msaa_state(s, 4, tile=32))
    r.render(s.frame_desc(abi.SHADE_PLAIN))
    out = np.zeros(32 * 32, np.uint64)  # one plane is not the whole buffer
    assert rosy_amd.lib().rsy_read_visibility(r.ctx, 0, out.ctypes.data, out.size) == abi.RSY_INVALID_ARGUMENT
    r.close()


def lib_update(r, s, rs):
    return rosy_amd.lib().rsy_update(r.ctx, C.byref(s.s.scene_data), C.byref(rs))
