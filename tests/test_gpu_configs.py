"""The batched, learned-decoder and combined workloads at their stated sizes, through the C ABI.

- Batched views: 1,048,576 triangles (log-uniform 64-65536 per mesh), one 2560x1440 camera view and
  four 2048^2 cascades with splits 5 / 15 / 40 / 100, as one frame: every view and cascade
  bit-exact, camera colour within 1e-3 of the FP64 oracle.
- Learned decoder: 262,144 triangles at 1920x1080, 64 materials of 1024^2 maps replaced by 256^2
  latent grids and the MLP (a short training run); the neural frame bit-exact in visibility and
  within 1e-3 of the FP64 oracle shading with the same decoder parameters.
- Combined: 2,097,152 triangles (64-131072 per mesh), four camera views of uneven size
  (3840x2160, 1920x1080, 1280x720, 640x360) and four 2048^2 cascades in one frame, the decoder with
  FP8 (E4M3) weights and FP32 accumulation: every view and cascade bit-exact, colour within 1e-3
  in every camera view.
The colour gate follows the plain parity tests: the GPU's PCF factor may differ from the oracle's
only at a shadow-compare tie, and the oracle is then re-run with the GPU's factor."""
import numpy as np
import pytest

import oracle
import rosy_amd
from rosy_amd import abi
from test_gpu_neural import inference_params
from test_gpu_parity import TIE, TOL, check_frame

pytestmark = pytest.mark.gpu


def check_views(r, scene, neural=None):
    """Every cascade and camera view against the oracle; colour of every camera view (with the
    decoder parameters `neural` when given).  Returns the colour errors per view."""
    r.set_debug_flags(abi.DEBUG_SHADOW_FACTOR)
    r.render(scene.frame_desc(abi.SHADE_NEURAL if neural is not None else abi.SHADE_PLAIN))
    rs = scene.s.raster_state
    maps = []
    for ci, cs in enumerate(scene.cascades()):
        g = r.shadow_map(ci)
        o, _ = oracle.raster_view(scene, cs.light_view_projection[:], cs.size, cs.size, True, rs)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), f"cascade {ci}: {(g != o).sum()} texels differ"
        maps.append(g)
    errs = []
    for vi, cv in enumerate(scene.camera_views()):
        g = r.visibility(vi)
        o, _ = oracle.raster_view(scene, cv.view_projection[:], cv.width, cv.height, False, rs)
        assert np.array_equal(g, o), f"view {vi}: {(g != o).sum()} visibility words differ"
        sf = r.shadow_factor(vi)
        ref, rsf, margin = oracle.shade_view(scene, cv, scene.cascades(), maps, g, neural=neural)
        flip = np.abs(sf - rsf) > 2e-3
        assert (margin[flip] < TIE).all(), f"view {vi}: shadow factor differs away from a compare tie"
        ref, _, _ = oracle.shade_view(scene, cv, scene.cascades(), maps, g, neural=neural, override_shadow=sf)
        err = float(np.abs(r.color(vi).astype(np.float64) - ref).max())
        assert err <= TOL, f"view {vi} ({cv.width}x{cv.height}): colour max abs error {err:.3e}"
        errs.append(err)
    return errs


def test_batched_views_at_full_size():
    s = rosy_amd.Scene(2, seed=1)
    r = rosy_amd.Renderer(device=0)
    try:
        assert s.n_triangles == 1048576
        cv = s.camera_views()
        assert len(cv) == 1 and (cv[0].width, cv[0].height) == (2560, 1440)
        cs = s.cascades()
        assert [c.size for c in cs] == [2048] * 4 and [c.split_far for c in cs] == [5.0, 15.0, 40.0, 100.0]
        r.load(s)
        errs = check_frame(r, s)
        st = r.stats()
        assert st.triangle_count == 5 * 1048576 and st.overflowed == 0
        print(f"batched views: colour max abs err {errs}, {st.tile_pairs} tile pairs, bin {st.tile_size}")
    finally:
        r.close()
        s.close()


def test_learned_decoder_workload_at_full_size():
    s = rosy_amd.Scene(3, seed=1)
    r = rosy_amd.Renderer(device=0)
    try:
        assert s.n_triangles == 262144 and s.s.n_materials == 64
        assert s.s.textures[s.s.materials[0].color_sampled_image_index].width == 1024
        cv = s.camera_views()[0]
        assert (cv.width, cv.height) == (1920, 1080)
        r.load(s)
        r.neural_init(latent_size=256, psnr_threshold=0.0)
        r.neural_train(300)
        psnr = r.neural_evaluate()
        st = r.neural_stats()
        assert st.n_eligible == 64 and st.n_active == 64
        errs = check_views(r, s, neural=inference_params(r))
        print(f"decoder workload: PSNR min/median {psnr.min():.2f}/{np.median(psnr):.2f} dB, colour err {errs}")
    finally:
        r.close()
        s.close()


def test_combined_workload_at_full_size_with_fp8_decoder():
    s = rosy_amd.Scene(4, seed=1)
    r = rosy_amd.Renderer(device=0)
    try:
        assert s.n_triangles == 2097152
        sizes = [(v.width, v.height) for v in s.camera_views()]
        assert sizes == [(3840, 2160), (1920, 1080), (1280, 720), (640, 360)]
        assert [c.size for c in s.cascades()] == [2048] * 4
        r.load(s)
        r.neural_init(latent_size=256, weights=abi.NEURAL_WEIGHTS_FP8)
        r.neural_train(300)
        psnr = r.neural_evaluate()
        st = r.neural_stats()
        # materials below 40 dB fall back to the plain maps; the oracle follows the same active set
        active = (psnr >= 40.0).astype(np.uint32)
        assert st.n_active == int(active.sum())
        errs = check_views(r, s, neural=inference_params(r, active=active))
        print(f"combined workload (FP8 decoder): PSNR min/median {psnr.min():.2f}/{np.median(psnr):.2f} dB, "
              f"{st.n_active}/64 decoded, colour err per view {errs}")
    finally:
        r.close()
        s.close()
