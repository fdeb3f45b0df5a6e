"""GPU tests of the neural material decoder through the C ABI, against the FP64 oracle.

- training targets: the plain sampler's values (FP32 on the GPU vs FP64), within 1e-5;
- training gradients: the GPU's exact loss and gradients of an explicit batch (the training kernels
  over the whole grid) vs the FP64 oracle's, componentwise within 2e-4 of the largest component;
- inference: decoder outputs (FP16 latents and weights, activations as FP16 hi/lo pairs into the
  MFMA, FP32 accumulation) vs the FP64 decoder of the same FP16 parameters, within DECODE_TOL;
- shading: visibility bit-exact; colour vs the FP64 oracle shading with the same decoder, within
  NEURAL_TOL, the plain path's bar;
- the PSNR gate and the plain-map fallback; and an ablation: training beats the decoder's own
  starting point, the bilinear interpolation of the maps' values at the latent cell centres."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rosy_amd
from rosy_amd import abi

pytestmark = pytest.mark.gpu

DECODE_TOL = 2e-5
NEURAL_TOL = 1e-3
TIE = 1e-5


@pytest.fixture(scope="module")
def scene():
    s = rosy_amd.Scene(3, seed=2, total_triangles=20000, n_meshes=24, width=320, height=192, shadow_size=512,
                       texture_size=64, n_materials=3)
    yield s
    s.close()


@pytest.fixture(scope="module")
def renderer(scene):
    r = rosy_amd.Renderer(device=0)
    r.load(scene)
    yield r
    r.close()


def f16(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def inference_params(r, active=None):
    """Oracle parameters equal to what the GPU decodes with: FP16 latents and the inference MLP
    (FP16 or scaled FP8 W1 / b1 / W2, FP32 b2)."""
    lat, _ = r.neural_params()
    n = r.scene.s.n_materials
    return oracle.NeuralParams(np.arange(n), f16(lat), r.neural_inference_mlp(), active)


def e4m3_table():
    """All finite non-negative OCP FP8 E4M3 values with their codes (bias 7, max 448)."""
    vals, codes = [], []
    for code in range(0x7f):  # 0x7f is NaN
        e, m = code >> 3, code & 7
        vals.append(m * 2.0 ** -9 if e == 0 else (1 + m / 8) * 2.0 ** (e - 7))
        codes.append(code)
    return np.array(vals), np.array(codes)


def e4m3_round(x):
    """Round to nearest E4M3 value, ties to the even code."""
    vals, codes = e4m3_table()
    a = np.abs(np.asarray(x, np.float64))
    idx = np.searchsorted(vals, a)
    lo = np.clip(idx - 1, 0, len(vals) - 1)
    hi = np.clip(idx, 0, len(vals) - 1)
    dl, dh = a - vals[lo], vals[hi] - a
    pick = np.where(dl < dh, lo, np.where(dh < dl, hi, np.where(codes[lo] % 2 == 0, lo, hi)))
    return np.sign(x) * vals[pick]


def reference_inference_mlp(mlp, fp8):
    """The inference MLP restated independently: FP16 rounding, or per-layer E4M3 with scale max/448."""
    mlp = np.asarray(mlp, np.float32)
    out = mlp.astype(np.float64).copy()
    layers = [np.concatenate([mlp[0:320], mlp[320:352]]), mlp[352:576]]
    spans = [(0, 352), (352, 576)]
    for w, (a, b) in zip(layers, spans):
        if fp8:
            scale = np.float32(np.abs(w).max()) / np.float32(448.0)
            q = e4m3_round((mlp[a:b] / scale).astype(np.float32))
            out[a:b] = (q.astype(np.float32) * scale).astype(np.float32)
        else:
            out[a:b] = f16(mlp[a:b])
    return out.astype(np.float32)


def maps(scene, m):
    md = scene.s.materials[m]
    t = scene.s.textures
    return t[md.color_sampled_image_index], t[md.normal_sampled_image_index], t[md.metallic_sampled_image_index]


def plain_targets(scene, material, uv):
    out = np.zeros((len(material), 7))
    for i, (m, (u, v)) in enumerate(zip(material, uv)):
        a, n, mr = maps(scene, int(m))
        out[i, 0:3] = oracle.sample_texture(a, None, float(u), float(v))[:3]
        out[i, 3:5] = oracle.sample_texture(n, None, float(u), float(v))[:2]
        out[i, 5:7] = oracle.sample_texture(mr, None, float(u), float(v))[:2]
    return out


def test_every_generated_material_is_eligible(renderer, scene):
    renderer.neural_init(latent_size=16)
    st = renderer.neural_stats()
    assert st.n_materials == 3 and st.n_eligible == 3 and st.n_active == 0


def test_training_targets_are_the_plain_sampler(renderer, scene):
    renderer.neural_init(latent_size=16)
    rng = np.random.default_rng(1)
    material = rng.integers(0, 3, 600)
    uv = rng.uniform(-1.5, 2.5, (600, 2)).astype(np.float32)
    got = renderer.neural_decode(material, uv, targets=True)
    want = plain_targets(scene, material, uv)
    err = np.abs(got - want).max()
    assert err < 1e-5, err


def test_gradients_match_the_fp64_oracle(renderer, scene):
    L = 16
    renderer.neural_init(latent_size=L, init=abi.NEURAL_INIT_RANDOM, seed=5)
    rng = np.random.default_rng(7)
    lat = rng.uniform(-0.8, 0.8, (3, L, L, 8)).astype(np.float32)
    mlp = rng.uniform(-0.5, 0.5, 583).astype(np.float32)
    renderer.neural_set_params(lat, mlp)
    n = 4000
    material = rng.integers(0, 3, n)
    uv = rng.uniform(-1, 2, (n, 2)).astype(np.float32)
    frac = np.modf(uv.astype(np.float64) * L - 0.5)[0] % 1.0
    keep = ((frac > 0.01) & (frac < 0.99)).all(axis=1)  # no cell-boundary ties between FP32 and FP64
    target = rng.uniform(0, 1, (n, 7)).astype(np.float32)
    p = oracle.NeuralParams(np.arange(3), lat, mlp)
    _, _, _, margin = oracle.nm_gradients(p, material, uv, target)
    keep &= margin > 1e-4  # no ReLU-mask ties
    material, uv, target = material[keep], uv[keep], target[keep]
    assert len(material) > 2500
    loss_o, gm_o, gl_o, _ = oracle.nm_gradients(p, material, uv, target)
    loss_g, gm_g, gl_g = renderer.neural_gradients(material, uv, target)
    assert abs(loss_g - loss_o) <= 1e-5 * loss_o, (loss_g, loss_o)
    em = np.abs(gm_g - gm_o).max() / np.abs(gm_o).max()
    el = np.abs(gl_g - gl_o).max() / np.abs(gl_o).max()
    print(f"gradient max error / max |g|: mlp {em:.2e} latents {el:.2e}")
    assert em < 2e-4 and el < 2e-4
    assert np.linalg.norm(gl_g - gl_o) / np.linalg.norm(gl_o) < 1e-4
    assert (np.abs(gl_o) > 0).sum() > 0.5 * gl_o.size  # the batch touches most of the grid


def test_inference_matches_fp64_decoder_of_fp16_parameters(renderer, scene):
    renderer.neural_init(latent_size=32)
    renderer.neural_train(40)
    renderer.neural_evaluate()
    rng = np.random.default_rng(11)
    material = rng.integers(0, 3, 5000)
    uv = rng.uniform(-2, 3, (5000, 2)).astype(np.float32)
    got = renderer.neural_decode(material, uv)
    want = oracle.nm_decode(inference_params(renderer), material, uv)
    err = np.abs(got - want).max()
    print(f"decoder max abs error vs FP64 (FP16 parameters): {err:.2e}")
    assert err < DECODE_TOL


@pytest.mark.parametrize("fp8", [False, True])
def test_inference_weights_are_the_stated_rounding(renderer, scene, fp8):
    renderer.neural_init(latent_size=16, weights=abi.NEURAL_WEIGHTS_FP8 if fp8 else abi.NEURAL_WEIGHTS_FP16)
    renderer.neural_train(20)
    renderer.neural_evaluate()
    _, mlp = renderer.neural_params()
    got = renderer.neural_inference_mlp()
    want = reference_inference_mlp(mlp, fp8)
    assert np.array_equal(got, want), np.abs(got - want).max()
    if fp8:
        assert len(np.unique(np.abs(got[0:352]))) <= 127  # at most the E4M3 magnitudes, times one scale


def test_fp8_decoder_matches_fp64_and_reports_its_psnr(renderer, scene):
    renderer.neural_init(latent_size=32, psnr_threshold=0.0)
    renderer.neural_train(60)
    p16 = renderer.neural_evaluate()
    renderer.neural_set_weights(abi.NEURAL_WEIGHTS_FP8)
    p8 = renderer.neural_evaluate()
    print("PSNR FP16", p16, "FP8", p8)
    assert (p8 <= p16 + 0.5).all() and (p8 > 20).all()
    rng = np.random.default_rng(13)
    material = rng.integers(0, 3, 3000)
    uv = rng.uniform(-2, 3, (3000, 2)).astype(np.float32)
    got = renderer.neural_decode(material, uv)
    want = oracle.nm_decode(inference_params(renderer), material, uv)
    assert np.abs(got - want).max() < DECODE_TOL
    vis, sm, col, sf, cv, cs = shaded(renderer, scene, abi.SHADE_NEURAL)
    ref, _, _ = oracle.shade_view(scene, cv, [cs], [sm], vis, neural=inference_params(renderer), override_shadow=sf)
    err = np.abs(col - ref).max()
    print(f"FP8 neural frame colour max abs error vs FP64: {err:.2e}")
    assert err <= NEURAL_TOL


def test_reported_psnr_matches_decoded_texel_centres(renderer, scene):
    renderer.neural_init(latent_size=16)
    renderer.neural_train(30)
    psnr = renderer.neural_evaluate()
    m = 1
    a, _, _ = maps(scene, m)
    w, h = a.width, a.height
    xs, ys = np.meshgrid((np.arange(w) + 0.5) / w, (np.arange(h) + 0.5) / h)
    uv = np.stack([xs.ravel(), ys.ravel()], axis=1).astype(np.float32)
    dec = renderer.neural_decode(np.full(len(uv), m), uv)
    mse = ((dec - plain_targets(scene, np.full(len(uv), m), uv)) ** 2).mean()
    assert abs(10 * np.log10(1 / mse) - psnr[m]) < 0.02, (10 * np.log10(1 / mse), psnr[m])


def test_training_beats_the_untrained_decoder(renderer, scene):
    """Ablation: the decoder starts as the bilinear interpolation of the maps at the latent cell
    centres (a 16^2 stand-in for 64^2 maps, the method without a trained network); Adam on the GPU
    must improve every material's PSNR over that starting point.  From a random start the loss must
    fall by an order of magnitude."""
    renderer.neural_init(latent_size=16, batch=16384)
    before = renderer.neural_evaluate()
    loss = renderer.neural_train(400)
    after = renderer.neural_evaluate()
    print("PSNR untrained", before, "trained", after, "loss", loss[0], loss[-1])
    assert (after > before + 0.5).all()
    renderer.neural_init(latent_size=16, batch=16384, init=abi.NEURAL_INIT_RANDOM, lr_latent=1e-2, lr_mlp=1e-2)
    loss = renderer.neural_train(400)
    assert loss[-20:].mean() < 0.1 * loss[:20].mean(), (loss[:20].mean(), loss[-20:].mean())


def shaded(renderer, scene, mode, active=None):
    renderer.set_debug_flags(abi.DEBUG_SHADOW_FACTOR)
    frame = scene.frame_desc(mode)
    renderer.render(frame)
    cv = scene.camera_views()[0]
    cs = scene.cascades()[0]
    vis = renderer.visibility(0)
    sm = renderer.shadow_map(0)
    rs = scene.s.raster_state
    ov, _ = oracle.raster_view(scene, cv.view_projection[:], cv.width, cv.height, False, rs)
    assert np.array_equal(vis, ov)
    return vis, sm, renderer.color(0).astype(np.float64), renderer.shadow_factor(0), cv, cs


def test_neural_frame_matches_the_fp64_oracle(renderer, scene):
    renderer.neural_init(latent_size=32, psnr_threshold=0.0)
    renderer.neural_train(60)
    renderer.neural_evaluate()
    assert renderer.neural_stats().n_active == 3
    vis, sm, col, sf, cv, cs = shaded(renderer, scene, abi.SHADE_NEURAL)
    params = inference_params(renderer)
    ref, rsf, margin = oracle.shade_view(scene, cv, [cs], [sm], vis, neural=params)
    flip = np.abs(sf - rsf) > 2e-3
    assert (margin[flip] < TIE).all()
    ref, _, _ = oracle.shade_view(scene, cv, [cs], [sm], vis, neural=params, override_shadow=sf)
    err = np.abs(col - ref).max()
    print(f"neural frame colour max abs error vs FP64: {err:.2e}")
    assert err <= NEURAL_TOL
    # the decoded frame differs from the plain one (it is not silently the plain path)
    plain, _, _ = oracle.shade_view(scene, cv, [cs], [sm], vis, override_shadow=sf)
    assert np.abs(col - plain).max() > 1e-3


def test_fallback_materials_use_the_plain_maps(renderer, scene):
    renderer.neural_init(latent_size=32, psnr_threshold=0.0)
    renderer.neural_train(20)
    renderer.neural_evaluate()
    renderer.neural_set_active([0, 1, 0])
    vis, sm, col, sf, cv, cs = shaded(renderer, scene, abi.SHADE_NEURAL)
    params = inference_params(renderer, active=[0, 1, 0])
    ref, _, _ = oracle.shade_view(scene, cv, [cs], [sm], vis, neural=params, override_shadow=sf)
    assert np.abs(col - ref).max() <= NEURAL_TOL
    renderer.neural_set_active([0, 0, 0])
    vis, sm, col, sf, cv, cs = shaded(renderer, scene, abi.SHADE_NEURAL)
    ref, _, _ = oracle.shade_view(scene, cv, [cs], [sm], vis, override_shadow=sf)
    assert np.abs(col - ref).max() <= 1e-3  # every material on the plain maps: the plain bar


def test_psnr_threshold_sets_the_active_set(renderer, scene):
    renderer.neural_init(latent_size=16, psnr_threshold=1000.0)
    psnr = renderer.neural_evaluate()
    assert np.isfinite(psnr).all() and renderer.neural_stats().n_active == 0
    renderer.render(scene.frame_desc(abi.SHADE_NEURAL))  # renders, every material on the plain maps
    renderer.neural_init(latent_size=16, psnr_threshold=0.0)
    renderer.neural_evaluate()
    assert renderer.neural_stats().n_active == 3


def test_graph_replays_of_neural_frames_match_eager(renderer, scene):
    renderer.neural_init(latent_size=32, psnr_threshold=0.0)
    renderer.neural_train(10)
    renderer.neural_evaluate()
    renderer.set_debug_flags(0)
    frame = scene.frame_desc(abi.SHADE_NEURAL)
    renderer.render(frame)
    eager = renderer.color(0)
    for k in range(3):
        renderer.render_graph(frame, 1 + k)
        assert np.array_equal(renderer.color(0), eager), k
        renderer.render(frame)
        renderer.synchronize()


def test_decoder_errors_cross_the_abi(renderer, scene):
    lib = rosy_amd.lib()
    cfg = abi.NeuralConfig()
    lib.rsy_neural_default_config(C.byref(cfg))
    cfg.latent_size = 48
    assert lib.rsy_neural_init(renderer.ctx, C.byref(cfg)) == abi.RSY_INVALID_ARGUMENT
    cfg.latent_size, cfg.batch = 64, 1000
    assert lib.rsy_neural_init(renderer.ctx, C.byref(cfg)) == abi.RSY_INVALID_ARGUMENT
    cfg.batch = 65536
    renderer.neural_init(latent_size=64)
    f = scene.frame_desc(abi.SHADE_NEURAL)
    assert lib.rsy_render(renderer.ctx, C.byref(f)) == abi.RSY_INVALID_STATE  # not evaluated yet
    renderer.neural_evaluate()
    renderer.neural_train(1)
    assert lib.rsy_render(renderer.ctx, C.byref(f)) == abi.RSY_INVALID_STATE  # trained since the evaluation
    m = np.zeros(4, np.uint32)
    uv = np.zeros((4, 2), np.float32)
    t = np.zeros((4, 7), np.float32)
    loss = C.c_float()
    assert lib.rsy_neural_gradients(renderer.ctx, m.ctypes.data, uv.ctypes.data, t.ctypes.data, 4, C.byref(loss),
                                    None, None, 0) == abi.RSY_INVALID_ARGUMENT  # latent_size 64 > 32
    fresh = rosy_amd.Renderer(device=0)
    try:
        assert lib.rsy_neural_train(fresh.ctx, 1, None) == abi.RSY_INVALID_STATE
        assert lib.rsy_neural_init(fresh.ctx, C.byref(cfg)) == abi.RSY_INVALID_STATE  # no materials
    finally:
        fresh.close()
