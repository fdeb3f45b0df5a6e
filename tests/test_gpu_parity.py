"""GPU parity through the C ABI: visibility buffers and shadow maps bit-exact against the CPU oracle,
shaded colour within 1e-3 (max abs, FP32 vs the FP64 oracle) on the same seeded inputs.

The colour gate: the oracle is re-run with the GPU's own per-pixel PCF factor, and the GPU's factor
may differ from the oracle's only where some shadow compare is a tie (|ref - texel| < 1e-5): a depth
compare is a step function, so an FP32 and an FP64 reference depth can land on either side of it."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rosy_amd
from rosy_amd import abi
from scenes import CustomScene, ndc_from_subpixel

pytestmark = pytest.mark.gpu

TOL = 1e-3
TIE = 1e-5


@pytest.fixture(scope="module")
def renderer():
    r = rosy_amd.Renderer(device=0)
    yield r
    r.close()


def check_frame(r, scene, frame=None, colour=True, colour_views=None, sd=None, rs=None, flags=0):
    """Render `frame` (already loaded scene) and compare everything with the oracle."""
    frame = frame if frame is not None else scene.frame_desc(abi.SHADE_PLAIN)
    r.set_debug_flags(abi.DEBUG_SHADOW_FACTOR | flags)
    r.render(frame)
    rs = rs if rs is not None else scene.s.raster_state
    cascades = scene.cascades() if frame.n_cascades else [None]
    maps = []
    for ci, cs in enumerate(scene.cascades()):
        g = r.shadow_map(ci)
        o, _ = oracle.raster_view(scene, cs.light_view_projection[:], cs.size, cs.size, True, rs)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), f"shadow map {ci}: {(g != o).sum()} texels differ"
        maps.append(g)
    errs = []
    for vi, cv in enumerate(scene.camera_views()):
        g = r.visibility(vi)
        o, _ = oracle.raster_view(scene, cv.view_projection[:], cv.width, cv.height, False, rs)
        assert np.array_equal(g, o), f"view {vi}: {(g != o).sum()} visibility words differ"
        if not colour or (colour_views is not None and vi not in colour_views):
            continue
        col = r.color(vi).astype(np.float64)
        sf = r.shadow_factor(vi)
        ref, rsf, margin = oracle.shade_view(scene, cv, scene.cascades(), maps, g, scene_data=sd, raster_state=rs)
        flip = np.abs(sf - rsf) > 2e-3
        assert (margin[flip] < TIE).all(), f"view {vi}: shadow factor differs away from a compare tie"
        ref2, _, _ = oracle.shade_view(scene, cv, scene.cascades(), maps, g, scene_data=sd, raster_state=rs,
                                       override_shadow=sf)
        err = float(np.abs(col - ref2).max())
        assert err <= TOL, f"view {vi}: colour max abs error {err:.3e}"
        errs.append(err)
    return errs


@pytest.fixture(scope="module")
def sphere():
    s = rosy_amd.Scene(0, seed=1)
    yield s
    s.close()


@pytest.fixture(scope="module")
def scene_scale():
    s = rosy_amd.Scene(1, seed=1)
    yield s
    s.close()


def test_sphere_workload(renderer, sphere):
    renderer.load(sphere)
    errs = check_frame(renderer, sphere)
    st = renderer.stats()
    assert st.triangle_count == 2 * 2048 and st.draw_call_count == 2 and st.overflowed == 0
    print("sphere colour max abs err", errs)


def test_scene_scale_workload(renderer, scene_scale):
    renderer.load(scene_scale)
    errs = check_frame(renderer, scene_scale)
    st = renderer.stats()
    assert st.triangle_count == 2 * 262144 and st.overflowed == 0
    print("scene-scale colour max abs err", errs)


def test_overflow_regrowth_matches_the_oracle(scene_scale):
    """A context whose frame buffers start at their minimum capacities (RSY_DEBUG_MIN_CAPACITY):
    the first frame overflows its bin lists, bin pairs and clip queue, grows them and is re-run;
    the result is still bit-exact and within the colour bar."""
    r = rosy_amd.Renderer(device=0)
    try:
        r.set_debug_flags(abi.DEBUG_MIN_CAPACITY)
        r.load(scene_scale)
        check_frame(r, scene_scale, flags=abi.DEBUG_MIN_CAPACITY)
        st = r.stats()
        assert st.overflowed >= 1, "the minimum capacities did not overflow"
        assert st.triangle_count == 2 * 262144
        print("re-runs after overflow:", st.overflowed, "tile pairs", st.tile_pairs)
        # the grown buffers hold the next frame without another re-run
        check_frame(r, scene_scale, flags=abi.DEBUG_MIN_CAPACITY)
        assert r.stats().overflowed == 0
    finally:
        r.close()


def test_bin_sizes_give_identical_frames(renderer, scene_scale):
    # 8-pixel bins give this frame 240 x 135 + 256 x 256 = 97,936 bins, past the 65,536 of the
    # one-launch bin scan: the two-launch scan must agree with it bit for bit
    cv, cs = scene_scale.camera_views()[0], scene_scale.cascades()[0]
    assert (cv.width // 8) * (cv.height // 8) + (cs.size // 8) ** 2 > 65536
    assert (cv.width // 16) * (cv.height // 16) + (cs.size // 16) ** 2 <= 65536
    renderer.load(scene_scale)
    frame = scene_scale.frame_desc(abi.SHADE_PLAIN)
    outs = []
    for tile in (8, 16, 32, 64):
        rs = abi.RasterState()
        C.memmove(C.byref(rs), C.byref(scene_scale.s.raster_state), C.sizeof(rs))
        rs.tile_size = tile
        renderer.update(scene_scale.s.scene_data, rs)
        renderer.render(frame)
        outs.append((renderer.visibility(0), renderer.shadow_map(0), renderer.color(0)))
        assert renderer.stats().tile_size == tile
    for v, s, c in outs[1:]:
        assert np.array_equal(v, outs[0][0]) and np.array_equal(s, outs[0][1]) and np.array_equal(c, outs[0][2])


def test_graph_replays_match_eager_frames(renderer, scene_scale):
    renderer.load(scene_scale)
    frame = scene_scale.frame_desc(abi.SHADE_PLAIN)
    renderer.render(frame)
    eager = (renderer.visibility(0), renderer.color(0), renderer.shadow_map(0))
    for replay in range(3):
        renderer.render_graph(frame, 1 + replay)
        assert np.array_equal(renderer.visibility(0), eager[0]), f"replay {replay}"
        assert np.array_equal(renderer.color(0), eager[1]), f"replay {replay}"
        assert np.array_equal(renderer.shadow_map(0), eager[2]), f"replay {replay}"
        renderer.render(frame)  # eager work between replays
        renderer.synchronize()
    # a moved object: graph replay after the update still matches the eager frame
    obj = abi.ObjectRecord()
    rosy_amd.lib().rsy_object_record_from_trs((C.c_float * 3)(2.0, 0.0, 1.0), 1.0, 0.4, C.byref(obj))
    assert rosy_amd.lib().rsy_update_graphic_objects(renderer.ctx, 5, C.byref(obj), 1) == 0
    renderer.render(frame)
    moved = renderer.visibility(0)
    renderer.render_graph(frame, 2)
    assert np.array_equal(renderer.visibility(0), moved)


def test_raster_state_variants(renderer, sphere):
    renderer.load(sphere)
    for cull, reverse, bias in ((0, 0, 1), (1, 1, 1), (1, 0, 0), (0, 0, 0)):
        rs = abi.RasterState()
        C.memmove(C.byref(rs), C.byref(sphere.s.raster_state), C.sizeof(rs))
        rs.cull_enabled, rs.reverse_winding_order_enabled, rs.depth_bias_enabled = cull, reverse, bias
        renderer.update(sphere.s.scene_data, rs)
        check_frame(renderer, sphere, rs=rs, colour=(cull, reverse, bias) == (0, 0, 1))


@pytest.mark.parametrize("mode", list(range(16)))
def test_debug_output_modes(renderer, sphere, mode):
    renderer.load(sphere)
    sd = abi.SceneData()
    C.memmove(C.byref(sd), C.byref(sphere.s.scene_data), C.sizeof(sd))
    sd.fragment_output = mode
    sd.light_enabled = 0 if mode % 2 == 0 else 1
    renderer.update(sd, sphere.s.raster_state)
    check_frame(renderer, sphere, sd=sd)


def test_lighting_switches(renderer, scene_scale):
    renderer.load(scene_scale)
    for kw in ({"brdf_lighting_enabled": 0}, {"shadows_enabled": 0}, {"tangent_space_enabled": 0},
               {"normal_maps_enabled": 0}, {"ensure_orthogonal_bitangent": 1, "ignore_asset_tangent_sign": 1}):
        sd = abi.SceneData()
        C.memmove(C.byref(sd), C.byref(scene_scale.s.scene_data), C.sizeof(sd))
        for k, v in kw.items():
            setattr(sd, k, v)
        renderer.update(sd, scene_scale.s.raster_state)
        check_frame(renderer, scene_scale, sd=sd)


def test_random_triangles_match_oracle(renderer):
    rng = np.random.default_rng(5)
    W = H = 64
    pts = rng.integers(-600, 64 * 256 + 600, size=(300, 2))
    pos = [(float(ndc_from_subpixel(x, W)), float(ndc_from_subpixel(y, H)), float(rng.uniform(0.05, 0.95)))
           for x, y in pts]
    tris = rng.integers(0, len(pos), size=(400, 3))
    sc = CustomScene(pos, tris.reshape(-1), W, H, draws=[(0, 200), (200, 200)], shadow_size=64)
    renderer.load(sc)
    check_frame(renderer, sc, colour=False)


def test_clipping_matches_oracle(renderer):
    W = H = 64
    # near-plane crossing, far-plane crossing and guard-band triangles (identity view-projection)
    pos = [(-0.9, -0.9, 0.4), (0.9, -0.2, 1.6), (0.0, 0.9, 1.001),
           (-0.7, 0.5, -0.3), (0.8, 0.7, 0.6), (0.2, -0.8, 0.2),
           (-100.0, -100.0, 0.5), (300.0, -100.0, 0.5), (-100.0, 300.0, 0.45)]
    sc = CustomScene(pos, list(range(9)), W, H, shadow_size=64)
    renderer.load(sc)
    check_frame(renderer, sc, colour=False)
    assert renderer.stats().triangles_clipped >= 3


def _translate(x, y, z):
    return [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, x, y, z, 1.0]


@pytest.mark.parametrize("perspective", [False, True])
def test_draws_beyond_a_side_of_the_view_match_oracle(renderer, perspective):
    # one mesh drawn through many objects: boxes just beyond each side of the view (skipped whole
    # by the draw test), touching an edge (within the test's rounding margin: not skipped),
    # crossing the guard band, straddling a corner, and (with perspective) crossing the near plane
    W, H = 96, 64
    g = np.linspace(0.0, 0.2, 5, dtype=np.float32)
    pos = [(float(x), float(y), 0.5 + 0.1 * float(x) - 0.05 * float(y)) for y in g for x in g]
    tris = []
    for j in range(4):
        for i in range(4):
            a = j * 5 + i
            tris += [a, a + 1, a + 5, a + 1, a + 6, a + 5]
    n_tris = len(tris) // 3
    if perspective:
        n, f, s = 0.5, 50.0, 1.5
        a = n / (f - n)
        vp = [s, 0, 0, 0, 0, s, 0, 0, 0, 0, a, -1.0, 0, 0, a * f, 0]
        d = 4.0  # eye depth of most objects: clip w = d - z_local, z_local in [0.49, 0.52]
        e = (d - 0.48) / s  # eye x (or y) beyond which a box at that depth is past the side
        offs = [(e - 0.01, 0.0, -d), (e + 0.02, 0.1, -d), (-e - 0.22, 0.0, -d), (0.0, e + 0.02, -d),
                (0.0, -e - 0.22, -d), (2.5 * e, 0.0, -d), (3.95 * e, 0.0, -d), (e - 0.1, e - 0.1, -d),
                (e + 0.1, 0.0, -0.2), (e + 0.02, 0.0, -60.0), (0.0, 0.0, -d)]
    else:
        vp = None
        offs = [(1.0, 0.0, 0.0), (1.0 + 2.0 ** -20, 0.1, 0.0), (1.001, -0.4, 0.0), (-1.201, 0.0, 0.0),
                (-1.2, 0.3, 0.0), (0.3, 1.001, 0.0), (0.3, -1.201, 0.0), (2.9, 0.0, 0.0), (3.9, 0.0, 0.0),
                (0.9, 0.9, 0.0), (-0.1, -0.1, 0.0)]
    objs = [_translate(*o) for o in offs]
    if perspective:  # the mesh's y extent turned into depth, from in front of the near plane to behind the eye
        objs.append([1.0, 0, 0, 0, 0, 0, -10.0, 0, 0, 1.0, 0, 0, 2.0, 0.0, -0.5, 1.0])
    draws = [(0, n_tris)] * len(objs)
    sc = CustomScene(pos, tris, W, H, draws=draws, view_projection=vp, shadow_projection=vp, shadow_size=64,
                     objects=objs, object_of_draw=list(range(len(objs))))
    renderer.load(sc)
    check_frame(renderer, sc, colour=False)
    assert renderer.stats().triangles_setup > 0
    # the same draws with 4 samples per pixel: sample points at 1/8 and 7/8 of a pixel lie inside
    # the image edge too, so every sample plane still matches
    rs = abi.RasterState()
    C.memmove(C.byref(rs), C.byref(sc.s.raster_state), C.sizeof(rs))
    rs.msaa_samples, rs.tile_size = 4, 32
    renderer.update(sc.s.scene_data, rs)
    renderer.render(sc.frame_desc(abi.SHADE_PLAIN))
    cam = sc.camera_views()[0]
    want = oracle.raster_view_msaa(sc, cam.view_projection[:], cam.width, cam.height, rs)
    got = renderer.visibility(0)
    assert np.array_equal(got, want), f"{int((got != want).sum())} sample words differ"
    renderer.update(sc.s.scene_data, sc.s.raster_state)


def test_four_cascades_and_multiple_views(renderer):
    s2 = rosy_amd.Scene(2, seed=3, total_triangles=120000, n_meshes=60, width=1280, height=720, shadow_size=1024,
                        texture_size=256)
    renderer.load(s2)
    check_frame(renderer, s2)
    s2.close()
    s4 = rosy_amd.Scene(4, seed=3, total_triangles=150000, n_meshes=60, shadow_size=512, texture_size=256,
                        n_materials=8)
    renderer.load(s4)
    check_frame(renderer, s4, colour_views=[2, 3])
    s4.close()


def test_frames_are_deterministic(renderer, scene_scale):
    renderer.load(scene_scale)
    frame = scene_scale.frame_desc(abi.SHADE_PLAIN)
    renderer.render(frame)
    a = (renderer.visibility(0), renderer.color(0))
    renderer.render(frame)
    b = (renderer.visibility(0), renderer.color(0))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_autotune_picks_a_bin_size_and_stays_exact(renderer, sphere):
    renderer.load(sphere)
    t = renderer.autotune(sphere.frame_desc(abi.SHADE_PLAIN), 2)
    assert t in (8, 16, 32, 64)
    check_frame(renderer, sphere)


def test_error_codes_cross_the_abi(renderer, sphere):
    lib = rosy_amd.lib()
    s = sphere.s
    bad = (C.c_uint32 * 3)(0, 1, 99999)
    mesh = (abi.MeshRange * 1)(abi.MeshRange(0, 0, 3, 3, 0))
    assert lib.rsy_set_geometry(renderer.ctx, s.vertices, s.n_vertices, bad, 3, mesh, 1) == abi.RSY_INVALID_ARGUMENT
    assert lib.rsy_set_graphic_objects(renderer.ctx, s.objects, 0, s.surfaces, s.n_surfaces) == abi.RSY_INVALID_ARGUMENT
    rs = abi.RasterState()
    rs.tile_size = 7
    assert lib.rsy_update(renderer.ctx, C.byref(s.scene_data), C.byref(rs)) == abi.RSY_INVALID_ARGUMENT
    fresh = rosy_amd.Renderer(device=0)
    try:
        assert lib.rsy_render(fresh.ctx, None) == abi.RSY_INVALID_STATE
        out = np.zeros(4, np.uint64)
        assert lib.rsy_read_visibility(fresh.ctx, 0, out.ctypes.data, 4) == abi.RSY_INVALID_STATE
    finally:
        fresh.close()
    renderer.load(sphere)
    f = sphere.frame_desc(abi.SHADE_PLAIN)
    renderer.render(f)
    out = np.zeros(7, np.uint64)
    assert lib.rsy_read_visibility(renderer.ctx, 0, out.ctypes.data, 7) == abi.RSY_INVALID_ARGUMENT
    assert lib.rsy_read_visibility(renderer.ctx, 5, out.ctypes.data, 7) == abi.RSY_INVALID_ARGUMENT


def test_view_limit_is_an_invalid_argument(renderer, sphere):
    # one camera view and RSY_MAX_VIEWS - 1 cascades fit a frame; one cascade more is refused
    renderer.load(sphere)
    cs = sphere.cascades()[0]
    many = (abi.ShadowCascade * abi.MAX_VIEWS)(*([cs] * abi.MAX_VIEWS))
    f = sphere.frame_desc(abi.SHADE_PLAIN)
    assert f.n_camera_views == 1
    f.n_cascades = abi.MAX_VIEWS
    f.cascades = C.cast(many, C.POINTER(abi.ShadowCascade))
    assert rosy_amd.lib().rsy_render(renderer.ctx, C.byref(f)) == abi.RSY_INVALID_ARGUMENT
    f.n_cascades = abi.MAX_VIEWS - 1
    assert rosy_amd.lib().rsy_render(renderer.ctx, C.byref(f)) == abi.RSY_OK
