"""The blended pass on the GPU against the FP64 oracle (oracle.blend_view): blended surfaces in draw
order over the opaque frame -- GEQUAL against the opaque depth, no depth write, both faces, alpha
mask discard, SRC_ALPHA / ONE_MINUS_SRC_ALPHA -- within 1e-3 max abs, at two view sizes.  The scene
holds up to six layers at a pixel; the visibility buffer stays the opaque one, bit-exact, and graph
replays give the eager frame."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rosy_amd
from rosy_amd import abi
from blend_scene import blend_scene

pytestmark = pytest.mark.gpu

TOL = 1e-3


@pytest.mark.parametrize("size", [64, 200])
def test_blended_pass_matches_the_oracle(size):
    s = blend_scene(size)
    r = rosy_amd.Renderer(device=0)
    r.load(s)
    frame = s.frame_desc(abi.SHADE_PLAIN)
    r.render(frame)
    vis, col = r.visibility(0), r.color(0).astype(np.float64)
    st = r.stats()
    cam = s.camera_views()[0]
    want_vis, _ = oracle.raster_view(s, cam.view_projection[:], size, size, False, s.s.raster_state)
    assert np.array_equal(vis, want_vis)
    maps = [r.shadow_map(0)]
    base, _, _ = oracle.shade_view(s, cam, s.cascades(), maps, vis)
    want, count = oracle.blend_view(s, cam, s.cascades(), maps, vis, base)
    assert count.max() >= 5
    err = float(np.abs(col - want).max())
    print(f"blended pass {size}x{size}: max abs err {err:.3e}, layers up to {count.max()}, re-runs {st.overflowed}")
    assert err <= TOL
    # the blended pass changed the pixels that hold a layer and no others
    changed = np.any(np.abs(col - base) > 1e-6, axis=-1)
    assert not changed[count == 0].any() and changed[count > 0].mean() > 0.5
    r.render_graph(frame, 2)
    assert np.array_equal(r.color(0), col.astype(np.float32))
    r.close()


def test_frames_without_blended_surfaces_are_unchanged():
    s = blend_scene(64)
    for i in range(s.s.n_surfaces):
        s.s.surfaces[i].blended = 0
    r = rosy_amd.Renderer(device=0)
    r.load(s)
    r.render(s.frame_desc(abi.SHADE_PLAIN))
    vis = r.visibility(0)
    cam = s.camera_views()[0]
    want_vis, _ = oracle.raster_view(s, cam.view_projection[:], 64, 64, False, s.s.raster_state)
    assert np.array_equal(vis, want_vis)
    base, _, _ = oracle.shade_view(s, cam, s.cascades(), [r.shadow_map(0)], vis)
    assert float(np.abs(r.color(0) - base).max()) <= TOL
    r.close()


def test_generated_workload_with_translucent_materials():
    """The decoder workload's generator with 16 translucent materials (BLEND and MASK over a varying
    texture alpha) at 320x180: opaque visibility and shadow map bit-exact, the blended frame within
    the tolerance of the FP64 oracle, in eager frames and graph replays."""
    s = rosy_amd.Scene(3, seed=3, width=320, height=180, shadow_size=512, total_triangles=16384, n_meshes=30,
                       texture_size=64, n_translucent=16)
    r = rosy_amd.Renderer(device=0)
    r.load(s)
    frame = s.frame_desc(abi.SHADE_PLAIN)
    r.render(frame)
    cam, cas = s.camera_views()[0], s.cascades()
    rs = s.s.raster_state
    vis = r.visibility(0)
    want_vis, _ = oracle.raster_view(s, cam.view_projection[:], cam.width, cam.height, False, rs)
    want_sm, _ = oracle.raster_view(s, cas[0].light_view_projection[:], cas[0].size, cas[0].size, True, rs)
    assert np.array_equal(vis, want_vis)
    assert np.array_equal(r.shadow_map(0).view(np.uint32), want_sm.view(np.uint32))
    base, _, _ = oracle.shade_view(s, cam, cas, [want_sm], vis)
    want, count = oracle.blend_view(s, cam, cas, [want_sm], vis, base, max_layers=64)
    col = r.color(0)
    err = float(np.abs(col.astype(np.float64) - want).max())
    print(f"generated blended frame: max abs err {err:.3e}, layers up to {count.max()}, "
          f"{(count > 0).mean():.3f} of pixels")
    assert count.max() >= 4 and err <= TOL
    r.render_graph(frame, 3)
    assert np.array_equal(r.color(0), col)
    # the same frame with 16- and 64-pixel bins (the blended pass collects 16x16 sub-tiles of a bin);
    # 8-pixel bins are refused with a blended pass
    for tile in (16, 64):
        rs2 = abi.RasterState()
        C.memmove(C.byref(rs2), C.byref(rs), C.sizeof(rs2))
        rs2.tile_size = tile
        r.update(s.s.scene_data, rs2)
        r.render(frame)
        assert np.array_equal(r.visibility(0), vis) and np.array_equal(r.color(0), col), tile
    rs2.tile_size = 8
    r.update(s.s.scene_data, rs2)
    assert rosy_amd.lib().rsy_render(r.ctx, C.byref(frame)) == abi.RSY_INVALID_ARGUMENT
    r.close()


def test_deep_stacks_grow_the_fragment_array():
    """Forty blended quads over the whole view: a 16x16 collection sub-tile holds 10,240 fragments,
    more than its LDS lists take (they go through global memory), and the frame more than the first
    fragment array holds, so the frame is re-run with a larger one; colours still match the oracle."""
    quads = [(-0.8, 0.8, -0.8, 0.8, 0.30, 0, False)]
    quads += [(-1.0, 1.0, -1.0, 1.0, 0.35 + 0.015 * k, 1 if k % 2 else 3, True) for k in range(40)]
    s = blend_scene(96, quads=quads)
    r = rosy_amd.Renderer(device=0)
    r.load(s)
    frame = s.frame_desc(abi.SHADE_PLAIN)
    r.render(frame)
    vis, col = r.visibility(0), r.color(0).astype(np.float64)
    cam = s.camera_views()[0]
    base, _, _ = oracle.shade_view(s, cam, s.cascades(), [r.shadow_map(0)], vis)
    want, count = oracle.blend_view(s, cam, s.cascades(), [r.shadow_map(0)], vis, base, max_layers=48)
    assert count.min() == 40 and r.stats().overflowed >= 1  # re-run with a larger fragment array
    err = float(np.abs(col - want).max())
    print(f"40-deep blended stacks: max abs err {err:.3e}")
    assert err <= TOL
    r.render_graph(frame, 2)
    assert np.array_equal(r.color(0), col.astype(np.float32))
    r.close()
