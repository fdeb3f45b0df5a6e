"""Pins for the oracle's raster half (visibility buffer, depth, shadow depth bias).

Coverage is checked against an exact rational statement of the rule (pixel centre strictly inside,
or on a top or left edge, y down), written here from the geometric definition rather than from the
oracle's edge-function form; watertightness, tie-breaking, culling, clipping and depth bias are
checked by invariants and closed forms."""
import math
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle
from rosy_amd import abi
from scenes import CustomScene, ndc_from_subpixel

W = H = 16


def state(cull=0, reverse=0, bias=0, const=0.0, clamp=0.0, slope=0.0):
    rs = abi.RasterState()
    rs.cull_enabled, rs.reverse_winding_order_enabled, rs.depth_bias_enabled = cull, reverse, bias
    rs.depth_bias_constant, rs.depth_bias_clamp, rs.depth_bias_slope_factor = const, clamp, slope
    rs.clear_color[:] = [0, 0.05, 0.1, 1]
    return rs


def positions_from_subpixels(pts, z=0.5):
    return [(float(ndc_from_subpixel(X, W)), float(ndc_from_subpixel(Y, H)), z) for X, Y in pts]


def raster(scene, rs=None, shadow=False, w=W, h=H):
    out, st = oracle.raster_view(scene, scene.camera_views()[0].view_projection[:], w, h, shadow, rs or state())
    return out, st


def brute_force_mask(pts, sample=(128, 128)):
    """Exact coverage of each pixel's sample point (the centre by default) by the triangle with
    integer sub-pixel vertices (y down)."""
    (x0, y0), (x1, y1), (x2, y2) = [(Fr(X), Fr(Y)) for X, Y in pts]
    area = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0)
    mask = np.zeros((H, W), dtype=bool)
    if area == 0:
        return mask
    V = [(x0, y0), (x1, y1), (x2, y2)]

    def top_or_left(a, b, c):
        (ax, ay), (bx, by), (cx, cy) = a, b, c
        if ay == by:  # horizontal: top edge when the third vertex lies below (y grows downwards)
            return cy > ay
        # non-horizontal: left edge when the third vertex lies to the right of the edge's line
        x_on_line = ax + (bx - ax) * (cy - ay) / (by - ay)
        return cx > x_on_line

    for py in range(H):
        for px in range(W):
            P = (Fr(px * 256 + sample[0]), Fr(py * 256 + sample[1]))
            ok = True
            for i in range(3):
                a, b, c = V[(i + 1) % 3], V[(i + 2) % 3], V[i]
                # barycentric of vertex i = signed area(P, a, b) / area(c, a, b)
                lam = ((a[0] - P[0]) * (b[1] - P[1]) - (b[0] - P[0]) * (a[1] - P[1])) / (
                    (a[0] - c[0]) * (b[1] - c[1]) - (b[0] - c[0]) * (a[1] - c[1]))
                if lam < 0 or (lam == 0 and not top_or_left(a, b, c)):
                    ok = False
                    break
            mask[py, px] = ok
    return mask


def random_triangles(n, seed):
    rng = np.random.default_rng(seed)
    tris = []
    for k in range(n):
        if k % 3 == 0:  # vertices on pixel centres and edges through them: exercises the tie rule
            pts = [(int(rng.integers(-2, 18)) * 256 + 128, int(rng.integers(-2, 18)) * 256 + 128) for _ in range(3)]
        elif k % 3 == 1:  # axis-aligned edges through centres (top/bottom/left/right edges)
            x0, y0 = int(rng.integers(0, 12)) * 256 + 128, int(rng.integers(0, 12)) * 256 + 128
            dx, dy = int(rng.integers(1, 6)) * 256, int(rng.integers(1, 6)) * 256
            pts = [(x0, y0), (x0 + dx, y0), (x0, y0 + dy)] if rng.random() < 0.5 else [(x0 + dx, y0 + dy), (x0, y0 + dy), (x0 + dx, y0)]
        else:
            pts = [(int(rng.integers(-300, 16 * 256 + 300)), int(rng.integers(-300, 16 * 256 + 300))) for _ in range(3)]
        tris.append(pts)
    return tris


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_coverage_matches_exact_top_left_rule(seed):
    for pts in random_triangles(60, seed):
        sc = CustomScene(positions_from_subpixels(pts), [0, 1, 2], W, H)
        vis, st = raster(sc)
        got = vis != 0
        want = brute_force_mask(pts)
        assert np.array_equal(got, want), (pts, np.argwhere(got != want)[:5])
        assert st.fragments == want.sum()


def test_shared_edges_are_watertight():
    """A perturbed grid mesh that covers the whole view: every pixel centre exactly once."""
    rng = np.random.default_rng(11)
    n = 9
    pts, index = [], {}
    for j in range(n):
        for i in range(n):
            X = int(round(-200 + i * (16 * 256 + 400) / (n - 1)))
            Y = int(round(-200 + j * (16 * 256 + 400) / (n - 1)))
            if 0 < i < n - 1 and 0 < j < n - 1:
                if rng.random() < 0.5:  # some interior vertices sit exactly on pixel centres
                    X, Y = (X // 256) * 256 + 128, (Y // 256) * 256 + 128
                else:
                    X += int(rng.integers(-100, 100))
                    Y += int(rng.integers(-100, 100))
            index[(i, j)] = len(pts)
            pts.append((X, Y))
    tris = []
    for j in range(n - 1):
        for i in range(n - 1):
            a, b, c, d = index[(i, j)], index[(i + 1, j)], index[(i + 1, j + 1)], index[(i, j + 1)]
            tris += [a, b, c, a, c, d] if (i + j) % 2 else [a, b, d, b, c, d]
    sc = CustomScene(positions_from_subpixels(pts), tris, W, H)
    vis, st = raster(sc)
    assert (vis != 0).all()                 # every pixel centre covered ...
    assert st.fragments == W * H            # ... exactly once


def test_depth_plane_and_later_draw_wins_ties():
    pts = [(128, 128), (15 * 256 + 128, 128), (128, 15 * 256 + 128)]
    pos = positions_from_subpixels(pts, z=0.375) * 2
    sc = CustomScene(pos, [0, 1, 2, 3, 4, 5], W, H, draws=[(0, 1), (1, 1)])
    vis, _ = raster(sc)
    covered = vis != 0
    assert covered.sum() > 100
    ids = (vis[covered] & 0xFFFFFFFF) - 1
    assert (ids == 1).all()  # equal depth: the later draw wins (GREATER_OR_EQUAL in draw order)
    depth = (vis[covered] >> np.uint64(32)).astype(np.uint32).view(np.float32)
    assert (depth == np.float32(0.375)).all()
    # sloped plane: compare with the exact rational interpolation
    zs = [0.25, 0.75, 0.5]
    pos2 = [(p[0], p[1], z) for p, z in zip(positions_from_subpixels(pts), zs)]
    vis2, _ = raster(CustomScene(pos2, [0, 1, 2], W, H))
    for py, px in np.argwhere(vis2 != 0):
        Px, Py = px * 256 + 128, py * 256 + 128
        exact = zs[0] + (zs[1] - zs[0]) * (Px - 128) / (15 * 256) + (zs[2] - zs[0]) * (Py - 128) / (15 * 256)
        got = np.uint32(int(vis2[py, px]) >> 32).view(np.float32)
        assert abs(float(got) - exact) <= 4 * np.spacing(np.float32(exact))


def test_nearer_depth_wins_regardless_of_draw_order():
    pts = [(128, 128), (15 * 256 + 128, 128), (128, 15 * 256 + 128)]
    near = positions_from_subpixels(pts, z=0.75)   # reverse-Z: larger depth is nearer
    far = positions_from_subpixels(pts, z=0.25)
    for order, nearer_id in (((near, far), 0), ((far, near), 1)):
        sc = CustomScene(order[0] + order[1], [0, 1, 2, 3, 4, 5], W, H, draws=[(0, 1), (1, 1)])
        vis, _ = raster(sc)
        ids = (vis[vis != 0] & 0xFFFFFFFF) - 1
        assert (ids == nearer_id).all()


def test_cull_front_with_front_clockwise():
    # area2 = (X1-X0)(Y2-Y0) - (X2-X0)(Y1-Y0) > 0 is front-facing (Vulkan front = CW, y down)
    front = [(128, 128), (10 * 256, 128), (128, 10 * 256)]
    back = [front[0], front[2], front[1]]
    for pts, rs, kept in [(front, state(cull=1), False), (back, state(cull=1), True),
                          (front, state(cull=1, reverse=1), True), (back, state(cull=1, reverse=1), False),
                          (front, state(cull=0), True), (back, state(cull=0), True)]:
        vis, _ = raster(CustomScene(positions_from_subpixels(pts), [0, 1, 2], W, H), rs)
        assert ((vis != 0).any()) == kept
    # both windings of the same triangle cover the same pixels when culling is off
    a, _ = raster(CustomScene(positions_from_subpixels(front), [0, 1, 2], W, H))
    b, _ = raster(CustomScene(positions_from_subpixels(back), [0, 1, 2], W, H))
    assert np.array_equal(a != 0, b != 0)


def test_near_plane_clipping_keeps_the_part_in_front():
    # identity view-projection: w = 1, reverse-Z near plane is z <= w = 1, far plane z >= 0
    pts = [(-512, -512), (16 * 256 + 512, 8 * 256), (8 * 256, 16 * 256 + 512)]
    zs = [0.4, 1.6, 1.0 + 1e-3]
    pos = [(p[0], p[1], z) for p, z in zip(positions_from_subpixels(pts), zs)]
    vis, st = raster(CustomScene(pos, [0, 1, 2], W, H))
    assert st.triangles_clipped == 1
    inside = brute_force_mask(pts)
    # exact depth of the unclipped plane at every centre
    (x0, y0), (x1, y1), (x2, y2) = pts
    A = np.array([[x0, y0, 1], [x1, y1, 1], [x2, y2, 1]], dtype=np.float64)
    coef = np.linalg.solve(A, np.array(zs))
    for py in range(H):
        for px in range(W):
            z = coef @ [px * 256 + 128, py * 256 + 128, 1]
            if abs(z - 1.0) < 1e-3 or not inside[py, px]:
                continue
            assert (vis[py, px] != 0) == (z <= 1.0), (px, py, z)


def test_guard_band_clipping_of_a_huge_triangle_covers_the_view():
    pos = [(-100.0, -100.0, 0.5), (300.0, -100.0, 0.5), (-100.0, 300.0, 0.5)]
    vis, st = raster(CustomScene(pos, [0, 1, 2], W, H))
    assert st.triangles_clipped == 1
    assert (vis != 0).all() and st.fragments == W * H


def test_clipping_through_guard_band_and_near_plane_corners():
    """Triangles whose vertices sit on, or one rounding step from, corners where two guard-band
    planes (|x| <= 4w, |y| <= 4w) meet, or where one meets the near plane (z <= w): plane distances
    of exactly 0 and of rounding size with either sign.  The clipped polygon keeps at most 9
    vertices (the exact bound for a triangle and six planes), so each clipped triangle gives at
    most 7 fan triangles."""
    rng = np.random.default_rng(11)
    corners = [(4.0, 4.0), (-4.0, 4.0), (4.0, -4.0), (-4.0, -4.0), (4.0, 0.3), (-0.7, 4.0), (-4.0, -0.2)]
    clipped = 0
    for trial in range(300):
        pos = []
        for k in range(3):
            cx, cy = corners[int(rng.integers(len(corners)))]
            jx, jy = (float(v) for v in rng.choice([0.0, 1e-7, -1e-7, 3e-7], size=2))
            z = float(rng.choice([1.0, 1.0 + 6e-8, 1.0 - 6e-8, 0.5]))  # on or next to the near plane
            pos.append((cx + jx, cy + jy, z))
        vis, st = raster(CustomScene(pos, [0, 1, 2], W, H))
        assert st.triangles_setup <= 7 * st.triangles_clipped + (1 - st.triangles_clipped)
        clipped += st.triangles_clipped
    assert clipped > 200


def test_triangle_fully_behind_the_far_plane_is_rejected():
    pos = [(-0.5, -0.5, -0.1), (0.5, -0.5, -0.2), (-0.5, 0.5, -0.3)]
    vis, st = raster(CustomScene(pos, [0, 1, 2], W, H))
    assert not (vis != 0).any() and st.triangles_setup == 0


def shadow_depth(pos, rs):
    sc = CustomScene(pos, [0, 1, 2], W, H, shadow_projection=None)
    out, _ = oracle.raster_view(sc, sc.cascades()[0].light_view_projection[:], W, H, True, rs)
    return out


def test_depth_bias_constant_term_uses_the_exponent_of_the_largest_depth():
    pts = [(128, 128), (15 * 256 + 128, 128), (128, 15 * 256 + 128)]
    pos = positions_from_subpixels(pts, z=0.375)
    d = shadow_depth(pos, state(bias=1, const=-21.882, slope=-3.0, clamp=0.0))
    covered = d > 0
    r = 2.0 ** (-2 - 23)  # 0.375 = 1.5 * 2^-2
    want = np.float32(np.float32(0.375) + np.float32(np.float32(r) * np.float32(-21.882)))
    assert (d[covered] == want).all()


def test_depth_bias_slope_term_and_clamp():
    # plane z = 0.25 + (x_pixels)/64: slope 1/64 per pixel, exact in FP32
    pts = [(128, 128), (128 + 8 * 256, 128), (128, 128 + 15 * 256)]
    zs = [0.25, 0.25 + 8 / 64, 0.25]
    pos = [(p[0], p[1], z) for p, z in zip(positions_from_subpixels(pts), zs)]
    d = shadow_depth(pos, state(bias=1, const=0.0, slope=-3.0, clamp=0.0))
    for py, px in np.argwhere(d > 0):
        assert d[py, px] == np.float32(0.25 + px / 64 - 3 / 64)
    d2 = shadow_depth(pos, state(bias=1, const=0.0, slope=-3.0, clamp=-0.01))
    for py, px in np.argwhere(d2 > 0):
        assert abs(float(d2[py, px]) - (0.25 + px / 64 - 0.01)) < 1e-7
    d3 = shadow_depth(pos, state(bias=0, const=-21.882, slope=-3.0))
    for py, px in np.argwhere(d3 > 0):
        assert d3[py, px] == np.float32(0.25 + px / 64)


def test_shadow_map_keeps_the_largest_depth():
    pts = [(128, 128), (15 * 256 + 128, 128), (128, 15 * 256 + 128)]
    a = positions_from_subpixels(pts, z=0.25)
    b = positions_from_subpixels(pts, z=0.625)
    sc = CustomScene(a + b, [0, 1, 2, 3, 4, 5], W, H)
    out, _ = oracle.raster_view(sc, sc.cascades()[0].light_view_projection[:], W, H, True, state())
    assert (out[out > 0] == np.float32(0.625)).all()


def test_tile_pair_count_matches_bounding_boxes():
    pts = [(128, 128), (9 * 256 + 128, 128), (128, 5 * 256 + 128)]
    sc = CustomScene(positions_from_subpixels(pts), [0, 1, 2], W, H)
    vp = sc.camera_views()[0].view_projection[:]
    # pixel bbox x 0..9, y 0..5: 8x8 bins -> 2 x 1
    assert oracle.tile_pairs(sc, vp, W, H, False, state(), 8) == 2
    assert oracle.tile_pairs(sc, vp, W, H, False, state(), 16) == 1
