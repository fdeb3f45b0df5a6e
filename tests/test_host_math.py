"""Closed forms of the camera, light and node matrices (Camera.cpp:62-98, Level.cpp:464-528,
Node.cpp:96-108, 225-247).  glm is not vendored in the reference, so these are pinned by what the
mathematics fixes, not by glm's own rounding."""
import ctypes as C
import math

import numpy as np
import pytest

import rosy_amd
from rosy_amd import abi


def mat(a):
    return np.array(list(a), dtype=np.float64).reshape(4, 4).T  # column-major -> matrix


def proj(aspect, n, f, fov):
    out = (C.c_float * 16)()
    rosy_amd.lib().rsy_camera_projection(aspect, n, f, fov, out)
    return mat(out)


def test_projection_is_reverse_z_with_vulkan_y_flip():
    P = proj(16 / 9, 0.1, 100.0, math.radians(60))
    near = P @ np.array([0, 0, -0.1, 1.0])
    far = P @ np.array([0, 0, -100.0, 1.0])
    assert abs(near[2] / near[3] - 1.0) < 1e-6      # depth 1 at the near plane
    assert abs(far[2] / far[3]) < 1e-6              # depth 0 at the far plane
    mid = P @ np.array([0, 0, -1.0, 1.0])
    assert 0.0 < mid[2] / mid[3] < 1.0
    up = P @ np.array([0, 1.0, -1.0, 1.0])
    assert up[1] / up[3] < 0                        # +y up maps to -y NDC (Vulkan flip)
    h = 1 / math.tan(math.radians(30))
    assert abs(P[1, 1] + h) < 1e-6 and abs(P[0, 0] - h / (16 / 9)) < 1e-6
    assert P[3, 2] == -1.0                          # w = -z_view


def test_view_matrix_moves_camera_to_origin_and_looks_down_minus_z():
    pos = (C.c_float * 3)(1.0, 2.0, 3.0)
    out = (C.c_float * 16)()
    rosy_amd.lib().rsy_camera_view_matrix(pos, 0.3, 1.1, out)
    V = mat(out)
    o = V @ np.array([1.0, 2.0, 3.0, 1.0])
    assert np.allclose(o[:3], 0, atol=1e-5)
    R = V[:3, :3]
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)
    # yaw = pi turns the camera round: a point at -z in front of the default camera is now behind
    out2 = (C.c_float * 16)()
    rosy_amd.lib().rsy_camera_view_matrix((C.c_float * 3)(0, 0, 0), 0.0, math.pi, out2)
    p = mat(out2) @ np.array([0, 0, 5.0, 1.0])
    assert p[2] < 0
    # positive pitch looks down: a point below the horizon ahead moves towards the view axis
    out3 = (C.c_float * 16)()
    rosy_amd.lib().rsy_camera_view_matrix((C.c_float * 3)(0, 0, 0), 0.5, 0.0, out3)
    q = mat(out3) @ np.array([0, -math.tan(0.5), -1.0, 1.0])
    assert abs(q[1]) < 1e-5 and q[2] < 0


def test_light_matrices_restate_level_cpp():
    sun = (C.c_float * 4)()
    sp = (C.c_float * 16)()
    rosy_amd.lib().rsy_light_matrices(23.776, 5.343, 6.414, 36.0, 76.838, sun, sp)
    s = np.array(list(sun))
    assert abs(np.linalg.norm(s[:3]) - 23.776) < 1e-4 and s[3] == 1.0
    S = mat(sp)
    # orthographic: last row (0,0,0,1)
    assert np.allclose(S[3], [0, 0, 0, 1])
    # the origin lies on the light axis at light-space depth d: depth = d / ortho_depth
    o = S @ np.array([0, 0, 0, 1.0])
    assert np.allclose(o[:2], 0, atol=1e-6) and abs(o[2] - 23.776 / 76.838) < 1e-6
    # points towards the sun get larger depth (the map keeps the largest depth)
    toward = S @ np.append(s[:3] * 0.1, 1.0)
    assert toward[2] > o[2]
    # x/y scale is 2 / cascade_level
    assert abs(np.linalg.norm(S[0, :3]) - 2 / 36) < 1e-7


@pytest.mark.parametrize("d,pitch,yaw,level,od", [(23.776, 5.343, 6.414, 36.0, 76.838), (30.0, 0.4, 1.1, 20.0, 50.0),
                                                  (15.0, -0.7, -2.3, 12.0, 40.0)])
def test_light_matrices_closed_forms(d, pitch, yaw, level, od):
    """Level.cpp:464-476, 508-528 written out.  light_line_rot = R_y(-yaw) R_x(-pitch) (the quaternions
    angleAxis(yaw, -y) and angleAxis(-pitch, x)), so
      sun_position = rot (0, 0, -d, 1) = (d sin(yaw) cos(pitch), -d sin(pitch), -d cos(yaw) cos(pitch), 1),
      rot e_x = (cos yaw, 0, sin yaw),  rot e_y = (sin yaw sin pitch, cos pitch, -cos yaw sin pitch),
    and shadowProjNear = diag(-2/c, -2/c, -1/od, 1) T(0, 0, -d) rot^T maps the origin to (0, 0, d/od),
    rot e_x to (-2/c, 0, d/od), rot e_y to (0, -2/c, d/od) and the sun to (0, 0, 2d/od)."""
    sun = (C.c_float * 4)()
    sp = (C.c_float * 16)()
    rosy_amd.lib().rsy_light_matrices(d, pitch, yaw, level, od, sun, sp)
    # the angles as the FP32 values the reference stores
    p, y = float(np.float32(pitch)), float(np.float32(yaw))
    want_sun = [d * math.sin(y) * math.cos(p), -d * math.sin(p), -d * math.cos(y) * math.cos(p), 1.0]
    assert np.allclose(list(sun), want_sun, rtol=0, atol=4e-6 * d)
    S = mat(sp)
    ex = np.array([math.cos(y), 0.0, math.sin(y), 1.0])
    ey = np.array([math.sin(y) * math.sin(p), math.cos(p), -math.cos(y) * math.sin(p), 1.0])
    tol = 4e-6 * max(1.0, d / od, 2 / level)
    assert np.allclose(S @ np.array([0, 0, 0, 1.0]), [0, 0, d / od, 1], atol=tol)
    assert np.allclose(S @ ex, [-2 / level, 0, d / od, 1], atol=tol)
    assert np.allclose(S @ ey, [0, -2 / level, d / od, 1], atol=tol)
    assert np.allclose(S @ np.array(want_sun), [0, 0, 2 * d / od, 1], atol=tol)


def test_object_record_inverse_and_normal_matrix():
    rec = abi.ObjectRecord()
    rosy_amd.lib().rsy_object_record_from_trs((C.c_float * 3)(3.0, -2.0, 5.0), 1.7, 0.8, C.byref(rec))
    W = mat(rec.transform)
    Wi = mat(rec.to_object_space_transform)
    assert np.allclose(Wi @ W, np.eye(4), atol=1e-5)
    N = np.array(list(rec.normal_transform)).reshape(3, 3).T
    assert np.allclose(N, np.linalg.inv(W[:3, :3]).T, atol=1e-6)
    # yaw about -y: +x rotates towards +z for positive yaw
    x = W[:3, :3] @ np.array([1.0, 0, 0])
    assert x[2] > 0 and abs(np.linalg.norm(x) - 1.7) < 1e-5


def test_mat4_mul_is_the_fma_chain():
    rng = np.random.default_rng(3)
    a = rng.standard_normal(16).astype(np.float32)
    b = rng.standard_normal(16).astype(np.float32)
    c = (C.c_float * 16)()
    rosy_amd.lib().rsy_mat4_mul((C.c_float * 16)(*a), (C.c_float * 16)(*b), c)
    ref = mat(a) @ mat(b)
    assert np.allclose(mat(c), ref, rtol=1e-6, atol=1e-6)
