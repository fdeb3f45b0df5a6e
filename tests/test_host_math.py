"""Closed forms of the camera, light and node matrices (Camera.cpp:62-98, Level.cpp:464-528,
Node.cpp:96-108, 225-247).  glm is not vendored in the reference, so these are pinned by what the
mathematics fixes, not by glm's own rounding."""
import ctypes as C
import math

import numpy as np

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
