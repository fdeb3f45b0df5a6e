"""Pins for the oracle's FP64 material decoder (oracle/neural.cpp).

The decoder has no counterpart in the reference, so these pins come from its definition and from
calculus: the analytic gradients match central finite differences of the oracle's own loss; an
MLP built as the identity on channels 0..6 turns the decoder into bilinear REPEAT interpolation of
the latent grid, computed here independently with numpy; zero weights leave the output bias; the
positional features are the triangle waves of the bilinear weights."""
import numpy as np
import pytest

import oracle

W1, B1, W2, B2 = 0, 320, 352, 576


def identity_mlp():
    p = np.zeros(583, np.float32)
    for c in range(7):
        p[W1 + (2 * c) * 10 + c] = 1
        p[W1 + (2 * c + 1) * 10 + c] = -1
        p[W2 + c * 32 + 2 * c] = 1
        p[W2 + c * 32 + 2 * c + 1] = -1
    return p


def bilinear_repeat(grid, u, v):
    L = grid.shape[0]
    x, y = u * L - 0.5, v * L - 0.5
    x0, y0 = int(np.floor(x)), int(np.floor(y))
    ax, ay = x - x0, y - y0
    g = lambda i, j: grid[j % L, i % L].astype(np.float64)
    return ((1 - ax) * (1 - ay) * g(x0, y0) + ax * (1 - ay) * g(x0 + 1, y0) + (1 - ax) * ay * g(x0, y0 + 1)
            + ax * ay * g(x0 + 1, y0 + 1))


def test_identity_mlp_is_bilinear_repeat_interpolation():
    rng = np.random.default_rng(3)
    L = 8
    lat = rng.uniform(0.05, 0.95, (2, L, L, 8)).astype(np.float32)
    p = oracle.NeuralParams([1, 0], lat, identity_mlp())  # material 0 -> slot 1, material 1 -> slot 0
    uv = np.array([[0.5 / L, 0.5 / L],     # a cell centre: exactly that cell
                   [0.0, 0.0],             # the corner: wraps to the average of 4 corner cells
                   [0.37, 0.81], [0.999, 0.002], [1.73, -2.4]], np.float32)
    for m in (0, 1):
        out = oracle.nm_decode(p, [m] * len(uv), uv)
        grid = lat[1 - m]
        for i, (u, v) in enumerate(uv):
            want = bilinear_repeat(grid, float(u), float(v))[:7]
            assert np.allclose(out[i], want, atol=1e-12), (m, i)
    assert np.allclose(oracle.nm_decode(p, [0], uv[:1])[0], lat[1, 0, 0, :7], atol=1e-12)
    corner = lat[1][[0, 0, L - 1, L - 1], [0, L - 1, 0, L - 1]].astype(np.float64).mean(axis=0)[:7]
    assert np.allclose(oracle.nm_decode(p, [0], uv[1:2])[0], corner, atol=1e-12)


def test_outputs_clamp_to_unit_range_and_zero_weights_leave_the_bias():
    lat = np.full((1, 4, 4, 8), 5.0, np.float32)
    p = identity_mlp()
    out = oracle.nm_decode(oracle.NeuralParams([0], lat, p), [0], [[0.3, 0.3]])[0]
    assert np.all(out == 1.0)
    q = np.zeros(583, np.float32)
    q[B2:B2 + 7] = [0.1, 0.2, 0.3, 0.4, 0.5, -0.5, 1.5]
    out = oracle.nm_decode(oracle.NeuralParams([0], lat, q), [0], [[0.3, 0.3]])[0]
    assert np.allclose(out, [0.1, 0.2, 0.3, 0.4, 0.5, 0.0, 1.0], atol=1e-7)


def test_positional_features_are_triangle_waves_of_the_bilinear_weights():
    L = 16
    lat = np.zeros((1, L, L, 8), np.float32)
    q = np.zeros(583, np.float32)
    q[W1 + 0 * 10 + 8] = 1.0  # hidden 0 = relu(tri(ax))
    q[W1 + 1 * 10 + 9] = 1.0  # hidden 1 = relu(tri(ay))
    q[W2 + 0 * 32 + 0] = 1.0
    q[W2 + 1 * 32 + 1] = 1.0
    p = oracle.NeuralParams([0], lat, q)
    for ax, ay in ((0.0, 0.5), (0.25, 0.75), (0.5, 0.1), (0.9, 0.0)):
        u, v = (3 + ax + 0.5) / L, (7 + ay + 0.5) / L
        out = oracle.nm_decode(p, [0], np.array([[u, v]], np.float32))[0]
        assert abs(out[0] - (1 - abs(2 * ax - 1))) < 1e-5 and abs(out[1] - (1 - abs(2 * ay - 1))) < 1e-5


def loss_of(lat, mlp, material, uv, target):
    return oracle.nm_gradients(oracle.NeuralParams([0, 1], lat, mlp), material, uv, target)[0]


@pytest.mark.parametrize("seed", [0, 1])
def test_analytic_gradients_match_central_differences(seed):
    rng = np.random.default_rng(seed)
    L, n = 8, 40
    lat = rng.uniform(-1, 1, (2, L, L, 8)).astype(np.float32)
    mlp = rng.uniform(-0.6, 0.6, 583).astype(np.float32)
    material = rng.integers(0, 2, n)
    uv = rng.uniform(-1, 2, (n, 2)).astype(np.float32)
    target = rng.uniform(0, 1, (n, 7)).astype(np.float32)
    loss, g_mlp, g_lat, margin = oracle.nm_gradients(oracle.NeuralParams([0, 1], lat, mlp), material, uv, target)
    keep = margin > 2e-3  # stay away from ReLU kinks, where the loss is not differentiable
    material, uv, target = material[keep], uv[keep], target[keep]
    loss, g_mlp, g_lat, _ = oracle.nm_gradients(oracle.NeuralParams([0, 1], lat, mlp), material, uv, target)
    checked = 0
    for idx in list(rng.choice(583, 40, replace=False)) + [B1 + 3, B2 + 2, W2 + 5]:
        hi, lo = mlp.copy(), mlp.copy()
        hi[idx] = np.float32(mlp[idx] + np.float32(1e-3))
        lo[idx] = np.float32(mlp[idx] - np.float32(1e-3))
        fd = (loss_of(lat, hi, material, uv, target) - loss_of(lat, lo, material, uv, target)) / \
            (float(hi[idx]) - float(lo[idx]))
        assert abs(fd - g_mlp[idx]) <= 1e-6 + 1e-4 * abs(g_mlp[idx]), (idx, fd, g_mlp[idx])
        checked += 1
    nz = np.argwhere(np.abs(g_lat) > 1e-6)
    for s, j, i, c in nz[rng.choice(len(nz), 30, replace=False)]:
        hi, lo = lat.copy(), lat.copy()
        hi[s, j, i, c] += np.float32(1e-3)
        lo[s, j, i, c] -= np.float32(1e-3)
        fd = (loss_of(hi, mlp, material, uv, target) - loss_of(lo, mlp, material, uv, target)) / \
            (float(hi[s, j, i, c]) - float(lo[s, j, i, c]))
        assert abs(fd - g_lat[s, j, i, c]) <= 1e-6 + 1e-4 * abs(g_lat[s, j, i, c]), ((s, j, i, c), fd)
        checked += 1
    assert checked == 73
    # the loss itself: mean squared error over samples and channels of the unclamped outputs
    assert loss > 0


def test_plain_sampler_export_matches_numpy_bilinear():
    from rosy_amd import abi
    import ctypes as C
    rng = np.random.default_rng(9)
    w, h = 5, 3
    px = rng.integers(0, 256, (h, w, 2), dtype=np.uint8)
    d = abi.TextureDesc()
    d.width, d.height, d.format = w, h, abi.TEX_RG8_UNORM
    d.pixels = px.ctypes.data_as(C.c_void_p)
    for u, v in ((0.1, 0.2), (0.93, 0.51), (1.37, -0.2)):
        got = oracle.sample_texture(d, None, u, v)
        x, y = u * w - 0.5, v * h - 0.5
        x0, y0 = int(np.floor(x)), int(np.floor(y))
        ax, ay = x - x0, y - y0
        t = lambda i, j: px[j % h, i % w].astype(np.float64) / 255.0
        want = (1 - ax) * (1 - ay) * t(x0, y0) + ax * (1 - ay) * t(x0 + 1, y0) + (1 - ax) * ay * t(x0, y0 + 1) + \
            ax * ay * t(x0 + 1, y0 + 1)
        assert np.allclose(got[:2], want, atol=1e-12)
