"""The allocation-free frame (rsy_render_noalloc) against rsy_render, through the C ABI.

Two calls in a row on one context, the second with a smaller scene and view (so the static scratch
still holds the first call's contents): visibility buffers and shadow maps bit-exact, colour within
1e-3 (the same kernels give the same bits) against rsy_render's outputs, with the caller's output
buffers filled with garbage first.  A frame whose working memory exceeds the static scratch is
rejected before it runs, and targets that do not match the frame are rejected."""
import ctypes as C

import numpy as np
import pytest
import torch

import rosy_amd
from rosy_amd import abi

pytestmark = pytest.mark.gpu


def render_noalloc(r, scene, frame, garbage=True):
    """rsy_render_noalloc into torch-allocated device buffers; returns (code, vis, colour, shadow maps)."""
    cams, cas = scene.camera_views(), scene.cascades()
    if frame.n_cascades:
        cas = [frame.cascades[i] for i in range(frame.n_cascades)]
    dev = torch.device("cuda", 0)
    vis = [torch.empty(v.height * v.width, dtype=torch.int64, device=dev) for v in cams]
    col = [torch.empty(v.height * v.width * 4, dtype=torch.float32, device=dev) for v in cams]
    sm = [torch.empty(cs.size * cs.size, dtype=torch.float32, device=dev) for cs in cas]
    if garbage:
        for t in vis + col + sm:
            t.view(torch.int32).fill_(-1)
    torch.cuda.synchronize()
    P = C.c_void_p
    tg = abi.FrameTargets()
    tg.n_camera_views, tg.n_cascades = len(cams), len(cas)
    va, ca, sa = (P * len(vis))(*[t.data_ptr() for t in vis]), (P * len(col))(*[t.data_ptr() for t in col]), \
        (P * max(1, len(sm)))(*[t.data_ptr() for t in sm])
    tg.visibility, tg.color, tg.shadow_maps = C.cast(va, C.POINTER(P)), C.cast(ca, C.POINTER(P)), C.cast(sa, C.POINTER(P))
    code = rosy_amd.lib().rsy_render_noalloc(r.ctx, C.byref(frame), C.byref(tg))
    out_v = [t.cpu().numpy().view(np.uint64).reshape(v.height, v.width) for t, v in zip(vis, cams)]
    out_c = [t.cpu().numpy().reshape(v.height, v.width, 4) for t, v in zip(col, cams)]
    out_s = [t.cpu().numpy().reshape(cs.size, cs.size) for t, cs in zip(sm, cas)]
    return code, out_v, out_c, out_s


def check_against_render(r, scene, mode=abi.SHADE_PLAIN):
    frame = scene.frame_desc(mode)
    r.set_debug_flags(0)
    r.render(frame)
    want_v = [r.visibility(i) for i in range(len(scene.camera_views()))]
    want_c = [r.color(i) for i in range(len(scene.camera_views()))] if mode != abi.SHADE_NONE else []
    want_s = [r.shadow_map(i) for i in range(len(scene.cascades()))]
    code, got_v, got_c, got_s = render_noalloc(r, scene, frame)
    assert code == abi.RSY_OK, rosy_amd.lib().rsy_last_error(r.ctx)
    for a, b in zip(got_v, want_v):
        assert np.array_equal(a, b), f"{(a != b).sum()} visibility words differ"
    for a, b in zip(got_s, want_s):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), "shadow map differs"
    for a, b in zip(got_c, want_c):
        err = float(np.abs(a.astype(np.float64) - b).max())
        assert err <= 1e-3, err
        assert np.array_equal(a, b)  # the same kernels on the same inputs
    return frame


def test_noalloc_frames_match_render_twice_the_second_smaller():
    big = rosy_amd.Scene(1, seed=4)
    small = rosy_amd.Scene(0, seed=4)
    r = rosy_amd.Renderer(device=0)
    try:
        print("no-alloc: scene-scale workload (262,144 triangles, 1920x1080, 2048^2 shadow map)", flush=True)
        r.load(big)
        check_against_render(r, big)
        print("no-alloc: sphere workload (2,048 triangles, 640x360, 1024^2 shadow map)", flush=True)
        r.load(small)
        check_against_render(r, small)
        # and the bigger one again after the smaller, without shading (visibility and shadow maps)
        r.load(big)
        check_against_render(r, big, abi.SHADE_NONE)
    finally:
        r.close()
        big.close()
        small.close()


def test_noalloc_after_a_material_change():
    # rsy_set_materials after a frame keeps the draw lists built (their sampler flags are refreshed
    # in place), so the allocation-free frame that follows needs no rebuild and matches rsy_render
    s = rosy_amd.Scene(1, seed=4)
    r = rosy_amd.Renderer(device=0)
    lib = rosy_amd.lib()
    try:
        r.load(s)
        frame = s.frame_desc(abi.SHADE_PLAIN)
        r.render(frame)
        want_v, want_c = r.visibility(0), r.color(0)
        assert lib.rsy_set_materials(r.ctx, s.s.materials, s.s.n_materials, s.s.textures, s.s.n_textures,
                                     s.s.samplers, s.s.n_samplers) == abi.RSY_OK
        code, got_v, got_c, _ = render_noalloc(r, s, frame)
        assert code == abi.RSY_OK, lib.rsy_last_error(r.ctx)
        assert np.array_equal(got_v[0], want_v) and np.array_equal(got_c[0], want_c)
    finally:
        r.close()
        s.close()


def test_noalloc_rejects_what_it_cannot_hold():
    s = rosy_amd.Scene(1, seed=4)
    r = rosy_amd.Renderer(device=0)
    lib = rosy_amd.lib()
    try:
        r.load(s)
        # five views of 262,144 triangles need more working memory than the static scratch holds
        f = s.frame_desc(abi.SHADE_PLAIN)
        cas = (abi.ShadowCascade * 4)(*([s.cascades()[0]] * 4))
        f.n_cascades, f.cascades = 4, C.cast(cas, C.POINTER(abi.ShadowCascade))
        code, *_ = render_noalloc(r, s, f)
        assert code == abi.RSY_OVERFLOW
        assert b"static scratch" in lib.rsy_last_error(r.ctx)
        # targets that do not match the frame
        tg = abi.FrameTargets()
        tg.n_camera_views, tg.n_cascades = 2, 1
        assert lib.rsy_render_noalloc(r.ctx, C.byref(s.frame_desc(abi.SHADE_PLAIN)), C.byref(tg)) == abi.RSY_INVALID_ARGUMENT
        assert lib.rsy_render_noalloc(r.ctx, C.byref(s.frame_desc(abi.SHADE_PLAIN)), None) == abi.RSY_INVALID_ARGUMENT
        # the ordinary frame still works after the rejected ones
        r.render(s.frame_desc(abi.SHADE_PLAIN))
        r.synchronize()
    finally:
        r.close()
        s.close()
