"""The C ABI library loads, exports every entry point the header declares, and keeps the reference's
record layouts; host-only entry points work without a GPU."""
import ctypes as C
import os
import re

import pytest

import rosy_amd
from rosy_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "rosy_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(rsy_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_every_declared_function_is_exported_and_bound():
    lib = rosy_amd.lib()
    names = header_functions()
    assert len(names) >= 25
    bound = {p[0] for p in abi.PROTOTYPES}
    for n in names:
        assert hasattr(lib, n), f"{n} declared in rosy_amd.h but not exported"
        assert n in bound, f"{n} has no ctypes prototype"


def test_record_sizes_match_reference_layouts():
    # Asset.h:92-99, Graphics.cpp:345-360, 422-427, 281-306, Types.h:39-48, Asset.h:48-54
    assert C.sizeof(abi.Position) == 64
    assert C.sizeof(abi.GpuMaterial) == 64
    assert C.sizeof(abi.ObjectRecord) == 164
    assert C.sizeof(abi.SceneData) == 408
    assert C.sizeof(abi.Draw) == 48
    assert C.sizeof(abi.Sampler) == 8
    assert abi.SceneData.brdf_lighting_enabled.offset == 404
    assert abi.SceneData.sun_position.offset == 256
    assert abi.GpuMaterial.mixmap_sampler_index.offset == 60


def test_result_codes_match_rosy_result():
    # Types.h:13-31
    assert abi.RESULT_NAMES[0] == "ok" and abi.RESULT_NAMES[3] == "invalid_argument"
    assert abi.RESULT_NAMES[11] == "read_failed" and abi.RESULT_NAMES[14] == "overflow"
    assert abi.RESULT_NAMES[15] == "invalid_state"


def test_abi_version():
    assert rosy_amd.lib().rsy_abi_version() == abi.RSY_ABI_VERSION


def test_null_arguments_are_rejected_without_a_device():
    lib = rosy_amd.lib()
    assert lib.rsy_create(None, None) == abi.RSY_INVALID_ARGUMENT
    assert lib.rsy_set_geometry(None, None, 0, None, 0, None, 0) == abi.RSY_INVALID_ARGUMENT
    assert lib.rsy_render(None, None) == abi.RSY_INVALID_ARGUMENT
    bad = abi.Config(99, 0, 0, 0)
    ctx = C.c_void_p()
    assert lib.rsy_create(C.byref(bad), C.byref(ctx)) == abi.RSY_INVALID_ARGUMENT
    bad_scene = abi.SceneDesc()
    bad_scene.workload = 7
    p = C.POINTER(abi.Scene)()
    assert lib.rsy_scene_generate(C.byref(bad_scene), C.byref(p)) == abi.RSY_INVALID_ARGUMENT


def test_create_without_gpu_reports_init_failure():
    try:
        import torch  # noqa: F401  (device count only; does not initialise HIP)
    except Exception:
        pass
    cfg = abi.Config(abi.RSY_ABI_VERSION, 0, 0, 0)
    ctx = C.c_void_p()
    rc = rosy_amd.lib().rsy_create(C.byref(cfg), C.byref(ctx))
    if rc == abi.RSY_OK:
        rosy_amd.lib().rsy_destroy(ctx)
        pytest.skip("a HIP device is present")
    assert rc == abi.RSY_GRAPHICS_INIT_FAILURE
