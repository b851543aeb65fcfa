"""C-ABI checks that need no GPU: the library loads, exports every function
include/smoke_abi.h declares, keeps the reference's struct layouts, and
reports errors through status codes (never by aborting)."""
import ctypes
import os
import re

import numpy as np
import pytest

import smokesim_amd
from smokesim_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "smoke_abi.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(smk_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_step_surface():
    names = declared_functions()
    # 1:1 with grid.h:36-84 / physics.h:50-67
    for stage in ("handle_emitters", "update_emitters_status", "advect_velocities",
                  "apply_buoyancy", "calculate_vorticity", "apply_vorticity", "pressure_solve",
                  "update_velocities", "advect_smoke", "obstacle_update_step",
                  "update_solid_map", "init_solid_map", "run_physics_step"):
        assert f"smk_{stage}" in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.lib_path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding types every export
    assert set(declared_functions()) <= set(_abi.EXPORTS)


def test_struct_layouts_match_reference():
    # physics_settings is 48 B (physics.h:22-32), emitter 48 B (emitters.h),
    # obstacle 32 B (obstacles.h), physics_step_settings 7 ints
    assert ctypes.sizeof(_abi._Phys) == 48
    assert ctypes.sizeof(_abi._Toggles) == 28
    assert _abi.EMITTER_DTYPE.itemsize == 48
    assert _abi.OBSTACLE_DTYPE.itemsize == 32
    assert _abi.BOX_DTYPE.itemsize == 24


def test_abi_version_and_errors_without_device():
    L = smokesim_amd.lib()
    assert L.smk_abi_version() == 1
    h = ctypes.c_void_p()
    # null descriptor
    assert L.smk_grid_create(None, ctypes.byref(h)) == 1
    assert b"null" in L.smk_last_error()
    # invalid sizes are rejected before any device work
    d = _abi._Desc()
    d.size[0], d.size[1], d.size[2] = 0, 16, 16
    d.cell_size = 0.1
    assert L.smk_grid_create(ctypes.byref(d), ctypes.byref(h)) == 1
    d.size[0] = 16
    d.cell_size = -1.0
    assert L.smk_grid_create(ctypes.byref(d), ctypes.byref(h)) == 1
    d.cell_size = 0.1
    d.storage = 7
    assert L.smk_grid_create(ctypes.byref(d), ctypes.byref(h)) == 1
    d.storage = 0
    d.size[0] = 2  # below the pyramid's minimum level size
    assert L.smk_grid_create(ctypes.byref(d), ctypes.byref(h)) == 1
    # null handles are invalid arguments, not crashes
    assert L.smk_run_physics_step(None) == 1
    assert L.smk_sync(None) == 1
    assert L.smk_grid_destroy(None) == 0


def test_grid_without_gpu_raises_cleanly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(smokesim_amd.SmokeError) as e:
        smokesim_amd.Grid((16, 16, 16))
    assert e.value.status == 5  # SMK_ERR_NO_DEVICE


def test_scenes_are_seeded_and_separated():
    from smokesim_amd import scenes
    a = scenes.config_3d_128()
    b = scenes.config_3d_128()
    assert np.array_equal(a["emitters"], b["emitters"])
    assert np.array_equal(a["boxes"], b["boxes"])
    assert np.array_equal(a["velocity"]["vx"], b["velocity"]["vx"])
    assert a["emitters"].shape == (2, 12) and a["boxes"].shape == (1, 6)
    assert np.all(np.abs(a["velocity"]["vy"]) <= 0.01)
    c = scenes.config_2d_256()
    assert c["emitters"].shape == (1, 12) and c["obstacles"].shape == (1, 8)
    # emitter in the bottom quarter, obstacle does not overlap it
    e, o = c["emitters"][0], c["obstacles"][0]
    assert e[1] < 64 + 12
    assert np.hypot(e[0] - o[0], e[1] - o[1]) > e[3] + o[3]


def test_header_is_plain_c():
    """include/smoke_abi.h compiles as C11 with warnings as errors (the
    boundary is a C ABI; C++ is more lenient about struct declarations)."""
    import subprocess
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-x", "c", HEADER],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
