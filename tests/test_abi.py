"""The C ABI library (runs without a GPU): it loads, exports every function include/svo.h declares,
the Python binding covers them, the record layouts match the header, and the host-only entry
points (scene generator, camera) behave."""
import math
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    with open(os.path.join(ROOT, "include", "svo.h")) as f:
        src = f.read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(svo_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import ctypes

    from sparsevoxeloctree_amd import LIB_PATH

    lib = ctypes.CDLL(LIB_PATH)
    names = header_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from sparsevoxeloctree_amd import exported_symbols

    assert sorted(exported_symbols()) == header_functions()


def test_record_layouts():
    import ctypes

    from sparsevoxeloctree_amd._capi import HIT_DTYPE, svo_camera, svo_timing

    assert HIT_DTYPE.itemsize == 32
    assert [HIT_DTYPE.fields[n][1] for n in ("pos", "iter", "vxy", "vz_flags", "leaf", "nfetch")] == [
        0, 12, 16, 20, 24, 28,
    ]
    assert ctypes.sizeof(svo_camera) == 64
    assert ctypes.sizeof(svo_timing) == 32


def test_scene_generator_sizes_and_determinism():
    import sparsevoxeloctree_amd as S

    for kind, param, T in (
        (1, 64, 258048),
        (2, 1, 990800),
        (1, 2, 8064),
        (2, 10, 6 * 20 * 20 * 2 + 40 * 12 + 20 * 2530),
        (4, 1000, 30 * 29),
        (4, 5000, 4032 + 30 * 29),
        (4, 20, 0),
    ):
        xyz, rgb = S.gen_scene(kind, 0x5EED0000 + kind, param)
        assert rgb.shape == (T,), (kind, param, rgb.shape)
        assert xyz.shape == (T, 9)
        xyz2, rgb2 = S.gen_scene(kind, 0x5EED0000 + kind, param)
        assert np.array_equal(xyz, xyz2) and np.array_equal(rgb, rgb2)
        assert np.all(rgb <= 0xFFFFFF)
    # kind 1: one exactly degenerate triangle per south-pole quad (64 per sphere)
    xyz, _ = S.gen_scene(1, 1, 1)
    p = xyz.reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert int(np.sum(np.all(n == 0, axis=1))) == 64


def test_camera_basis():
    import sparsevoxeloctree_amd as S

    cam = S.Camera((1.5, 1.5, 1.5), 0.7, -0.3, math.pi / 3, 16 / 9).uniform_array()
    pos, look, side, up = cam[0:3], cam[4:7], cam[8:11], cam[12:15]
    assert np.allclose(pos, 1.5)
    tg = math.tan(math.pi / 6)
    assert abs(np.linalg.norm(look) - 1) < 1e-6
    assert abs(np.linalg.norm(side) - tg * 16 / 9) < 1e-5
    assert abs(np.linalg.norm(up) - tg) < 1e-5
    assert abs(np.dot(look, side)) < 1e-5 and abs(np.dot(look, up)) < 1e-5
    # yaw = pitch = 0 looks along +z with side = +x (src/Camera.cpp:77-95)
    c0 = S.Camera((1, 1, 1), 0.0, 0.0, math.pi / 3, 1.0).uniform_array()
    assert np.allclose(c0[4:7], [0, 0, 1]) and np.allclose(c0[8:11] / tg, [1, 0, 0], atol=1e-6)


def test_no_device_is_reported():
    # this container has no GPU: creating a context must fail cleanly, not crash
    import ctypes

    from sparsevoxeloctree_amd._capi import lib

    h = ctypes.c_void_p()
    rc = lib.svo_ctx_create(1 << 20, ctypes.byref(h))
    assert rc == 6 and not h.value


def test_batch_and_server_argument_checks():
    # argument errors are reported before any device work (no GPU needed)
    import ctypes

    from sparsevoxeloctree_amd._capi import lib, svo_server_stats

    assert ctypes.sizeof(svo_server_stats) == 10 * 8 + 7 * 8
    h = ctypes.c_void_p()
    assert lib.svo_server_create(1 << 20, 1 << 30, 0, ctypes.byref(h)) == 1  # max_octrees 0
    assert lib.svo_server_create(1 << 20, 1 << 30, 8, ctypes.byref(h)) == 6 and not h.value  # no such device
    rid = ctypes.c_uint64()
    assert lib.svo_server_submit_build(None, None, 5, ctypes.byref(rid)) == 1
    assert lib.svo_server_submit_rays(None, 1, None, 0, 0, None, ctypes.byref(rid)) == 1
    assert lib.svo_server_wait(None, 1, None, None) == 1
    assert lib.svo_server_destroy(None) == 1
    assert lib.svo_march_mode(None, 0) == 1
    assert lib.svo_autotune(None, 1) == 1
    assert lib.svo_voxelize_batch(None, None, None, 0, None) == 1
    assert lib.svo_build_batch(None, None, 0, None) == 1
    assert lib.svo_march_batch(None, None, None, 0, 8, 8, 0, None) == 1
    assert lib.svo_march_plan_destroy(None) == 1
    assert lib.svo_march_plan_set_shape(None, 0) == 1


def test_path_trace_and_obj_argument_checks():
    # argument errors before any device work (no GPU needed)
    import ctypes

    from sparsevoxeloctree_amd._capi import lib

    sky = (ctypes.c_float * 3)(1.0, 1.0, 1.0)
    assert lib.svo_path_trace(None, None, None, 8, 8, 4, 0, sky, None, None, None) == 1
    assert lib.svo_bounce_rays(None, None, 8, 8, 0, 1, None) == 1
    xyz = ctypes.POINTER(ctypes.c_float)()
    rgb = ctypes.POINTER(ctypes.c_uint32)()
    T = ctypes.c_uint64()
    assert lib.svo_obj_load(None, ctypes.byref(xyz), ctypes.byref(rgb), ctypes.byref(T), None, 0) == 1
    msg = ctypes.create_string_buffer(128)
    assert lib.svo_obj_load(b"/nonexistent/x.obj", ctypes.byref(xyz), ctypes.byref(rgb), ctypes.byref(T), msg, 128) == 1
    assert b"cannot read" in msg.value and not xyz and not rgb and T.value == 0
    lib.svo_obj_free(None, None)  # freeing nothing is allowed
    env = ctypes.c_void_p()
    rgba = (ctypes.c_float * 32)()
    assert lib.svo_envmap_create(None, 4, 2, rgba, ctypes.byref(env)) == 1 and not env
    assert lib.svo_envmap_download(None, None, None, None) == 1
    assert lib.svo_envmap_destroy(None) == 1
    assert lib.svo_path_trace_env(None, None, None, 8, 8, 4, 0, None, 0.0, 1.0, None, None, None) == 1


def test_build_plan_and_scene_update_argument_checks():
    # argument errors before any device work (no GPU needed)
    import ctypes

    from sparsevoxeloctree_amd._capi import lib

    plan = ctypes.c_void_p()
    lv = (ctypes.c_uint32 * 1)(5)
    assert lib.svo_build_plan_create(None, None, lv, 1, None, None, ctypes.byref(plan)) == 1 and not plan
    assert lib.svo_build_plan_use_graph(None, 1) == 1
    assert lib.svo_build_plan_run(None) == 1
    assert lib.svo_build_plan_finish(None, None, None, None) == 1
    assert lib.svo_build_plan_get(None, 0, None, None) == 1
    assert lib.svo_build_plan_destroy(None) == 1
    assert lib.svo_scene_update(None, None, None, None, 1) == 1
    assert lib.svo_scene_update_device(None, None, None, None, 1) == 1
