"""The allocation-free primary march (svo_march_noalloc) through the C ABI: its hit records equal
svo_march's with the beam pre-pass (beam image in pool memory) on two calls in a row, the second
with a smaller frame, without touching the context's allocator; a frame whose beam image exceeds
the static array is refused on the host before any launch."""
import ctypes as C

import numpy as np
import pytest

from svo_testutil import hit_mask, hits_equal

pytestmark = pytest.mark.gpu


def _octree(ctx):
    import sparsevoxeloctree_amd as S

    xyz, rgb = S.gen_scene(1, 0x5EED0001, 24)
    sc = S.Scene.Create(ctx, xyz, rgb)
    vox = S.Voxelizer.Create(sc, 8)
    vox.CmdVoxelize()
    b = S.OctreeBuilder.Create(vox)
    b.CmdBuild()
    return b.GetOctree(), (b, vox, sc)


def test_noalloc_march_equals_svo_march(ctx, no_device_errors):
    import sparsevoxeloctree_amd as S
    from sparsevoxeloctree_amd._capi import lib

    octree, keep = _octree(ctx)
    cams = [S.camera_look_at((2.9, 2.1, 0.6), (1.5, 1.5, 1.5), fov=1.0472, aspect=640 / 480),
            S.camera_look_at((1.5, 1.5, 3.2), (1.5, 1.5, 1.5), fov=1.0472, aspect=200 / 136)]
    for (W, H), cam in zip([(640, 480), (200, 136)], cams):  # the second call's frame is smaller
        print("noalloc march %dx%d" % (W, H), flush=True)
        want = ctx.alloc(W * H * 32)
        got = ctx.alloc(W * H * 32)
        ctx.check(lib.svo_march(ctx.h, octree.h, C.byref(cam.uniform()), W, H, 1, want.ptr, None))
        ctx.synchronize()
        before = ctx.memory()
        ctx.check(lib.svo_march_noalloc(ctx.h, octree.h, C.byref(cam.uniform()), W, H, got.ptr))
        ctx.synchronize()
        after = ctx.memory()
        assert (after["in_use"], after["cached"]) == (before["in_use"], before["cached"]), "the pool was used"
        ref = want.download(S.HIT_DTYPE, W * H)
        ok, msg = hits_equal(got.download(S.HIT_DTYPE, W * H), ref)
        assert ok, "%dx%d: %s" % (W, H, msg)
        assert hit_mask(ref).sum() > 100


def test_noalloc_march_refuses_oversized_frame(ctx, no_device_errors):
    import sparsevoxeloctree_amd as S
    from sparsevoxeloctree_amd._capi import lib

    octree, keep = _octree(ctx)
    cam = S.camera_look_at((2.9, 2.1, 0.6), (1.5, 1.5, 1.5), fov=1.0472, aspect=1.0)
    hits = ctx.alloc(64)
    # (8200/8 + 1)^2 = 1026^2 beam cells > 2^20: refused on the host, nothing is launched
    rc = lib.svo_march_noalloc(ctx.h, octree.h, C.byref(cam.uniform()), 8200, 8200, hits.ptr)
    assert rc == 1  # SVO_ERR_INVALID_ARGUMENT
    assert b"beam cells" in lib.svo_last_error(ctx.h)
    ctx.synchronize()
