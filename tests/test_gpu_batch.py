"""Batched voxelize / build / march (svo_*_batch, march plans) against the single-scene entry points
and the oracle: every fragment list, octree word and hit record bit-exact.

The batched forms are new (the reference handles one scene per Voxelizer / OctreeBuilder /
OctreeTracer, src/Voxelizer.cpp:5-179, src/OctreeBuilder.cpp:142-214, src/OctreeTracer.cpp:183-246);
their contract is that every scene's result equals the single-scene result.
"""
import math

import numpy as np
import pytest

import oracle
from svo_testutil import hits_equal, pack_frags, random_frags

pytestmark = pytest.mark.gpu

# (scene kind 4 triangle budget, level): uneven sizes and depths, one scene without triangles
SCENES = [(1000, 3), (30000, 6), (9000, 8), (5000, 1), (20, 4), (20000, 7), (4032, 5)]


@pytest.fixture(scope="module")
def batch(ctx):
    import sparsevoxeloctree_amd as S

    scenes = []
    for i, (T, level) in enumerate(SCENES):
        xyz, rgb = S.gen_scene(4, 0x5EED0400 + i, T)
        scenes.append(S.Scene.Create(ctx, xyz, rgb))
    levels = [lv for _, lv in SCENES]
    print("batch voxelize %d scenes" % len(scenes), flush=True)
    voxs = S.voxelize_batch(scenes, levels)
    print("batch build", flush=True)
    octs = S.build_batch(voxs)
    return scenes, levels, voxs, octs


def single(ctx, scene, level):
    import sparsevoxeloctree_amd as S

    vox = S.Voxelizer.Create(scene, level)
    vox.CmdVoxelize()
    b = S.OctreeBuilder.Create(vox)
    b.CmdBuild()
    return vox, b, b.GetOctree()


def test_voxelize_batch_equals_single(ctx, no_device_errors, batch):
    scenes, levels, voxs, _ = batch
    for i, (sc, lv) in enumerate(zip(scenes, levels)):
        vox, _, _ = single(ctx, sc, lv)
        a, b = voxs[i].GetVoxelFragmentList(), vox.GetVoxelFragmentList()
        assert a.shape == b.shape, (i, a.shape, b.shape)
        assert np.array_equal(a, b), i


def test_build_batch_equals_single_and_oracle(ctx, no_device_errors, batch):
    scenes, levels, voxs, octs = batch
    for i, (sc, lv) in enumerate(zip(scenes, levels)):
        _, _, o1 = single(ctx, sc, lv)
        w_batch, w_single = octs[i].words(), o1.words()
        assert np.array_equal(w_batch, w_single), i
        assert np.array_equal(w_batch, oracle.build(voxs[i].GetVoxelFragmentList(), lv)), i


def test_build_batch_host_fragments(ctx, no_device_errors):
    # host fragment lists at every level, sparse 2^12, hot voxels, an empty list, the 2^12 overflow
    import sparsevoxeloctree_amd as S

    lists = []
    for L in range(1, 13):
        lists.append((random_frags(900 + L, L, 400 * L, hot=30 if L > 2 else 0), L))
    lists.append((np.zeros((0, 2), np.uint32), 6))
    lists.append((pack_frags([4096, 17, 4095], [7, 4095, 4095], [4096, 0, 4095], [0xABCDEF, 1, 0xFFFFFF]), 12))
    hot = pack_frags(np.full(300, 5), np.full(300, 9), np.full(300, 2), np.arange(300) * 7919 % (1 << 24))
    lists.append((np.concatenate([random_frags(5, 6, 2000), hot]), 6))
    voxs = [S.Voxelizer.FromFragments(ctx, fr, L) for fr, L in lists]
    print("batch build of %d host lists" % len(voxs), flush=True)
    octs = S.build_batch(voxs)
    for (fr, L), o in zip(lists, octs):
        assert np.array_equal(o.words(), oracle.build(fr, L)), L


def _cams(n, phase=0.0, aspect=1.0):
    import sparsevoxeloctree_amd as S

    out = []
    for i in range(n):
        a = 0.7 * i + phase
        eye = (1.5 + 1.6 * math.cos(a), 1.5 + 0.5 * math.sin(1.3 * a), 1.5 + 1.6 * math.sin(a))
        out.append(S.camera_look_at(eye, (1.5, 1.5, 1.5), fov=math.pi / 3, aspect=aspect).uniform())
    return out


def _single_frames(octs, cams, W, H, beam):
    import sparsevoxeloctree_amd as S

    frames = []
    for o, c in zip(octs, cams):
        tr = S.OctreeTracer.Create(o, c)
        tr.Resize(W, H)
        tr.m_beam_enable = beam
        tr.CmdDrawPipeline()
        frames.append(tr.GetHits())
    return frames


@pytest.mark.parametrize("beam", [False, True])
def test_march_batch_equals_single(ctx, no_device_errors, batch, beam):
    import sparsevoxeloctree_amd as S

    _, _, _, octs = batch
    W, H = 72, 40  # not a multiple of any tile
    cams = _cams(len(octs), aspect=W / H)
    print("march batch beam=%d" % beam, flush=True)
    got = S.march_batch(octs, cams, W, H, beam=beam)
    ref = _single_frames(octs, cams, W, H, beam)
    for i in range(len(octs)):
        ok, msg = hits_equal(got[i], ref[i])
        assert ok, "octree %d: %s" % (i, msg)


def test_march_plan_shapes_and_graph(ctx, no_device_errors, batch):
    import sparsevoxeloctree_amd as S

    _, _, _, octs = batch
    W, H = 48, 56
    cams = _cams(len(octs), aspect=W / H)
    ref = _single_frames(octs, cams, W, H, True)
    plan = S.MarchPlan(octs, cams, W, H, beam=True)
    hits = plan.alloc_hits()
    for shape in (0, 1, 2, 3):
        print("plan shape %d" % shape, flush=True)
        plan.set_shape(shape)
        plan.run(hits)
        got = hits.download(S.HIT_DTYPE, len(octs) * W * H).reshape(len(octs), H, W)
        for i in range(len(octs)):
            ok, msg = hits_equal(got[i], ref[i])
            assert ok, "shape %d octree %d: %s" % (shape, i, msg)
    # graph replays with new cameras between runs (the table update is ordered before the replay)
    plan.set_shape(-1)
    plan.use_graph(True)
    for rep in range(3):
        cams = _cams(len(octs), phase=0.4 * rep + 0.1, aspect=W / H)
        print("plan graph replay %d" % rep, flush=True)
        plan.set_cameras(cams)
        plan.run(hits)
        got = hits.download(S.HIT_DTYPE, len(octs) * W * H).reshape(len(octs), H, W)
        ref = _single_frames(octs, cams, W, H, True)
        for i in range(len(octs)):
            ok, msg = hits_equal(got[i], ref[i])
            assert ok, "replay %d octree %d: %s" % (rep, i, msg)


def test_autotuned_plan(ctx, no_device_errors, batch):
    import sparsevoxeloctree_amd as S

    _, _, _, octs = batch
    W, H = 64, 64
    cams = _cams(len(octs), phase=2.0)
    S.autotune(ctx, True)
    try:
        got = S.march_batch(octs, cams, W, H, beam=True)
        report = S.autotune(ctx, True)
    finally:
        S.autotune(ctx, False)
    assert "march_frame_batch S=%d 64x64 beam=1" % len(octs) in report
    ref = _single_frames(octs, cams, W, H, True)
    for i in range(len(octs)):
        ok, msg = hits_equal(got[i], ref[i])
        assert ok, "octree %d: %s" % (i, msg)


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_march_rays_batch_equals_single(ctx, no_device_errors, batch, variant):
    import sparsevoxeloctree_amd as S

    _, _, _, octs = batch
    rng = np.random.default_rng(40 + variant)
    sets = []
    for i in range(len(octs)):
        n = [0, 1, 300, 77, 1024, 5, 640][i % 7]
        o = rng.uniform(0.6, 2.4, (n, 3))
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        sets.append(np.concatenate([o, d], axis=1).astype(np.float32))
    print("ray batch variant=%d" % variant, flush=True)
    got = S.march_rays_batch(octs, sets, variant)
    for i, (o, rs) in enumerate(zip(octs, sets)):
        if rs.shape[0] == 0:
            assert got[i].shape[0] == 0
            continue
        ref = S.march_rays(ctx, o, rs, variant)
        ok, msg = hits_equal(got[i], ref)
        assert ok, "octree %d: %s" % (i, msg)
