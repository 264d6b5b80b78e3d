"""Captured voxelize + build (svo_build_plan_*) against the plain batched entry points and the oracle:
every fragment list, octree word and hit record bit-exact, on replays of one captured graph with
new geometry between them, and with plain launches of the same passes.

The plan is new (the reference voxelizes and builds each scene once and reads the fragment count and
the octree size back in between, src/Voxelizer.cpp:134-179, src/OctreeBuilder.cpp:142-214); its
contract is that a run that fits its capacities equals svo_voxelize_batch + svo_build_batch of the
same scenes, and that a run that does not fit writes nothing past them and says so.
"""
import math

import numpy as np
import pytest

import oracle
from svo_testutil import hits_equal

pytestmark = pytest.mark.gpu

# (scene kind 4 triangle budget, level): uneven sizes and depths, level 1, a scene with no triangles
SCENES = [(1000, 3), (30000, 6), (9000, 8), (5000, 1), (20, 4), (20000, 7), (0, 5), (4032, 10)]


def _rotate(xyz, a, b):
    """The triangle soup turned by a about z and b about x (then normalised again by the scene)."""
    ca, sa, cb, sb = math.cos(a), math.sin(a), math.cos(b), math.sin(b)
    rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    rx = np.array([[1, 0, 0], [0, cb, -sb], [0, sb, cb]])
    v = xyz.reshape(-1, 3).astype(np.float64) @ (rx @ rz).T
    return v.astype(np.float32).reshape(xyz.shape)


@pytest.fixture(scope="module")
def geometry():
    import sparsevoxeloctree_amd as S

    out = []
    for i, (T, _) in enumerate(SCENES):
        if T == 0:
            out.append((np.zeros((0, 9), np.float32), np.zeros(0, np.uint32)))
            continue
        xyz, rgb = S.gen_scene(4, 0x5EED0500 + i, T)
        out.append((np.asarray(xyz, np.float32).reshape(-1, 9), np.asarray(rgb, np.uint32)))
    return out


def _variant(geometry, k):
    """Variant k of every scene: k = 0 as generated, k > 0 turned by a per-scene angle."""
    if k == 0:
        return geometry
    return [(_rotate(x, 0.37 * k + 0.11 * i, 0.23 * k), c) for i, (x, c) in enumerate(geometry)]


def _reference(ctx, geo, levels):
    """Fragment lists and octree words of the plain batched entry points."""
    import sparsevoxeloctree_amd as S

    scenes = [S.Scene.Create(ctx, x, c) for x, c in geo]
    voxs = S.voxelize_batch(scenes, levels)
    octs = S.build_batch(voxs)
    return [v.GetVoxelFragmentList() for v in voxs], [o.words() for o in octs]


def _check(plan, frags, words, what):
    for i in range(plan.n):
        got_f = plan.fragments(i).GetVoxelFragmentList()
        assert np.array_equal(got_f, frags[i]), "%s: fragments of scene %d" % (what, i)
        assert plan.F[i] == frags[i].shape[0], (what, i)
        got_w = plan.octree(i).words()
        assert got_w.shape == words[i].shape, "%s: W of scene %d: %s vs %s" % (what, i, got_w.shape, words[i].shape)
        assert np.array_equal(got_w, words[i]), "%s: octree of scene %d" % (what, i)
        assert plan.W[i] == words[i].shape[0], (what, i)


def test_plan_equals_batch_and_oracle(ctx, no_device_errors, geometry):
    import sparsevoxeloctree_amd as S

    levels = [lv for _, lv in SCENES]
    scenes = [S.Scene.Create(ctx, x, c) for x, c in geometry]
    print("plan over %d scenes, measured capacities" % len(scenes), flush=True)
    plan = S.BuildPlan(scenes, levels)
    plan.finish()  # before any run: no counts, no error
    assert plan.F == [0] * plan.n and plan.W == [0] * plan.n
    frags, words = _reference(ctx, geometry, levels)
    for rep in range(3):  # replays of one graph, with eager work and waits between them
        plan.run()
        ms = plan.finish()
        assert ms > 0.0
        _check(plan, frags, words, "replay %d" % rep)
        ctx.synchronize()
    for i, lv in enumerate(levels):
        assert np.array_equal(words[i], oracle.build(frags[i], lv)), i


@pytest.mark.parametrize("graph", [True, False])
def test_plan_steady_state_updates(ctx, no_device_errors, geometry, graph):
    """Scene.update between runs (new geometry, the same triangle counts): every run equals a fresh
    batched voxelize + build of the updated scenes.  Capacities are the largest counts of the
    variants, so some scenes fill their list or octree exactly."""
    import sparsevoxeloctree_amd as S

    levels = [lv for _, lv in SCENES]
    variants = [_variant(geometry, k) for k in range(3)]
    refs = [_reference(ctx, v, levels) for v in variants]
    fcap = [max(r[0][i].shape[0] for r in refs) for i in range(len(SCENES))]
    wcap = [max(r[1][i].shape[0] for r in refs) for i in range(len(SCENES))]
    scenes = [S.Scene.Create(ctx, x, c) for x, c in variants[0]]
    plan = S.BuildPlan(scenes, levels, fcap, wcap)
    plan.use_graph(graph)
    for k in [0, 1, 2, 1, 0]:
        print("graph=%d variant %d" % (graph, k), flush=True)
        for sc, (x, c) in zip(scenes, variants[k]):
            if sc.T:
                sc.update(x, c)
        plan.run()
        plan.finish()
        _check(plan, refs[k][0], refs[k][1], "graph=%d variant %d" % (graph, k))


def test_plan_overflow_then_fit(ctx, no_device_errors, geometry):
    """A run whose fragment or word counts pass the capacities reports SVO_ERR_OVERFLOW with the
    counts it needed and leaves the plan usable: the next run that fits is exact again."""
    import sparsevoxeloctree_amd as S

    levels = [lv for _, lv in SCENES]
    frags, words = _reference(ctx, geometry, levels)
    F = [f.shape[0] for f in frags]
    W = [w.shape[0] for w in words]
    scenes = [S.Scene.Create(ctx, x, c) for x, c in geometry]
    big = int(np.argmax(F))
    # fragments: one list one short of its count
    fcap = list(F)
    fcap[big] = F[big] - 1
    plan = S.BuildPlan(scenes, levels, fcap, W)
    print("plan fragment overflow", flush=True)
    plan.run()
    with pytest.raises(S.SvoError) as ei:
        plan.finish()
    assert ei.value.code == 5 and "fragments" in str(ei.value)
    assert plan.F == F
    del plan
    # words: one octree one block short, the fragment lists fit
    wcap = list(W)
    wcap[big] = W[big] - 8
    plan = S.BuildPlan(scenes, levels, F, wcap)
    print("plan word overflow", flush=True)
    plan.run()
    with pytest.raises(S.SvoError) as ei:
        plan.finish()
    assert ei.value.code == 5 and "words" in str(ei.value)
    assert plan.W[big] == W[big]
    del plan
    # capacities of the generated geometry exactly; a turned variant that needs more overflows, and
    # the same plan is exact again once the generated geometry is back
    plan = S.BuildPlan(scenes, levels, F, W)
    over = None
    for k in range(1, 7):
        fk, wk = _reference(ctx, _variant(geometry, k), levels)
        if any(f.shape[0] > F[i] or w.shape[0] > W[i] for i, (f, w) in enumerate(zip(fk, wk))):
            over = k
            break
    assert over is not None, "no turned variant needs more than the generated geometry"
    for sc, (x, c) in zip(scenes, _variant(geometry, over)):
        if sc.T:
            sc.update(x, c)
    print("plan overflow with variant %d" % over, flush=True)
    plan.run()
    with pytest.raises(S.SvoError) as ei:
        plan.finish()
    assert ei.value.code == 5
    for sc, (x, c) in zip(scenes, geometry):
        if sc.T:
            sc.update(x, c)
    print("plan fits again", flush=True)
    plan.run()
    plan.finish()
    _check(plan, frags, words, "after overflow")


def test_plan_octrees_march(ctx, no_device_errors, geometry):
    """Hit records of the plan's octrees equal those of the batched build's octrees."""
    import sparsevoxeloctree_amd as S

    levels = [lv for _, lv in SCENES]
    scenes = [S.Scene.Create(ctx, x, c) for x, c in geometry]
    plan = S.BuildPlan(scenes, levels)
    plan.run()
    plan.finish()
    voxs = S.voxelize_batch(scenes, levels)
    octs = S.build_batch(voxs)
    W, H = 72, 40
    cams = []
    for i in range(len(scenes)):
        a = 0.7 * i
        eye = (1.5 + 1.6 * math.cos(a), 1.5 + 0.5 * math.sin(1.3 * a), 1.5 + 1.6 * math.sin(a))
        cams.append(S.camera_look_at(eye, (1.5, 1.5, 1.5), fov=math.pi / 3, aspect=W / H).uniform())
    got = S.march_batch([plan.octree(i) for i in range(plan.n)], cams, W, H, beam=True)
    ref = S.march_batch(octs, cams, W, H, beam=True)
    for i in range(plan.n):
        ok, msg = hits_equal(got[i], ref[i])
        assert ok, "octree %d: %s" % (i, msg)


def test_plan_then_march_plan_steady_state(ctx, no_device_errors, geometry):
    """The whole steady-state loop from two graphs: new geometry, build plan run, then a march plan
    made once over the build plan's octrees, with no finish in between; every frame equals a
    fresh batched build + march of the same geometry."""
    import sparsevoxeloctree_amd as S

    levels = [lv for _, lv in SCENES]
    variants = [_variant(geometry, k) for k in range(3)]
    refs = [_reference(ctx, v, levels) for v in variants]
    fcap = [max(r[0][i].shape[0] for r in refs) for i in range(len(SCENES))]
    wcap = [max(r[1][i].shape[0] for r in refs) for i in range(len(SCENES))]
    scenes = [S.Scene.Create(ctx, x, c) for x, c in variants[0]]
    plan = S.BuildPlan(scenes, levels, fcap, wcap)
    plan.run()
    plan.finish()
    W, H = 48, 32
    cams = []
    for i in range(len(scenes)):
        a = 0.9 * i + 0.3
        eye = (1.5 + 1.6 * math.cos(a), 1.5 + 0.4 * math.sin(1.7 * a), 1.5 + 1.6 * math.sin(a))
        cams.append(S.camera_look_at(eye, (1.5, 1.5, 1.5), fov=math.pi / 3, aspect=W / H).uniform())
    mp = S.MarchPlan([plan.octree(i) for i in range(plan.n)], cams, W, H, beam=True)
    mp.use_graph(True)
    hits = mp.alloc_hits()
    for k in [1, 2, 0, 2]:
        print("steady state variant %d" % k, flush=True)
        for sc, (x, c) in zip(scenes, variants[k]):
            if sc.T:
                sc.update(x, c)
        plan.run()
        mp.run(hits)
        plan.finish()
        got = hits.download(S.HIT_DTYPE, plan.n * W * H).reshape(plan.n, H, W)
        ref_scenes = [S.Scene.Create(ctx, x, c) for x, c in variants[k]]
        ref_octs = S.build_batch(S.voxelize_batch(ref_scenes, levels))
        ref = S.march_batch(ref_octs, cams, W, H, beam=True)
        for i in range(plan.n):
            ok, msg = hits_equal(got[i], ref[i])
            assert ok, "variant %d octree %d: %s" % (k, i, msg)


def test_scene_update_from_device_buffers(ctx, no_device_errors, geometry):
    """svo_scene_update_device: new triangles from device buffers feed the next run exactly as a
    scene created from them does."""
    import sparsevoxeloctree_amd as S
    from sparsevoxeloctree_amd._capi import lib

    levels = [lv for _, lv in SCENES]
    turned = _variant(geometry, 2)
    frags, words = _reference(ctx, turned, levels)
    scenes = [S.Scene.Create(ctx, x, c) for x, c in geometry]
    plan = S.BuildPlan(scenes, levels, [f.shape[0] + 64 for f in frags], [w.shape[0] + 64 for w in words])
    keep = []
    for sc, (x, c) in zip(scenes, turned):
        if not sc.T:
            continue
        xd, cd = ctx.upload(np.ascontiguousarray(x, np.float32)), ctx.upload(np.ascontiguousarray(c, np.uint32))
        keep.append((xd, cd))
        ctx.check(lib.svo_scene_update_device(ctx.h, sc.h, xd.ptr, cd.ptr, 1), "scene update from the device")
    plan.run()
    plan.finish()
    _check(plan, frags, words, "device update")
