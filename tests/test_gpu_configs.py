"""Every benchmark configuration at its stated size on the GPU, against the oracle.

The inputs are the benchmark's own (bench.CONFIGS, the same generators, seeds and cameras):
  c1  64 UV spheres (258,048 triangles) at 2^8, 512x512 primary rays from the seeded orbit camera:
      fragment list, octree words, beam image and every hit record against the oracle, in full;
  c2  the 990,800-triangle room at 2^10, 1920x1080 primary rays with the beam, then bounce 1 of
      sample 0 and its 5-argument march: everything against the oracle, in full;
  c3  7,999,488 triangles at 2^12 (about 1.8e8 fragments, 9e8 octree words): the fragment list
      against the oracle in full; the octree against an independent vectorised builder (unique
      Morton prefixes per depth give every internal word and every leaf's count; sampled leaves'
      colour folds by tests/pyref.py); the 3840x2160 frame's beam image in full and row bands of
      hit records against the oracle marching the GPU-built octree;
  c4  all 256 scenes batched against one scene per launch (octree words and hit records), and one
      scene of every level 6..11 against the oracle;
  c5  the serving loop with builds of sphere + sliver scenes overlapping ray queries from several
      client threads; every served octree and every hit record against the oracle.
"""
import ctypes as C
import math
import os
import sys
import threading
import time

import numpy as np
import pytest

import oracle
import pyref
from svo_testutil import hit_mask, hits_equal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the benchmark's workloads: configs, generators, cameras)

pytestmark = pytest.mark.gpu


def _gpu_build(ctx, S, xyz, rgb, level):
    sc = S.Scene.Create(ctx, xyz, rgb)
    vox = S.Voxelizer.Create(sc, level)
    vox.CmdVoxelize()
    b = S.OctreeBuilder.Create(vox)
    b.CmdBuild()
    octree = b.GetOctree()
    return octree, vox, (sc, vox, b)


def _march(ctx, S, octree, cam, W, H, beam):
    from sparsevoxeloctree_amd._capi import lib

    hits = ctx.alloc(W * H * 32)
    bimg = ctx.alloc(((W + 7) // 8 + 1) * ((H + 7) // 8 + 1) * 4)
    ctx.check(lib.svo_march(ctx.h, octree.h, C.byref(cam.uniform()), W, H, 1 if beam else 0, hits.ptr,
                            bimg.ptr if beam else None))
    bw, bh = (W + 7) // 8 + 1, (H + 7) // 8 + 1
    return hits, hits.download(S.HIT_DTYPE, W * H), (bimg.download(np.float32, bw * bh) if beam else None)


def _full_case(ctx, S, cfg_id):
    cfg = bench.CONFIGS[cfg_id]
    level, W, H = cfg["level"], cfg["width"], cfg["height"]
    xyz, rgb = S.gen_scene(cfg["kind"], cfg["seed"], cfg["param"])
    cam = bench.make_camera(S, cfg, xyz)
    cam_arr = cam.uniform_array()
    print("c%d: T=%d level=%d %dx%d" % (cfg_id, rgb.shape[0], level, W, H), flush=True)
    octree, vox, keep = _gpu_build(ctx, S, xyz, rgb, level)
    frags = vox.GetVoxelFragmentList()
    ref_frags = oracle.voxelize(oracle.normalize(xyz), rgb, level)
    assert np.array_equal(frags, ref_frags), "fragment list differs from the oracle"
    words = octree.words()
    del frags
    t0 = time.time()
    ref_words = oracle.build(ref_frags, level)
    print("c%d: F=%d W=%d oracle build %.1f s" % (cfg_id, ref_frags.shape[0], ref_words.shape[0], time.time() - t0),
          flush=True)
    assert np.array_equal(words, ref_words), "octree differs from the oracle"
    del ref_frags, words
    d_hits, hits, beam = _march(ctx, S, octree, cam, W, H, True)
    ref_beam, _ = oracle.beam(ref_words, level, cam_arr, W, H)
    assert np.array_equal(beam.view(np.uint32), ref_beam.reshape(-1).view(np.uint32)), "beam image differs"
    ok, msg = hits_equal(hits, oracle.march_frame(ref_words, level, cam_arr, W, H, ref_beam))
    assert ok, "primary hits, beam on: " + msg
    _, hits0, _ = _march(ctx, S, octree, cam, W, H, False)
    ok, msg = hits_equal(hits0, oracle.march_frame(ref_words, level, cam_arr, W, H))
    assert ok, "primary hits, beam off: " + msg
    assert hit_mask(hits).mean() > 0.2
    return cfg, octree, ref_words, d_hits, hits, keep


def test_c1_full_size(ctx, no_device_errors):
    import sparsevoxeloctree_amd as S

    cfg, *_ = _full_case(ctx, S, 1)
    assert cfg["param"] == 64


def test_c2_full_size(ctx, no_device_errors):
    # the room at 2^10: primary frame with the beam, then bounce 1 of path-tracer sample 0
    # (shader/path_tracer.comp:42) and the 5-argument march of the bounce rays
    import sparsevoxeloctree_amd as S
    from sparsevoxeloctree_amd._capi import lib

    cfg, octree, ref_words, d_hits, hits, keep = _full_case(ctx, S, 2)
    W, H, level = cfg["width"], cfg["height"], cfg["level"]
    assert ref_words.shape[0] > 7e7
    d_rays = S.bounce_rays(ctx, d_hits, W, H, 0, 1)
    rays = d_rays.download(np.float32, W * H * 6).reshape(-1, 6)
    ref_rays, mask = oracle.bounce_rays(hits, W, H, 0, 1)
    assert np.array_equal(rays.view(np.uint32), ref_rays.view(np.uint32)), "bounce rays differ"
    d_h2 = ctx.alloc(W * H * 32)
    ctx.check(lib.svo_march_rays(ctx.h, octree.h, d_rays.ptr, W * H, 1, d_h2.ptr))
    ok, msg = hits_equal(d_h2.download(S.HIT_DTYPE, W * H), oracle.march_rays(ref_words, level, ref_rays, 1))
    assert ok, "bounce hits: " + msg
    assert mask.mean() > 0.9


def test_c3_full_size(ctx, no_device_errors):
    import sparsevoxeloctree_amd as S

    cfg = bench.CONFIGS[3]
    level, W, H = cfg["level"], cfg["width"], cfg["height"]
    xyz, rgb = S.gen_scene(cfg["kind"], cfg["seed"], cfg["param"])
    assert rgb.shape[0] == 7999488
    cam = bench.make_camera(S, cfg, xyz)
    cam_arr = cam.uniform_array()
    print("c3: T=%d" % rgb.shape[0], flush=True)
    octree, vox, keep = _gpu_build(ctx, S, xyz, rgb, level)
    frags = vox.GetVoxelFragmentList()
    print("c3: F=%d W=%d" % (frags.shape[0], octree.W), flush=True)
    assert frags.shape[0] > 1.5e8
    t0 = time.time()
    ref = oracle.voxelize(oracle.normalize(xyz), rgb, level)
    assert np.array_equal(frags, ref), "fragment list differs from the oracle"
    del ref
    print("c3: fragments identical (oracle %.1f s)" % (time.time() - t0), flush=True)
    words = octree.words()
    t0 = time.time()
    pyref.check_octree_by_prefixes(frags, words, level)
    print("c3: octree checked by prefixes (%.1f s)" % (time.time() - t0), flush=True)
    del frags
    _, hits, beam = _march(ctx, S, octree, cam, W, H, True)
    ref_beam, _ = oracle.beam(words, level, cam_arr, W, H)
    assert np.array_equal(beam.view(np.uint32), ref_beam.reshape(-1).view(np.uint32)), "beam image differs"
    hits = hits.reshape(H, W)
    assert hit_mask(hits).mean() > 0.1
    for y0 in (0, 517, 1080, 1601, H - 16):
        ref_rows = oracle.march_frame_rows(words, level, cam_arr, W, H, y0, y0 + 16, ref_beam)
        ok, msg = hits_equal(hits[y0:y0 + 16], ref_rows)
        assert ok, "rows %d..%d: %s" % (y0, y0 + 16, msg)
    print("c3: beam and 5 bands of 16 rows identical", flush=True)
    # the persistent schedule (the one the benchmark's autotuner picks at this size) on the same
    # frame: every record identical to the one-ray-per-lane march checked above
    from sparsevoxeloctree_amd._capi import lib

    try:
        ctx.check(lib.svo_march_mode(ctx.h, 1))
        _, other, _ = _march(ctx, S, octree, cam, W, H, True)
        ok, msg = hits_equal(other.reshape(H, W), hits)
        assert ok, "persistent schedule: " + msg
    finally:
        ctx.check(lib.svo_march_mode(ctx.h, -1))
    print("c3: persistent schedule identical", flush=True)
    # the same voxelize + build as a captured build plan replay: counts and every octree word
    F = vox.GetVoxelFragmentCount()
    plan = S.BuildPlan([keep[0]], [level])
    plan.run()
    plan.finish()
    assert plan.F == [F] and plan.W == [octree.W], (plan.F, plan.W, F, octree.W)
    assert np.array_equal(plan.octree(0).words(), words), "captured plan's octree differs"
    print("c3: captured build plan identical", flush=True)


def test_c4_all_scenes_batched_single_and_oracle(ctx, no_device_errors):
    # all 256 scenes: batched voxelize / build / march and a captured build plan against one scene
    # per launch, every octree word and hit record; then the smallest scene of every level 6..11
    # against the oracle
    import sparsevoxeloctree_amd as S
    from sparsevoxeloctree_amd._capi import lib

    cfg = bench.CONFIGS[4]
    Ts, levels = bench.c4_scenes(cfg)
    n, W, H = cfg["scenes"], cfg["width"], cfg["height"]
    sample = {}  # the smallest scene of every level: checked against the oracle
    for lv in range(6, 12):
        cand = [i for i in range(n) if levels[i] == lv]
        if cand:
            sample[min(cand, key=lambda i: Ts[i])] = True
    scenes, cams, host = [], [], {}
    for i in range(n):
        xyz, rgb = S.gen_scene(4, cfg["seed"] + (i << 16), Ts[i])
        scenes.append(S.Scene.Create(ctx, xyz, rgb))
        cams.append(bench.orbit_camera(S, xyz, cfg["seed"] + i, W / H))
        if i in sample:
            host[i] = (xyz, rgb)
    print("c4: %d scenes, %d triangles" % (n, sum(Ts)), flush=True)
    voxs = S.voxelize_batch(scenes, levels)
    octs = S.build_batch(voxs)
    bhits = S.march_batch(octs, cams, W, H, beam=True)
    print("c4: batched done", flush=True)
    # the same voxelize + build as one captured build plan over all 256 scenes
    plan = S.BuildPlan(scenes, levels)
    plan.run()
    plan.finish()
    print("c4: captured plan done", flush=True)
    frame = ctx.alloc(W * H * 32)
    for i in range(n):
        vox = S.Voxelizer.Create(scenes[i], levels[i])
        vox.CmdVoxelize()
        b = S.OctreeBuilder.Create(vox)
        b.CmdBuild()
        o = b.GetOctree()
        words = o.words()
        assert np.array_equal(words, octs[i].words()), "scene %d: batched octree differs" % i
        assert np.array_equal(words, plan.octree(i).words()), "scene %d: captured plan's octree differs" % i
        ctx.check(lib.svo_march(ctx.h, o.h, C.byref(cams[i].uniform()), W, H, 1, frame.ptr, None))
        single = frame.download(S.HIT_DTYPE, W * H)
        ok, msg = hits_equal(bhits[i].reshape(-1), single)
        assert ok, "scene %d: batched hits differ: %s" % (i, msg)
        if i in sample:
            xyz, rgb = host[i]
            ref_frags = oracle.voxelize(oracle.normalize(xyz), rgb, levels[i])
            assert np.array_equal(vox.GetVoxelFragmentList(), ref_frags), "scene %d fragments" % i
            ref_words = oracle.build(ref_frags, levels[i])
            assert np.array_equal(words, ref_words), "scene %d octree" % i
            cam_arr = cams[i].uniform_array()
            ref_beam, _ = oracle.beam(ref_words, levels[i], cam_arr, W, H)
            ok, msg = hits_equal(single, oracle.march_frame(ref_words, levels[i], cam_arr, W, H, ref_beam))
            assert ok, "scene %d vs oracle: %s" % (i, msg)
            print("c4: scene %d (level %d, T=%d) identical to the oracle" % (i, levels[i], Ts[i]), flush=True)
    assert len(sample) == 6


def _camera_rays(rng, n):
    v = rng.normal(size=3)
    eye = 1.5 + 1.6 * v / np.linalg.norm(v)
    look = 1.5 - eye + rng.uniform(-0.2, 0.2, 3)
    look /= np.linalg.norm(look)
    side = np.cross(look, [0.0, 1.0, 0.0])
    side /= np.linalg.norm(side)
    up = np.cross(side, look)
    u, w = rng.uniform(-0.6, 0.6, (2, n))
    d = look[None, :] + u[:, None] * side[None, :] + w[:, None] * up[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([np.broadcast_to(eye, d.shape), d], axis=1).astype(np.float32)


def _sphere_rays(rng, n):
    o = rng.uniform(1.0, 2.0, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


def test_c5_served_mix_with_overlapping_builds_matches_oracle(ctx, no_device_errors):
    # sphere + sliver scenes (the c3 generator) built by the server while three client threads keep
    # submitting ray queries (half camera fans, half uniform-sphere rays from in-volume origins,
    # log-uniform sizes) against the octrees already resident; every served octree equals the oracle's
    # build and every served hit record equals the oracle marching that octree
    import sparsevoxeloctree_amd as S
    from sparsevoxeloctree_amd._capi import lib

    rng = np.random.default_rng(bench.CONFIGS[5]["seed"])
    specs = [(int(rng.integers(4, 9 if lv == 11 else 13)), lv) for lv in (8, 9, 10, 8, 9, 10, 11, 9)]
    host, scenes = [], []
    for i, (nsph, lv) in enumerate(specs):
        xyz, rgb = S.gen_scene(3, bench.CONFIGS[5]["seed"] + (i << 16), nsph)
        host.append((xyz, rgb, lv))
        scenes.append(S.Scene.Create(ctx, xyz, rgb))
    ref_words = {}

    def oracle_words(i):
        if i not in ref_words:
            xyz, rgb, lv = host[i]
            ref_words[i] = oracle.build(oracle.voxelize(oracle.normalize(xyz), rgb, lv), lv)
        return ref_words[i]

    srv = S.Server(0, budget_bytes=32 << 30, max_octrees=64)
    lock = threading.Lock()
    ctx_lock = threading.Lock()  # one context: its calls are made from one thread at a time
    resident = {}  # request id -> scene index
    queries = []  # (octree id, scene index, rays, variant, hits buffer, status, t_submit, t_done)
    build_windows = []
    errors = []
    try:
        first = [srv.submit_build(scenes[i], specs[i][1]) for i in range(3)]
        for i, rid in enumerate(first):
            st, _ = srv.wait(rid)
            assert st == 0
            resident[rid] = i
        stop = threading.Event()

        def builder():
            try:
                for i in range(3, len(specs)):
                    t0 = time.perf_counter()
                    rid = srv.submit_build(scenes[i], specs[i][1])
                    st, _ = srv.wait(rid)
                    with lock:
                        build_windows.append((t0, time.perf_counter()))
                        assert st == 0
                        resident[rid] = i
            except Exception as e:  # noqa: BLE001
                errors.append(e)
            finally:
                stop.set()

        def client(k):
            crng = np.random.default_rng(1000 + k)
            try:
                j = 0
                while not stop.is_set() or j < 6:
                    with lock:
                        ids = sorted(resident.items())
                    oid, si = ids[int(crng.integers(len(ids)))]
                    R = int(np.exp(crng.uniform(math.log(4096), math.log(262144))))
                    rays = _camera_rays(crng, R) if j % 2 == 0 else _sphere_rays(crng, R)
                    variant = int(crng.integers(0, 3))
                    with ctx_lock:
                        d_rays, d_hits = ctx.upload(rays), ctx.alloc(R * 32)
                    t0 = time.perf_counter()
                    qid = srv.submit_rays(oid, d_rays, R, d_hits, variant)
                    st, _ = srv.wait(qid)
                    with lock:
                        queries.append((oid, si, rays, variant, d_hits, d_rays, st, t0, time.perf_counter()))
                    j += 1
            except Exception as e:  # noqa: BLE001
                errors.append(e)

        threads = [threading.Thread(target=builder)] + [threading.Thread(target=client, args=(k,)) for k in range(3)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=600)
        assert not errors, errors
        assert all(not t.is_alive() for t in threads)
        print("c5: %d builds, %d queries served" % (len(resident), len(queries)), flush=True)
        # the mix really overlapped: queries submitted while a build was in flight
        overlapped = sum(1 for q in queries for (b0, b1) in build_windows if b0 < q[7] < b1)
        assert overlapped > 0
        for rid, i in resident.items():
            W, h = srv.resident(rid)
            got = np.empty(W, np.uint32)
            ctx.check(lib.svo_octree_download(ctx.h, h, got.ctypes.data_as(C.c_void_p), W))
            assert np.array_equal(got, oracle_words(i)), "served octree of scene %d" % i
        for oid, si, rays, variant, d_hits, d_rays, st, _, _ in queries:
            assert st == 0
            got = d_hits.download(S.HIT_DTYPE, rays.shape[0])
            ok, msg = hits_equal(got, oracle.march_rays(oracle_words(si), host[si][2], rays, variant))
            assert ok, "query on scene %d, variant %d: %s" % (si, variant, msg)
        print("c5: every served octree and hit record identical to the oracle", flush=True)
    finally:
        srv.close()
