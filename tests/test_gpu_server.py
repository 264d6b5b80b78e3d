"""Serving loop (svo_server_*): builds and ray queries from several client threads, resident
octrees, admission by memory and octree count, latency statistics.

Every served result must equal the single-scene path bit for bit: the served octree equals
svo_build of the same scene, and every served hit record equals svo_march_rays of the same rays.
The server is new (the reference's nearest analogue is its loader thread,
src/LoaderThread.cpp:15-116).
"""
import threading

import numpy as np
import pytest

from svo_testutil import hits_equal

pytestmark = pytest.mark.gpu

SVO_ERR_INVALID_ARGUMENT, SVO_ERR_OUT_OF_MEMORY = 1, 3


def _scenes(ctx, n, seed=0x5EED0500):
    import sparsevoxeloctree_amd as S

    out = []
    for i in range(n):
        xyz, rgb = S.gen_scene(4, seed + i, [3000, 12000, 800, 6000, 20000, 4032][i % 6])
        out.append(S.Scene.Create(ctx, xyz, rgb))
    return out


def _single_words(ctx, scene, level):
    import sparsevoxeloctree_amd as S

    vox = S.Voxelizer.Create(scene, level)
    vox.CmdVoxelize()
    b = S.OctreeBuilder.Create(vox)
    b.CmdBuild()
    o = b.GetOctree()
    return o.words(), (vox, b, o)


def _rays(seed, n):
    rng = np.random.default_rng(seed)
    o = rng.uniform(0.7, 2.3, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


def test_server_builds_and_queries_match_single_path(ctx, no_device_errors):
    import sparsevoxeloctree_amd as S

    levels = [6, 8, 5, 7]
    scenes = _scenes(ctx, len(levels))
    srv = S.Server(0, budget_bytes=8 << 30, max_octrees=16)
    try:
        print("server builds", flush=True)
        ids = [srv.submit_build(sc, lv) for sc, lv in zip(scenes, levels)]
        for rid in ids:
            st, lat = srv.wait(rid)
            assert st == 0 and lat > 0.0
        keep = []
        octs = []
        for rid, sc, lv in zip(ids, scenes, levels):
            W, h = srv.resident(rid)
            served = S.Octree(ctx, h, lv, W)
            ref, k = _single_words(ctx, sc, lv)
            keep.append(k)
            assert np.array_equal(served.words(), ref)
            octs.append(k[2])
        # ray queries from several client threads, uneven sizes, every variant, one empty query
        jobs = []
        for q in range(24):
            i = q % len(ids)
            n = [0, 1, 513, 4096, 20000, 77][q % 6]
            rays = _rays(100 + q, n)
            d_rays = ctx.upload(rays if n else np.zeros((1, 6), np.float32))
            d_hits = ctx.alloc(max(n, 1) * 32)
            jobs.append((i, n, q % 3, rays, d_rays, d_hits))
        results = [None] * len(jobs)

        def client(lo, hi):
            for k in range(lo, hi):
                i, n, variant, _, d_rays, d_hits = jobs[k]
                rid = srv.submit_rays(ids[i], d_rays, n, d_hits, variant)
                results[k] = srv.wait(rid)

        print("server ray queries", flush=True)
        threads = [threading.Thread(target=client, args=(t * 6, t * 6 + 6)) for t in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for k, (i, n, variant, rays, _, d_hits) in enumerate(jobs):
            assert results[k][0] == 0, (k, results[k])
            if n == 0:
                continue
            got = d_hits.download(S.HIT_DTYPE, n)
            ref = S.march_rays(ctx, octs[i], rays, variant)
            ok, msg = hits_equal(got, ref)
            assert ok, "query %d: %s" % (k, msg)
        # an unknown octree is refused at submission
        d = ctx.alloc(32 * 4)
        with pytest.raises(S.SvoError):
            srv.submit_rays(987654, d, 1, d, 0)
        stats = srv.stats()
        assert stats["builds_done"] == len(ids)
        assert stats["ray_queries_done"] == len(jobs)
        assert stats["ray_queries_failed"] == 0
        assert stats["rays_p99_ms"] >= stats["rays_p50_ms"] > 0.0
        assert stats["resident_octrees"] == len(ids)
    finally:
        srv.close()


def test_server_admission_and_eviction(ctx, no_device_errors):
    import sparsevoxeloctree_amd as S

    scenes = _scenes(ctx, 5, seed=0x5EED0510)
    # at most 2 resident octrees: the least recently used one is evicted for each new build
    srv = S.Server(0, budget_bytes=8 << 30, max_octrees=2)
    try:
        a, b = srv.submit_build(scenes[0], 6), srv.submit_build(scenes[1], 6)
        assert srv.wait(a)[0] == 0 and srv.wait(b)[0] == 0
        d = ctx.upload(_rays(7, 64))
        h = ctx.alloc(64 * 32)
        assert srv.wait(srv.submit_rays(a, d, 64, h, 0))[0] == 0  # a is now more recent than b
        c = srv.submit_build(scenes[2], 6)
        assert srv.wait(c)[0] == 0
        assert srv.resident(b) is None and srv.resident(a) is not None
        with pytest.raises(S.SvoError) as e:  # evicted: refused at submission
            srv.submit_rays(b, d, 64, h, 0)
        assert e.value.code == SVO_ERR_INVALID_ARGUMENT
        assert srv.stats()["evictions"] == 1
    finally:
        srv.close()
    # a budget smaller than any octree: the build is rejected, nothing stays resident
    srv = S.Server(0, budget_bytes=4096, max_octrees=8)
    try:
        r = srv.submit_build(scenes[3], 7)
        assert srv.wait(r)[0] == SVO_ERR_OUT_OF_MEMORY
        assert srv.resident(r) is None
        st = srv.stats()
        assert st["builds_rejected"] == 1 and st["resident_octrees"] == 0
    finally:
        srv.close()
