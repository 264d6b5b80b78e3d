"""Benchmark of the MI355X SVO hot path (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2] [--no-cpu]

A step is one pass of the hot path over the resident scene: voxelize (count pass, allocation,
fill), level-by-level octree build, beam pre-pass + primary-ray march of the whole frame, and for
the room scene one cosine-weighted diffuse bounce (bounce generation + 5-argument march).
The triangles are uploaded and normalised before timing (the reference loads and normalises on
the host once, src/Scene.cpp:36-147).

Prints one JSON line.  `value` is the primary-ray throughput (Mrays/s, beam pre-pass included)
of the timed steps; with the default configuration the line also carries `batched` (c4: 256
scenes through the batched API against one scene per launch) and `served` (c5: the serving loop
under Poisson arrivals, p50/p99 latency) summaries (skip them with --no-extra); the SVO build time is reported in the reference's scope (fill voxelize + all
level passes, device time, src/LoaderThread.cpp:57-75) and end to end (count + allocation + fill +
build, wall clock).  Per-kernel device times come from HIP events on the library's stream
(svo_prof_*); `roofline` prices the dominant kernel's algorithmic bytes against 8 TB/s.
`cpu_baseline` times the CPU oracle (a C++ port of the reference's algorithm) on the host cores.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SVO build ms @2^10/2^12; primary Mrays/s; batched/served p50/p99 latency"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    1: dict(
        workload="c1: 64 UV spheres (258,048 tris) voxelized at 2^8, SVO build, 512x512 primary rays",
        kind=1, seed=0x5EED0001, param=64, level=8, width=512, height=512, bounce=False, camera="orbit",
    ),
    2: dict(
        workload="c2: box room (walls 0.5% tessellation) + 400 rotated boxes + 200 spheres (990,800 tris) "
        "at 2^10, full build, 1920x1080 primary rays + 1 cosine diffuse bounce",
        kind=2, seed=0x5EED0002, param=1, level=10, width=1920, height=1080, bounce=True, camera="room",
    ),
    # the spec's sizes (radius U[0.02,0.1], sliver edge U[0.002,0.02]) give 1.3e9 fragments and an
    # octree above 2^30 words, which the node format's 30-bit child pointers cannot address; both
    # are scaled by 1/4 (about 1.8e8 fragments, 9.0e8 words)
    3: dict(
        workload="c3: 992 UV spheres (radius U[0.005,0.025]) + 3,999,744 slivers (edge U[0.0005,0.005], "
        "aspect log-U[1,1000]) = 7,999,488 tris at 2^12, full build, 3840x2160 primary rays",
        kind=3, seed=0x5EED0003, param=992 | (25 << 32) | (25 << 48), level=12, width=3840, height=2160,
        bounce=False, camera="orbit",
    ),
    # many small and large scenes: batched launches against one scene per launch
    4: dict(
        workload="c4: 256 scenes of UV spheres (T log-U[1K,2M], level U{6..11}), voxelize + SVO build + "
        "256x256 primary rays each; batched launches vs one scene per launch",
        kind=4, seed=0x5EED0004, scenes=256, width=256, height=256,
    ),
    # serving: Poisson arrivals, 10% builds of c3-type scenes, 90% ray queries (4K..2M rays)
    5: dict(
        workload="c5: serving loop, Poisson arrivals at 1000 QPS for 5 s; 10% builds of c3-type scenes (the c3 "
        "generator and scale: spheres + slivers, 64K..8M tris, log-uniform sphere count 8..992, level "
        "U{9..12}), 90% ray queries of log-U[4K,2M] rays (half camera, half uniform-sphere directions from "
        "in-volume origins) against up to 512 resident octrees",
        seed=0x5EED0005, qps=1000.0, seconds=5.0, initial=64, pool_scenes=24, budget_gib=160, max_octrees=512,
    ),
}


def make_camera(S, cfg, xyz):
    aspect = cfg["width"] / cfg["height"]
    if cfg["camera"] == "room":
        # inside the room near the -z wall, looking along +z (the reference's default yaw/pitch = 0)
        return S.Camera((1.5, 1.45, 1.04), 0.0, 0.0, math.pi / 3.0, aspect)
    # eye on a sphere of radius 1.5 about the world point (0.5,0.5,0.5), mapped to octree space by
    # the scene normalisation (src/Scene.cpp:90-99) and v -> 1.5 + 0.5 v
    p = xyz.reshape(-1, 3).astype(np.float64)
    mn, mx = p.min(axis=0), p.max(axis=0)
    centre, ext = (mn + mx) * 0.5, (mx - mn).max() * 0.5
    rng = np.random.default_rng(cfg["seed"])
    v = rng.normal(size=3)
    v /= np.linalg.norm(v)
    eye_w = np.array([0.5, 0.5, 0.5]) + 1.5 * v
    to_oct = lambda w: 1.5 + 0.5 * (w - centre) / ext  # noqa: E731
    return S.camera_look_at(tuple(to_oct(eye_w)), tuple(to_oct(np.array([0.5, 0.5, 0.5]))), math.pi / 3.0, aspect)


def run_step(S, ctx, scene, cfg, cam, bufs):
    from sparsevoxeloctree_amd._capi import lib

    t0 = time.perf_counter()
    vox = S.Voxelizer.Create(scene, cfg["level"])  # count pass + read-back of F + allocation
    vox.CmdVoxelize()  # fill
    builder = S.OctreeBuilder.Create(vox)
    builder.CmdBuild()
    W = builder.GetOctreeRange() // 4  # waits for the build
    t_build_e2e = time.perf_counter() - t0
    octree = builder.GetOctree()
    w, h = cfg["width"], cfg["height"]
    R = w * h
    ctx.check(lib.svo_march(ctx.h, octree.h, C_cam(cam), w, h, 1, bufs["hits"].ptr, bufs["beam"].ptr), "march")
    if cfg["bounce"]:
        # bounce 1 of path-tracer sample 0 (Sobol dimensions 2-3 + blue noise, shader/path_tracer.comp:42)
        ctx.check(lib.svo_bounce_rays(ctx.h, bufs["hits"].ptr, w, h, 0, 1, bufs["rays"].ptr), "bounce")
        ctx.check(lib.svo_march_rays(ctx.h, octree.h, bufs["rays"].ptr, R, 1, bufs["hits2"].ptr), "bounce march")
    ctx.synchronize()
    wall = time.perf_counter() - t0
    stage = ctx.timing()
    F = vox.GetVoxelFragmentCount()
    return dict(wall=wall, build_e2e=t_build_e2e, stage=stage, F=F, W=W), (vox, builder, octree)


_CAM_CACHE = {}


def C_cam(cam):
    import ctypes

    key = id(cam)
    if key not in _CAM_CACHE:
        _CAM_CACHE[key] = cam.uniform()
    return ctypes.byref(_CAM_CACHE[key])


def cpu_baseline(S, cfg, cam, words, budget_s=20.0):
    """The CPU oracle (port of the reference's traversal) on the host cores: beam pre-pass + primary
    march of a band of rows of the same frame, sized to a bounded amount of work."""
    import oracle

    cores = max(1, min(16, os.cpu_count() or 1))
    level, w, h = cfg["level"], cfg["width"], cfg["height"]
    cam_arr = cam.uniform_array()
    t0 = time.perf_counter()
    beam, _ = oracle.beam(words, level, cam_arr, w, h, threads=cores)
    t_beam = time.perf_counter() - t0
    # probe a small band to size the sample
    y0 = h // 2
    t0 = time.perf_counter()
    oracle.march_frame_rows(words, level, cam_arr, w, h, y0, y0 + 8, beam, threads=cores)
    per_row = (time.perf_counter() - t0) / 8
    rows = int(max(8, min(h, budget_s / max(per_row, 1e-6))))
    y0 = max(0, h // 2 - rows // 2)
    t0 = time.perf_counter()
    oracle.march_frame_rows(words, level, cam_arr, w, h, y0, y0 + rows, beam, threads=cores)
    t_rows = time.perf_counter() - t0
    t = t_rows + t_beam * rows / h  # the beam pass's share for those rows
    reps = 1
    if rows == h and t < budget_s:
        # the whole frame took less than the budget: march it again (beam pre-pass included) as
        # many times as fit
        reps = int(max(1, min(1000, budget_s / max(t, 1e-6))))
        t0 = time.perf_counter()
        for _ in range(reps):
            beam, _ = oracle.beam(words, level, cam_arr, w, h, threads=cores)
            oracle.march_frame_rows(words, level, cam_arr, w, h, 0, h, beam, threads=cores)
        t = time.perf_counter() - t0
    return {
        "value": round(reps * rows * w / t / 1e6, 4),
        "unit": "Mrays/s",
        "cores": cores,
        "kind": "port",
        "sample": "CPU oracle, beam pre-pass + Octree_RayMarchLeaf for rows %d..%d of the %dx%d frame, "
        "%d pass(es) (%d rays, %.1f s)" % (y0, y0 + rows, w, h, reps, reps * rows * w, t),
    }


def c4_scenes(cfg):
    """Triangle budgets and levels of the batch: T log-uniform in [1e3, 2e6], level uniform in 6..11."""
    rng = np.random.default_rng(cfg["seed"])
    n = cfg["scenes"]
    T = np.exp(rng.uniform(math.log(1e3), math.log(2e6), n)).astype(np.int64)
    levels = rng.integers(6, 12, n)
    return [int(t) for t in T], [int(x) for x in levels]


def orbit_camera(S, xyz, seed, aspect):
    """Eye on a sphere of radius 1.5 about the world point (0.5,0.5,0.5), mapped to octree space by the
    scene normalisation (src/Scene.cpp:90-99) and v -> 1.5 + 0.5 v, looking at that point."""
    p = xyz.reshape(-1, 3).astype(np.float64)
    if p.shape[0] == 0:
        return S.camera_look_at((1.5, 1.5, 3.0), (1.5, 1.5, 1.5), math.pi / 3.0, aspect)
    mn, mx = p.min(axis=0), p.max(axis=0)
    centre, ext = (mn + mx) * 0.5, max((mx - mn).max() * 0.5, 1e-30)
    rng = np.random.default_rng(seed)
    v = rng.normal(size=3)
    v /= np.linalg.norm(v)
    to_oct = lambda w: 1.5 + 0.5 * (w - centre) / ext  # noqa: E731
    eye_w = np.array([0.5, 0.5, 0.5]) + 1.5 * v
    return S.camera_look_at(tuple(to_oct(eye_w)), tuple(to_oct(np.array([0.5, 0.5, 0.5]))), math.pi / 3.0, aspect)


def bench_batch(args, cfg):
    """c4: the same 256 scenes voxelized, built and ray-marched (a) one scene per launch through the
    single-scene API and (b) with the batched API (one launch per pass for all scenes, autotuned march
    shape, march launches replayed from a HIP graph).  A step is one pass of (b) over all scenes."""
    import sparsevoxeloctree_amd as S

    Ts, levels = c4_scenes(cfg)
    n, W, H = cfg["scenes"], cfg["width"], cfg["height"]
    ctx = S.Context(0)
    scenes, cams = [], []
    for i in range(n):
        xyz, rgb = S.gen_scene(4, cfg["seed"] + (i << 16), Ts[i])
        scenes.append(S.Scene.Create(ctx, xyz, rgb, normalize=True))
        cams.append(orbit_camera(S, xyz, cfg["seed"] + i, W / H))
    ctx.synchronize()
    Ttot = sum(sc.GetTriangleCount() for sc in scenes)
    R = n * W * H
    hits = ctx.alloc(R * 32)

    def one_by_one():
        octs, keep = [], []
        t0 = time.perf_counter()
        for sc, lv in zip(scenes, levels):
            vox = S.Voxelizer.Create(sc, lv)
            vox.CmdVoxelize()
            b = S.OctreeBuilder.Create(vox)
            b.CmdBuild()
            b.GetOctreeRange()
            keep.append((vox, b))
            octs.append(b.GetOctree())
        ctx.synchronize()
        t_build = time.perf_counter() - t0
        import ctypes

        from sparsevoxeloctree_amd._capi import lib

        beam = ctx.alloc(((W + 7) // 8 + 1) * ((H + 7) // 8 + 1) * 4)
        ucams = [c.uniform() for c in cams]
        t0 = time.perf_counter()
        for i, (o, c) in enumerate(zip(octs, ucams)):
            frame = ctypes.c_void_p(hits.ptr.value + i * W * H * 32)  # frames one after another
            ctx.check(lib.svo_march(ctx.h, o.h, ctypes.byref(c), W, H, 1, frame, beam.ptr), "march")
        ctx.synchronize()
        t_march = time.perf_counter() - t0
        return t_build, t_march, keep

    def batched(plan_cache):
        t0 = time.perf_counter()
        voxs = S.voxelize_batch(scenes, levels)
        ctx.synchronize()
        t_vox = time.perf_counter() - t0
        octs = S.build_batch(voxs, wait=False)
        ctx.synchronize()
        t_build = time.perf_counter() - t0
        split.append((t_vox, t_build - t_vox, sum(v.GetVoxelFragmentCount() for v in voxs)))
        plan = S.MarchPlan(octs, cams, W, H, beam=True)
        plan.use_graph(True)
        plan.run(hits)  # autotune (first time) + graph capture
        ctx.synchronize()
        t0 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            plan.run(hits)
        ctx.synchronize()
        t_march = (time.perf_counter() - t0) / reps
        plan_cache.append((plan, voxs, octs))
        return t_build, t_march, octs

    split = []
    S.autotune(ctx, True)
    # both forms warmed the same way: allocator pools filled and the autotuner's trials for every
    # shape class both use done before any timed pass
    for _ in range(max(1, args.warmup)):
        keep = []
        batched(keep)
        del keep
        one_by_one()
    sb, sm = [], []
    for _ in range(args.steps):  # as many timed passes as the batched form; medians of both
        b, m, keep_seq = one_by_one()
        sb.append(b)
        sm.append(m)
        del keep_seq
    seq_build, seq_march = float(np.median(sb)), float(np.median(sm))
    seq_hits = hits.download(S.HIT_DTYPE, R)
    tb, tm = [], []
    keep = []
    split.clear()
    for _ in range(args.steps):
        keep.clear()
        b, m, octs = batched(keep)
        tb.append(b)
        tm.append(m)
    # per-kernel device times of one more batched pass
    keep.clear()
    ctx.prof_reset()
    ctx.prof_enable(True)
    batched(keep)
    ctx.synchronize()
    ctx.prof_enable(False)
    prof = ctx.prof_stats()
    bat_hits = hits.download(S.HIT_DTYPE, R)
    same = bool(np.array_equal(seq_hits.view(np.uint32), bat_hits.view(np.uint32)))
    # (c) the steady-state form: the same voxelize + build captured once as a HIP graph
    # (svo_build_plan, capacities from one plain pass) and replayed; checked by the counts and by
    # marching its octrees into the same hit records
    bplan = S.BuildPlan(scenes, levels)
    for _ in range(max(1, args.warmup)):
        bplan.run()
        bplan.finish()
    tc, tc_dev = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        bplan.run()
        tc_dev.append(bplan.finish())
        tc.append(time.perf_counter() - t0)
    _, voxs_b, octs_b = keep[-1]
    import ctypes

    from sparsevoxeloctree_amd._capi import lib

    def words_of(o):
        Wo = ctypes.c_uint64()
        lib.svo_octree_info(o.h, None, ctypes.byref(Wo), None)
        return Wo.value

    captured_same = bplan.F == [v.GetVoxelFragmentCount() for v in voxs_b] and bplan.W == [words_of(o) for o in octs_b]
    cplan = S.MarchPlan([bplan.octree(i) for i in range(n)], cams, W, H, beam=True)
    cplan.run(hits)
    ctx.synchronize()
    cap_hits = hits.download(S.HIT_DTYPE, R)
    captured_same = bool(captured_same and np.array_equal(cap_hits.view(np.uint32), bat_hits.view(np.uint32)))
    del cplan, bplan
    t_cap, t_cap_dev = float(np.median(tc)), float(np.median(tc_dev)) / 1e3
    assert ctx.device_errors() == 0
    report = S.autotune(ctx, True)
    Wtot = sum(o.W for o in octs) if octs and octs[0].W else None
    t_build, t_march = float(np.median(tb)), float(np.median(tm))
    line = {
        "metric": METRIC,
        "value": round(R / t_march / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round((t_build + t_march) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": cfg["workload"], "scenes": n, "triangles": Ttot, "levels": levels,
                   "width": W, "height": H, "beam": True},
        "batched_builds_per_s": round(n / t_build, 2),
        "batched_build_ms": round(t_build * 1e3, 3),
        "batched_march_ms": round(t_march * 1e3, 3),
        "sequential_builds_per_s": round(n / seq_build, 2),
        "sequential_build_ms": round(seq_build * 1e3, 3),
        "sequential_march_ms": round(seq_march * 1e3, 3),
        "sequential_mrays_per_s": round(R / seq_march / 1e6, 3),
        "speedup_build": round(seq_build / t_build, 3),
        "speedup_march": round(seq_march / t_march, 3),
        "speedup_total": round((seq_build + seq_march) / (t_build + t_march), 3),
        "hits_identical_to_sequential": same,
        "captured_build_ms": round(t_cap * 1e3, 3),
        "captured_build_device_ms": round(t_cap_dev * 1e3, 3),
        "captured_builds_per_s": round(n / t_cap, 2),
        "speedup_build_captured": round(seq_build / t_cap, 3),
        "speedup_total_captured": round((seq_build + seq_march) / (t_cap + t_march), 3),
        "captured_identical_to_batched": captured_same,
        "fragments": split[-1][2],
        "batched_voxelize_ms": round(float(np.median([x[0] for x in split])) * 1e3, 3),
        "batched_octree_ms": round(float(np.median([x[1] for x in split])) * 1e3, 3),
        "kernel_ms_total": {k: round(v[1], 4) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])},
        "autotune": report.strip().splitlines(),
        "peak_pool_bytes": ctx.memory()["peak_in_use"],
        "cpu_baseline": None,
        "roofline": None,
    }
    ctx.close()
    return line


def bench_serve(args, cfg):
    """c5: a client submits requests at Poisson arrival times to the serving loop (svo_server_*) and a
    collector waits for them; the server batches them on the GPU.  Latencies are submit -> completion
    on the host clock, as the server measures them."""
    import queue
    import threading

    import sparsevoxeloctree_amd as S

    rng = np.random.default_rng(cfg["seed"])
    ctx = S.Context(0)
    # scene pool for build requests, uploaded once: the c3 generator at the c3 scale (spheres of
    # radius U[0.005, 0.025] plus 4032 slivers per sphere), log-uniform sphere counts 8..992
    pool = []
    c3_scale = CONFIGS[3]["param"] & ~0xFFFFFFFF
    for i in range(cfg["pool_scenes"]):
        n = int(np.exp(rng.uniform(math.log(8), math.log(992))))
        xyz, rgb = S.gen_scene(3, cfg["seed"] + (i << 16), n | c3_scale)
        pool.append(S.Scene.Create(ctx, xyz, rgb))
        del xyz, rgb
    # ray pools in HBM: camera fans and uniform-sphere rays, 2M rays each
    RMAX = 2 * 1024 * 1024
    ray_sets = []
    for i in range(8):
        if i % 2 == 0:
            v = rng.normal(size=3)
            eye = 1.5 + 1.6 * v / np.linalg.norm(v)
            look = 1.5 - eye + rng.uniform(-0.2, 0.2, 3)
            look /= np.linalg.norm(look)
            side = np.cross(look, [0.0, 1.0, 0.0])
            side /= np.linalg.norm(side)
            up = np.cross(side, look)
            g = int(np.ceil(np.sqrt(RMAX)))
            u, v = np.meshgrid(np.linspace(-0.6, 0.6, g), np.linspace(-0.6, 0.6, g))
            d = look[None, :] + u.reshape(-1, 1) * side[None, :] + v.reshape(-1, 1) * up[None, :]
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            o = np.broadcast_to(eye, d.shape)
        else:
            o = rng.uniform(1.0, 2.0, (RMAX, 3))
            d = rng.normal(size=(RMAX, 3))
            d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays = np.concatenate([o, d], axis=1).astype(np.float32)
        ray_sets.append(ctx.upload(rays[:RMAX]))
    free_hits = queue.Queue()
    for _ in range(48):
        free_hits.put(ctx.alloc(RMAX * 32))
    srv = S.Server(0, budget_bytes=cfg["budget_gib"] << 30, max_octrees=cfg["max_octrees"])
    levels = list(range(9, 13))
    # initial resident octrees
    resident = []
    init = [srv.submit_build(pool[i % len(pool)], levels[i % len(levels)]) for i in range(cfg["initial"])]
    for rid in init:
        st, _ = srv.wait(rid)
        if st == 0:
            resident.append(rid)
    srv.reset_stats()
    # arrivals
    T, qps = cfg["seconds"], cfg["qps"]
    t_arr = np.cumsum(rng.exponential(1.0 / qps, int(qps * T * 1.2)))
    t_arr = t_arr[t_arr < T]
    kinds = rng.uniform(size=t_arr.size) < 0.1
    Rs = np.exp(rng.uniform(np.log(4096), np.log(RMAX), t_arr.size)).astype(np.int64)
    pending = queue.Queue()
    done = {"lat_b": [], "lat_r": [], "fail": 0, "rej": 0, "retargeted": 0}
    lock = threading.Lock()

    def collector():
        while True:
            item = pending.get()
            if item is None:
                return
            rid, is_build, buf = item
            st, lat = srv.wait(rid)
            with lock:
                if is_build:
                    if st == 0:
                        resident.append(rid)
                    else:
                        done["rej"] += 1
                elif st != 0:
                    done["fail"] += 1
            if buf is not None:
                free_hits.put(buf)

    col = threading.Thread(target=collector)
    col.start()
    t0 = time.perf_counter()
    for k in range(t_arr.size):
        while time.perf_counter() - t0 < t_arr[k]:
            time.sleep(0.0002)
        if kinds[k]:
            rid = srv.submit_build(pool[int(rng.integers(len(pool)))], int(levels[int(rng.integers(4))]))
            pending.put((rid, True, None))
        else:
            buf = free_hits.get()
            rs = ray_sets[int(rng.integers(len(ray_sets)))]
            while True:
                with lock:
                    target = resident[int(rng.integers(len(resident)))]
                try:
                    rid = srv.submit_rays(target, rs, int(Rs[k]), buf, 0)
                    break
                except S.SvoError:  # evicted since the client learnt of it: forget it, pick another
                    with lock:
                        if target in resident:
                            resident.remove(target)
                        done["retargeted"] += 1
            pending.put((rid, False, buf))
    pending.put(None)
    col.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    nb, nr = int(kinds.sum()), int((~kinds).sum())
    # the capacity comparison starts from an empty device: the serving phase's resident octrees
    # (up to the budget) would leave no room for the plain form's octrees
    srv.close()
    capacity = serve_capacity(S, ctx, cfg, rng, pool, levels, ray_sets, free_hits)
    line = {
        "metric": METRIC,
        "value": round(st["rays_traced"] / wall / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generators; scenes and ray pools resident in HBM)",
        "config": {"workload": cfg["workload"], "qps": qps, "seconds": T, "requests": int(t_arr.size),
                   "build_requests": nb, "ray_requests": nr, "initial_octrees": cfg["initial"],
                   "budget_gib": cfg["budget_gib"], "max_octrees": cfg["max_octrees"]},
        "ray_p50_ms": round(st["rays_p50_ms"], 4),
        "ray_p99_ms": round(st["rays_p99_ms"], 4),
        "build_p50_ms": round(st["build_p50_ms"], 4),
        "build_p99_ms": round(st["build_p99_ms"], 4),
        "builds_per_s": round(st["builds_done"] / wall, 2),
        "ray_queries_per_s": round(st["ray_queries_done"] / wall, 2),
        "builds_rejected": st["builds_rejected"],
        "ray_queries_failed": st["ray_queries_failed"],
        "ray_queries_retargeted": done["retargeted"],
        "evictions": st["evictions"],
        "resident_octrees": st["resident_octrees"],
        "resident_gib": round(st["resident_bytes"] / 2**30, 3),
        "gpu_busy_build_s": round(st["busy_build_ms"] / 1e3, 3),
        "gpu_busy_rays_s": round(st["busy_rays_ms"] / 1e3, 3),
        "server_batches": st["batches"],
        "capacity": capacity,
        "cpu_baseline": None,
        "roofline": None,
    }
    ctx.close()
    return line


def serve_capacity(S, ctx, cfg, rng, pool, levels, ray_sets, free_hits, n=1000, n_targets=64, passes=2):
    """Aggregate throughput of the same request list (10% builds, 90% ray queries of log-U[4K,2M] rays
    against n_targets octrees) in the plain form, one request per launch with a wait after each (the
    single-scene API), and in the served form, all requests submitted at once to a fresh serving loop
    (same budget), after the plain form's octrees are freed.  Targets are built beforehand in both
    forms from the same scenes and levels.  Each form runs the list `passes` times in a row; the first
    pass meets the one-time costs (first allocations of each buffer size, code-object loads), so the
    speedup compares the last passes and the first passes are reported beside it."""
    from sparsevoxeloctree_amd._capi import lib

    RMAX = 2 * 1024 * 1024
    kinds = rng.uniform(size=n) < 0.1
    Rs = np.exp(rng.uniform(np.log(4096), np.log(RMAX), n)).astype(np.int64)
    scene_i = rng.integers(len(pool), size=n)
    level_i = rng.integers(len(levels), size=n)
    target_i = rng.integers(n_targets, size=n)
    set_i = rng.integers(len(ray_sets), size=n)
    tgt = [(int(i % len(pool)), int(levels[i % len(levels)])) for i in range(n_targets)]
    bufs = [free_hits.get() for _ in range(8)]
    # plain form: octrees through the single-scene API, then every request alone, waited for
    keep = []
    plain_octs = []
    for sc, lv in tgt:
        vox = S.Voxelizer.Create(pool[sc], lv)
        vox.CmdVoxelize()
        b = S.OctreeBuilder.Create(vox)
        b.CmdBuild()
        keep.append((vox, b))
        plain_octs.append(b.GetOctree())
    ctx.synchronize()
    t_plain = []
    for _ in range(passes):
        rays_plain = 0
        t0 = time.perf_counter()
        for k in range(n):
            if kinds[k]:
                vox = S.Voxelizer.Create(pool[int(scene_i[k])], int(levels[int(level_i[k])]))
                vox.CmdVoxelize()
                b = S.OctreeBuilder.Create(vox)
                b.CmdBuild()
                b.GetOctree()
                ctx.synchronize()
                del vox, b
            else:
                o = plain_octs[int(target_i[k])]
                ctx.check(lib.svo_march_rays(ctx.h, o.h, ray_sets[int(set_i[k])].ptr, int(Rs[k]),
                                             0, bufs[k % len(bufs)].ptr), "march rays")
                ctx.synchronize()
                rays_plain += int(Rs[k])
        t_plain.append(time.perf_counter() - t0)
    del keep, plain_octs
    ctx.trim()
    # served form: the same requests, submitted at once
    srv = S.Server(0, budget_bytes=cfg["budget_gib"] << 30, max_octrees=cfg["max_octrees"])
    ids = [srv.submit_build(pool[sc], lv) for sc, lv in tgt]
    targets = []
    for rid in ids:
        st, _ = srv.wait(rid)
        targets.append(rid if st == 0 else None)
    t_served = []
    failed, refused = 0, 0
    for _ in range(passes):
        srv.reset_stats()
        rays_served = 0
        t0 = time.perf_counter()
        reqs = []
        for k in range(n):
            if kinds[k]:
                reqs.append(srv.submit_build(pool[int(scene_i[k])], int(levels[int(level_i[k])])))
            elif targets[int(target_i[k])] is not None:
                try:
                    reqs.append(srv.submit_rays(targets[int(target_i[k])], ray_sets[int(set_i[k])], int(Rs[k]),
                                                bufs[k % len(bufs)], 0))
                    rays_served += int(Rs[k])
                except S.SvoError:  # the target was evicted by an earlier build of the list
                    targets[int(target_i[k])] = None
                    refused += 1
        for rid in reqs:
            st, _ = srv.wait(rid)
            failed += st != 0
        t_served.append(time.perf_counter() - t0)
    sst = srv.stats()
    srv.close()
    for b in bufs:
        free_hits.put(b)
    tp, ts = t_plain[-1], t_served[-1]
    return {
        "requests": n, "build_requests": int(kinds.sum()), "ray_requests": int((~kinds).sum()),
        "passes": passes, "plain_s": round(tp, 4), "served_s": round(ts, 4),
        "plain_first_pass_s": round(t_plain[0], 4), "served_first_pass_s": round(t_served[0], 4),
        "plain_requests_per_s": round(n / tp, 2), "served_requests_per_s": round(n / ts, 2),
        "plain_mrays_per_s": round(rays_plain / tp / 1e6, 3),
        "served_mrays_per_s": round(rays_served / ts / 1e6, 3),
        "speedup": round(tp / ts, 3), "first_pass_speedup": round(t_plain[0] / t_served[0], 3),
        "served_failed": int(failed), "served_refused": refused,
        "served_batches": sst["batches"], "served_busy_build_s": round(sst["busy_build_ms"] / 1e3, 4),
        "served_busy_rays_s": round(sst["busy_rays_ms"] / 1e3, 4), "served_evictions": sst["evictions"],
        "served_builds_rejected": sst["builds_rejected"],
    }


def copy_gbs(S, ctx, mib=1024, reps=5):
    """Achievable HBM bandwidth of a plain device copy of mib MiB (read + write bytes / time)."""
    from sparsevoxeloctree_amd._capi import lib

    n = mib << 20
    a, b = ctx.alloc(n), ctx.alloc(n)
    ctx.check(lib.svo_copy_device(ctx.h, b.ptr, a.ptr, n))  # warm-up
    ctx.synchronize()
    ctx.prof_reset()
    ctx.prof_enable(True)
    for _ in range(reps):
        ctx.check(lib.svo_copy_device(ctx.h, b.ptr, a.ptr, n))
    ctx.synchronize()
    ctx.prof_enable(False)
    cnt, ms = ctx.prof_stats()["copy_u128"]
    del a, b
    return 2 * n / (ms / cnt * 1e-3) / 1e9


def copy_calibration(mib):
    import sparsevoxeloctree_amd as S
    from sparsevoxeloctree_amd._capi import lib

    n = mib << 20
    with S.Context(0) as ctx:
        a, b = ctx.alloc(n), ctx.alloc(n)
        ctx.prof_enable(True)
        for _ in range(5):
            ctx.check(lib.svo_copy_device(ctx.h, b.ptr, a.ptr, n))
        ctx.synchronize()
        stats = ctx.prof_stats()
        cnt, ms = stats["copy_u128"]
        print(json.dumps({"copy_bytes_per_launch": n, "read_bytes": n, "write_bytes": n, "launches": cnt,
                          "avg_ms": ms / cnt, "gbs": 2 * n / (ms / cnt * 1e-3) / 1e9}))


def bench_primary(args, cfg, cpu=True):
    """One single-scene configuration (c1, c2 or c3): args.warmup untimed and args.steps timed steps
    of voxelize + build + march (+ the bounce for c2); returns the result line."""
    import sparsevoxeloctree_amd as S

    xyz, rgb = S.gen_scene(cfg["kind"], cfg["seed"], cfg["param"])
    T = rgb.shape[0]
    cam = make_camera(S, cfg, xyz)
    ctx = S.Context(0)
    # launch schedules (one ray per lane / persistent refill) picked per shape class by the
    # library's autotuner during the warm-up steps
    S.autotune(ctx, True)
    scene = S.Scene.Create(ctx, xyz, rgb, normalize=True)
    R = cfg["width"] * cfg["height"]
    bw, bh = (cfg["width"] + 7) // 8 + 1, (cfg["height"] + 7) // 8 + 1
    bufs = {"hits": ctx.alloc(R * 32), "beam": ctx.alloc(bw * bh * 4)}
    if cfg["bounce"]:
        bufs["rays"] = ctx.alloc(R * 24)
        bufs["hits2"] = ctx.alloc(R * 32)

    for _ in range(args.warmup):
        run_step(S, ctx, scene, cfg, cam, bufs)
    ctx.synchronize()
    ctx.prof_reset()
    ctx.prof_enable(True)
    results = []
    keep = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r, keep = run_step(S, ctx, scene, cfg, cam, bufs)
        results.append(r)
    ctx.synchronize()
    total = time.perf_counter() - t0
    ctx.prof_enable(False)
    prof = ctx.prof_stats()
    assert ctx.device_errors() == 0

    st = {k: float(np.mean([r["stage"][k] for r in results])) for k in results[0]["stage"]}
    F, W = results[-1]["F"], results[-1]["W"]
    peak_plain = ctx.memory()["peak_in_use"]  # the allocator's peak over the plain steps
    # the same voxelize + build in its steady-state form: captured once as a HIP graph with every
    # count kept on the device (svo_build_plan) and replayed; device ms of a replay and host ms of
    # run + finish, medians over args.steps replays after one warm-up.  Checked below by its
    # counts and by the hit records of its octree.
    bplan = S.BuildPlan([scene], [cfg["level"]])
    bplan.run()
    bplan.finish()
    cap_dev, cap_host = [], []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        bplan.run()
        cap_dev.append(bplan.finish())
        cap_host.append(time.perf_counter() - t1)
    march_ms = st["beam_ms"] + st["march_ms"]
    value = R / (march_ms * 1e-3) / 1e6
    build_ref = st["vox_fill_ms"] + st["build_ms"]
    build_e2e = float(np.mean([r["build_e2e"] for r in results])) * 1e3

    # algorithmic bytes of the kernels (fixed formulas, implementation independent)
    _vox, _builder, octree = keep
    hits = bufs["hits"].download(S.HIT_DTYPE, R)
    nfetch_primary = int(hits["nfetch"].astype(np.uint64).sum())
    from sparsevoxeloctree_amd._capi import lib

    cap_hits = ctx.alloc(R * 32)
    ctx.check(lib.svo_march(ctx.h, bplan.octree(0).h, C_cam(cam), cfg["width"], cfg["height"], 1, cap_hits.ptr, None),
              "march of the plan's octree")
    captured_same = bool(bplan.F == [F] and bplan.W == [W] and np.array_equal(
        cap_hits.download(S.HIT_DTYPE, R).view(np.uint32), hits.view(np.uint32)))
    del cap_hits, bplan
    nf_beam_buf = ctx.alloc(bw * bh * 4)
    beam_tmp = ctx.alloc(bw * bh * 4)
    ctx.check(lib.svo_march_beam(ctx.h, octree.h, C_cam(cam), cfg["width"], cfg["height"], beam_tmp.ptr, nf_beam_buf.ptr))
    nfetch_beam = int(nf_beam_buf.download(np.uint32, bw * bh).astype(np.uint64).sum())
    L = cfg["level"]
    alg = {
        "march_frame": 4 * nfetch_primary + R * 32,
        "march_beam": 4 * nfetch_beam + bw * bh * 4,
        "vox_fill": 40 * T + 4 * T + 8 * F,
        "vox_count": 40 * T + 4 * T,
        # occupancy marks: every fragment read once
        "bm_mark": 8 * F,
        # leaf-block index of every fragment: fragment read, 4-byte bin written
        "build_tag_leaf": 12 * F,
        # bin index and fragment read, fragment written into its bin's bucket
        "build_scatter": 20 * F,
        # the canonical colour fold: every bucketed fragment read once (leaf blocks written: not counted)
        "build_fold": 8 * F,
    }
    nb = None
    if cfg["bounce"]:
        hits2 = bufs["hits2"].download(S.HIT_DTYPE, R)
        nb = int(((hits["vz_flags"] >> 31) & 1).sum())
        alg["march_rays5"] = 4 * int(hits2["nfetch"].astype(np.uint64).sum()) + R * (24 + 32)
        alg["bounce"] = R * (32 + 24)
    kernel_ms = {k: v[1] / max(v[0], 1) for k, v in prof.items()}
    kernel_total = {k: v[1] for k, v in prof.items()}
    dominant = max(kernel_total, key=kernel_total.get)
    copy = copy_gbs(S, ctx)
    # algorithmic GB/s of every kernel with a byte formula, against the peak and the measured copy
    kernel_gbs = {k: {"gbs": round(alg[k] / (kernel_ms[k] * 1e-3) / 1e9, 1),
                      "frac_peak": round(alg[k] / (kernel_ms[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "frac_copy": round(alg[k] / (kernel_ms[k] * 1e-3) / 1e9 / copy, 4)}
                  for k in alg if k in kernel_ms and kernel_ms[k] > 0}
    roofline = None
    if dominant in alg:
        achieved = alg[dominant] / (kernel_ms[dominant] * 1e-3) / 1e9
        traffic = valu_busy = None
        cid = next(k for k, v in CONFIGS.items() if v is cfg)
        tpath = os.path.join(ROOT, "profiles", "traffic_%s_c%d.json" % (dominant, cid))
        if os.path.exists(tpath):
            with open(tpath) as f:
                traffic = json.load(f).get("bytes_per_launch")
        # VALU busy of the same kernel from the SQ counter pass (a march is bound by VALU issue,
        # so its HBM fraction alone says little)
        vpath = os.path.join(ROOT, "profiles", "valu_%s_c%d.json" % (dominant, cid))
        if os.path.exists(vpath):
            with open(vpath) as f:
                valu_busy = json.load(f).get("valu_busy")
        roofline = {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "valu_busy": valu_busy,
            "alg_bytes_per_launch": alg[dominant],
            "avg_launch_ms": round(kernel_ms[dominant], 5),
            "copy_gbs": round(copy, 1),
            "frac_of_copy": round(achieved / copy, 5),
        }
    build_bytes = 8 * F * L + 12 * W  # B_build
    vox_bytes = 40 * T + 8 * F  # B_vox

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {
            "workload": cfg["workload"],
            "triangles": T,
            "level": L,
            "width": cfg["width"],
            "height": cfg["height"],
            "fragments": F,
            "octree_words": W,
            "beam": True,
            "bounce": cfg["bounce"],
        },
        "primary_mrays_per_s": round(value, 3),
        "primary_ms": round(march_ms, 4),
        "build_ms_ref_scope": round(build_ref, 4),
        "build_ms_end_to_end": round(build_e2e, 4),
        "build_ms_captured": round(float(np.median(cap_dev)), 4),
        "build_ms_captured_host": round(float(np.median(cap_host)) * 1e3, 4),
        "captured_identical": captured_same,
        "build_gbs_ref_scope": round((build_bytes + vox_bytes) / (build_ref * 1e-3) / 1e9, 2),
        "stage_ms": {k: round(v, 4) for k, v in st.items()},
        "kernel_ms": {k: round(v, 5) for k, v in sorted(kernel_ms.items(), key=lambda kv: -kernel_total[kv[0]])},
        "roofline": roofline,
        "kernel_gbs": kernel_gbs,
    }
    if cfg["bounce"]:
        line["bounce_rays"] = nb
        line["bounce_mrays_per_s"] = round(nb / (st["rays_ms"] * 1e-3) / 1e6, 3)
        line["path_trace"] = path_trace_rate(S, ctx, octree, cam, cfg)
        env = synthetic_envmap(S, ctx)
        line["path_trace_env"] = dict(path_trace_rate(S, ctx, octree, cam, cfg, env=env), envmap="2048x1024")
        env.close()
    line["peak_pool_bytes"] = peak_plain
    line["autotune"] = S.autotune(ctx, True).strip().splitlines()
    line["reference_lavapipe"] = "unavailable: no Vulkan loader, lavapipe, Xvfb or glslc on the host"
    if cpu:
        words = octree.words()
        line["cpu_baseline"] = cpu_baseline(S, cfg, cam, words, args.cpu_budget)
    else:
        line["cpu_baseline"] = None
    del keep, octree, _vox, _builder
    ctx.close()
    return line


def bench_primary_extra(args, cfg):
    """A single-scene configuration summarised into the c2 line: 3 timed steps after 1 warm-up, no
    CPU leg."""
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup = 3, 1
    return bench_primary(sub, cfg, cpu=False)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2, choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument(
        "--no-extra", action="store_true", help="with --config 2: skip the batched (c4) and served (c5) summaries"
    )
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument(
        "--copy-calibration", type=int, default=0, metavar="MIB",
        help="only run svo_copy_device over MIB MiB, 5 times (the known-size copy used to scale PMC byte counters)",
    )
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(
            "bench.py: this build measures one GPU only (no multi-GPU path); run with --gpus 1",
            file=sys.stderr,
        )
        sys.exit(2)

    import sparsevoxeloctree_amd as S

    cfg = CONFIGS[args.config]
    if args.config in (4, 5):
        line = (bench_batch if args.config == 4 else bench_serve)(args, cfg)
        print(json.dumps(line))
        sys.stdout.flush()
        return
    line = bench_primary(args, cfg, cpu=not args.no_cpu)
    if args.config == 2 and not args.no_extra:
        # the metric's other headline figures: the 2^12 build and 3840x2160 frame (c3), the small
        # plain case (c1), batched (c4) and served (c5), each at its own configuration
        keys = ("build_ms_ref_scope", "build_ms_end_to_end", "build_ms_captured", "captured_identical",
                "build_gbs_ref_scope", "primary_mrays_per_s", "primary_ms", "peak_pool_bytes", "stage_ms", "roofline")
        line["c3"] = summarise_extra(bench_primary_extra, args, CONFIGS[3], keys + ("config",))
        line["c1"] = summarise_extra(bench_primary_extra, args, CONFIGS[1], keys + ("config",))
        line["batched"] = summarise_extra(bench_batch, args, CONFIGS[4], (
            "value", "batched_builds_per_s", "sequential_builds_per_s", "batched_march_ms", "sequential_march_ms",
            "speedup_build", "speedup_march", "speedup_total", "hits_identical_to_sequential", "fragments",
            "captured_build_ms", "speedup_build_captured", "speedup_total_captured", "captured_identical_to_batched"))
        line["served"] = summarise_extra(bench_serve, args, CONFIGS[5], (
            "value", "ray_p50_ms", "ray_p99_ms", "build_p50_ms", "build_p99_ms", "builds_per_s",
            "ray_queries_per_s", "builds_rejected", "ray_queries_failed", "evictions", "capacity"))
    print(json.dumps(line))
    sys.stdout.flush()


def synthetic_envmap(S, ctx, width=2048, height=1024, seed=0x5EED0E):
    """An equirectangular sky of smooth random radiance with a small bright sun (float32 RGBA)."""
    rng = np.random.default_rng(seed)
    rgba = np.ones((height, width, 4), np.float32)
    rgba[..., :3] = rng.uniform(0.3, 1.0, (height, width, 3))
    rgba[height // 5:height // 5 + 4, width // 3:width // 3 + 4, :3] *= 500.0
    return S.EnvironmentMap.Create(ctx, rgba)


def path_trace_rate(S, ctx, octree, cam, cfg, bounce=4, samples=8, env=None):
    """Progressive path tracing of the same frame (PathTracer::CmdRender, 4 bounces, constant sky or
    the environment-map light): full samples per pixel timed on the host clock around synchronised
    launches."""
    pt = S.PathTracer.Create(octree, cam)
    pt.Resize(cfg["width"], cfg["height"])
    pt.m_bounce = bounce
    pt.m_env_map = env
    for _ in range(2):
        pt.CmdRender()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(samples):
        pt.CmdRender()
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / samples
    paths = cfg["width"] * cfg["height"]
    return {"frame": "%dx%d" % (cfg["width"], cfg["height"]), "bounce": bounce, "samples": samples,
            "ms_per_sample": round(dt * 1e3, 4), "mpaths_per_s": round(paths / dt / 1e6, 2)}


def summarise_extra(fn, args, cfg, keys):
    """Runs the c4 / c5 benchmark and keeps its headline fields; an error is recorded, not raised,
    so the primary line is always printed."""
    try:
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup = 2, 1
        full = fn(sub, cfg)
        out = {"workload": cfg["workload"], "unit": full["unit"]}
        out.update({k: full[k] for k in keys})
        return out
    except Exception as exc:  # noqa: BLE001
        return {"workload": cfg["workload"], "error": "%s: %s" % (type(exc).__name__, exc)}


if __name__ == "__main__":
    main()
