"""ESVO marching on the GPU against the oracle: every field of every hit record bit-exact.

Reference: shader/octree.glsl (Octree_RayMarchCoarse / RayMarchLeaf 6-arg and 5-arg /
RayMarchOcclude), shader/octree_tracer_beam.frag, shader/octree_tracer.frag, shader/camera.glsl,
shader/diffuse.glsl, shader/util.glsl, src/OctreeTracer.cpp:183-246.
"""
import math

import numpy as np
import pytest

import oracle
from svo_testutil import hit_mask, hits_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene_octree(ctx):
    import sparsevoxeloctree_amd as S

    xyz, rgb = S.gen_scene(1, 0x5EED0001, 24)
    level = 8
    sc = S.Scene.Create(ctx, xyz, rgb)
    vox = S.Voxelizer.Create(sc, level)
    vox.CmdVoxelize()
    b = S.OctreeBuilder.Create(vox)
    b.CmdBuild()
    octree = b.GetOctree()
    words = octree.words()
    return octree, words, level, (b, vox, sc)


def cameras():
    import sparsevoxeloctree_amd as S

    out = []
    for k, (eye, fov) in enumerate(
        [
            ((2.9, 2.1, 0.6), math.pi / 3),
            ((1.5, 1.5, 3.2), math.pi / 3),
            ((0.2, 1.9, 1.1), math.pi / 4),
            ((1.45, 1.55, 1.5), math.pi / 2),  # inside the octree
        ]
    ):
        out.append(S.camera_look_at(eye, (1.5, 1.5, 1.5), fov=fov, aspect=1.25).uniform_array())
    return out


@pytest.mark.parametrize("cam_idx", [0, 1, 2, 3])
@pytest.mark.parametrize("beam", [False, True])
def test_march_frame_matches_oracle(ctx, no_device_errors, scene_octree, cam_idx, beam):
    import sparsevoxeloctree_amd as S

    octree, words, level, _ = scene_octree
    cam = cameras()[cam_idx]
    W, H = 160, 128
    print("march frame cam=%d beam=%d" % (cam_idx, beam), flush=True)
    tr = S.OctreeTracer.Create(octree, _camera_struct(cam))
    tr.Resize(W, H)
    tr.m_beam_enable = beam
    tr.CmdDrawPipeline()
    h_gpu = tr.GetHits()
    b_orc = None
    if beam:
        b_orc, _ = oracle.beam(words, level, cam, W, H)
        assert np.array_equal(tr.GetBeam().view(np.uint32), b_orc.view(np.uint32))
    oracle.reset_counters()
    h_orc = oracle.march_frame(words, level, cam, W, H, b_orc)
    assert oracle.stack_uninit_reads() == 0
    ok, msg = hits_equal(h_gpu, h_orc)
    assert ok, msg
    if cam_idx != 3:
        assert hit_mask(h_gpu).sum() > 100


def _camera_struct(cam):
    from sparsevoxeloctree_amd._capi import svo_camera

    c = svo_camera()
    for i, name in enumerate(("pos", "look", "side", "up")):
        for k in range(4):
            getattr(c, name)[k] = float(cam[4 * i + k])
    return c


def _random_rays(seed, n):
    rng = np.random.default_rng(seed)
    o = np.concatenate(
        [rng.uniform(0.5, 2.5, (n // 2, 3)), rng.uniform(1.0, 2.0, (n - n // 2, 3))]
    ).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    # axis-aligned and tiny components exercise the EPS clamp (>= vs > in the two variants)
    d[: n // 16, 1:] = 0.0
    d[n // 16 : n // 8, 0] = np.float32(3.552713678800501e-15)
    d[n // 8 : n // 8 + n // 32, 2] = -0.0
    return np.concatenate([o, d], axis=1).astype(np.float32)


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_march_rays_match_oracle(ctx, no_device_errors, scene_octree, variant):
    import sparsevoxeloctree_amd as S

    octree, words, level, _ = scene_octree
    rays = _random_rays(77 + variant, 20000)
    h_gpu = S.march_rays(ctx, octree, rays, variant)
    h_orc = oracle.march_rays(words, level, rays, variant)
    ok, msg = hits_equal(h_gpu, h_orc)
    assert ok, msg


@pytest.mark.parametrize("spp,cur", [(0, 1), (5, 1), (1021, 3)])
def test_bounce_rays_and_march_match_oracle(ctx, no_device_errors, scene_octree, spp, cur):
    # bounce `cur` of path-tracer sample spp (shader/path_tracer.comp:42-44): Sobol state after spp
    # advances, fract(Sobol_GetVec2(cur) + blue noise of the pixel), then the 5-argument march
    import sparsevoxeloctree_amd as S
    from sparsevoxeloctree_amd._capi import lib

    octree, words, level, _ = scene_octree
    cam = cameras()[0]
    W, H = 300, 96  # wider than the 256-texel noise tile: the noise wraps at pixel & 255
    tr = S.OctreeTracer.Create(octree, _camera_struct(cam))
    tr.Resize(W, H)
    tr.m_beam_enable = True
    tr.CmdDrawPipeline()
    R = W * H
    d_rays = S.bounce_rays(ctx, tr.hits, W, H, spp, cur)
    rays_gpu = d_rays.download(np.float32, R * 6).reshape(R, 6)
    rays_orc, mask = oracle.bounce_rays(tr.GetHits(), W, H, spp, cur)
    assert np.array_equal(rays_gpu.view(np.uint32), rays_orc.view(np.uint32))
    assert mask.sum() > 100
    d_hits = ctx.alloc(R * 32)
    ctx.check(lib.svo_march_rays(ctx.h, octree.h, d_rays.ptr, R, 1, d_hits.ptr))
    h_gpu = d_hits.download(S.HIT_DTYPE, R)
    ok, msg = hits_equal(h_gpu, oracle.march_rays(words, level, rays_orc, 1))
    assert ok, msg


def test_march_uploaded_noncanonical_octree(ctx, no_device_errors, scene_octree):
    # an octree in a different block order (as the reference's atomics produce) marches to the
    # same hits once canonicalised, and the marcher itself does not care about the order
    import sparsevoxeloctree_amd as S

    octree, words, level, _ = scene_octree
    shuffled = _shuffle_blocks(words, level, seed=5)
    assert not np.array_equal(shuffled, words)
    assert np.array_equal(oracle.canonicalize(shuffled, level), words)
    o2 = S.Octree.FromWords(ctx, shuffled, level)
    rays = _random_rays(3, 4000)
    h_a = S.march_rays(ctx, o2, rays, 0)
    h_b = oracle.march_rays(shuffled, level, rays, 0)
    ok, msg = hits_equal(h_a, h_b)
    assert ok, msg
    h_c = S.march_rays(ctx, octree, rays, 0)
    assert np.array_equal(h_a["vxy"], h_c["vxy"]) and np.array_equal(h_a["iter"], h_c["iter"])


def _shuffle_blocks(words, level, seed):
    """Relayout the non-root blocks in a random order, rewriting child pointers."""
    rng = np.random.default_rng(seed)
    nblk = words.shape[0] // 8
    perm = np.concatenate([[0], 1 + rng.permutation(nblk - 1)])  # new position of old block b
    out = np.zeros_like(words)
    for b in range(nblk):
        blk = words[8 * b : 8 * b + 8].copy()
        internal = ((blk & 0x80000000) != 0) & ((blk & 0x40000000) == 0)
        ptrs = (blk & 0x3FFFFFFF) // 8
        blk[internal] = 0x80000000 | (perm[ptrs[internal]] * 8).astype(np.uint32)
        out[8 * perm[b] : 8 * perm[b] + 8] = blk
    return out


@pytest.mark.parametrize("beam", [False, True])
def test_march_schedules_identical(ctx, no_device_errors, scene_octree, beam):
    # one ray per lane (mode 0, 16x16 frame tiles; mode 2, 8x8 tiles) and persistent waves with
    # per-lane refill (mode 1): the same records
    import sparsevoxeloctree_amd as S
    from sparsevoxeloctree_amd._capi import lib

    octree, words, level, _ = scene_octree
    W, H = 203, 117  # partial 8x8 tiles on both edges
    outs = []
    try:
        for mode in (0, 1, 2):
            ctx.check(lib.svo_march_mode(ctx.h, mode))
            per = []
            for cam_idx in range(4):
                tr = S.OctreeTracer.Create(octree, _camera_struct(cameras()[cam_idx]))
                tr.Resize(W, H)
                tr.m_beam_enable = beam
                tr.CmdDrawPipeline()
                per.append(tr.GetHits())
            rng = np.random.default_rng(5)
            o = rng.uniform(0.5, 2.5, (3001, 3))
            d = rng.normal(size=(3001, 3))
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            rays = np.concatenate([o, d], axis=1).astype(np.float32)
            for variant in (0, 1, 2):
                per.append(S.march_rays(ctx, octree, rays, variant))
            outs.append(per)
    finally:
        ctx.check(lib.svo_march_mode(ctx.h, -1))
    for other in (1, 2):
        for k, (a, b) in enumerate(zip(outs[0], outs[other])):
            ok, msg = hits_equal(a, b)
            assert ok, "mode %d case %d: %s" % (other, k, msg)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("bounce", [1, 2, 4, 16])
def test_path_trace_matches_oracle(ctx, no_device_errors, scene_octree, bounce, mode):
    # progressive path tracing (PathTracer::CmdRender), one path per lane (mode 0) or persistent
    # waves with per-lane refill (mode 1): three samples per pixel accumulated in place, colour,
    # first-hit albedo and normal bit-exact against the oracle after every sample
    import sparsevoxeloctree_amd as S
    from sparsevoxeloctree_amd._capi import lib

    ctx.check(lib.svo_march_mode(ctx.h, mode))
    try:
        _path_trace_case(scene_octree, bounce)
    finally:
        ctx.check(lib.svo_march_mode(ctx.h, -1))


def _pt_buffers(W, H):
    return [np.zeros((W * H, 4), np.float32), np.zeros(W * H, np.uint32), np.zeros(W * H, np.uint32)]


def _pt_compare(pt, ref, s):
    got = [pt.GetColor().reshape(-1, 4), pt.GetAlbedo().reshape(-1), pt.GetNormal().reshape(-1)]
    for g, r, name in zip(got, ref, ("colour", "albedo", "normal")):
        assert np.array_equal(g.view(np.uint32), r.view(np.uint32)), "%s after sample %d" % (name, s)


def _path_trace_case(scene_octree, bounce):
    import sparsevoxeloctree_amd as S

    octree, words, level, _ = scene_octree
    W, H = 99, 61  # partial 8x8 tiles on both edges
    cam = cameras()[1]
    sky = (5.0, 4.0, 3.0)
    pt = S.PathTracer.Create(octree, _camera_struct(cam))
    pt.Resize(W, H)
    pt.m_bounce = bounce
    pt.m_sun_radiance = sky
    ref = _pt_buffers(W, H)
    for s in range(3):
        pt.CmdRender()
        oracle.path_trace(words, level, cam, W, H, bounce, s, np.array(sky, np.float32), *ref)
        _pt_compare(pt, ref, s)
    assert pt.GetSampleCount() == 3
    pt.spp = 1021  # a later sample index: most Sobol columns set
    pt.CmdRender()
    oracle.path_trace(words, level, cam, W, H, bounce, 1021, np.array(sky, np.float32), *ref)
    _pt_compare(pt, ref, 1021)
    pt.Reset()
    pt.CmdRender()  # a restart overwrites the means
    first = oracle.path_trace(words, level, cam, W, H, bounce, 0, np.array(sky, np.float32))
    _pt_compare(pt, first, 0)


def _env_rgba(h, w, seed):
    rng = np.random.default_rng(seed)
    rgba = np.zeros((h, w, 4), np.float32)
    rgba[..., :3] = rng.uniform(0.1, 0.6, (h, w, 3))
    rgba[h // 4, w // 3, :3] *= 30.0  # a sun-like spot
    return rgba


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("bounce", [1, 2, 4])
def test_path_trace_env_matches_oracle(ctx, no_device_errors, scene_octree, bounce, mode):
    # EnvMapPathTrace (shader/path_tracer.comp:62-106), one path per lane (mode 0) or persistent
    # waves (mode 1): the environment map's tables built by the library equal the oracle's, and
    # three accumulated samples of colour, albedo and normal are bit-exact against the oracle after
    # every sample (light samples with occlusion rays, MIS)
    from sparsevoxeloctree_amd._capi import lib

    ctx.check(lib.svo_march_mode(ctx.h, mode))
    try:
        _path_trace_env_case(ctx, scene_octree, bounce)
    finally:
        ctx.check(lib.svo_march_mode(ctx.h, -1))


def _path_trace_env_case(ctx, scene_octree, bounce):
    import sparsevoxeloctree_amd as S

    octree, words, level, _ = scene_octree
    rgba = _env_rgba(24, 48, 7)
    env = S.EnvironmentMap.Create(ctx, rgba)
    env.m_rotation, env.m_multiplier = 0.75, 1.25
    tex, alias = env.tables()
    rtex, ralias = oracle.envmap_tables(rgba)
    assert np.array_equal(tex, rtex) and np.array_equal(alias, ralias)
    W, H = 99, 61
    cam = cameras()[1]
    pt = S.PathTracer.Create(octree, _camera_struct(cam))
    pt.Resize(W, H)
    pt.m_bounce = bounce
    pt.m_env_map = env
    ref = _pt_buffers(W, H)
    for s in range(3):
        pt.CmdRender()
        oracle.path_trace_env(words, level, cam, W, H, bounce, s, rtex, ralias, env.m_rotation, env.m_multiplier,
                              *ref)
        _pt_compare(pt, ref, s)
    pt.spp = 77
    pt.CmdRender()
    oracle.path_trace_env(words, level, cam, W, H, bounce, 77, rtex, ralias, env.m_rotation, env.m_multiplier, *ref)
    _pt_compare(pt, ref, 77)
    assert np.isfinite(ref[0]).all() and ref[0][:, :3].any()
    env.close()


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("shape,bounce,rotation,multiplier", [((1, 1), 3, 0.0, 1.0), ((3, 5), 16, -2.5, 2.0),
                                                             ((4, 8), 2, 7.0, 0.0)])
def test_path_trace_env_edge_cases(ctx, no_device_errors, scene_octree, mode, shape, bounce, rotation, multiplier):
    # a one-texel map (one alias entry, every direction the same texel), an odd-sized map, the
    # reference's largest bounce count (kMaxBounce = 16, src/Config.hpp:26), rotations beyond
    # +-pi and a zero multiplier: colour, albedo and normal bit-exact against the oracle
    import sparsevoxeloctree_amd as S
    from sparsevoxeloctree_amd._capi import lib

    octree, words, level, _ = scene_octree
    rng = np.random.default_rng(sum(shape) + bounce)
    rgba = np.ones(shape + (4,), np.float32)
    rgba[..., :3] = rng.uniform(0.05, 3.0, shape + (3,))
    env = S.EnvironmentMap.Create(ctx, rgba)
    env.m_rotation, env.m_multiplier = rotation, multiplier
    rtex, ralias = oracle.envmap_tables(rgba)
    tex, alias = env.tables()
    assert np.array_equal(tex, rtex) and np.array_equal(alias, ralias)
    W, H = 67, 45
    cam = cameras()[2]
    ctx.check(lib.svo_march_mode(ctx.h, mode))
    try:
        pt = S.PathTracer.Create(octree, _camera_struct(cam))
        pt.Resize(W, H)
        pt.m_bounce = bounce
        pt.m_env_map = env
        ref = _pt_buffers(W, H)
        for s in range(2):
            pt.CmdRender()
            oracle.path_trace_env(words, level, cam, W, H, bounce, s, rtex, ralias, rotation, multiplier, *ref)
            _pt_compare(pt, ref, s)
    finally:
        ctx.check(lib.svo_march_mode(ctx.h, -1))
    assert np.isfinite(ref[0]).all()
    env.close()


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("frame", [(1, 1), (9, 1), (1, 13)])
def test_path_trace_tiny_frames(ctx, no_device_errors, scene_octree, mode, frame):
    # frames narrower than one 8x8 tile: only the lanes inside the frame write, for both lights
    import sparsevoxeloctree_amd as S
    from sparsevoxeloctree_amd._capi import lib

    octree, words, level, _ = scene_octree
    W, H = frame
    cam = cameras()[0]
    rgba = _env_rgba(8, 16, 3)
    env = S.EnvironmentMap.Create(ctx, rgba)
    rtex, ralias = oracle.envmap_tables(rgba)
    ctx.check(lib.svo_march_mode(ctx.h, mode))
    try:
        for use_env in (False, True):
            pt = S.PathTracer.Create(octree, _camera_struct(cam))
            pt.Resize(W, H)
            pt.m_env_map = env if use_env else None
            pt.CmdRender()
            if use_env:
                ref = oracle.path_trace_env(words, level, cam, W, H, pt.m_bounce, 0, rtex, ralias, 0.0, 1.0)
            else:
                ref = oracle.path_trace(words, level, cam, W, H, pt.m_bounce, 0,
                                        np.array(pt.m_sun_radiance, np.float32))
            _pt_compare(pt, ref, 0)
    finally:
        ctx.check(lib.svo_march_mode(ctx.h, -1))
    env.close()


def test_envmap_rejects_a_dark_map(ctx):
    import sparsevoxeloctree_amd as S

    with pytest.raises(RuntimeError):
        S.EnvironmentMap.Create(ctx, np.zeros((4, 8, 4), np.float32))


@pytest.mark.parametrize("light", ["const", "env"])
@pytest.mark.parametrize("view", [0, 1, 2, 3])
def test_shade_view_types_match_oracle(ctx, no_device_errors, scene_octree, view, light):
    # OctreeTracer::m_view_type (shader/octree_tracer.frag:37-47): normal, position and the
    # iteration heat map bit-exact; diffuse (pow(colour, 1/2.2): device powf against the host
    # libm's) within one step of the 8-bit channel, and exact for nearly every pixel
    import sparsevoxeloctree_amd as S

    octree, words, level, _ = scene_octree
    cam = cameras()[0]
    W, H = 160, 128
    tr = S.OctreeTracer.Create(octree, _camera_struct(cam))
    tr.Resize(W, H)
    tr.m_view_type = view
    tr.m_light = (0.8, 0.35, 2.0)
    env = None
    if light == "env":
        rgba = _env_rgba(16, 32, 9)
        env = S.EnvironmentMap.Create(ctx, rgba)
        env.m_rotation = 1.1
        tr.m_env_map = env
    tr.CmdDrawPipeline()
    img = tr.GetImage().reshape(-1)
    hits = tr.GetHits()
    assert hit_mask(hits).any() and (~hit_mask(hits)).any()
    tex = oracle.envmap_tables(rgba)[0] if env is not None else None
    ref = oracle.shade(hits, cam, W, H, view, tr.m_light, tex, 1.1)
    if view == 0:
        a = (img[:, None] >> np.arange(0, 32, 8, dtype=np.uint32)) & 0xFF
        b = (ref[:, None] >> np.arange(0, 32, 8, dtype=np.uint32)) & 0xFF
        assert np.abs(a.astype(int) - b.astype(int)).max() <= 1
        assert (img == ref).mean() > 0.99
    else:
        assert np.array_equal(img, ref)
    assert np.all(img >> 24 == 0xFF)
    if env is not None:
        env.close()


def test_octree_transfer_orders_the_main_stream(ctx, no_device_errors):
    # OctreeBuilder::CmdTransferOctreeOwnership (src/LoaderThread.cpp:77-105): build on a loader
    # context, hand the octree to the main context without a host wait, march it there at once;
    # the hits equal the oracle's
    import sparsevoxeloctree_amd as S

    loader = S.Context(0)
    try:
        xyz, rgb = S.gen_scene(1, 0x5EED0011, 30)
        level = 9
        sc = S.Scene.Create(loader, xyz, rgb)
        vox = S.Voxelizer.Create(sc, level)
        vox.CmdVoxelize()
        b = S.OctreeBuilder.Create(vox)
        b.CmdBuild()
        b.CmdTransferOctreeOwnership(ctx)
        octree = S.Octree(ctx, b.h, level, 0)  # the builder's octree, used on the main context
        cam = cameras()[1]
        W, H = 192, 160
        tr = S.OctreeTracer.Create(octree, _camera_struct(cam))
        tr.Resize(W, H)
        tr.m_beam_enable = False
        tr.CmdDrawPipeline()  # enqueued on the main stream right after the hand-off
        hits = tr.GetHits()
        words = oracle.build(oracle.voxelize(oracle.normalize(xyz), rgb, level), level)
        ok, msg = hits_equal(hits, oracle.march_frame(words, level, cam, W, H))
        assert ok, msg
        assert np.array_equal(b.GetOctree().words(), words)
    finally:
        loader.synchronize()
