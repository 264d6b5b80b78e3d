"""Voxelization + normalisation on the GPU against the oracle.

Reference: src/Scene.cpp:62-99 (normalisation), shader/voxelizer.geom:15-54,
shader/voxelizer.frag:17-44, src/Voxelizer.cpp:112-179.  The GPU list must equal the oracle's list
exactly (same order: triangle order, row-major inside a triangle), which also makes the sorted
multisets equal.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def gpu_voxelize(ctx, xyz, rgb, level, normalize=True):
    import sparsevoxeloctree_amd as S

    sc = S.Scene.Create(ctx, xyz, rgb, normalize=normalize)
    vox = S.Voxelizer.Create(sc, level)
    vox.CmdVoxelize()
    return sc, vox, vox.GetVoxelFragmentList()


def test_normalize_matches_oracle(ctx, no_device_errors):
    import sparsevoxeloctree_amd as S

    for kind, param in ((1, 8), (2, 20), (3, 2)):
        xyz, rgb = S.gen_scene(kind, 0x5EED0000 + kind, param)
        sc = S.Scene.Create(ctx, xyz, rgb, normalize=True)
        assert np.array_equal(sc.download().view(np.uint32), oracle.normalize(xyz).view(np.uint32))


@pytest.mark.parametrize("kind,param,level", [(1, 4, 5), (1, 4, 8), (1, 16, 7), (2, 10, 6), (2, 20, 8), (3, 1, 7)])
def test_voxelize_scenes_match_oracle(ctx, no_device_errors, kind, param, level):
    import sparsevoxeloctree_amd as S

    xyz, rgb = S.gen_scene(kind, 0x5EED0100 + kind, param)
    print("voxelize kind=%d param=%d level=%d T=%d" % (kind, param, level, rgb.shape[0]), flush=True)
    _, vox, fr_gpu = gpu_voxelize(ctx, xyz, rgb, level)
    fr_orc = oracle.voxelize(oracle.normalize(xyz), rgb, level)
    assert vox.GetVoxelFragmentCount() == fr_orc.shape[0]
    assert np.array_equal(fr_gpu, fr_orc)


def test_voxelize_room_face_case(ctx, no_device_errors):
    # the +x/+y/+z walls lie on the +1 face: depth range = res, voxel coordinate = res (kept as is)
    import sparsevoxeloctree_amd as S

    xyz, rgb = S.gen_scene(2, 0x5EED0002, 40)
    level = 6
    _, _, fr = gpu_voxelize(ctx, xyz, rgb, level)
    res = 1 << level
    x = fr[:, 0] & 0xFFF
    y = (fr[:, 0] >> 12) & 0xFFF
    z = (fr[:, 0] >> 24) | ((fr[:, 1] >> 28) << 8)
    assert (x == res).any() and (y == res).any() and (z == res).any()
    assert np.array_equal(fr, oracle.voxelize(oracle.normalize(xyz), rgb, level))


def test_voxelize_level12_random_triangles(ctx, no_device_errors):
    # small random triangles at the maximum level, including the packing overflow at 4096
    rng = np.random.default_rng(12)
    T = 3000
    c = rng.uniform(-1.0, 1.0, (T, 1, 3))
    tri = (c + rng.normal(0.0, 0.002, (T, 3, 3))).clip(-1.0, 1.0).astype(np.float32)
    tri[:50, :, 0] = 1.0  # on the +x face
    rgb = rng.integers(0, 1 << 24, T).astype(np.uint32)
    xyz = tri.reshape(T, 9)
    _, _, fr = gpu_voxelize(ctx, xyz, rgb, 12, normalize=False)
    assert np.array_equal(fr, oracle.voxelize(xyz, rgb, 12))


def test_voxelize_row_spans_axis_parallel_and_long(ctx, no_device_errors):
    # the per-row span trackers: right triangles whose legs are parallel to the rows or columns of
    # the projection (edges that only narrow the row range), and thin slivers crossing hundreds of
    # rows at level 12 (many tracker steps, large edge constants)
    rng = np.random.default_rng(31)
    tris = []
    for _ in range(1500):
        o = rng.uniform(-0.9, 0.9, 3)
        a, b = rng.uniform(0.001, 0.02, 2)
        ax = rng.permutation(3)
        p1, p2 = o.copy(), o.copy()
        p1[ax[0]] += a
        p2[ax[1]] += b
        tris.append(np.concatenate([o, p1, p2]))
    for _ in range(1500):
        o = rng.uniform(-0.8, 0.8, 3)
        d = rng.normal(size=3)
        d *= rng.uniform(0.05, 0.3) / np.linalg.norm(d)
        w = rng.normal(size=3) * 2e-4
        tris.append(np.concatenate([o, o + d, o + d * 0.5 + w]))
    xyz = np.clip(np.array(tris), -1.0, 1.0).astype(np.float32)
    rgb = rng.integers(0, 1 << 24, xyz.shape[0]).astype(np.uint32)
    for level in (10, 12):
        print("row spans level=%d T=%d" % (level, xyz.shape[0]), flush=True)
        _, _, fr = gpu_voxelize(ctx, xyz, rgb, level, normalize=False)
        assert np.array_equal(fr, oracle.voxelize(xyz, rgb, level))


def test_voxelize_degenerate_and_empty(ctx, no_device_errors):
    xyz = np.zeros((3, 9), np.float32)
    xyz[1] = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.5, 0.5, 0.5]  # collinear -> n == 0
    xyz[2] = [0.0, 0.0, 0.0, 0.3, 0.0, 0.0, 0.0, 0.3, 0.0]  # a real triangle
    rgb = np.array([1, 2, 3], np.uint32)
    _, vox, fr = gpu_voxelize(ctx, xyz, rgb, 5, normalize=False)
    ref = oracle.voxelize(xyz, rgb, 5)
    assert np.array_equal(fr, ref)
    assert np.all((fr[:, 1] & 0xFFFFFF) == 3)


def test_voxelize_then_build_end_to_end(ctx, no_device_errors):
    import sparsevoxeloctree_amd as S

    xyz, rgb = S.gen_scene(1, 0x5EED0001, 8)
    level = 8
    _, vox, fr = gpu_voxelize(ctx, xyz, rgb, level)
    b = S.OctreeBuilder.Create(vox)
    b.CmdBuild()
    w = b.GetOctree().words()
    assert np.array_equal(w, oracle.build(oracle.voxelize(oracle.normalize(xyz), rgb, level), level))


def test_obj_file_voxelize_and_build(ctx, no_device_errors, tmp_path):
    # an OBJ scene (svo_obj_load, Scene::load_meshes) through normalisation, voxelization and the
    # build: 16 spheres written with 4 materials, and the polygon fixture
    import os

    import sparsevoxeloctree_amd as S

    xyz, rgb = S.gen_scene(1, 0x5EED0007, 16)
    kd = np.array([[0.8, 0.1, 0.1], [0.1, 0.7, 0.2], [0.5, 0.5, 0.5], [0.2, 0.3, 1.0]])
    (tmp_path / "s.mtl").write_text("".join("newmtl m%d\nKd %g %g %g\n" % ((m,) + tuple(kd[m])) for m in range(4)))
    lines = ["mtllib s.mtl"]
    for t in range(xyz.shape[0]):
        if t % 4032 == 0:
            lines.append("usemtl m%d" % ((t // 4032) % 4))
        for c in range(3):
            lines.append("v %.9g %.9g %.9g" % tuple(xyz[t, 3 * c:3 * c + 3]))
        lines.append("f -3 -2 -1")
    (tmp_path / "s.obj").write_text("\n".join(lines) + "\n")
    here = os.path.dirname(os.path.abspath(__file__))
    for path, level in ((str(tmp_path / "s.obj"), 9), (os.path.join(here, "golden", "obj", "shapes.obj"), 6)):
        oxyz, orgb = S.load_obj(path)
        print("obj %s: %d triangles, level %d" % (os.path.basename(path), oxyz.shape[0], level), flush=True)
        _, vox, fr = gpu_voxelize(ctx, oxyz, orgb, level)
        want = oracle.voxelize(oracle.normalize(oxyz), orgb, level)
        assert np.array_equal(fr, want)
        b = S.OctreeBuilder.Create(vox)
        b.CmdBuild()
        assert np.array_equal(b.GetOctree().words(), oracle.build(want, level))
