"""Octree build on the GPU (svo_build through the C ABI) against the oracle: bit-exact words.

Reference: shader/octree_{init,tag,alloc}_node.comp, octree_modify_arg.comp,
src/OctreeBuilder.cpp:142-214.  The oracle restates those passes with a serial allocator and the
canonical colour-fold order, so canonical GPU output must equal it word for word.
"""
import numpy as np
import pytest

import oracle
from svo_testutil import pack_frags, random_frags

pytestmark = pytest.mark.gpu


def gpu_build(ctx, frags, level):
    import sparsevoxeloctree_amd as S

    vox = S.Voxelizer.FromFragments(ctx, frags, level)
    b = S.OctreeBuilder.Create(vox)
    b.CmdBuild()
    octree = b.GetOctree()
    return octree.words()


@pytest.mark.parametrize("level", [1, 2, 3, 4, 5, 6, 8, 10])
def test_build_random_matches_oracle(ctx, no_device_errors, level):
    n = min(20000, 6 * (1 << (2 * level)))
    fr = random_frags(1000 + level, level, n)
    print("build random level=%d F=%d" % (level, fr.shape[0]), flush=True)
    w_gpu = gpu_build(ctx, fr, level)
    w_orc = oracle.build(fr, level)
    assert w_gpu.shape == w_orc.shape
    assert np.array_equal(w_gpu, w_orc)


def test_build_empty(ctx, no_device_errors):
    w = gpu_build(ctx, np.zeros((0, 2), np.uint32), 5)
    assert w.shape == (8,) and not w.any()


def test_build_single_fragment_closed_form(ctx, no_device_errors):
    # one fragment -> the root-children block plus L-1 allocated blocks: W = 8L
    L = 7
    fr = pack_frags([93], [5], [127], [0x123456])
    w = gpu_build(ctx, fr, L)
    assert w.shape == (8 * L,)
    assert np.array_equal(w, oracle.build(fr, L))
    assert (w != 0).sum() == L
    assert w[w != 0][-1] == 0xC1123456


def test_build_hot_voxels_big_buckets(ctx, no_device_errors):
    # one voxel with 40 fragments (block sorter in LDS) and one with 5000 (global-memory sorter)
    L = 6
    base = random_frags(7, L, 3000)
    hot1 = pack_frags(np.full(40, 3), np.full(40, 4), np.full(40, 5), np.arange(40) * 977 % (1 << 24))
    rng = np.random.default_rng(9)
    hot2 = pack_frags(np.full(5000, 60), np.full(5000, 2), np.full(5000, 33), rng.integers(0, 1 << 24, 5000))
    fr = np.concatenate([base, hot1, hot2])
    fr = np.ascontiguousarray(fr[rng.permutation(fr.shape[0])])
    assert np.array_equal(gpu_build(ctx, fr, L), oracle.build(fr, L))


def test_build_saturating_and_overflowing_coordinates(ctx, no_device_errors):
    # coordinate == res saturates in the descent (L <= 11); at L = 12, x = 4096 spills into y's bit 0
    for L in (3, 9):
        res = 1 << L
        fr = pack_frags([res, res - 1, 0, res], [0, res, res, 5], [res, 2, res, res - 1], [1, 2, 3, 4])
        assert np.array_equal(gpu_build(ctx, fr, L), oracle.build(fr, L))
    fr = pack_frags([4096, 17, 4095], [7, 4095, 4095], [4096, 0, 4095], [0xABCDEF, 0x010203, 0xFFFFFF])
    assert np.array_equal(gpu_build(ctx, fr, 12), oracle.build(fr, 12))


def test_build_dense_cube_closed_form(ctx, no_device_errors):
    L = 4
    res = 1 << L
    g = np.arange(res)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    fr = pack_frags(x.ravel(), y.ravel(), z.ravel(), (x * 7 + y * 11 + z * 13).ravel())
    w = gpu_build(ctx, fr, L)
    assert w.shape[0] == 8 * (1 + sum(8**k for k in range(1, L)))
    assert np.all(w & 0x80000000)
    assert np.array_equal(w, oracle.build(fr, L))


@pytest.mark.parametrize("level", [7, 9, 11, 12])
def test_build_sparse_deep_levels(ctx, no_device_errors, level):
    # sparse scatter at the deepest levels: Morton prefixes beyond 32 bits at 2^12, 8^(L-1)-bit
    # occupancy bitmaps, one fragment per leaf block
    fr = random_frags(31 + level, level, 30000)
    print("build sparse level=%d" % level, flush=True)
    assert np.array_equal(gpu_build(ctx, fr, level), oracle.build(fr, level))


def test_build_repeatable(ctx, no_device_errors):
    fr = random_frags(55, 9, 50000, hot=300)
    w1 = gpu_build(ctx, fr, 9)
    w2 = gpu_build(ctx, fr[::-1].copy(), 9)  # input order must not matter
    assert np.array_equal(w1, w2)
    assert np.array_equal(w1, oracle.build(fr, 9))


def test_build_bin_size_classes(ctx, no_device_errors):
    # leaf blocks holding 1..4, 5..8, 9..16 (thread sorter, three networks), 17..64 (wave shuffle
    # sorter), 65..512 (wave LDS sorter), 513..2048 (block LDS sorter) and more (in-place HBM sort)
    # fragments, every class boundary exactly once, spread over the 8 leaves with repeated
    # coordinates and repeated full fragments
    L = 7
    rng = np.random.default_rng(77)
    nodes = rng.choice(1 << (3 * (L - 1)), size=252, replace=False)
    edges = [1, 4, 5, 8, 9, 16, 17, 64, 65, 512, 513, 2048, 2049]
    classes = [(1, 9), (9, 17), (17, 65), (65, 513), (513, 2049), (2049, 2400)]  # [lo, hi)
    parts = []
    for i, P in enumerate(nodes):
        n = edges[i] if i < len(edges) else int(rng.integers(*classes[i % len(classes)]))
        X = Y = Z = 0
        for b in range(L - 1):  # de-interleave the Morton index of the depth-(L-1) node
            X |= ((int(P) >> (3 * b)) & 1) << b
            Y |= ((int(P) >> (3 * b + 1)) & 1) << b
            Z |= ((int(P) >> (3 * b + 2)) & 1) << b
        s = rng.integers(0, 8, n)
        col = rng.integers(0, 1 << 24, n)
        col[: n // 4] = col[0]  # ties on the full 64-bit fragment
        parts.append(pack_frags(2 * X + (s & 1), 2 * Y + ((s >> 1) & 1), 2 * Z + (s >> 2), col))
    fr = np.concatenate(parts)
    fr = np.ascontiguousarray(fr[rng.permutation(fr.shape[0])])
    print("build bin classes F=%d" % fr.shape[0], flush=True)
    assert np.array_equal(gpu_build(ctx, fr, L), oracle.build(fr, L))
