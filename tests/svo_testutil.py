"""Shared helpers for the tests: fragment packing, seeded inputs, hit-record comparison."""
import numpy as np


def pack_frags(x, y, z, rgb):
    """Fragment packing of shader/voxelizer.frag:38-42, with uint32 wrap-around (the 2^12 overflow)."""
    x = np.asarray(x, dtype=np.uint64)
    y = np.asarray(y, dtype=np.uint64)
    z = np.asarray(z, dtype=np.uint64)
    rgb = np.asarray(rgb, dtype=np.uint64)
    fx = (x | (y << 12) | ((z & 0xFF) << 24)) & 0xFFFFFFFF
    fy = (((z >> 8) << 28) | (rgb & 0xFFFFFF)) & 0xFFFFFFFF
    return np.stack([fx, fy], axis=1).astype(np.uint32)


def random_frags(seed, level, n, extent=None, hot=0):
    """n fragments with coordinates in [0, res] (res itself exercises the saturation of the descent),
    random colours; `hot` extra fragments piled into a handful of voxels."""
    rng = np.random.default_rng(seed)
    res = 1 << level
    ext = res if extent is None else extent
    x = rng.integers(0, ext + 1, n)
    y = rng.integers(0, ext + 1, n)
    z = rng.integers(0, ext + 1, n)
    c = rng.integers(0, 1 << 24, n)
    fr = pack_frags(np.minimum(x, 4095), np.minimum(y, 4095), np.minimum(z, 4095), c)
    if hot:
        hx, hy, hz = rng.integers(0, res, 3)
        hc = rng.integers(0, 1 << 24, hot)
        fr = np.concatenate([fr, pack_frags(np.full(hot, hx), np.full(hot, hy), np.full(hot, hz), hc)])
    perm = rng.permutation(fr.shape[0])
    return np.ascontiguousarray(fr[perm])


def hits_equal(a, b):
    """Bitwise equality of two hit-record arrays; returns (ok, message)."""
    a = a.reshape(-1)
    b = b.reshape(-1)
    if a.shape != b.shape:
        return False, "shape %s vs %s" % (a.shape, b.shape)
    av = a.view(np.uint32).reshape(-1, 8)
    bv = b.view(np.uint32).reshape(-1, 8)
    diff = np.any(av != bv, axis=1)
    if not diff.any():
        return True, ""
    idx = np.nonzero(diff)[0]
    k = idx[0]
    names = ["pos.x", "pos.y", "pos.z", "iter", "vxy", "vz_flags", "leaf", "nfetch"]
    cols = [names[c] for c in range(8) if av[k, c] != bv[k, c]]
    return False, "%d of %d rays differ; first at %d (%s): %s vs %s" % (
        idx.size,
        a.size,
        k,
        ",".join(cols),
        av[k].tolist(),
        bv[k].tolist(),
    )


def hit_mask(h):
    return (h["vz_flags"] >> 31) & 1 == 1
