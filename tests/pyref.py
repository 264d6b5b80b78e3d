"""Independent pure-Python implementations used to pin the CPU oracle (test infrastructure).

Each one reaches the reference's result by a different route than the oracle's restatement:
  build_canonical  - sort-based: per-depth unique Morton prefixes give the breadth-first layout
                     directly (the oracle re-descends level by level with a serial allocator);
  covers_exact     - exact rational clipping of the triangle against the pixel square (the oracle
                     uses an integer separating-axis test);
  dda_first_hit    - float64 3-D DDA over a dense occupancy grid (the oracle runs ESVO);
  check_octree_by_prefixes - the same sort-based layout, vectorised, as a checker of an octree
                     of 1e8+ fragments (W, every internal word, leaf counts, sampled colours).
"""
from fractions import Fraction

import numpy as np


def effective_coords(frags, level):
    """Voxel coordinates as the reference's descent sees them (octree_tag_node.comp:18-39)."""
    res = 1 << level
    fx = frags[:, 0].astype(np.int64)
    fy = frags[:, 1].astype(np.int64)
    x = np.minimum(fx & 0xFFF, res - 1)
    y = np.minimum((fx >> 12) & 0xFFF, res - 1)
    z = np.minimum((fx >> 24) | ((fy >> 28) << 8), res - 1)
    return x, y, z


def fold(colours):
    """Running average of octree_tag_node.comp:43-57 applied in the given order."""
    word = 0
    for c in colours:
        if word == 0:
            word = 0xC1000000 | c
            continue
        n = (word >> 24) & 0x3F
        w = n + 1
        ch = [(((word >> s) & 0xFF) * n + ((c >> s) & 0xFF)) // w for s in (0, 8, 16)]
        word = (min(w, 63) << 24) | ch[0] | (ch[1] << 8) | (ch[2] << 16) | 0xC0000000
    return word


def morton_prefix(x, y, z, level, depth):
    """Slots of depths 1..depth as a base-8 number (slot = xbit | ybit << 1 | zbit << 2)."""
    code = 0
    for d in range(1, depth + 1):
        s = level - d
        code = (code << 3) | ((x >> s) & 1) | (((y >> s) & 1) << 1) | (((z >> s) & 1) << 2)
    return code


def build_canonical(frags, level):
    frags = np.asarray(frags, dtype=np.uint32).reshape(-1, 2)
    x, y, z = effective_coords(frags, level)
    keys = (frags[:, 0].astype(np.uint64) << np.uint64(32)) | frags[:, 1].astype(np.uint64)
    leaves = {}
    for k in range(frags.shape[0]):
        code = morton_prefix(int(x[k]), int(y[k]), int(z[k]), level, level)
        leaves.setdefault(code, []).append(int(keys[k]))
    colour = {c: fold([k & 0xFFFFFF for k in sorted(ks)]) for c, ks in leaves.items()}
    nodes = {d: sorted({c >> (3 * (level - d)) for c in leaves}) for d in range(1, level + 1)}
    rank = {d: {p: r for r, p in enumerate(nodes[d])} for d in nodes}
    first_block = {}
    acc = 1
    for d in range(1, level):
        first_block[d] = acc  # children blocks of depth-d nodes
        acc += len(nodes[d])
    W = 8 * acc
    words = np.zeros(W, dtype=np.uint32)

    def word_index(d, p):
        if d == 1:
            return p & 7
        parent = p >> 3
        return 8 * (first_block[d - 1] + rank[d - 1][parent]) + (p & 7)

    for d in range(1, level + 1):
        for p in nodes[d]:
            idx = word_index(d, p)
            if d < level:
                words[idx] = 0x80000000 | (8 * (first_block[d] + rank[d][p]))
            else:
                words[idx] = colour[p]
    return words


# ---- exact coverage ------------------------------------------------------------------------------
def _clip(poly, axis, bound, keep_ge):
    out = []
    n = len(poly)
    for k in range(n):
        a, b = poly[k], poly[(k + 1) % n]
        ina = a[axis] >= bound if keep_ge else a[axis] <= bound
        inb = b[axis] >= bound if keep_ge else b[axis] <= bound
        if ina:
            out.append(a)
        if ina != inb:
            t = (bound - a[axis]) / (b[axis] - a[axis])
            out.append((a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])))
    return out


def covers_exact(P, i, j):
    """Does the closed triangle P (3 points, Fractions) meet the half-open square [i,i+1) x [j,j+1)?"""
    poly = [tuple(p) for p in P]
    for axis, bound, ge in ((0, i, True), (0, i + 1, False), (1, j, True), (1, j + 1, False)):
        poly = _clip(poly, axis, Fraction(bound), ge)
        if not poly:
            return False
    if all(p[0] == i + 1 for p in poly) or all(p[1] == j + 1 for p in poly):
        return False
    return True


# ---- DDA ----------------------------------------------------------------------------------------
def dda_first_hit(occ, level, o, d, tie_eps=1e-7):
    """First occupied voxel along the ray in [1,2]^3 (float64 DDA).  Returns (voxel or None, ambiguous):
    ambiguous when the path comes within tie_eps of a voxel edge or corner."""
    res = 1 << level
    o = np.asarray(o, dtype=np.float64)
    d = np.asarray(d, dtype=np.float64)
    tmin, tmax = 0.0, np.inf
    for a in range(3):
        if d[a] == 0.0:
            if not (1.0 <= o[a] <= 2.0):
                return None, False
            continue
        t0, t1 = (1.0 - o[a]) / d[a], (2.0 - o[a]) / d[a]
        if t0 > t1:
            t0, t1 = t1, t0
        tmin, tmax = max(tmin, t0), min(tmax, t1)
    if tmin > tmax:
        return None, False
    p = o + d * tmin
    v = np.clip(np.floor((p - 1.0) * res).astype(np.int64), 0, res - 1)
    step = np.sign(d).astype(np.int64)
    tnext = np.empty(3)
    tdelta = np.empty(3)
    for a in range(3):
        if d[a] > 0:
            tnext[a] = (1.0 + (v[a] + 1) / res - o[a]) / d[a]
            tdelta[a] = 1.0 / (res * d[a])
        elif d[a] < 0:
            tnext[a] = (1.0 + v[a] / res - o[a]) / d[a]
            tdelta[a] = -1.0 / (res * d[a])
        else:
            tnext[a] = np.inf
            tdelta[a] = np.inf
    ambiguous = False
    # entry point near an edge of the first voxel
    # an entry through a face is exact; an entry near an edge, or an origin inside the cube near a
    # voxel face, is ambiguous
    frac = (p - 1.0) * res - v
    near = np.minimum(np.abs(frac), np.abs(1 - frac)) < 1e-6 * res
    if near.sum() >= (2 if tmin > 0 else 1):
        ambiguous = True
    while True:
        if occ[v[0], v[1], v[2]]:
            return (int(v[0]), int(v[1]), int(v[2])), ambiguous
        order = np.sort(tnext)
        if order[1] - order[0] < tie_eps:
            ambiguous = True
        a = int(np.argmin(tnext))
        if tnext[a] > tmax:
            return None, ambiguous
        v[a] += step[a]
        if v[a] < 0 or v[a] >= res:
            return None, ambiguous
        tnext[a] += tdelta[a]


def _spread3(v):
    """Bit i of a 12-bit coordinate to bit 3i (uint64)."""
    v = v.astype(np.uint64)
    out = np.zeros_like(v)
    for i in range(12):
        out |= ((v >> np.uint64(i)) & np.uint64(1)) << np.uint64(3 * i)
    return out


def check_octree_by_prefixes(frags, words, level, nsample=3000, seed=0):
    """The canonical octree of a fragment list, checked without the oracle's builder: the occupied
    Morton prefixes of every depth (coordinates as the reference's descent saturates them,
    shader/octree_tag_node.comp:18-39) determine the breadth-first layout, so every internal word,
    every leaf's count and the total size W follow; sampled leaves' colours are folded by fold().  Vectorised: it
    handles the 1.8e8 fragments of a 2^12 scene.  Raises AssertionError on the first difference."""
    x, y, z = effective_coords(frags, level)
    codes = _spread3(x) | (_spread3(y) << np.uint64(1)) | (_spread3(z) << np.uint64(2))
    del x, y, z
    scodes = np.sort(codes)
    starts = np.flatnonzero(np.concatenate([[True], scodes[1:] != scodes[:-1]]))
    leaves = scodes[starts]
    counts = np.diff(np.append(starts, scodes.size))
    del scodes, starts
    L = level
    nodes = {L: leaves}
    for d in range(L - 1, 0, -1):
        p = nodes[d + 1] >> np.uint64(3)
        nodes[d] = p[np.concatenate([[True], p[1:] != p[:-1]])]
    first_block, acc = {}, 1
    for d in range(1, L):
        first_block[d] = acc
        acc += nodes[d].size
    W = 8 * acc
    assert words.shape[0] == W, "W %d, prefixes give %d" % (words.shape[0], W)
    nonzero_expected = 0
    for d in range(1, L + 1):
        p = nodes[d]
        slot = (p & np.uint64(7)).astype(np.int64)
        if d == 1:
            idx = slot
        else:
            parent_rank = np.searchsorted(nodes[d - 1], p >> np.uint64(3))
            idx = 8 * (first_block[d - 1] + parent_rank) + slot
        w = words[idx]
        if d < L:
            want = (0x80000000 | (8 * (first_block[d] + np.arange(p.size, dtype=np.int64)))).astype(np.uint32)
            assert np.array_equal(w, want), "internal words of depth %d" % d
        else:
            assert np.all((w >> 30) == 3), "leaf flags"
            assert np.array_equal((w >> 24) & 0x3F, np.minimum(counts, 63).astype(np.uint32)), "leaf counts"
            leaf_idx = idx
        nonzero_expected += p.size
    assert int(np.count_nonzero(words)) == nonzero_expected, "words outside the canonical layout"
    # colour folds of sampled leaves, in the canonical (fragment.x, fragment.y) order
    rng = np.random.default_rng(seed)
    pick = rng.choice(leaves.size, size=min(nsample, leaves.size), replace=False)
    want_codes = leaves[pick]
    sel = np.flatnonzero(np.isin(codes, want_codes))
    keys = (frags[sel, 0].astype(np.uint64) << np.uint64(32)) | frags[sel, 1].astype(np.uint64)
    by_code = {}
    for c, k in zip(codes[sel].tolist(), keys.tolist()):
        by_code.setdefault(c, []).append(k)
    for j, c in zip(pick.tolist(), want_codes.tolist()):
        want = fold([k & 0xFFFFFF for k in sorted(by_code[c])])
        assert int(words[leaf_idx[j]]) == want, "leaf colour fold at %d" % j
    return W
