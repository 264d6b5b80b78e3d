"""OBJ + MTL ingestion (svo_obj_load) against the reference's own reader.

The reference loads scenes with tinyobj::LoadObj and packs Kd with glm (src/Scene.cpp:36-164).
oracle/_ref/objref is that reader, built from the reference tree and driven as Scene::load_meshes
does; tests/golden/obj/*.expected hold its output for the fixtures, so the fixture test also runs
where objref is absent.  Positions are compared as bit patterns, colours exactly.
"""
import os
import subprocess

import numpy as np
import pytest

import sparsevoxeloctree_amd as S

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "obj")
OBJREF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "objref")
FIXTURES = ["shapes", "crlf"]


def as_lines(xyz, rgb):
    return [" ".join("%08x" % v for v in xyz[t].view(np.uint32)) + " %06x" % rgb[t] for t in range(len(rgb))]


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_matches_reference_reader_output(name):
    # polygons (concave hexagon, pentagon with collinear first corners), index forms, number
    # formats, unreadable tokens, CRLF, a missing mtllib file, a repeated material name, Kd ties
    xyz, rgb = S.load_obj(os.path.join(GOLD, name + ".obj"))
    with open(os.path.join(GOLD, name + ".expected")) as f:
        expected = f.read().splitlines()
    assert as_lines(xyz, rgb) == expected


def _number(rng):
    # decimal spellings the reader accepts: signs, leading/trailing dots, exponents, long fractions
    v = rng.uniform(-50, 50)
    form = rng.integers(0, 6)
    if form == 0:
        return "%.*f" % (int(rng.integers(0, 12)), v)
    if form == 1:
        return "%.*e" % (int(rng.integers(0, 9)), v)
    if form == 2:
        return ("%.*E" % (int(rng.integers(0, 9)), v)).replace("E+0", "E+")
    if form == 3:
        s = "%.6f" % abs(v)
        return ("-" if v < 0 else "+") + s.lstrip("0")
    if form == 4:
        return "%d." % int(v)
    return repr(float(np.float32(v)))


def _random_obj(tmp_path, seed):
    rng = np.random.default_rng(seed)
    nmat = int(rng.integers(1, 5))
    lines_m = []
    for m in range(nmat):
        lines_m.append("newmtl m%d" % m)
        lines_m.append("Kd %s %s %s" % tuple("%.*f" % (int(rng.integers(1, 5)), rng.uniform(-0.2, 1.2)) for _ in range(3)))
    (tmp_path / "r.mtl").write_text("\n".join(lines_m) + "\n")
    lines = ["mtllib r.mtl"]
    nv = 0
    for face in range(int(rng.integers(5, 60))):
        if rng.random() < 0.3:
            lines.append("usemtl m%d" % rng.integers(0, nmat))
        elif face == 0:
            lines.append("usemtl m0")
        # a star-shaped polygon (simple, often concave) in a random plane
        k = int(rng.integers(3, 9))
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        rad = rng.uniform(0.3, 2.0, k)
        c = rng.uniform(-5, 5, 3)
        u, w = np.linalg.qr(rng.normal(size=(3, 3)))[0][:, :2].T
        if rng.random() < 0.5:  # axis-aligned plane
            ax = rng.permutation(3)[:2]
            u, w = np.eye(3)[ax[0]], np.eye(3)[ax[1]]
        first = nv
        for a, r in zip(ang, rad):
            p = c + r * np.cos(a) * u + r * np.sin(a) * w
            lines.append("v " + " ".join(_number(rng) if rng.random() < 0.1 else repr(float(np.float32(q))) for q in p))
            nv += 1
        idx = []
        for j in range(k):
            i = first + j + 1
            tok = str(i) if rng.random() < 0.7 else str(i - nv - 1)  # 1-based or relative
            form = rng.integers(0, 3)
            idx.append(tok if form == 0 else tok + "/1" if form == 1 else tok + "//1")
        if rng.random() < 0.5:
            idx = idx[::-1]
        lines.append("f " + " ".join(idx))
    lines.insert(1, "vt 0 0")
    lines.insert(1, "vn 0 0 1")
    path = tmp_path / "r.obj"
    path.write_text("\n".join(lines) + "\n")
    return str(path)


@pytest.mark.skipif(not os.path.isfile(OBJREF), reason="oracle/_ref/objref is built only where the reference tree is present")
@pytest.mark.parametrize("seed", range(40))
def test_random_files_match_reference_reader(tmp_path, seed):
    path = _random_obj(tmp_path, seed)
    ref = subprocess.run([OBJREF, path], capture_output=True, text=True)
    assert ref.returncode == 0, ref.stdout
    xyz, rgb = S.load_obj(path)
    assert as_lines(xyz, rgb) == ref.stdout.splitlines()


def test_numbers_are_the_readers_not_correctly_rounded(tmp_path):
    # closed forms: short decimals are exact under the reader's accumulation; exponents move the
    # value by exact powers of 10 for these magnitudes
    (tmp_path / "n.mtl").write_text("newmtl a\nKd 0 0 0\n")
    (tmp_path / "n.obj").write_text("mtllib n.mtl\nusemtl a\nv 0.5 -2 1e2\nv .25 +4. 3E-1\nv 1.5e+1 -0.125 2\nf 1 2 3\n")
    xyz, _ = S.load_obj(str(tmp_path / "n.obj"))
    assert xyz.tolist() == [[0.5, -2.0, 100.0, 0.25, 4.0, np.float32(0.3), 15.0, -0.125, 2.0]]


@pytest.mark.parametrize("body,why", [
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n", "no material"),
    ("mtllib e.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n", "no material"),
    ("mtllib e.mtl\nusemtl a\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", "0 is not an index"),
    ("mtllib e.mtl\nusemtl a\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n", "out of range"),
    ("mtllib e.mtl\nusemtl a\nv 0 0 0\n", "empty mesh"),
])
def test_errors(tmp_path, body, why):
    (tmp_path / "e.mtl").write_text("newmtl a\nKd 1 1 1\n")
    (tmp_path / "e.obj").write_text(body)
    with pytest.raises(S.SvoError) as e:
        S.load_obj(str(tmp_path / "e.obj"))
    assert why in str(e.value)
    with pytest.raises(S.SvoError):
        S.load_obj(str(tmp_path / "missing.obj"))


def test_lone_cr_ends_a_line(tmp_path):
    # a file with CR-only line endings reads like its LF form (tinyobj's safeGetline ends a line at
    # "\n", "\r\n" or a lone "\r", dep/tiny_obj_loader.h:731-760), and a face line followed by a lone
    # CR and more text ends at the CR instead of being read without end
    src = os.path.join(GOLD, "shapes.obj")
    with open(src, "rb") as f:
        lf = f.read().replace(b"\r\n", b"\n")
    (tmp_path / "lf.obj").write_bytes(lf)
    (tmp_path / "cr.obj").write_bytes(lf.replace(b"\n", b"\r"))
    for mtl in os.listdir(GOLD):
        if mtl.endswith(".mtl"):
            with open(os.path.join(GOLD, mtl), "rb") as f:
                (tmp_path / mtl).write_bytes(f.read())
    a = S.load_obj(str(tmp_path / "lf.obj"))
    b = S.load_obj(str(tmp_path / "cr.obj"))
    assert a[1].shape[0] > 0
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)) and np.array_equal(a[1], b[1])
    (tmp_path / "m.mtl").write_bytes(b"newmtl a\nKd 0.5 0.25 1\n")
    (tmp_path / "tail.obj").write_bytes(b"mtllib m.mtl\nusemtl a\nv 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\n"
                                        b"f 1 2 3\r4\nf 1 2 4\n")
    xyz, rgb = S.load_obj(str(tmp_path / "tail.obj"))
    assert rgb.shape[0] == 2  # "4" alone on its line is not a statement
    if os.path.exists(OBJREF):
        for name in ("cr.obj", "tail.obj"):
            r = subprocess.run([OBJREF, str(tmp_path / name)], capture_output=True, text=True, timeout=60)
            assert r.returncode == 0
            assert as_lines(*S.load_obj(str(tmp_path / name))) == r.stdout.splitlines()
