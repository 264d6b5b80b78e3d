"""CPU oracle of the SVO hot path -- TEST INFRASTRUCTURE ONLY.

A C++ restatement of the reference's algorithm (oracle.cpp, built into _build/liboracle.so by
build_oracle.py).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package; the product (sparsevoxeloctree_amd, libsvo.so) never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

HIT_DTYPE = np.dtype(
    [
        ("pos", "<f4", (3,)),
        ("iter", "<u4"),
        ("vxy", "<u4"),
        ("vz_flags", "<u4"),
        ("leaf", "<u4"),
        ("nfetch", "<u4"),
    ]
)

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build_oracle

            build_oracle.build()
        lib = C.CDLL(LIB_PATH)
        P, u64, u32, i32, f32 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_float
        sigs = {
            "orc_normalize": (None, [P, u64]),
            "orc_voxelize": (u64, [P, P, u64, u32, P, u64, i32]),
            "orc_build": (u64, [P, u64, u32, P, u64]),
            "orc_canonicalize": (u64, [P, u64, u32, P, u64]),
            "orc_gen_ray": (None, [P, f32, f32, P]),
            "orc_beam": (None, [P, u64, u32, P, u32, u32, P, P, i32]),
            "orc_march_frame": (None, [P, u64, u32, P, u32, u32, P, P, i32]),
            "orc_march_frame_rows": (None, [P, u64, u32, P, u32, u32, u32, u32, P, P, i32]),
            "orc_march_rays": (None, [P, u64, u32, P, u64, i32, P, i32]),
            "orc_bounce_rays": (None, [P, u32, u32, u32, u32, P, P, P, P]),
            "orc_sobol_state": (None, [P, u32, u32, P]),
            "orc_sincos": (None, [f32, P, P]),
            "orc_path_trace": (None, [P, u64, u32, P, u32, u32, u32, u32, P, P, P, P, P, P, i32]),
            "orc_envmap_tables": (i32, [u32, u32, P, P, P]),
            "orc_path_trace_env": (None, [P, u64, u32, P, u32, u32, u32, u32, P, P, u32, u32, f32, f32, P, P, P, P,
                                          P, i32]),
            "orc_env_sample": (None, [P, P, u32, u32, f32, f32, f32, P, P, P]),
            "orc_env_lookup": (None, [P, u32, u32, f32, P, P, P]),
            "orc_atan2": (f32, [f32, f32]),
            "orc_acos": (f32, [f32]),
            "orc_shade": (None, [P, P, u32, u32, i32, P, P, u32, u32, f32, P]),
            "orc_stack_uninit_reads": (u64, []),
            "orc_reset_counters": (None, []),
        }
        for n, (r, a) in sigs.items():
            fn = getattr(lib, n)
            fn.restype = r
            fn.argtypes = a
        _lib = lib
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def nthreads():
    return max(1, min(16, os.cpu_count() or 1))


def normalize(xyz):
    """src/Scene.cpp:62-99 normalisation; returns a new (T, 9) float32 array."""
    out = np.ascontiguousarray(xyz, dtype=np.float32).copy()
    load().orc_normalize(_p(out), out.size // 3)
    return out


def voxelize(xyz, rgb, level, threads=None):
    """Fragment list (F, 2) uint32 in triangle order, row-major inside a triangle."""
    lib = load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1)
    rgb = np.ascontiguousarray(rgb, dtype=np.uint32).reshape(-1)
    T = rgb.shape[0]
    th = threads or nthreads()
    F = lib.orc_voxelize(_p(xyz), _p(rgb), T, level, None, 0, th)
    out = np.empty(2 * max(F, 1), dtype=np.uint32)
    lib.orc_voxelize(_p(xyz), _p(rgb), T, level, _p(out), F, th)
    return out[: 2 * F].reshape(-1, 2)


def build(frags, level):
    """Canonical octree words (uint32)."""
    lib = load()
    frags = np.ascontiguousarray(frags, dtype=np.uint32).reshape(-1)
    F = frags.shape[0] // 2
    W = lib.orc_build(_p(frags), F, level, None, 0)
    out = np.empty(W, dtype=np.uint32)
    lib.orc_build(_p(frags), F, level, _p(out), W)
    return out


def canonicalize(words, level):
    lib = load()
    words = np.ascontiguousarray(words, dtype=np.uint32)
    n = lib.orc_canonicalize(_p(words), words.shape[0], level, None, 0)
    if n == 0:
        raise ValueError("malformed octree")
    out = np.empty(n, dtype=np.uint32)
    lib.orc_canonicalize(_p(words), words.shape[0], level, _p(out), n)
    return out


def gen_ray(cam, cx, cy):
    cam = np.ascontiguousarray(cam, dtype=np.float32)
    d = np.empty(3, dtype=np.float32)
    load().orc_gen_ray(_p(cam), cx, cy, _p(d))
    return d


def beam(words, level, cam, width, height, threads=None):
    lib = load()
    words = np.ascontiguousarray(words, dtype=np.uint32)
    cam = np.ascontiguousarray(cam, dtype=np.float32)
    bw, bh = (width + 7) // 8 + 1, (height + 7) // 8 + 1
    out = np.empty(bw * bh, dtype=np.float32)
    nf = np.empty(bw * bh, dtype=np.uint32)
    lib.orc_beam(_p(words), words.shape[0], level, _p(cam), width, height, _p(out), _p(nf), threads or nthreads())
    return out.reshape(bh, bw), nf.reshape(bh, bw)


def march_frame(words, level, cam, width, height, beam_img=None, threads=None):
    lib = load()
    words = np.ascontiguousarray(words, dtype=np.uint32)
    cam = np.ascontiguousarray(cam, dtype=np.float32)
    hits = np.empty(width * height, dtype=HIT_DTYPE)
    b = None if beam_img is None else np.ascontiguousarray(beam_img, dtype=np.float32)
    lib.orc_march_frame(_p(words), words.shape[0], level, _p(cam), width, height, _p(b), _p(hits), threads or nthreads())
    return hits.reshape(height, width)


def march_frame_rows(words, level, cam, width, height, y0, y1, beam_img=None, threads=None):
    lib = load()
    words = np.ascontiguousarray(words, dtype=np.uint32)
    cam = np.ascontiguousarray(cam, dtype=np.float32)
    hits = np.empty(width * (y1 - y0), dtype=HIT_DTYPE)
    b = None if beam_img is None else np.ascontiguousarray(beam_img, dtype=np.float32)
    lib.orc_march_frame_rows(
        _p(words), words.shape[0], level, _p(cam), width, height, y0, y1, _p(b), _p(hits), threads or nthreads()
    )
    return hits.reshape(y1 - y0, width)


def march_rays(words, level, rays, variant=0, threads=None):
    lib = load()
    words = np.ascontiguousarray(words, dtype=np.uint32)
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
    hits = np.empty(rays.shape[0], dtype=HIT_DTYPE)
    lib.orc_march_rays(_p(words), words.shape[0], level, _p(rays), rays.shape[0], variant, _p(hits), threads or nthreads())
    return hits


_TABLES = None
DATA_DIR = os.path.join(os.path.dirname(HERE), "sparsevoxeloctree_amd", "data")


def sampler_tables():
    """(Sobol generator matrices uint32[64, 32], blue noise uint8[256 * 256 * 2]): the data tables of the
    reference's path-tracer sampler, read from the files the product library is built from (their
    bytes are checked against the reference's own tables by tests/test_oracle_pins.py)."""
    global _TABLES
    if _TABLES is None:
        sob = np.fromfile(os.path.join(DATA_DIR, "sobol_matrices.bin"), dtype="<u4").reshape(64, 32)
        noise = np.fromfile(os.path.join(DATA_DIR, "blue_noise_rg8.bin"), dtype=np.uint8)
        assert noise.size == 256 * 256 * 2
        _TABLES = (np.ascontiguousarray(sob), np.ascontiguousarray(noise))
    return _TABLES


def sobol_state(n, dims=64, mat=None):
    """The Sobol state after n advances from zero (sequential, as shader/sobol.comp)."""
    mat = sampler_tables()[0] if mat is None else np.ascontiguousarray(mat, dtype=np.uint32)
    x = np.zeros(dims, np.uint32)
    load().orc_sobol_state(_p(mat), dims, n, _p(x))
    return x


def bounce_rays(hits, width, height, spp, cur=1):
    """Bounce `cur` of sample spp for a width x height frame of primary hits: (rays (R, 6), hit mask)."""
    lib = load()
    hits = np.ascontiguousarray(hits.reshape(-1))
    R = hits.shape[0]
    assert R == width * height
    sob, noise = sampler_tables()
    rays = np.empty((R, 6), dtype=np.float32)
    mask = np.empty(R, dtype=np.uint8)
    lib.orc_bounce_rays(_p(hits), width, height, spp, cur, _p(sob), _p(noise), _p(rays), _p(mask))
    return rays, mask.astype(bool)


def _pt_buffers(width, height, color, albedo, normal):
    return (np.zeros((width * height, 4), np.float32) if color is None else color,
            np.zeros(width * height, np.uint32) if albedo is None else albedo,
            np.zeros(width * height, np.uint32) if normal is None else normal)


def path_trace(words, level, cam, width, height, bounce, spp, sky, color=None, albedo=None, normal=None,
               threads=None):
    """Sample spp of every pixel accumulated into color (float32 (height*width, 4)), albedo (RGBA8
    UNORM, uint32 per pixel) and normal (RGBA8 SNORM); created when None; returns the three."""
    words = np.ascontiguousarray(words, dtype=np.uint32)
    cam = np.ascontiguousarray(cam, dtype=np.float32)
    bufs = _pt_buffers(width, height, color, albedo, normal)
    sky = np.ascontiguousarray(sky, dtype=np.float32)
    sob, noise = sampler_tables()
    load().orc_path_trace(_p(words), words.shape[0], level, _p(cam), width, height, bounce, spp, _p(sky), _p(sob),
                          _p(noise), _p(bufs[0]), _p(bufs[1]), _p(bufs[2]), threads or nthreads())
    return bufs


def envmap_tables(rgba):
    """(texels (H, W, 4) rgb + pdf, alias (H, W, 2) uint32) for an (H, W, 4) float32 map, or None."""
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w = rgba.shape[:2]
    t = np.zeros((h, w, 4), np.float32)
    a = np.zeros((h, w, 2), np.uint32)
    if not load().orc_envmap_tables(w, h, _p(rgba), _p(t), _p(a)):
        return None
    return t, a


def path_trace_env(words, level, cam, width, height, bounce, spp, texels, alias, rotation, multiplier,
                   color=None, albedo=None, normal=None, threads=None):
    words = np.ascontiguousarray(words, dtype=np.uint32)
    cam = np.ascontiguousarray(cam, dtype=np.float32)
    texels = np.ascontiguousarray(texels, dtype=np.float32)
    alias = np.ascontiguousarray(alias, dtype=np.uint32)
    eh, ew = texels.shape[:2]
    bufs = _pt_buffers(width, height, color, albedo, normal)
    sob, noise = sampler_tables()
    load().orc_path_trace_env(_p(words), words.shape[0], level, _p(cam), width, height, bounce, spp, _p(texels),
                              _p(alias), ew, eh, rotation, multiplier, _p(sob), _p(noise), _p(bufs[0]), _p(bufs[1]),
                              _p(bufs[2]), threads or nthreads())
    return bufs


def env_sample(texels, alias, rotation, u, v):
    """(direction, radiance, pdf) of EnvMap_Sample at the 2-D point (u, v)."""
    texels = np.ascontiguousarray(texels, dtype=np.float32)
    alias = np.ascontiguousarray(alias, dtype=np.uint32)
    h, w = texels.shape[:2]
    d, l, p = np.empty(3, np.float32), np.empty(3, np.float32), np.empty(1, np.float32)
    load().orc_env_sample(_p(texels), _p(alias), w, h, rotation, u, v, _p(d), _p(l), _p(p))
    return d, l, p[0]


def env_lookup(texels, rotation, d):
    """(rgb + pdf texel filter, EnvMap_PDF) of one direction."""
    texels = np.ascontiguousarray(texels, dtype=np.float32)
    h, w = texels.shape[:2]
    d = np.ascontiguousarray(d, dtype=np.float32)
    out, p = np.empty(4, np.float32), np.empty(1, np.float32)
    load().orc_env_lookup(_p(texels), w, h, rotation, _p(d), _p(out), _p(p))
    return out, p[0]


def atan2(y, x):
    return load().orc_atan2(y, x)


def acos(x):
    return load().orc_acos(x)


def sincos(phi):
    s, c = C.c_float(), C.c_float()
    load().orc_sincos(phi, C.byref(s), C.byref(c))
    return s.value, c.value


def shade(hits, cam, width, height, view, light=(5.0, 5.0, 5.0), texels=None, rotation=0.0):
    """RGBA8 UNORM texels (uint32 per pixel) of the tracer's view-type shading of a hit frame."""
    hits = np.ascontiguousarray(hits.reshape(-1))
    cam = np.ascontiguousarray(cam, dtype=np.float32)
    light = np.ascontiguousarray(light, dtype=np.float32)
    out = np.empty(width * height, np.uint32)
    eh, ew = (0, 0) if texels is None else texels.shape[:2]
    tex = None if texels is None else np.ascontiguousarray(texels, dtype=np.float32)
    load().orc_shade(_p(hits), _p(cam), width, height, view, _p(light), _p(tex), ew, eh, rotation, _p(out))
    return out


def stack_uninit_reads():
    return load().orc_stack_uninit_reads()


def reset_counters():
    load().orc_reset_counters()
