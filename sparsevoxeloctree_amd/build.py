"""Builds libsvo.so (HIP kernels for gfx950 + the C ABI of include/svo.h) in-tree.

Usage: python -m sparsevoxeloctree_amd.build   (or sparsevoxeloctree_amd.build.build_library())

Sources are compiled in parallel with hipcc into _build/ and linked into lib/libsvo.so.
Floating-point flags are part of the contract: no contraction, IEEE division / sqrt (hipcc's
default), no fast-math; packed fp32 instructions are disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libsvo.so")

SOURCES = [
    "svo_capi.cpp",
    "svo_voxelize.hip",
    "svo_build.hip",
    "svo_march.hip",
    "svo_util.hip",
    "svo_server.cpp",
    "scenegen.cpp",
    "svo_obj.cpp",
]

DATA = os.path.join(HERE, "data")
GEN = os.path.join(BUILD, "gen")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SVO_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wall",
    "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + GEN,
]
DEVICE = [
    "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fgpu-rdc" if False else "-fno-gpu-rdc",
]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC] + lang + COMMON + DEVICE + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def write_tables():
    """The sampler's data tables (data/*.bin, see data/make_tables.py) as C initializer lists in
    _build/gen/, included by svo_march.hip: the Sobol matrices as 2048 uint32 and the blue-noise
    texture as 65536 uint16 texels (R | G << 8)."""
    import numpy as np

    os.makedirs(GEN, exist_ok=True)
    sob = np.fromfile(os.path.join(DATA, "sobol_matrices.bin"), dtype="<u4")
    noise = np.fromfile(os.path.join(DATA, "blue_noise_rg8.bin"), dtype=np.uint8)
    if sob.size != 64 * 32 or noise.size != 256 * 256 * 2:
        raise RuntimeError("sampler tables have the wrong size")
    texels = noise[0::2].astype(np.uint32) | (noise[1::2].astype(np.uint32) << 8)
    for name, vals, fmt in (("svo_sobol_matrices.inc", sob, "0x%08xu"), ("svo_blue_noise.inc", texels, "0x%04xu")):
        lines = [",".join(fmt % int(v) for v in vals[i:i + 16]) + "," for i in range(0, vals.size, 16)]
        text = "\n".join(lines) + "\n"
        path = os.path.join(GEN, name)
        old = None
        if os.path.exists(path):
            with open(path) as f:
                old = f.read()
        if old != text:
            with open(path, "w") as f:
                f.write(text)


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    write_tables()
    os.makedirs(LIB_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-6000:]))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_cli(verbose)
    return LIB


CLI = os.path.join(LIB_DIR, "svo_cli")


def build_cli(verbose=False):
    """The C++ command-line driver over the class surface (host/svo.hpp), linked to libsvo.so."""
    tmp = CLI + ".tmp"
    cmd = [
        "g++", "-O2", "-std=c++17", "-Wall",
        "-I" + os.path.join(ROOT, "include"),
        os.path.join(HERE, "host", "svo_cli.cpp"),
        "-L" + LIB_DIR, "-lsvo", "-Wl,-rpath,$ORIGIN",
        "-o", tmp,
    ]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("svo_cli build failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-6000:]))
    os.replace(tmp, CLI)
    if verbose:
        print("built", CLI)
    return CLI


if __name__ == "__main__":
    build_library(verbose=True)
    sys.exit(0)
