"""Writes the two data tables of the progressive path tracer's sampler as little-endian binaries.

  sobol_matrices.bin  64 x 32 uint32: the Sobol generator matrices, one row of 32 columns per
                      dimension (column k = the direction number XORed in when bit k of the Gray
                      code changes); the table of shader/sobol.comp:4-86 in the reference.
  blue_noise_rg8.bin  256 x 256 x 2 bytes: the blue-noise texture (two 8-bit UNORM channels,
                      row-major, texel (x, y) at (y * 256 + x) * 2), the table kNoise of
                      src/Noise.inl:3-4, from http://momentsingraphics.de/BlueNoise.html.

Both tables are data the reference's per-sample values depend on, so they are used as they are
(the Sobol matrices are the Joe-Kuo direction numbers; the noise texture is a published asset).
Usage: python make_tables.py [REFERENCE_ROOT]   (default /root/reference); --check compares the
committed files with the reference's tables instead of writing them.
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SOBOL_BIN = os.path.join(HERE, "sobol_matrices.bin")
NOISE_BIN = os.path.join(HERE, "blue_noise_rg8.bin")


def _hex_values(text):
    return [int(v, 16) for v in re.findall(r"0x([0-9a-fA-F]+)", text)]


def read_reference(root):
    with open(os.path.join(root, "shader", "sobol.comp")) as f:
        src = f.read()
    body = src[src.index("kMatrices[64][32]"):]
    body = body[: body.index("};")]
    sob = np.array(_hex_values(body), dtype=np.uint64)
    assert sob.size == 64 * 32 and sob.max() <= 0xFFFFFFFF, sob.size
    with open(os.path.join(root, "src", "Noise.inl")) as f:
        src = f.read()
    body = src[src.index("kNoise["):]
    body = body[body.index("{") + 1: body.index("};")]
    noise = np.array(_hex_values(body), dtype=np.uint64)
    assert noise.size == 256 * 256 * 2 and noise.max() <= 0xFF, noise.size
    return sob.astype("<u4").reshape(64, 32), noise.astype(np.uint8)


def load():
    """(sobol uint32[64, 32], noise uint8[256 * 256 * 2]) from the committed files."""
    sob = np.fromfile(SOBOL_BIN, dtype="<u4").reshape(64, 32)
    noise = np.fromfile(NOISE_BIN, dtype=np.uint8)
    assert noise.size == 256 * 256 * 2
    return sob, noise


def main(argv):
    check = "--check" in argv
    args = [a for a in argv if not a.startswith("--")]
    root = args[0] if args else "/root/reference"
    sob, noise = read_reference(root)
    if check:
        s2, n2 = load()
        ok = np.array_equal(sob, s2) and np.array_equal(noise, n2)
        print("tables match the reference" if ok else "tables DIFFER from the reference")
        return 0 if ok else 1
    sob.tofile(SOBOL_BIN)
    noise.tofile(NOISE_BIN)
    print("wrote", SOBOL_BIN, NOISE_BIN)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
