"""Builds the CPU oracle (test infrastructure) into oracle/_build/liboracle.so.

The reference itself (GLSL run by a Vulkan driver) cannot be built here: there is no glslc and no
Vulkan implementation, so there is no oracle/_ref build; this restatement is the checker.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "liboracle.so")

FLAGS = ["-O2", "-ffp-contract=off", "-fno-fast-math", "-std=c++17", "-shared", "-fPIC", "-Wall", "-pthread"]


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = ["g++"] + FLAGS + ["-o", tmp, os.path.join(HERE, "oracle.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stderr[-4000:])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
