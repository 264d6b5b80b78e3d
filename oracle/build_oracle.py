"""Builds the CPU oracle (test infrastructure) into oracle/_build/liboracle.so, and, where the
reference tree is present, the reference's OBJ reader checker into oracle/_ref/objref.

The reference's GPU path (GLSL run by a Vulkan driver) cannot be built here: there is no glslc and
no Vulkan implementation, so the restatement in oracle.cpp is the checker for it.  Its OBJ reader
(tinyobj and glm, header-only C++ under dep/) compiles with g++ where it lies: obj_ref.cpp drives it
as Scene::load_meshes does.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "liboracle.so")

FLAGS = ["-O2", "-ffp-contract=off", "-fno-fast-math", "-std=c++17", "-shared", "-fPIC", "-Wall", "-pthread"]


REF_DEP = "/root/reference/dep"
REF_OUT = os.path.join(HERE, "_ref", "objref")


def build_ref():
    """oracle/_ref/objref from obj_ref.cpp and the reference's dep/ sources (None without them)."""
    if not os.path.isfile(os.path.join(REF_DEP, "tiny_obj_loader.cpp")):
        return None
    os.makedirs(os.path.dirname(REF_OUT), exist_ok=True)
    tmp = REF_OUT + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-I" + REF_DEP, "-I" + os.path.join(REF_DEP, "glm"), "-o", tmp,
           os.path.join(HERE, "obj_ref.cpp"), os.path.join(REF_DEP, "tiny_obj_loader.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("objref build failed:\n" + r.stderr[-4000:])
    os.replace(tmp, REF_OUT)
    return REF_OUT


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
    print(build_ref())
