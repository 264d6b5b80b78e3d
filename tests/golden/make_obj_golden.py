"""Writes tests/golden/obj/<name>.expected for every <name>.obj there: the reference's own OBJ
reader (tinyobj + glm, driven as Scene::load_meshes does) run through oracle/_ref/objref, which
oracle/build_oracle.py builds from the reference tree.  One line per triangle: 9 position bit
patterns and the packed colour, in hexadecimal.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OBJREF = os.path.join(ROOT, "oracle", "_ref", "objref")


def main():
    if not os.path.isfile(OBJREF):
        sys.exit("oracle/_ref/objref is not built (python oracle/build_oracle.py, with the reference tree present)")
    for path in sorted(glob.glob(os.path.join(HERE, "obj", "*.obj"))):
        r = subprocess.run([OBJREF, path], capture_output=True, text=True)
        with open(path[:-4] + ".expected", "w") as f:
            f.write(r.stdout)
        print(path, r.returncode, len(r.stdout.splitlines()))


if __name__ == "__main__":
    main()
