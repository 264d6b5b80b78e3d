"""Writes tests/golden/known_answers.json: hand-derived known answers (no code under test is run).

fold: the leaf colour running average of shader/octree_tag_node.comp:43-57 applied by hand,
  avg' = floor((avg * n + c) / (n + 1)) per channel, n' = min(n + 1, 63), first fragment
  0xC1000000 | c, fragments taken in ascending (fragment.x, fragment.y) order, which for one voxel
  is ascending colour:
  - three primaries 0x0000FF, 0x00FF00, 0xFF0000:
      (255,0,0) n=1 -> (255*1+0)/2, (0+255)/2, 0 = (127,127,0) n=2
      -> (127*2+0)/3, (127*2+0)/3, (0*2+255)/3 = (84,84,85) n=3          => 0xC3555454
  - truncation 0x000001, 0x000002: (1*1 + 2) / 2 = 1                   => 0xC2000001
  - saturation: 70 x 0x102030 keeps the colour, count stops at 63       => 0xFF102030
  - saturated divisor: 63 x 0x000000 then 0x000040: n = 63, divisor 64,
      R = (0*63 + 64) / 64 = 1                                          => 0xFF000001

march (level 1, octree words 0..7 = the root's children, one leaf; Octree_RayMarchLeaf 6-arg,
shader/octree.glsl:179-340, EPS = 2^-48):
  A. leaf in slot 1 (cube [1.5,2]x[1,1.5]x[1,1.5]); o = (0.5,1.25,1.25), d = (1,0,0) -> d = (1,EPS,EPS)
     t_coef = (-1,-2^48,-2^48); all three axes mirrored (oct_mask 7); t_bias.x = 3*-1 + 0.5 = -2.5
     t_min = 0.5, t_max = 1.5, idx = 7, pos = (1.5,1.5,1.5)
     iter 1: load word[7^7 = 0] = 0 -> ADVANCE along x: pos.x = 1.0, t_min = 1.0, idx = 6, no pop
     iter 2: load word[6^7 = 1] = leaf -> hit.  2 iterations, 2 loads.
     normal -x (t_corner.x = 1.0 is the largest), unmirrored cube (1.5,1,1) -> voxel (1,0,0), depth 1
     o_pos = (1.5 - 2^-48, 1.25, 1.25) = (1.5, 1.25, 1.25) in binary32
  B. leaf in slot 0, same ray: iter 1 loads word[0] = leaf -> hit.  voxel (0,0,0), normal -x,
     o_pos = (1.0 - 2^-48 -> 1.0, 1.25, 1.25)
  C. leaf in slot 1, o = (2.5,1.25,1.25), d = (-1,0,0): y,z mirrored (oct_mask 6), t_min = 0.5,
     idx = 7; iter 1 loads word[7^6 = 1] = leaf -> hit.  normal +x (x not mirrored), voxel (1,0,0),
     o_pos.x = (1.5 + 0.5) + 2^-47 -> 2.0
"""
import json
import os

LEAF = 0xC1AABBCC


def words_with(slot):
    w = [0] * 8
    w[slot] = LEAF
    return ["%08x" % v for v in w]


def main():
    data = {
        "fold": [
            {"name": "three primaries", "colours": [0x0000FF, 0x00FF00, 0xFF0000], "word": "C3555454"},
            {"name": "truncation", "colours": [1, 2], "word": "C2000001"},
            {"name": "saturation", "colours": [0x102030] * 70, "word": "FF102030"},
            {"name": "saturated divisor", "colours": [0] * 63 + [0x40], "word": "FF000001"},
        ],
        "march": [
            {
                "name": "A: advance then hit",
                "level": 1,
                "words": words_with(1),
                "ray": [0.5, 1.25, 1.25, 1.0, 0.0, 0.0],
                "expect": {"iter": 2, "nfetch": 2, "vxy": 1, "vz_flags": "80010000", "pos": [1.5, 1.25, 1.25]},
            },
            {
                "name": "B: first child is the leaf",
                "level": 1,
                "words": words_with(0),
                "ray": [0.5, 1.25, 1.25, 1.0, 0.0, 0.0],
                "expect": {"iter": 1, "nfetch": 1, "vxy": 0, "vz_flags": "80010000", "pos": [1.0, 1.25, 1.25]},
            },
            {
                "name": "C: negative direction, +x face",
                "level": 1,
                "words": words_with(1),
                "ray": [2.5, 1.25, 1.25, -1.0, 0.0, 0.0],
                "expect": {"iter": 1, "nfetch": 1, "vxy": 1, "vz_flags": "81010000", "pos": [2.0, 1.25, 1.25]},
            },
        ],
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
