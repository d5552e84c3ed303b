"""Converts the reference's sample mesh data/cow.obj (2,903 vertices, 5,804
triangles; the input of its README usage `python decompose.py data/cow.obj`)
into tests/golden/cow.npz: float64 vertices and int32 triangles, parsed from
the OBJ's `v` and `f` lines (1-based indices, texture/normal indices after
'/' ignored).  Run where the reference is present:

    python tests/golden/make_cow.py [path/to/cow.obj]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data/cow.obj"
    V, F = [], []
    with open(src) as fh:
        for line in fh:
            if line.startswith("v "):
                V.append([float(x) for x in line.split()[1:4]])
            elif line.startswith("f "):
                F.append([int(x.split("/")[0]) - 1 for x in line.split()[1:4]])
    np.savez_compressed(os.path.join(HERE, "cow.npz"), vertices=np.array(V, np.float64),
                        triangles=np.array(F, np.int32))


if __name__ == "__main__":
    main()
