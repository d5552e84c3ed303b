"""Hand-built test meshes with closed-form or degenerate-heavy geometry."""
from __future__ import annotations

import numpy as np

# unit-cube face (axis, side) -> 4 corner offsets, counter-clockwise seen from outside
_FACES = {
    (0, 0): [(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0)],
    (0, 1): [(1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)],
    (1, 0): [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)],
    (1, 1): [(0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)],
    (2, 0): [(0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0)],
    (2, 1): [(0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)],
}


def polycube(cells, scale: float = 0.25, offset=(0.0, 0.0, 0.0)):
    """Closed surface of a union of unit voxels (integer cell coordinates).

    Every boundary unit square becomes two triangles with outward normals;
    vertices sit on the integer grid (times ``scale``, a power of two keeps
    them exact in float32), so many segments run along faces, through edges
    and vertices -- the degenerate cases the exact predicates must decide.
    """
    cells = {tuple(c) for c in cells}
    verts = {}
    tris = []

    def vid(p):
        if p not in verts:
            verts[p] = len(verts)
        return verts[p]

    for c in sorted(cells):
        for (ax, side), quad in _FACES.items():
            nb = list(c)
            nb[ax] += 1 if side else -1
            if tuple(nb) in cells:
                continue
            ids = [vid((c[0] + q[0], c[1] + q[1], c[2] + q[2])) for q in quad]
            tris.append((ids[0], ids[1], ids[2]))
            tris.append((ids[0], ids[2], ids[3]))
    V = np.zeros((len(verts), 3), dtype=np.float64)
    for p, i in verts.items():
        V[i] = p
    V = V * scale + np.asarray(offset)
    return V.astype(np.float32), np.asarray(tris, dtype=np.uint32)


def l_shape():
    """3x3x1 block minus its 2x2 corner: an L with one reflex edge line."""
    cells = [(x, y, 0) for x in range(3) for y in range(3) if not (x >= 1 and y >= 1)]
    return polycube(cells)


def u_channel():
    """U-shaped 3x2x1 block (middle column of the top row removed)."""
    cells = [(x, y, 0) for x in range(3) for y in range(2) if not (x == 1 and y == 1)]
    return polycube(cells)


def box(lo=(0, 0, 0), hi=(1, 1, 1)):
    cells = [(x, y, z) for x in range(lo[0], hi[0]) for y in range(lo[1], hi[1])
             for z in range(lo[2], hi[2])]
    return polycube(cells)


def cage_of(V, factor: float = 1.08):
    """Radially scaled copy around the bbox centre (an enclosing cage for star-shaped meshes)."""
    V = np.asarray(V, dtype=np.float64)
    c = (V.min(0) + V.max(0)) / 2
    return ((V - c) * factor + c).astype(np.float32)
