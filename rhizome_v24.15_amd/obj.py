"""Wavefront OBJ loading for the mesh-to-SDF entry points.

Reads ``v x y z`` and ``f`` records (1-based or negative indices, ``a/b/c``
forms, polygons fanned into triangles), as the reference's ObjParser does
(native/src/engine/io/obj_parser.cpp:36-120).  The reference's winding
pre-pass after loading (obj_parser.cpp:8-34) is the library's
``orient_mesh`` (rhz_orient_mesh).
"""
from __future__ import annotations

import numpy as np


def load_obj(path: str):
    """Returns (vertices float32[N, 3], faces uint32[M, 3])."""
    verts = []
    faces = []
    with open(path, "r", encoding="utf-8", errors="replace") as fh:
        for line in fh:
            if line.startswith("v "):
                p = line.split()
                verts.append((float(p[1]), float(p[2]), float(p[3])))
            elif line.startswith("f "):
                idx = []
                for tok in line.split()[1:]:
                    k = int(tok.split("/")[0])
                    idx.append(k - 1 if k > 0 else len(verts) + k)
                for i in range(1, len(idx) - 1):
                    faces.append((idx[0], idx[i], idx[i + 1]))
    v = np.asarray(verts, dtype=np.float32).reshape(-1, 3)
    f = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    if f.size and (f.min() < 0 or f.max() >= len(v)):
        raise ValueError(f"{path}: face index out of range")
    return v, f.astype(np.uint32)


def save_obj(path: str, verts, faces) -> None:
    """Writes ``v x y z`` records (nine significant digits: binary32 values
    read back exactly) and 1-based ``f a b c`` triangles, the records
    load_obj reads (the reference's ObjExporter writes the same two kinds)."""
    v = np.asarray(verts, dtype=np.float32).reshape(-1, 3)
    f = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    with open(path, "w", encoding="utf-8") as fh:
        for x, y, z in v.tolist():
            fh.write(f"v {x:.9g} {y:.9g} {z:.9g}\n")
        for a, b, c in (f + 1).tolist():
            fh.write(f"f {a} {b} {c}\n")

