"""Mesh and solve-record files (host-side helpers, not on the step path).

  * load_obj / save_obj: Wavefront OBJ with "v" and triangular "f" lines, as
    io/obj_io.cpp:12-102 reads and writes them ("f a/b/c" entries keep the
    vertex index; indices are 1-based in the file, 0-based in memory).
  * save_solve / load_solve: one debug_dump_last_solve() 6-tuple
    (diag, row_offsets, col_indices, values, rhs, dx) as .npz, the fixture
    format of the regression files under tests/golden.
"""

from __future__ import annotations

from typing import Tuple

import numpy as np

SOLVE_KEYS = ("diag", "row_offsets", "col_indices", "values", "rhs", "dx")


def load_obj(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """(positions (N, 3) float32, triangles (M, 3) int32).  Faces with more
    than three vertices keep their first three, as the reference does."""
    pos, tris = [], []
    with open(path, "r") as f:
        for line in f:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                pos.append([float(t) for t in tok[1:4]])
            elif tok[0] == "f":
                idx = [int(t.split("/")[0]) - 1 for t in tok[1:4]]
                if len(idx) == 3:
                    tris.append(idx)
    return np.asarray(pos, np.float32).reshape(-1, 3), np.asarray(tris, np.int32).reshape(-1, 3)


def save_obj(path: str, positions, triangles) -> None:
    p = np.asarray(positions, np.float32).reshape(-1, 3)
    t = np.asarray(triangles, np.int64).reshape(-1, 3)
    with open(path, "w") as f:
        f.write(f"# ChysX OBJ export  verts={p.shape[0]} tris={t.shape[0]}\n")
        for x, y, z in p:  # 9 significant digits round-trip FP32 exactly
            f.write("v %.9g %.9g %.9g\n" % (x, y, z))
        for a, b, c in t + 1:
            f.write(f"f {a} {b} {c}\n")


def save_solve(path: str, dump) -> None:
    """Write one debug_dump_last_solve() result."""
    if len(dump) != 6:
        raise ValueError("save_solve: expected the 6-tuple of debug_dump_last_solve()")
    np.savez_compressed(path, **{k: np.asarray(v) for k, v in zip(SOLVE_KEYS, dump)})


def load_solve(path: str):
    with np.load(path, allow_pickle=False) as z:
        return tuple(z[k] for k in SOLVE_KEYS)
