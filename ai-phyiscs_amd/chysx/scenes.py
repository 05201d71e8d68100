"""Synthetic cloth workloads (seeded, deterministic).

Grid cloths follow the reference demo's layout (scenes/cloth_freefall.cu:38-62):
particle ``j*W + i``; per quad the triangles ``(v00, v10, v11)`` and
``(v00, v11, v01)``.  Here the cloth lies in the xz plane (``x = i*h``,
``y = 0``, ``z = j*h``) with gravity along -y; the "top" row is ``j = W-1``
(the demo pins ``(H-1)*W`` and ``(H-1)*W + W-1``, cloth_freefall.cu:104).

Every generator returns a :class:`Scene` holding host (numpy) arrays only.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

GRAVITY = (0.0, -9.8, 0.0)


@dataclass
class Scene:
    name: str
    pos: np.ndarray            # (N,3) float32, rest = initial positions
    vel: np.ndarray            # (N,3) float32
    inv_mass: np.ndarray       # (N,)  float32
    tris: np.ndarray           # (T,3) int32
    pin_idx: np.ndarray        # (P,)  int32
    pin_targets: np.ndarray    # (P,3) float32
    pin_k: float = 1.0e6
    spring_k: object = 1.0e3   # float, or (E,) float32 in edge order
    stretch_k: object = 1.0e3  # float, or (T,) float32
    force: Optional[np.ndarray] = None  # (N,3) float32 external force [N]
    dt: float = 1.0 / 60.0
    steps: int = 200
    gravity: tuple = GRAVITY
    damping: float = 0.0
    grid_w: int = 0
    meta: dict = field(default_factory=dict)
    shear_k: Optional[float] = None     # FEM shear stiffness (None: no shear elements)
    bend_k: Optional[float] = None      # dihedral bending stiffness (None: no bending)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])


def grid(w: int, size: float = 1.0):
    """W x W grid in the xz plane: positions (N,3) f32, triangles (T,3) i32."""
    h = size / (w - 1)
    j, i = np.meshgrid(np.arange(w), np.arange(w), indexing="ij")
    pos = np.zeros((w * w, 3), np.float32)
    pos[:, 0] = (i.reshape(-1) * h).astype(np.float32)
    pos[:, 2] = (j.reshape(-1) * h).astype(np.float32)
    jj, ii = np.meshgrid(np.arange(w - 1), np.arange(w - 1), indexing="ij")
    v00 = (jj * w + ii).reshape(-1)
    v10 = v00 + 1
    v01 = v00 + w
    v11 = v00 + w + 1
    tris = np.empty((2 * (w - 1) * (w - 1), 3), np.int32)
    tris[0::2] = np.stack([v00, v10, v11], 1)
    tris[1::2] = np.stack([v00, v11, v01], 1)
    return pos, tris


def edges_of(tris: np.ndarray) -> np.ndarray:
    """Unique undirected edges (min,max), lexicographically sorted.

    Same order as the simulator's edge list (cloth_simulator.cu:331-349), so
    per-edge arrays line up with it.
    """
    t = tris.astype(np.int64)
    a = np.concatenate([t[:, 0], t[:, 1], t[:, 2]])
    b = np.concatenate([t[:, 1], t[:, 2], t[:, 0]])
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    key = np.unique(lo * (1 << 32) + hi)
    return np.stack([key >> 32, key & 0xFFFFFFFF], 1).astype(np.int32)


def vertex_areas(pos: np.ndarray, tris: np.ndarray) -> np.ndarray:
    """A_i = sum over incident triangles of area/3 (float64)."""
    p = pos.astype(np.float64)
    e1 = p[tris[:, 1]] - p[tris[:, 0]]
    e2 = p[tris[:, 2]] - p[tris[:, 0]]
    area = 0.5 * np.linalg.norm(np.cross(e1, e2), axis=1)
    out = np.zeros(pos.shape[0])
    for c in range(3):
        np.add.at(out, tris[:, c], area / 3.0)
    return out


def _pins(w: int, kind: str, pos: np.ndarray):
    top = (w - 1) * w
    if kind == "corner0":
        idx = np.array([0], np.int32)
    elif kind == "one_corner":
        idx = np.array([top], np.int32)
    elif kind == "two_corners":
        idx = np.array([top, top + w - 1], np.int32)
    elif kind == "top_edge":
        idx = np.arange(top, top + w, dtype=np.int32)
    else:
        raise ValueError(kind)
    return idx, pos[idx].copy()


def config1() -> Scene:
    """10x10, 1 m, corner particle 0 pinned, m = 0.01 kg, springs + stretch
    k = 1e3, dt = 1/60, 200 steps from rest."""
    w = 10
    pos, tris = grid(w, 1.0)
    idx, tgt = _pins(w, "corner0", pos)
    return Scene("cloth_10x10", pos, np.zeros_like(pos), np.full(w * w, 100.0, np.float32),
                 tris, idx, tgt, 1.0e6, 1.0e3, 1.0e3, None, 1.0 / 60.0, 200, grid_w=w)


def config2(w: int = 256, steps: int = 1000, seed: int = 2) -> Scene:
    """256x256, 2 m, two top corners pinned, rest positions jittered by
    U(+-1% of edge length), per-element k log-uniform in [1e2, 1e4],
    dt = 1/120, 1000 steps."""
    pos, tris = grid(w, 2.0)
    h = 2.0 / (w - 1)
    rng = np.random.default_rng(seed)
    pos = (pos + rng.uniform(-0.01 * h, 0.01 * h, pos.shape)).astype(np.float32)
    e = edges_of(tris)
    ks = (10.0 ** rng.uniform(2.0, 4.0, e.shape[0])).astype(np.float32)
    kt = (10.0 ** rng.uniform(2.0, 4.0, tris.shape[0])).astype(np.float32)
    idx, tgt = _pins(w, "two_corners", pos)
    return Scene(f"cloth_{w}x{w}_jitter", pos, np.zeros_like(pos), np.full(w * w, 100.0, np.float32),
                 tris, idx, tgt, 1.0e6, ks, kt, None, 1.0 / 120.0, steps, grid_w=w)


def config3(w: int = 1024, steps: int = 300, seed: int = 3) -> Scene:
    """1024x1024, 1 m, top edge pinned, gravity plus a seeded constant
    horizontal wind of U(0,5) N/m^2 (f_i = p * A_i * d), m = 0.01 kg,
    springs + stretch k = 1e3, dt = 1/60, 300 steps."""
    pos, tris = grid(w, 1.0)
    rng = np.random.default_rng(seed)
    p = float(rng.uniform(0.0, 5.0))
    ang = float(rng.uniform(0.0, 2.0 * np.pi))
    d = np.array([np.cos(ang), 0.0, np.sin(ang)])
    force = (vertex_areas(pos, tris)[:, None] * p * d[None, :]).astype(np.float32)
    idx, tgt = _pins(w, "top_edge", pos)
    return Scene(f"cloth_{w}x{w}_wind", pos, np.zeros_like(pos), np.full(w * w, 100.0, np.float32),
                 tris, idx, tgt, 1.0e6, 1.0e3, 1.0e3, force, 1.0 / 60.0, steps, grid_w=w,
                 meta={"wind_pressure": p, "wind_dir": d.tolist()})


def config4(count: int = 512, wmin: int = 16, wmax: int = 128, steps: int = 500, seed: int = 4) -> List[Scene]:
    """Batch of uneven cloths: W ~ U{16..128}, per-cloth k ~ 10^U(2,4),
    mass ~ U(0.005, 0.02) kg, pins ~ U{one corner, two top corners, top edge}."""
    rng = np.random.default_rng(seed)
    kinds = ["one_corner", "two_corners", "top_edge"]
    out = []
    for c in range(count):
        w = int(rng.integers(wmin, wmax + 1))
        k = float(10.0 ** rng.uniform(2.0, 4.0))
        m = float(rng.uniform(0.005, 0.02))
        kind = kinds[int(rng.integers(0, 3))]
        pos, tris = grid(w, 1.0)
        idx, tgt = _pins(w, kind, pos)
        out.append(Scene(f"batch{c}_{w}", pos, np.zeros_like(pos), np.full(w * w, 1.0 / m, np.float32),
                         tris, idx, tgt, 1.0e6, k, k, None, 1.0 / 60.0, steps, grid_w=w,
                         meta={"pins": kind}))
    return out


def config5_train(count: int = 16, w: int = 256, seed: int = 5) -> List[Scene]:
    """Warm-start training cloths: W x W (default 256), 1 m, pins ~
    U{one corner, two top corners, top edge}, seeded wind U(0,5) N/m^2 in a
    random horizontal direction, k ~ 10^U(2,4) (springs and stretch),
    m ~ U(0.005, 0.02) kg, dt = 1/60."""
    rng = np.random.default_rng(seed)
    kinds = ["one_corner", "two_corners", "top_edge"]
    out = []
    pos0, tris = grid(w, 1.0)
    areas = vertex_areas(pos0, tris)
    for c in range(count):
        k = float(10.0 ** rng.uniform(2.0, 4.0))
        m = float(rng.uniform(0.005, 0.02))
        kind = kinds[int(rng.integers(0, 3))]
        p = float(rng.uniform(0.0, 5.0))
        ang = float(rng.uniform(0.0, 2.0 * np.pi))
        d = np.array([np.cos(ang), 0.0, np.sin(ang)])
        force = (areas[:, None] * p * d[None, :]).astype(np.float32)
        pos = pos0.copy()
        idx, tgt = _pins(w, kind, pos)
        out.append(Scene(f"train{c}_{w}", pos, np.zeros_like(pos), np.full(w * w, 1.0 / m, np.float32),
                         tris, idx, tgt, 1.0e6, k, k, force, 1.0 / 60.0, 0, grid_w=w,
                         meta={"pins": kind, "k": k, "mass": m, "wind": p}))
    return out


def drape(w: int = 21, size: float = 0.8, seed: Optional[int] = None) -> Scene:
    """The reference demo's cloth (scenes/cloth_freefall.cu:36-110): W x W grid
    of side 0.8 m, FEM stretch k = 5000, shear k = 500, bending k = 0.01,
    area-lumped mass of density 0.1 kg/m^2, damping 2, the two top corners
    pinned; springs are built but, as in the reference step, not active.
    With a seed the rest shape is crumpled out of plane (+-30% of the grid
    spacing), so rest dihedral angles are non-zero.  dt = 1/60."""
    pos, tris = grid(w, size)
    if seed is not None:
        h = size / (w - 1)
        rng = np.random.default_rng(seed)
        pos = pos.copy()
        pos[:, 1] += rng.uniform(-0.3 * h, 0.3 * h, pos.shape[0]).astype(np.float32)
    mass = 0.1 * vertex_areas(pos, tris)
    inv_mass = (1.0 / mass).astype(np.float32)
    idx, tgt = _pins(w, "two_corners", pos)
    return Scene(f"drape_{w}x{w}" + ("" if seed is None else f"_crumpled{seed}"), pos, np.zeros_like(pos), inv_mass,
                 tris, idx, tgt, 1.0e6, 5000.0, 5000.0, None, 1.0 / 60.0, 100, damping=2.0, grid_w=w,
                 shear_k=500.0, bend_k=0.01)


def degenerate(w: int = 5) -> Scene:
    """A W x W grid cloth (1 m, top edge pinned, springs + stretch k = 1e3,
    m = 0.01 kg) with four extra particles hung from its first rows by
    degenerate triangles, one per guard of the reference's element code:

    * ``p1``, ``p2``: a millimetre triangle (0, p1, p2) with edge 2^-10 m and
      height 2^-11 m, whose rest determinant lab * v_c = 2^-21 is below 1e-6,
      so the rest frame falls back to the identity with a non-zero area
      (triangle_stretch_constraint.cu:95-99);
    * ``q``: on the segment (1, 2), a zero-area triangle (determinant 0);
    * ``r``: at the position of particle W+1 and joined to it by a spring of
      rest length 0, so the spring length is clamped to 1e-6
      (spring_constraint.cu:35) and the triangle (W+1, r, W+2) has two
      coincident corners.

    With W = 5 every coordinate is a short binary fraction, so the FP32 rest
    frame (v_c^2 = |ac|^2 - u_c^2, as the reference computes it) is exact or
    free of cancellation and takes the same branch as the FP64 oracle.  A
    sliver with long edges would not: its v_c^2 cancels to rounding noise in
    FP32, in the reference's kernel as here, and the branch taken depends on
    that noise.
    """
    pos, tris = grid(w, 1.0)
    h = 1.0 / (w - 1)
    n0 = w * w
    p1, p2, q, r = n0, n0 + 1, n0 + 2, n0 + 3
    extra = np.array([[2.0 ** -10, 0.0, 0.0],
                      [2.0 ** -11, 0.0, -(2.0 ** -11)],
                      [1.5 * h, 0.0, 0.0],
                      pos[w + 1]], np.float32)
    pos = np.concatenate([pos, extra]).astype(np.float32)
    tris = np.concatenate([tris, np.array([[0, p1, p2], [1, q, 2], [w + 1, r, w + 2]], np.int32)])
    idx, tgt = _pins(w, "top_edge", pos)
    return Scene(f"degenerate_{w}", pos, np.zeros_like(pos), np.full(pos.shape[0], 100.0, np.float32),
                 tris, idx, tgt, 1.0e6, 1.0e3, 1.0e3, None, 1.0 / 60.0, 20, grid_w=w,
                 meta={"small": (p1, p2), "collinear": q, "coincident": r, "coincident_with": w + 1})


def freefall(w: int = 6) -> Scene:
    """Unpinned cloth with no elastic terms active beyond rest state: every
    particle follows x_n = x_0 + dt^2 g n(n+1)/2 exactly."""
    pos, tris = grid(w, 1.0)
    return Scene(f"freefall_{w}", pos, np.zeros_like(pos), np.full(w * w, 100.0, np.float32),
                 tris, np.zeros(0, np.int32), np.zeros((0, 3), np.float32), 1.0e6, 1.0e3, 1.0e3,
                 None, 1.0 / 60.0, 10, grid_w=w)
