"""Writes tests/golden/cloth10_golden.npz from the FP64 oracle.

Workload: the 10x10 cloth (1 m square in xz, particle 0 pinned, m = 0.01 kg,
springs and stretch k = 1e3, dt = 1/60, PCG to 1e-10 true relative residual).
Contents: pattern, edges, slot maps, the step-1 linear system (diag, values,
rhs, dx) and positions after steps 1, 10 and 200.

The reference repository holds no golden vectors for this path; these are the
oracle's own outputs, committed so later changes to the oracle or the GPU path
are compared against a fixed record.  Run from the repository root:
    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "ai-phyiscs_amd"))

from chysx import scenes  # noqa: E402
from oracle.oracle import oracle_from_scene  # noqa: E402

TOL = 1e-10


def main():
    s = scenes.config1()
    oc, pos, vel, im, f = oracle_from_scene(s, precision=64, tol=TOL, max_iter=100000)
    edges = oc.edges()
    out = {"edges": edges}
    oc.step(pos, vel, im, f, s.dt, 1)
    ro, co = oc.pattern()
    ts, ss, ps = oc.slots(s.tris.shape[0], edges.shape[0], s.pin_idx.size)
    diag, vals, rhs, dx = oc.system()
    out.update(row_offsets=ro, col_indices=co, tri_slots=ts, spring_slots=ss, pin_slots=ps,
               diag1=diag, values1=vals, rhs1=rhs, dx1=dx, pos1=pos.copy(), vel1=vel.copy())
    oc.step(pos, vel, im, f, s.dt, 9)
    out["pos10"] = pos.copy()
    oc.step(pos, vel, im, f, s.dt, 190)
    out["pos200"] = pos.copy()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cloth10_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
