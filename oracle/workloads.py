"""Oracle: synthetic cantilever inputs (test infrastructure only).

Independent restatement of the workload definition (seeded density, tip-face
load cases) used to cross-check the product's generator, vox8.workloads.
Load cases: LC1 Fz=-10 N; LC2 Mx=0.5 N m with f_i = Mx(-dz_i, dy_i)/sum(dy^2+dz^2)
(the exact-moment form of sims/make_lc45.py:90-111); LC3 (0, 5, -8.660254) N;
LC4 Fy=+3 N; LC5 = 0.6 LC1 + 0.4 LC2.  Equal share per tip-face node.
"""
from __future__ import annotations

import numpy as np


def tip_face(geom):
    out = []
    for iz in range(geom.nz + 1):
        for iy in range(geom.ny + 1):
            out.append((geom.node_index(geom.nx, iy, iz), iy * geom.dy, iz * geom.dz))
    return out


def load_vector(geom, name):
    F = np.zeros(geom.n_dofs)
    face = tip_face(geom)
    n = len(face)
    if name in ("LC1", "LC3", "LC4"):
        f = {"LC1": (0.0, 0.0, -10.0), "LC3": (0.0, 5.0, -8.660254), "LC4": (0.0, 3.0, 0.0)}[name]
        for node, _, _ in face:
            F[3 * node:3 * node + 3] += np.array(f) / n
        return F
    if name == "LC2":
        yc = np.mean([y for _, y, _ in face])
        zc = np.mean([z for _, _, z in face])
        s = sum((y - yc) ** 2 + (z - zc) ** 2 for _, y, z in face)
        for node, y, z in face:
            F[3 * node + 1] += -0.5 * (z - zc) / s
            F[3 * node + 2] += 0.5 * (y - yc) / s
        return F
    if name == "LC5":
        return 0.6 * load_vector(geom, "LC1") + 0.4 * load_vector(geom, "LC2")
    raise KeyError(name)


def seeded_density(n, base, amp, seed):
    x = base + np.random.default_rng(seed).uniform(-amp, amp, n)
    return np.minimum(np.maximum(x, 1e-3), 1.0)
