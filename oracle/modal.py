"""Oracle: modal effective mass (test infrastructure only).

Restates sims/modal_participation.py:34-94 (EXCITABLE_FRAC, effective_mass_voxel,
first_excitable).  ``tip_nodes`` replaces the reference's hard call of
geom.motor_mount_top_nodes() (sims/modal_participation.py:44-45) so the
synthetic cantilevers, which have no tip mass, can use it too.
"""
from __future__ import annotations

import numpy as np

EXCITABLE_FRAC = 0.05


def effective_mass_voxel(fem, geom, rho_active, mode_shapes, tip_mass_kg, direction=2,
                         tip_nodes=None):
    if tip_nodes is None:
        tip_nodes = geom.motor_mount_top_nodes()
    M = fem.assemble_M(rho_active, tip_mass_kg=tip_mass_kg, tip_mass_nodes=tip_nodes)
    r = np.zeros(geom.n_dofs)
    r[direction::3] = 1.0
    fd = fem.free_dofs
    Mf, rf = M[fd][:, fd], r[fd]
    total = float(rf @ (Mf @ rf))
    out = []
    for k in range(mode_shapes.shape[1]):
        p = mode_shapes[fd, k]
        gm = float(p @ (Mf @ p))
        L = float(p @ (Mf @ rf))
        out.append(L * L / max(gm, 1e-30) / total)
    return np.array(out)


def first_excitable(freqs, eff, thresh=EXCITABLE_FRAC):
    for k, (f, e) in enumerate(zip(freqs, eff)):
        if e >= thresh:
            return float(f), float(e), k
    return None, None, None
