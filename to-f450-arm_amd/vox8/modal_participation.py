"""Modal effective mass (sims/modal_participation.py:34-94) on the device.

M_eff,k = (phi_k^T M r)^2 / (phi_k^T M phi_k) / (r^T M r), r a unit rigid
translation in `direction`, restricted to free DOFs, with the same mass matrix
(tip lump included) the eigensolve used.  M products run in libvox8; the k+1
dot products of the result are reduced on the device.
"""
from __future__ import annotations

import numpy as np
import torch

EXCITABLE_FRAC = 0.05


def effective_mass_voxel(fem, geom, rho_active, mode_shapes, tip_mass_kg, direction=2,
                         tip_nodes=None):
    if tip_nodes is None:
        tip_nodes = (geom.motor_mount_top_nodes() if hasattr(geom, "motor_mount_top_nodes")
                     else np.array([], dtype=np.int64))
    M = fem.assemble_M(rho_active, tip_mass_kg=tip_mass_kg, tip_mass_nodes=tip_nodes)
    free = torch.zeros(geom.n_dofs, dtype=torch.float64, device=fem.device)
    free[torch.as_tensor(fem.free_dofs, device=fem.device)] = 1.0
    r = torch.zeros(geom.n_dofs, dtype=torch.float64, device=fem.device)
    r[direction::3] = 1.0
    r = r * free
    P, back = fem._to_dev(mode_shapes)
    P = P.reshape(geom.n_dofs, -1) * free[:, None]
    Mr = M @ r
    MP = M @ P
    total = float(torch.dot(r, Mr))
    gm = (P * MP).sum(dim=0)
    L = P.T @ Mr
    out = (L * L / torch.clamp(gm, min=1e-30) / total).cpu().numpy()
    return out


def first_excitable(freqs, eff, thresh=EXCITABLE_FRAC):
    """(frequency, fraction, index) of the lowest mode carrying >= thresh of the mass."""
    for k, (f, e) in enumerate(zip(freqs, eff)):
        if e >= thresh:
            return float(f), float(e), k
    return None, None, None
