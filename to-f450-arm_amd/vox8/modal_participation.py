"""Modal effective mass (sims/modal_participation.py:34-94) on the device.

M_eff,k = (phi_k^T M r)^2 / (phi_k^T M phi_k) / (r^T M r), r a unit rigid
translation in `direction`, restricted to free DOFs, with the same mass matrix
(tip lump included) the eigensolve used.  The M products and the FP64
reductions run in libvox8 (``vox8_effective_mass``); the host receives the k
fractions.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import ptr

EXCITABLE_FRAC = 0.05

__all__ = ["EXCITABLE_FRAC", "effective_mass_voxel", "effective_mass_device", "rank_by_effective_mass",
           "first_excitable"]


def effective_mass_device(fem, X, direction=2):
    """Fractions, L = X_j^T M r, g = X_j^T M X_j and r^T M r of the k modes
    in X (a (k, n_dofs) device block, zero at fixed DOFs) with fem's current
    density and tip mass: (frac, L, g, total), NumPy arrays / float."""
    k, n = X.shape
    Xc = X if X.is_contiguous() else X.contiguous()
    out = [], [], []
    tot = C.c_double()
    for j0 in range(0, k, 16):             # the entry point takes up to 16 modes per call
        kk = min(16, k - j0)
        frac, L, g = (C.c_double * kk)(), (C.c_double * kk)(), (C.c_double * kk)()
        fem._call("vox8_effective_mass", kk, ptr(Xc[j0:j0 + kk]), n, int(direction), frac, L, g,
                  C.byref(tot))
        for lst, arr in zip(out, (frac, L, g)):
            lst.extend(arr[:])
    return np.array(out[0]), np.array(out[1]), np.array(out[2]), tot.value


def effective_mass_voxel(fem, geom, rho_active, mode_shapes, tip_mass_kg, direction=2,
                         tip_nodes=None):
    """sims/modal_participation.py:37-57: fraction of the total translational
    mass per mode (mode_shapes (n_dofs, k), NumPy or device tensor)."""
    if tip_nodes is None:
        tip_nodes = (geom.motor_mount_top_nodes() if hasattr(geom, "motor_mount_top_nodes")
                     else np.array([], dtype=np.int64))
    fem._set_density(rho_active)
    fem._set_tip(tip_mass_kg, tip_nodes)
    P, _ = fem._to_dev(mode_shapes)
    X = P.reshape(geom.n_dofs, -1).T.contiguous()
    free = torch.zeros(geom.n_dofs, dtype=torch.float64, device=fem.device)
    free[torch.as_tensor(fem.free_dofs, device=fem.device)] = 1.0
    X = X * free          # the reference restricts the modes to the free DOFs
    return effective_mass_device(fem, X, direction)[0]


def rank_by_effective_mass(frac):
    """Mode indices by decreasing effective-mass fraction (ties: lower mode first)."""
    f = np.asarray(frac, dtype=float)
    return np.lexsort((np.arange(len(f)), -f))


def first_excitable(freqs, eff, thresh=EXCITABLE_FRAC):
    """(frequency, fraction, index) of the lowest mode carrying >= thresh of the mass."""
    for k, (f, e) in enumerate(zip(freqs, eff)):
        if e >= thresh:
            return float(f), float(e), k
    return None, None, None
