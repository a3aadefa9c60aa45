"""Test helpers: small random problems shared by the oracle and GPU tests."""
from __future__ import annotations

import numpy as np


class MaskedBox:
    """Tiny box with a random set of inactive elements and a clamped x=0 face.

    Duck-typed geometry (same interface as the workload cantilever); used to
    exercise inactive elements, ghost nodes and fixed DOFs together.
    """

    def __init__(self, nx, ny, nz, h=1e-3, hole_frac=0.15, seed=0, clamp=True):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx = self.dy = self.dz = h
        self.n_elements = nx * ny * nz
        self.n_nodes = (nx + 1) * (ny + 1) * (nz + 1)
        self.n_dofs = 3 * self.n_nodes
        rng = np.random.default_rng(seed)
        self._in = rng.random((nx, ny, nz)) >= hole_frac
        self._in[0, :, :] = True
        self._clamp = clamp

    def node_index(self, ix, iy, iz):
        return iz * (self.nx + 1) * (self.ny + 1) + iy * (self.nx + 1) + ix

    def element_nodes(self, ex, ey, ez):
        offs = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))
        return np.array([self.node_index(ex + a, ey + b, ez + c) for a, b, c in offs])

    def build_masks(self):
        design = self._in.copy()
        design[0, :, :] = False          # a keep-solid layer at the root
        return dict(in_arm=self._in.copy(), design=design, keep_solid=self._in & ~design)

    def fixed_nodes(self):
        if not self._clamp:
            return np.array([], dtype=np.int64)
        iy, iz = np.meshgrid(np.arange(self.ny + 1), np.arange(self.nz + 1), indexing="ij")
        return np.sort(self.node_index(0, iy, iz).ravel()).astype(np.int64)


def random_density(n, seed=1, lo=0.05, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, n)


def rel(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


class TipBox(MaskedBox):
    """MaskedBox with the load/tip-mass interface the SIMP loops read
    (load_dofs_for_force, motor_mount_top_nodes): forces are spread evenly
    over the x = nx face nodes, the tip mass over that face's top edge."""

    def _tip_face(self):
        iy, iz = np.meshgrid(np.arange(self.ny + 1), np.arange(self.nz + 1), indexing="ij")
        return np.sort(self.node_index(self.nx, iy, iz).ravel()).astype(np.int64)

    def load_dofs_for_force(self, F_xyz):
        nodes = self._tip_face()
        dofs, mags = [], []
        for c, f in enumerate(F_xyz):
            if abs(f) > 1e-12:
                dofs.append(3 * nodes + c)
                mags.append(np.full(len(nodes), f / len(nodes)))
        if not dofs:
            return np.array([], dtype=np.int64), np.array([])
        return np.concatenate(dofs), np.concatenate(mags)

    def load_dofs_for_moment_Mz(self, Mz):
        return np.array([], dtype=np.int64), np.array([])

    def motor_mount_top_nodes(self):
        return np.sort(self.node_index(self.nx, np.arange(self.ny + 1), self.nz)).astype(np.int64)


# inner-solve tolerances of the SIMP fixture (the loop itself does not forward
# them: src/fem.py:558-561, src/topopt.py:315-321)
TIPBOX_SOLVE_RTOL = 1e-12
TIPBOX_MODAL_TOL = 1e-5
TIPBOX_MODAL_MAXITER = 4000


def tipbox_problem():
    """The SIMP fixture problem (tests/golden/simp_tipbox.npz)."""
    geom = TipBox(12, 4, 4, h=1e-3, hole_frac=0.08, seed=11)
    loads = [{"F": (0.0, 0.0, -1.0)}, {"F": (0.0, 0.6, 0.0)}]
    opts = dict(vol_frac=0.3, r_min_mm=1.5, move=0.2, max_iter=8, weights=[0.7, 0.3],
                min_freq_Hz=300.0, tip_mass_kg=2e-3, n_modes_track=3)
    mat = dict(E=1.7e9, nu=0.39, rho=930.0)
    return geom, loads, opts, mat
