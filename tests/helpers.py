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
