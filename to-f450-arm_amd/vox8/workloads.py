"""Synthetic voxel cantilevers: geometry, seeded densities and load cases.

``Cantilever`` implements the duck-typed geometry interface FEM3D and
SIMPOptimizer read (the surface of src/mesh3d.py:42-146: nx, ny, nz, dx, dy,
dz, n_elements, n_nodes, n_dofs, node_index, element_nodes, build_masks,
fixed_nodes).  The x = 0 face is clamped; loads act on the x = nx face.  It has
no ``motor_mount_top_nodes``, so modal analysis adds no tip mass (as the
reference does for such a geometry, src/fem.py:426-432).

Workload definitions (sizes, material, density law, load cases):

=====  =====================  ========  =============================  ======
name   elements               voxel     density x (seeded uniform)     loads
=====  =====================  ========  =============================  ======
c1     110 x 16 x 12          2 mm      0.30 +- 0.05, unfiltered       LC1
c2     220 x 32 x 24          1 mm      0.30 +- 0.05, cone r=1.5       LC1
c3     220 x 32 x 24          1 mm      0.30 +- 0.05, cone r=1.5       LC1-5
c4     nx 80..220 x 32 x 24   1 mm      V_f +- 0.25 V_f, cone r=1.5    LC1
c5     440 x 64 x 48          0.5 mm    0.08 +- 0.02, cone r=1.5       LC1-5
c5t    as c5, 6 modes instead of 12 (the 5-LC + 6-mode SIMP iteration)
=====  =====================  ========  =============================  ======

Material PA12-like: E 1.7 GPa, nu 0.4, rho 1010 kg/m^3, SIMP p = 3.
Densities: x = clip(base + U(-a, a), 1e-3, 1) from numpy default_rng(seed), in
element order e = ex + nx (ey + ny ez).
Load cases (equal share per tip node, tip face centroid (yc, zc)):
LC1 Fz = -10 N; LC2 Mx = 0.5 N m as f_i = Mx (-dz_i, dy_i) / sum(dy^2 + dz^2) in
(y, z); LC3 (0, 5, -8.660254) N; LC4 Fy = +3 N; LC5 = 0.6 LC1 + 0.4 LC2.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

E_PA, NU, RHO_MAT = 1.7e9, 0.4, 1010.0


class Cantilever:
    def __init__(self, nx: int, ny: int, nz: int, h: float):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx = self.dy = self.dz = float(h)
        self.n_elements = self.nx * self.ny * self.nz
        self.n_nodes = (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        self.n_dofs = 3 * self.n_nodes

    def node_index(self, ix, iy, iz):
        return iz * (self.nx + 1) * (self.ny + 1) + iy * (self.nx + 1) + ix

    def element_index(self, ex, ey, ez):
        return ez * self.nx * self.ny + ey * self.nx + ex

    def element_nodes(self, ex, ey, ez):
        offs = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))
        return np.array([self.node_index(ex + a, ey + b, ez + c) for a, b, c in offs], dtype=np.int64)

    def build_masks(self):
        full = np.ones((self.nx, self.ny, self.nz), dtype=bool)
        return dict(in_arm=full, design=full.copy(), keep_solid=np.zeros_like(full))

    def _face(self, ix):
        iy, iz = np.meshgrid(np.arange(self.ny + 1), np.arange(self.nz + 1), indexing="ij")
        return np.sort(self.node_index(ix, iy, iz).ravel()).astype(np.int64)

    def fixed_nodes(self):
        return self._face(0)

    def tip_nodes(self):
        return self._face(self.nx)

    def tip_force(self, F_xyz):
        """Full-length load vector of force F spread evenly over the tip face."""
        F = np.zeros(self.n_dofs)
        t = self.tip_nodes()
        for c in range(3):
            F[3 * t + c] += F_xyz[c] / len(t)
        return F

    def tip_torque_x(self, Mx):
        """Moment about the x axis through the tip-face centroid, f ~ r_perp."""
        F = np.zeros(self.n_dofs)
        t = self.tip_nodes()
        rem = t % ((self.nx + 1) * (self.ny + 1))
        iy = rem // (self.nx + 1)
        iz = t // ((self.nx + 1) * (self.ny + 1))
        dy = iy * self.dy - 0.5 * self.ny * self.dy
        dz = iz * self.dz - 0.5 * self.nz * self.dz
        s = float(np.sum(dy * dy + dz * dz))
        F[3 * t + 1] += -Mx * dz / s
        F[3 * t + 2] += Mx * dy / s
        return F


def load_cases(geom: Cantilever, names=("LC1",)) -> np.ndarray:
    """(n_lc, n_dofs) load block for the named cases."""
    lc1 = geom.tip_force((0.0, 0.0, -10.0))
    lc2 = geom.tip_torque_x(0.5)
    table = {
        "LC1": lc1,
        "LC2": lc2,
        "LC3": geom.tip_force((0.0, 5.0, -8.660254)),
        "LC4": geom.tip_force((0.0, 3.0, 0.0)),
        "LC5": 0.6 * lc1 + 0.4 * lc2,
    }
    return np.stack([table[n] for n in names])


def seeded_density(geom: Cantilever, base: float, amp: float, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.clip(base + rng.uniform(-amp, amp, size=geom.n_elements), 1e-3, 1.0)


@dataclass
class Workload:
    name: str
    nx: int
    ny: int
    nz: int
    h: float
    base: float
    amp: float
    filtered: bool
    loads: tuple
    n_modes: int
    pcg_rtol: float
    modal_tol: float
    seed: int = 0
    r_min_vox: float = 1.5
    eff_dir: int = None         # translation direction of the effective-mass ranking (c5: z)

    def geometry(self) -> Cantilever:
        return Cantilever(self.nx, self.ny, self.nz, self.h)


ALL_LC = ("LC1", "LC2", "LC3", "LC4", "LC5")
WORKLOADS = {
    "c1": Workload("c1", 110, 16, 12, 2e-3, 0.30, 0.05, False, ("LC1",), 6, 1e-8, 1e-6),
    "c2": Workload("c2", 220, 32, 24, 1e-3, 0.30, 0.05, True, ("LC1",), 6, 1e-8, 1e-6),
    "c3": Workload("c3", 220, 32, 24, 1e-3, 0.30, 0.05, True, ALL_LC, 6, 1e-8, 1e-6),
    # 12 modes ranked by their z-direction effective mass
    "c5": Workload("c5", 440, 64, 48, 0.5e-3, 0.08, 0.02, True, ALL_LC, 12, 1e-10, 1e-6, eff_dir=2),
    # the C5 grid with the 6-mode SIMP iteration (5 load cases + 6 modes + adjoint)
    "c5t": Workload("c5t", 440, 64, 48, 0.5e-3, 0.08, 0.02, True, ALL_LC, 6, 1e-10, 1e-6),
}


def c4_designs():
    """The 16 uneven designs: nx = rint(linspace(80, 220, 16)), V_f = linspace(0.08, 0.5, 16)."""
    nxs = np.rint(np.linspace(80, 220, 16)).astype(int)
    vfs = np.linspace(0.08, 0.50, 16)
    return [Workload(f"c4_{d}", int(nx), 32, 24, 1e-3, float(vf), 0.25 * float(vf), True,
                     ("LC1",), 6, 1e-8, 1e-6, seed=d) for d, (nx, vf) in enumerate(zip(nxs, vfs))]
