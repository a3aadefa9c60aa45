"""Oracle: voxel geometries (test infrastructure only, see oracle/__init__.py).

Two geometries with the duck-typed interface FEM3D reads
(``nx, ny, nz, dx, dy, dz, n_elements, n_nodes, n_dofs, node_index,
element_nodes, build_masks, fixed_nodes``):

* ``Cantilever`` -- the synthetic full-box cantilevers the workloads are defined on:
  x=0 face clamped, loads spread evenly over the x=nx face.  It has no
  ``motor_mount_top_nodes`` so modal analysis adds no tip mass
  (src/fem.py:426-432).
* ``ArmGeometry3D`` -- restatement of the reference production geometry
  (src/mesh3d.py:42-365), used only to reproduce the reference's committed
  modal cross-check (notes/solver_crosscheck.json:6-12).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


class _Grid:
    """Node/element numbering shared by both geometries.

    Node index ix + (nx+1)(iy + (ny+1) iz)      (src/mesh3d.py:134-135)
    Element local node order                      (src/mesh3d.py:140-146)
    """

    def _finish(self):
        self.n_elements = self.nx * self.ny * self.nz
        self.n_nodes = (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        self.n_dofs = 3 * self.n_nodes

    def node_index(self, ix, iy, iz):
        return iz * (self.nx + 1) * (self.ny + 1) + iy * (self.nx + 1) + ix

    def element_index(self, ex, ey, ez):
        return ez * self.nx * self.ny + ey * self.nx + ex

    def element_nodes(self, ex, ey, ez):
        out = []
        for dz_ in (0, 1):
            for dy_, dx_ in ((0, 0), (0, 1), (1, 1), (1, 0)):
                out.append(self.node_index(ex + dx_, ey + dy_, ez + dz_))
        return np.array(out, dtype=np.int64)

    def node_coord(self, ix, iy, iz):
        return (ix * self.dx, iy * self.dy, iz * self.dz)


class Cantilever(_Grid):
    """Full-box cantilever nx x ny x nz of cubic voxels of edge h [m]."""

    def __init__(self, nx: int, ny: int, nz: int, h: float):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx = self.dy = self.dz = float(h)
        self._finish()

    def build_masks(self):
        full = np.ones((self.nx, self.ny, self.nz), dtype=bool)
        return dict(in_arm=full, design=full.copy(),
                    keep_solid=np.zeros_like(full))

    def fixed_nodes(self):
        iy, iz = np.meshgrid(np.arange(self.ny + 1), np.arange(self.nz + 1), indexing="ij")
        return np.sort(self.node_index(0, iy, iz).ravel()).astype(np.int64)

    def tip_nodes(self):
        iy, iz = np.meshgrid(np.arange(self.ny + 1), np.arange(self.nz + 1), indexing="ij")
        return np.sort(self.node_index(self.nx, iy, iz).ravel()).astype(np.int64)


@dataclass
class ArmGeometry3D(_Grid):
    """Restatement of src/mesh3d.py:42-123 (production 216x42x54 mm arm)."""

    L_x: float = 216e-3
    L_y: float = 42e-3
    L_z: float = 54e-3
    dx: float = 2e-3
    dy: float = 2e-3
    dz: float = 2e-3
    screw_dia: float = 2.5e-3
    screws: tuple = (
        (25.8e-3, 9.5e-3, 33.2e-3, 53.2e-3),
        (25.8e-3, 32.5e-3, 33.2e-3, 53.2e-3),
        (5.8e-3, 12.7e-3, 46.9e-3, 53.2e-3),
        (5.8e-3, 29.2e-3, 46.9e-3, 53.2e-3),
        (25.5e-3, 10.5e-3, 15.2e-3, 35.2e-3),
        (25.5e-3, 31.5e-3, 15.2e-3, 35.2e-3),
    )                                                   # src/mesh3d.py:72-87
    motor_dia: float = 10e-3
    motor_centre: tuple = (195.6e-3, 21e-3)             # src/mesh3d.py:92-93
    keep_solid_ring: float = 4.0e-3
    motor_keep_ring: float = 8.0e-3
    motor_pad_depth: float = 4e-3                       # src/mesh3d.py:96-102
    nx: int = field(init=False)
    ny: int = field(init=False)
    nz: int = field(init=False)

    def __post_init__(self):
        self.nx = int(round(self.L_x / self.dx))
        self.ny = int(round(self.L_y / self.dy))
        self.nz = int(round(self.L_z / self.dz))
        self._finish()

    def build_masks(self):
        """src/mesh3d.py:149-200."""
        xs = (np.arange(self.nx) + 0.5) * self.dx
        ys = (np.arange(self.ny) + 0.5) * self.dy
        zs = (np.arange(self.nz) + 0.5) * self.dz
        XC, YC, ZC = np.meshgrid(xs, ys, zs, indexing="ij")
        in_screw = np.zeros_like(XC, dtype=bool)
        for cx, cy, z0, z1 in self.screws:
            in_screw |= (((XC - cx) ** 2 + (YC - cy) ** 2) <= (self.screw_dia / 2.0) ** 2) \
                & (ZC >= z0) & (ZC <= z1)
        z_top = self.L_z - self.motor_pad_depth
        in_motor = (((XC - self.motor_centre[0]) ** 2 + (YC - self.motor_centre[1]) ** 2)
                    <= (self.motor_dia / 2.0) ** 2) & (ZC > z_top)
        in_arm = ~in_screw & ~in_motor
        keep_screw = np.zeros_like(XC, dtype=bool)
        r_keep = (self.screw_dia / 2.0 + self.keep_solid_ring) ** 2
        for cx, cy, z0, z1 in self.screws:
            keep_screw |= (((XC - cx) ** 2 + (YC - cy) ** 2) <= r_keep) \
                & (ZC >= z0 - self.dz) & (ZC <= z1 + self.dz)
        keep_screw &= in_arm
        keep_motor = ((((XC - self.motor_centre[0]) ** 2 + (YC - self.motor_centre[1]) ** 2)
                       <= (self.motor_dia / 2.0 + self.motor_keep_ring) ** 2) & (ZC > z_top))
        keep_motor &= in_arm
        keep_solid = keep_screw | keep_motor
        return dict(in_arm=in_arm, keep_solid=keep_solid, design=in_arm & ~keep_solid,
                    keep_screw=keep_screw, keep_motor=keep_motor)

    def fixed_nodes(self):
        """Screw bearing annuli over each screw's span (src/mesh3d.py:203-228)."""
        r_out = self.screw_dia / 2.0 + self.keep_solid_ring
        r_in = self.screw_dia / 2.0
        IX, IY = np.meshgrid(np.arange(self.nx + 1), np.arange(self.ny + 1), indexing="ij")
        X, Y = IX * self.dx, IY * self.dy
        iz = np.arange(self.nz + 1)
        Z = iz * self.dz
        fixed = []
        for cx, cy, z0, z1 in self.screws:
            d = np.hypot(X - cx, Y - cy)
            sel = (r_in <= d) & (d <= r_out)
            zsel = (z0 - self.dz <= Z) & (Z <= z1 + self.dz)
            for ix, iy in zip(IX[sel], IY[sel]):
                fixed.extend(self.node_index(ix, iy, iz[zsel]).tolist())
        return np.unique(np.asarray(fixed, dtype=np.int64))

    def motor_mount_top_nodes(self):
        """Top-face nodes within the motor pad radius (src/mesh3d.py:231-251)."""
        cx, cy = self.motor_centre
        r_out = self.motor_dia / 2.0 + self.motor_keep_ring
        out = []
        for ix in range(self.nx + 1):
            for iy in range(self.ny + 1):
                if np.hypot(ix * self.dx - cx, iy * self.dy - cy) <= r_out:
                    out.append(self.node_index(ix, iy, self.nz))
        return np.asarray(out, dtype=np.int64)

    def _ring_nodes(self):
        cx, cy = self.motor_centre
        r_in = self.motor_dia / 2.0
        r_out = self.motor_dia / 2.0 + self.motor_keep_ring
        iz0 = int(round((self.L_z - self.motor_pad_depth) / self.dz))
        rows = []
        for ix in range(self.nx + 1):
            for iy in range(self.ny + 1):
                x, y = ix * self.dx, iy * self.dy
                d = np.hypot(x - cx, y - cy)
                if r_in <= d <= r_out:
                    rows.append((ix, iy, x, y, d))
        return rows, iz0

    def load_dofs_for_force(self, F_xyz):
        """Fz over motor-pad top nodes, Fx/Fy cosine-weighted on the hole wall
        (src/mesh3d.py:273-337)."""
        Fx, Fy, Fz = F_xyz
        dofs, mags = [], []
        if abs(Fz) > 1e-12:
            top = self.motor_mount_top_nodes()
            dofs.append(3 * top + 2)
            mags.append((Fz / len(top)) * np.ones(len(top)))
        if abs(Fx) + abs(Fy) > 1e-12:
            cx, cy = self.motor_centre
            ld = np.asarray((Fx, Fy), dtype=float)
            ld /= np.linalg.norm(ld)
            rows, iz0 = self._ring_nodes()
            nodes, w = [], []
            for ix, iy, x, y, d in rows:
                cos_t = -(((x - cx) / d) * ld[0] + ((y - cy) / d) * ld[1])
                if cos_t <= 1e-6:
                    continue
                for iz in range(iz0, self.nz + 1):
                    nodes.append(self.node_index(ix, iy, iz))
                    w.append(cos_t)
            if nodes:
                nodes = np.asarray(nodes, dtype=np.int64)
                w = np.asarray(w)
                w /= w.sum()
                dofs += [3 * nodes, 3 * nodes + 1]
                mags += [Fx * w, Fy * w]
        if not dofs:
            return np.array([], dtype=np.int64), np.array([])
        return np.concatenate(dofs), np.concatenate(mags)

    def load_dofs_for_moment_Mz(self, Mz):
        """Tangential forces on the hole wall, F_t = Mz / sum(r) (src/mesh3d.py:340-365)."""
        cx, cy = self.motor_centre
        rows, iz0 = self._ring_nodes()
        nodes, r, c, s = [], [], [], []
        for ix, iy, x, y, d in rows:
            for iz in range(iz0, self.nz + 1):
                nodes.append(self.node_index(ix, iy, iz))
                r.append(d)
                c.append((x - cx) / d)
                s.append((y - cy) / d)
        nodes = np.asarray(nodes, dtype=np.int64)
        Ft = Mz / np.asarray(r).sum()
        return (np.concatenate([3 * nodes, 3 * nodes + 1]),
                np.concatenate([-Ft * np.asarray(s), Ft * np.asarray(c)]))
