"""Oracle: FEM3D restatement with assembled CSR matrices (test infrastructure only).

Follows src/fem.py:191-625 (class FEM3D) and the SciPy calls it makes:

* active-element order ez -> ey -> ex, index tables     src/fem.py:244-270
* fixed DOFs, ghost DOFs treated as fixed               src/fem.py:281-305
* assemble_K: f = e_min + (1-e_min) rho^p               src/fem.py:311-325
* Pedersen mass factor and its gradient                 src/fem.py:331-360
* assemble_M (+ tip lump over every returned tip node)  src/fem.py:362-398
* modal_analysis: Jacobi-preconditioned LOBPCG          src/fem.py:400-470
* eigenvalue_sensitivity                                src/fem.py:472-526
* apply_BCs_and_solve: Jacobi-PCG, spsolve fallback     src/fem.py:529-555
* solve (rtol not forwarded -> 1e-5)                    src/fem.py:558-561
* compute_stress / von_mises / principal_stress_max     src/fem.py:564-602
* elem_compliance / total_mass                          src/fem.py:605-625

The index tables are built vectorised instead of with the reference's Python
triple loop; the order and values are the same.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .element import element_B_centre, element_K, element_M, iso_D

# Local node offsets (dx, dy, dz) in reference order (src/mesh3d.py:140-146)
LOCAL_OFFSETS = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                          (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], dtype=np.int64)


@dataclass
class Material:
    """src/fem.py:182-188."""
    E: float = 1700e6
    nu: float = 0.39
    sigma_y: float = 38e6
    sigma_u: float = 48e6
    rho: float = 930.0


def tip_nodes_of(geom):
    """Tip-mass node lookup order of src/fem.py:426-432."""
    if hasattr(geom, "motor_mount_top_nodes"):
        return geom.motor_mount_top_nodes()
    if hasattr(geom, "_motor_top_boundary_nodes"):
        return geom._motor_top_boundary_nodes()
    return np.array([], dtype=np.int64)


class FEM3D:
    SIMP_PENALTY = 3.0
    E_MIN_FRAC = 1e-9
    MASS_THRESHOLD = 0.1
    MASS_PENALTY_Q = 6.0

    def __init__(self, geom, mat: Material, build_pattern: bool = True):
        self.geom, self.mat = geom, mat
        self.D_solid = iso_D(mat.E, mat.nu)
        self.Ke_solid = element_K(mat.E, mat.nu, geom.dx, geom.dy, geom.dz, D=self.D_solid)
        self.B_centre = element_B_centre(geom.dx, geom.dy, geom.dz)
        self.Me_solid = element_M(mat.rho, geom.dx, geom.dy, geom.dz)
        in_arm = geom.build_masks()["in_arm"].astype(bool)
        # (nx,ny,nz) grid -> active list in ez-outer, ex-fastest order
        ez, ey, ex = np.nonzero(in_arm.transpose(2, 1, 0))
        self.elem_grid_idx = np.stack([ex, ey, ez], axis=1).astype(np.int64)
        self.n_active = len(ex)
        nxp, nyp = geom.nx + 1, geom.ny + 1
        ix = ex[:, None] + LOCAL_OFFSETS[None, :, 0]
        iy = ey[:, None] + LOCAL_OFFSETS[None, :, 1]
        iz = ez[:, None] + LOCAL_OFFSETS[None, :, 2]
        self.elem_node_ids = (iz * nyp + iy) * nxp + ix
        self.elem_dof_ids = (3 * self.elem_node_ids[:, :, None]
                             + np.arange(3)[None, None, :]).reshape(self.n_active, 24)
        self.Ke_flat = self.Ke_solid.ravel()
        self.Me_flat = self.Me_solid.ravel()
        fixed_nodes = np.asarray(geom.fixed_nodes(), dtype=np.int64)
        fixed = np.concatenate([3 * fixed_nodes, 3 * fixed_nodes + 1, 3 * fixed_nodes + 2])
        used = np.zeros(geom.n_nodes, dtype=bool)
        used[self.elem_node_ids.ravel()] = True
        ghost = np.nonzero(~used)[0]
        fixed = np.concatenate([fixed, 3 * ghost, 3 * ghost + 1, 3 * ghost + 2])
        self.fixed_dofs = np.unique(fixed)
        free = np.ones(geom.n_dofs, dtype=bool)
        free[self.fixed_dofs] = False
        self.free_dofs = np.nonzero(free)[0]
        self.rho_active = np.ones(self.n_active)
        self._csr_pattern = None
        if build_pattern:
            self._build_pattern()

    # -- assembly ------------------------------------------------------------
    def _build_pattern(self):
        """CSR pattern plus, per COO entry of the (n_active*576) stream, its slot
        in the CSR data array, so that K(rho) is assembled with one bincount.
        Equivalent to coo_matrix((vals,(rows,cols))).tocsr() of src/fem.py:322-325."""
        d = self.elem_dof_ids
        rows = np.repeat(d, 24, axis=1).ravel()
        cols = np.tile(d, (1, 24)).ravel()
        n = self.geom.n_dofs
        key = rows * n + cols
        uniq, slot = np.unique(key, return_inverse=True)
        self._slot = slot.astype(np.int64)
        r = uniq // n
        c = uniq % n
        indptr = np.zeros(n + 1, dtype=np.int64)
        np.add.at(indptr, r + 1, 1)
        self._csr_pattern = (np.cumsum(indptr), c.astype(np.int64), len(uniq))

    def _assemble(self, factor, flat):
        if self._csr_pattern is None:
            self._build_pattern()
        indptr, indices, nnz = self._csr_pattern
        vals = (factor[:, None] * flat[None, :]).ravel()
        data = np.bincount(self._slot, weights=vals, minlength=nnz)
        n = self.geom.n_dofs
        return sp.csr_matrix((data, indices, indptr), shape=(n, n))

    def simp_factor(self, rho):
        return self.E_MIN_FRAC + (1.0 - self.E_MIN_FRAC) * rho ** self.SIMP_PENALTY

    def assemble_K(self, rho_active):
        return self._assemble(self.simp_factor(rho_active), self.Ke_flat)

    def mass_factor(self, rho):
        t, q = self.MASS_THRESHOLD, self.MASS_PENALTY_Q
        return np.where(rho >= t, rho, t * (np.maximum(rho, 0.0) / t) ** q)

    def mass_factor_grad(self, rho):
        t, q = self.MASS_THRESHOLD, self.MASS_PENALTY_Q
        return np.where(rho >= t, np.ones_like(rho), q * (np.maximum(rho, 0.0) / t) ** (q - 1.0))

    def assemble_M(self, rho_active, tip_mass_kg=0.0, tip_mass_nodes=None):
        M = self._assemble(self.mass_factor(rho_active), self.Me_flat)
        if tip_mass_kg > 0.0 and tip_mass_nodes is not None and len(tip_mass_nodes) > 0:
            nodes = np.asarray(tip_mass_nodes, dtype=np.int64)
            dofs = np.concatenate([3 * nodes, 3 * nodes + 1, 3 * nodes + 2])
            n = self.geom.n_dofs
            lump = sp.coo_matrix((np.full(len(dofs), tip_mass_kg / len(nodes)), (dofs, dofs)),
                                 shape=(n, n)).tocsr()
            M = (M + lump).tocsr()
        return M

    # -- static --------------------------------------------------------------
    def apply_BCs_and_solve(self, K, F, method="cg", x0=None, rtol=1e-5, info_out=None):
        free = self.free_dofs
        Kff = K[free][:, free].tocsc()
        Ff = F[free]
        if method == "cg":
            d = Kff.diagonal()
            Minv = sp.diags(1.0 / np.where(d != 0, d, 1.0))
            uf, info = spla.cg(Kff, Ff, M=Minv, rtol=rtol, maxiter=5000,
                               x0=None if x0 is None else x0[free])
            if info_out is not None:
                info_out.append(info)
            if info != 0:
                uf = spla.spsolve(Kff, Ff)
        else:
            uf = spla.spsolve(Kff, Ff)
        u = np.zeros(self.geom.n_dofs)
        u[free] = uf
        return u

    def solve(self, rho_active, F, method="cg", x0=None):
        return self.apply_BCs_and_solve(self.assemble_K(rho_active), F, method=method, x0=x0)

    # -- modal ---------------------------------------------------------------
    def modal_analysis(self, rho_active, n_modes=6, tip_mass_kg=0.060, tip_mass_nodes=None,
                       X0_free=None, tol=1e-4, maxiter=500):
        K = self.assemble_K(rho_active)
        if tip_mass_nodes is None:
            tip_mass_nodes = tip_nodes_of(self.geom)
        M = self.assemble_M(rho_active, tip_mass_kg=tip_mass_kg, tip_mass_nodes=tip_mass_nodes)
        free = self.free_dofs
        Kff = K[free][:, free].tocsc()
        Mff = M[free][:, free].tocsc()
        nf = Kff.shape[0]
        d = Kff.diagonal()
        T = sp.diags(1.0 / np.where(d > 0, d, 1.0))
        if X0_free is not None and X0_free.shape == (nf, n_modes):
            X0 = X0_free.copy()
        else:
            X0 = np.random.default_rng(0).standard_normal((nf, n_modes))
        try:
            lam, V = spla.lobpcg(Kff, X0, B=Mff, M=T, tol=tol, largest=False, maxiter=maxiter)
        except Exception:
            lam, V = spla.eigsh(Kff, k=n_modes, M=Mff, sigma=1.0, which="LM",
                                mode="normal", tol=1e-5, maxiter=2000)
        order = np.argsort(lam)
        lam = np.maximum(lam[order], 0.0)
        V = V[:, order]
        modes = np.zeros((self.geom.n_dofs, n_modes))
        modes[free, :] = V
        return np.sqrt(lam) / (2.0 * np.pi), modes, lam

    def eigenvalue_sensitivity(self, rho_active, omega_sq_k, v_k, tip_mass_kg=0.060,
                               tip_mass_nodes=None):
        p, emin = self.SIMP_PENALTY, self.E_MIN_FRAC
        dK = (1.0 - emin) * p * np.maximum(rho_active, 1e-9) ** (p - 1.0)
        dM = self.mass_factor_grad(rho_active)
        ve = v_k[self.elem_dof_ids]
        vKv = np.einsum("ij,jk,ik->i", ve, self.Ke_solid, ve)
        vMv_e = np.einsum("ij,jk,ik->i", ve, self.Me_solid, ve)
        if tip_mass_nodes is None:
            tip_mass_nodes = tip_nodes_of(self.geom)
        vMv = float(np.sum(self.mass_factor(rho_active) * vMv_e))
        if tip_mass_kg > 0.0 and len(tip_mass_nodes) > 0:
            t = np.asarray(tip_mass_nodes, dtype=np.int64)
            tip = np.concatenate([3 * t, 3 * t + 1, 3 * t + 2])
            vMv += float(tip_mass_kg / len(t) * np.sum(v_k[tip] ** 2))
        if vMv <= 0.0:
            vMv = 1.0
        return (dK * vKv - omega_sq_k * dM * vMv_e) / vMv

    # -- post-processing -----------------------------------------------------
    def compute_stress(self, u, rho_active):
        ue = u[self.elem_dof_ids]
        return rho_active[:, None] * ((ue @ self.B_centre.T) @ self.D_solid.T)

    @staticmethod
    def von_mises(s):
        vm2 = 0.5 * ((s[..., 0] - s[..., 1]) ** 2 + (s[..., 1] - s[..., 2]) ** 2
                     + (s[..., 2] - s[..., 0]) ** 2) \
            + 3.0 * (s[..., 3] ** 2 + s[..., 4] ** 2 + s[..., 5] ** 2)
        return np.sqrt(np.maximum(vm2, 0.0))

    @staticmethod
    def principal_stress_max(s):
        T = np.empty((s.shape[0], 3, 3))
        T[:, 0, 0], T[:, 1, 1], T[:, 2, 2] = s[:, 0], s[:, 1], s[:, 2]
        T[:, 0, 1] = T[:, 1, 0] = s[:, 3]
        T[:, 1, 2] = T[:, 2, 1] = s[:, 4]
        T[:, 0, 2] = T[:, 2, 0] = s[:, 5]
        return np.linalg.eigvalsh(T).max(axis=1)

    def elem_compliance(self, u, rho_active=None):
        ue = u[self.elem_dof_ids]
        return np.einsum("ij,jk,ik->i", ue, self.Ke_solid, ue)

    def total_mass(self, rho_active):
        return float(np.sum(rho_active) * self.geom.dx * self.geom.dy * self.geom.dz * self.mat.rho)
