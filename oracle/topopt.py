"""Oracle: density filter, OC update and the SIMP loop (test infrastructure only).

* ``cone_kernel``          -- src/topopt.py:32-41
* ``filter_field``         -- src/topopt.py:44-58 (scipy.ndimage.convolve, constant mode)
* ``SIMPOptimizer``        -- src/topopt.py:64-427 (__init__, _density_to_grid,
                              _select_mode_by_mac, _oc_update, run)
"""
from __future__ import annotations

import time

import numpy as np
import scipy.ndimage as ndi

from .fem import tip_nodes_of


def cone_kernel(r_min_vox: float) -> np.ndarray:
    """w(d) = max(0, r - d), Euclidean d in voxels (src/topopt.py:32-41)."""
    r = int(np.ceil(r_min_vox))
    c = np.arange(-r, r + 1)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    return np.maximum(0.0, r_min_vox - np.sqrt(X ** 2 + Y ** 2 + Z ** 2))


def filter_field(field_grid, mask_grid, kernel):
    """Masked normalised convolution (src/topopt.py:44-58)."""
    num = ndi.convolve(field_grid * mask_grid, kernel, mode="constant", cval=0.0)
    den = ndi.convolve(mask_grid.astype(float), kernel, mode="constant", cval=0.0)
    out = np.zeros_like(field_grid)
    out[mask_grid] = num[mask_grid] / np.maximum(den[mask_grid], 1e-12)
    return out


def oc_update(x, dc, V_e, target_volume, move):
    """Bisection on the OC multiplier (src/topopt.py:190-214)."""
    l1, l2 = 1e-9, 1e9
    x_new = np.empty_like(x)
    sens = np.abs(dc)
    while (l2 - l1) / (l1 + l2) > 1e-4:
        lmid = 0.5 * (l1 + l2)
        tmp = x * np.sqrt(sens / (lmid * V_e))
        x_new = np.maximum(0.001, np.maximum(x - move, np.minimum(1.0, np.minimum(x + move, tmp))))
        if x_new.sum() * V_e > target_volume:
            l1 = lmid
        else:
            l2 = lmid
    return x_new


def select_mode_by_mac(modes_new, v_prev):
    """MAC argmax (src/topopt.py:178-188)."""
    num = modes_new.T @ v_prev
    a2 = (modes_new ** 2).sum(axis=0)
    b2 = float((v_prev ** 2).sum())
    return int(np.argmax(num ** 2 / np.maximum(a2 * b2, 1e-24)))


class SIMPOptimizer:
    """Restatement of src/topopt.py:64-427.

    ``F_list`` may be passed directly (the synthetic workloads build their
    load vectors themselves); otherwise load dicts are turned into vectors with
    the geometry's ``load_dofs_for_force`` / ``load_dofs_for_moment_Mz``
    (src/topopt.py:148-158).  ``solve_rtol`` defaults to the reference's
    effective 1e-5 (src/fem.py:558-561 does not forward rtol).
    """

    def __init__(self, geom, fem, load_cases, vol_frac=0.30, r_min_mm=2.0, move=0.2,
                 max_iter=100, tol_change=0.01, weights=None, min_freq_Hz=None,
                 tip_mass_kg=0.060, n_modes_track=4, mu_freq_init=0.10, mu_freq_max=50.0,
                 mu_freq_ramp_every=5, mu_freq_ramp_factor=1.5, F_list=None,
                 solve_rtol=None, modal_tol=1e-4, modal_maxiter=500):
        self.geom, self.fem, self.load_cases = geom, fem, load_cases
        self.vol_frac = vol_frac
        self.r_min_vox = max(1.0, r_min_mm / (geom.dx * 1e3))
        self.move, self.max_iter, self.tol_change = move, max_iter, tol_change
        self.min_freq_Hz, self.tip_mass_kg = min_freq_Hz, tip_mass_kg
        self.n_modes_track = max(1, int(n_modes_track))
        self.mu_freq, self.mu_freq_max = mu_freq_init, mu_freq_max
        self.mu_freq_ramp_every = max(1, int(mu_freq_ramp_every))
        self.mu_freq_ramp_factor = mu_freq_ramp_factor
        self.solve_rtol, self.modal_tol, self.modal_maxiter = solve_rtol, modal_tol, modal_maxiter
        n_lc = len(F_list) if F_list is not None else len(load_cases)
        w = np.asarray(weights if weights is not None else [1.0 / n_lc] * n_lc, dtype=float)
        self.weights = w / w.sum()
        masks = geom.build_masks()
        self.in_arm, self.design, self.keep_solid = masks["in_arm"], masks["design"], masks["keep_solid"]
        self.active_idx = fem.elem_grid_idx
        ex, ey, ez = self.active_idx.T
        self.is_design_active = self.design[ex, ey, ez]
        self.n_design = int(self.is_design_active.sum())
        self.kernel = cone_kernel(self.r_min_vox)
        self.design_mask_grid = self.design.astype(bool)
        if F_list is None:
            F_list = []
            for lc in load_cases:
                F = np.zeros(geom.n_dofs)
                if lc.get("F") is not None:
                    d, m = geom.load_dofs_for_force(lc["F"])
                    F[d] += m
                if lc.get("Mz") is not None:
                    d, m = geom.load_dofs_for_moment_Mz(lc["Mz"])
                    F[d] += m
                F_list.append(F)
        self.F_list = list(F_list)
        self.V_e = geom.dx * geom.dy * geom.dz
        self.design_volume = self.n_design * self.V_e
        self.target_volume = vol_frac * self.design_volume

    def _to_grid(self, v):
        g = np.zeros((self.geom.nx, self.geom.ny, self.geom.nz))
        ex, ey, ez = self.active_idx.T
        g[ex, ey, ez] = v
        return g

    def _to_active(self, g):
        ex, ey, ez = self.active_idx.T
        return g[ex, ey, ez]

    def _solve(self, rho, F, x0):
        fem = self.fem
        if self.solve_rtol is None:
            return fem.solve(rho, F, method="cg", x0=x0)
        return fem.apply_BCs_and_solve(fem.assemble_K(rho), F, method="cg", x0=x0,
                                       rtol=self.solve_rtol)

    def run(self, verbose=False, rho_init=None):
        geom, fem = self.geom, self.fem
        if rho_init is None:
            rho_active = np.where(self.is_design_active, self.vol_frac, 1.0)
        else:
            rho_active = np.asarray(rho_init, dtype=float).copy()
        rho_grid = self._to_grid(rho_active)
        rho_phys_grid = rho_grid.copy()
        u_prev = [None] * len(self.F_list)
        hist = {k: [] for k in ("compliance", "mass", "change", "vol_frac",
                                "omega_1_Hz", "mu_freq", "freq_violation", "time_s")}
        modal = self.min_freq_Hz is not None
        tip = tip_nodes_of(geom) if modal else np.array([], dtype=np.int64)
        X0 = None
        v_prev = None
        w_target = (2.0 * np.pi * self.min_freq_Hz) ** 2 if modal else None
        p, emin = fem.SIMP_PENALTY, fem.E_MIN_FRAC
        u_list = []
        for it in range(1, self.max_iter + 1):
            t0 = time.time()
            phys = filter_field(rho_grid, self.design_mask_grid, self.kernel)
            phys[~self.design_mask_grid] = 0.0
            full = np.where(self.design, phys, np.where(self.in_arm, 1.0, 0.0))
            rho_phys = self._to_active(full)
            sens = np.zeros((geom.nx, geom.ny, geom.nz))
            c_total = 0.0
            u_list = []
            for i, F in enumerate(self.F_list):
                u = self._solve(rho_phys, F, u_prev[i])
                u_list.append(u)
                u_prev[i] = u
                ce = fem.elem_compliance(u, rho_phys)
                c_total += self.weights[i] * float(((emin + (1.0 - emin) * rho_phys ** p) * ce).sum())
                sens += self.weights[i] * self._to_grid(-p * (1.0 - emin) * rho_phys ** (p - 1) * ce)
            sens_f = filter_field(sens, self.design_mask_grid, self.kernel)
            if modal:
                freqs, modes, lam = fem.modal_analysis(
                    rho_phys, n_modes=self.n_modes_track, tip_mass_kg=self.tip_mass_kg,
                    tip_mass_nodes=tip, X0_free=X0, tol=self.modal_tol, maxiter=self.modal_maxiter)
                X0 = modes[fem.free_dofs, :].copy()
                k = 0 if v_prev is None else select_mode_by_mac(modes, v_prev)
                lam1, f1, v1 = float(lam[k]), float(freqs[k]), modes[:, k]
                v_prev = v1.copy()
                viol = max(0.0, 1.0 - lam1 / w_target)
                if viol > 0.0:
                    dlam = fem.eigenvalue_sensitivity(rho_phys, lam1, v1,
                                                      tip_mass_kg=self.tip_mass_kg,
                                                      tip_mass_nodes=tip)
                    dlam_f = filter_field(self._to_grid(dlam), self.design_mask_grid, self.kernel)
                    s_c = float(np.mean(np.abs(sens_f[self.design_mask_grid])))
                    s_m = float(np.mean(np.abs(dlam_f[self.design_mask_grid])))
                    sens_f = sens_f - self.mu_freq * viol * (s_c / max(s_m, 1e-30)) * dlam_f
                hist["omega_1_Hz"].append(f1)
                hist["mu_freq"].append(self.mu_freq)
                hist["freq_violation"].append(viol)
                if it % self.mu_freq_ramp_every == 0 and viol > 0.0 and self.mu_freq < self.mu_freq_max:
                    self.mu_freq = min(self.mu_freq_max, self.mu_freq * self.mu_freq_ramp_factor)
            else:
                for k in ("omega_1_Hz", "mu_freq", "freq_violation"):
                    hist[k].append(np.nan)
            x_new = oc_update(rho_grid[self.design], sens_f[self.design], self.V_e,
                              self.target_volume, self.move)
            rho_grid_new = rho_grid.copy()
            rho_grid_new[self.design] = x_new
            check = filter_field(rho_grid_new, self.design_mask_grid, self.kernel)
            change = float(np.max(np.abs(check - rho_phys_grid)))
            rho_grid, rho_phys_grid = rho_grid_new, check
            m = x_new.sum() * self.V_e
            hist["compliance"].append(c_total)
            hist["mass"].append(m)
            hist["change"].append(change)
            hist["vol_frac"].append(m / self.design_volume)
            hist["time_s"].append(time.time() - t0)
            if verbose:
                print(f"  iter {it:3d} c={c_total:.6e} Vf={m / self.design_volume:.4f} change={change:.4f}")
            if change < self.tol_change and it > 10:
                break
        final = filter_field(rho_grid, self.design_mask_grid, self.kernel)
        full = np.where(self.design, final, np.where(self.in_arm, 1.0, 0.0))
        return {"rho_active": self._to_active(full), "rho_grid": full, "history": hist,
                "u_per_lc": u_list}
