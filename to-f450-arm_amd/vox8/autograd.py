"""torch.autograd Functions over the voxel FEM hot path.

The reference differentiates by hand inside its optimiser loop
(src/topopt.py:152-179 with src/fem.py:472-526 and 605-615); here the same
adjoint kernels back three autograd Functions, so any torch objective of the
physical densities can be differentiated:

* ``compliance(rho, fem, F, weights)``: c = sum_j w_j F_j . U_j with
  K(rho) U_j = F_j;  dc/drho_e = -p (1 - e_min) rho_e^(p-1) sum_j w_j u_je^T k0 u_je.
* ``eigenvalues(rho, fem, k, ...)``: the k lowest lam of K phi = lam M phi;
  dlam/drho_e = phi_e^T (dK_e - lam dM_e) phi_e / phi^T M phi (simple eigenvalues).
* ``DensityFilter``: rho = H x / Hs on the design mask, backward H^T (g / Hs),
  the exact transpose (the reference's loop filters sensitivities with the
  forward filter, src/topopt.py:162; its SIMPOptimizer here does the same).

Densities are float64 device tensors in active-element order (element-grid
order for the filter); every forward and backward runs in libvox8.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import ptr
from .fem import FEM3D

__all__ = ["compliance", "eigenvalues", "DensityFilter"]


class _Compliance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, fem: FEM3D, F, weights, rtol, maxiter, x0):
        rho_d = rho.detach()
        U, info = fem.solve_block(rho_d, F, x0=x0, rtol=rtol, maxiter=maxiter)
        if any(i != 0 for i in info["info"]):
            raise RuntimeError(f"vox8 PCG did not converge: {info}")
        s = fem.sensitivities(rho_d, U=U, weights=weights)
        ctx.save_for_backward(fem.active_from_grid(s["dc"]))
        fem.last_U = U
        return torch.tensor(float(np.dot(weights, s["comp"])), dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, g):
        (dc,) = ctx.saved_tensors
        return g * dc, None, None, None, None, None, None


def compliance(rho, fem: FEM3D, F, weights=None, rtol=1e-10, maxiter=20000, x0=None):
    """Weighted compliance sum_j w_j F_j . U_j (scalar tensor), differentiable in rho.

    F: (n_lc, n_dofs) loads; the solution block is left in ``fem.last_U``.
    """
    Ft, _ = fem._to_dev(F)
    Ft = Ft.reshape(-1, fem.geom.n_dofs)
    w = np.ones(Ft.shape[0]) if weights is None else np.asarray(weights, dtype=float)
    if w.shape != (Ft.shape[0],):
        raise ValueError(f"expected {Ft.shape[0]} weights, got {w.shape}")
    return _Compliance.apply(rho, fem, Ft, w, rtol, maxiter, x0)


class _Eigenvalues(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, fem: FEM3D, k, tip_mass_kg, tip_nodes, tol, maxiter, X0):
        rho_d = rho.detach()
        lam, X, info = fem.modal_block(rho_d, k, tip_mass_kg, tip_nodes, X0=X0, tol=tol,
                                       maxiter=maxiter)
        if info["info"] != 0:
            raise RuntimeError(f"vox8 LOBPCG did not converge: {info}")
        order = np.argsort(lam)
        lam = lam[order]
        X = X[torch.as_tensor(order, device=X.device)].contiguous()
        ctx.fem, ctx.tip = fem, (tip_mass_kg, tip_nodes)
        ctx.lam, ctx.X = lam, X
        ctx.save_for_backward(rho_d)
        fem.last_modes = X
        return torch.as_tensor(lam, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, g):
        (rho,) = ctx.saved_tensors
        fem = ctx.fem
        fem._set_tip(*ctx.tip)
        grad = torch.zeros_like(rho)
        gh = g.detach().cpu().numpy()
        for j in np.nonzero(gh)[0]:
            s = fem.sensitivities(rho, phi=ctx.X[j], lam=float(ctx.lam[j]))
            grad += float(gh[j]) * fem.active_from_grid(s["dlam"])
        return grad, None, None, None, None, None, None, None


def eigenvalues(rho, fem: FEM3D, k=6, tip_mass_kg=0.0, tip_nodes=None, tol=1e-6, maxiter=2000,
                X0=None):
    """The k lowest eigenvalues lam = omega^2 (ascending), differentiable in rho.

    Gradients are exact for simple eigenvalues; the modes (k, n_dofs) are
    left in ``fem.last_modes``.
    """
    if tip_nodes is None:
        tip_nodes = np.array([], dtype=np.int64)
    return _Eigenvalues.apply(rho, fem, int(k), float(tip_mass_kg), np.asarray(tip_nodes, dtype=np.int64),
                              tol, maxiter, X0)


class DensityFilter:
    """Masked cone filter of radius r_min_vox (voxels) as a differentiable map on
    element-grid tensors (src/topopt.py:32-58)."""

    def __init__(self, fem: FEM3D, r_min_vox: float, design_mask=None):
        from .topopt import _GridContext
        g = fem.geom
        self.n = g.nx * g.ny * g.nz
        self.ctx = _GridContext(g.nx, g.ny, g.nz)
        m = (np.ones(self.n, dtype=np.uint8) if design_mask is None
             else np.asarray(design_mask, dtype=np.uint8).reshape(-1))
        self.mask = torch.as_tensor(m, device=self.ctx.device)
        self.ctx._call("vox8_filter_setup", float(r_min_vox), ptr(self.mask))

    def _run(self, name, x):
        x = x.detach().to(torch.float64).contiguous()
        if x.numel() != self.n:
            raise ValueError(f"expected {self.n} element values, got {x.numel()}")
        out = torch.empty_like(x)
        self.ctx._call(name, ptr(x), ptr(out))
        return out

    def __call__(self, x):
        return _Filter.apply(x, self)


class _Filter(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, flt: DensityFilter):
        ctx.flt = flt
        return flt._run("vox8_filter", x)

    @staticmethod
    def backward(ctx, g):
        return ctx.flt._run("vox8_filter_adjoint", g), None
