"""GPU: autograd Functions (vox8.autograd) against the adjoint kernels, the
oracle's filter and central finite differences.

The finite-difference test is the 220x32x24 one-SIMP-iteration workload
(1 mm voxels, seeded density 0.3 +- 0.05, cone filter r = 1.5 voxels, LC1):
gradients of compliance and of the mode-1 eigenvalue with respect to the
design variables x (through the filter) on 64 seeded random elements, central
differences with h = 1e-4.  Compliance is differenced through the stationary
functional J = 2 F.u - u.K u with u re-solved at each perturbed density (its
error is quadratic in the solve error); the eigenvalue is re-solved by
warm-started LOBPCG (Ritz-value error quadratic in the residual).
"""
import numpy as np
import pytest
import torch

from helpers import MaskedBox, random_density, rel

pytestmark = pytest.mark.gpu

MAT = dict(E=1.7e9, nu=0.4, rho=1010.0)


def test_filter_forward_and_adjoint(gpu):
    from oracle.topopt import cone_kernel, filter_field as ofilter
    from vox8.autograd import DensityFilter
    from vox8.fem import FEM3D, Material
    geom = MaskedBox(9, 5, 4, hole_frac=0.2, seed=3)
    fem = FEM3D(geom, Material(**MAT))
    mask_grid = geom.build_masks()["design"]                 # (nx, ny, nz)
    mask = mask_grid.transpose(2, 1, 0).reshape(-1)          # element-grid order
    flt = DensityFilter(fem, 1.5, mask)
    rng = np.random.default_rng(0)
    x = torch.tensor(rng.random(mask.size), device=gpu)
    y = torch.tensor(rng.random(mask.size), device=gpu)
    fx = flt(x)
    ref = ofilter(x.cpu().numpy().reshape(geom.nz, geom.ny, geom.nx).transpose(2, 1, 0),
                  mask_grid, cone_kernel(1.5))
    assert rel(fx.cpu().numpy(), ref.transpose(2, 1, 0).reshape(-1)) < 1e-14
    xr = x.clone().requires_grad_(True)
    (flt(xr) * y).sum().backward()
    # <H x, y> = <x, H^T y>
    assert abs(float((fx * y).sum()) - float((x * xr.grad).sum())) < 1e-12 * float((fx * y).sum())
    assert float(xr.grad[torch.as_tensor(~mask, device=gpu)].abs().max()) == 0.0


def test_autograd_equals_adjoint_kernels(gpu):
    from vox8.autograd import compliance, eigenvalues
    from vox8.fem import FEM3D, Material
    geom = MaskedBox(10, 4, 3, hole_frac=0.1, seed=5)
    fem = FEM3D(geom, Material(**MAT))
    rho0 = random_density(fem.n_active, 7, 0.05, 1.0)
    F = np.zeros((2, geom.n_dofs))
    tip = geom.node_index(10, np.arange(5)[:, None], np.arange(4)[None, :]).ravel()
    F[0, 3 * tip + 2] = -1.0
    F[1, 3 * tip + 1] = 0.5
    w = np.array([0.7, 0.3])
    rho = torch.tensor(rho0, device=gpu, requires_grad=True)
    c = compliance(rho, fem, F, w, rtol=1e-12)
    (g,) = torch.autograd.grad(c, rho)
    U = fem.last_U
    s = fem.sensitivities(rho0, U=U, weights=w)
    assert rel(g.cpu().numpy(), fem.active_from_grid(s["dc"]).cpu().numpy()) < 1e-14
    assert abs(float(c.detach()) / float((torch.tensor(F, device=gpu) * U).sum(1) @ torch.tensor(w, device=gpu)) - 1) < 1e-12
    nodes = tip[:4]
    lam = eigenvalues(rho, fem, 4, tip_mass_kg=1e-5, tip_nodes=nodes, tol=1e-9 * 1.7e9 * 1e-3)
    (gl,) = torch.autograd.grad(lam[0] + 0.5 * lam[2], rho)
    X = fem.last_modes
    ref = (fem.eigenvalue_sensitivity(rho0, float(lam[0]), X[0], 1e-5, nodes)
           + 0.5 * fem.eigenvalue_sensitivity(rho0, float(lam[2]), X[2], 1e-5, nodes))
    assert rel(gl.cpu().numpy(), ref.cpu().numpy()) < 1e-12


def test_c2_gradients_vs_central_fd_64_elements(gpu):
    from vox8.autograd import DensityFilter, compliance, eigenvalues
    from vox8.fem import FEM3D, Material
    from vox8.workloads import E_PA, NU, RHO_MAT, WORKLOADS, load_cases, seeded_density
    wl = WORKLOADS["c2"]
    geom = wl.geometry()
    fem = FEM3D(geom, Material(E=E_PA, nu=NU, rho=RHO_MAT), precision="mixed")
    F = load_cases(geom, wl.loads)
    x0 = seeded_density(geom, wl.base, wl.amp, wl.seed)
    flt = DensityFilter(fem, wl.r_min_vox)
    x = torch.tensor(x0, device=gpu, requires_grad=True)
    rho = flt(x)
    c = compliance(rho, fem, F, rtol=1e-12)
    (gc,) = torch.autograd.grad(c, x)
    U0 = fem.last_U.clone()
    ltol = 1e-7
    lam = eigenvalues(flt(x), fem, 6, tol=ltol, maxiter=4000)
    (gl,) = torch.autograd.grad(lam[0], x)
    X0 = fem.last_modes.clone()
    Ft = torch.tensor(F, device=gpu)
    h = 1e-4
    idx = np.random.default_rng(2024).choice(geom.n_elements, 64, replace=False)
    err_c, err_l = [], []
    gc_h, gl_h = gc.cpu().numpy(), gl.cpu().numpy()
    for e in idx:
        Js, Ls = [], []
        for sgn in (1.0, -1.0):
            xp = x.detach().clone()
            xp[e] += sgn * h
            r = flt(xp)
            U, info = fem.solve_block(r, Ft, x0=U0, rtol=1e-12, maxiter=20000)
            assert info["info"] == [0]
            u = U[0]
            Ku = fem.assemble_K(r) @ u
            Js.append(float(2 * Ft[0] @ u - u @ Ku))
            lm, _, linfo = fem.modal_block(r, 6, 0.0, np.array([], dtype=np.int64), X0=X0, tol=ltol,
                                           maxiter=4000)
            assert linfo["info"] == 0
            Ls.append(float(np.min(lm)))
        err_c.append(((Js[0] - Js[1]) / (2 * h) - gc_h[e]))
        err_l.append(((Ls[0] - Ls[1]) / (2 * h) - gl_h[e]))
    err_c, err_l = np.abs(err_c), np.abs(err_l)
    sc, sl = np.abs(gc_h[idx]), np.abs(gl_h[idx])
    print(f"compliance FD: max rel err {np.max(err_c / sc):.2e}; eigenvalue FD: max rel err "
          f"{np.max(err_l / sl):.2e} (max |err| / max |grad| {err_l.max() / sl.max():.2e})")
    assert np.all(err_c <= 1e-5 * sc)
    assert np.all(err_l <= 1e-5 * sl + 1e-5 * sl.max())
