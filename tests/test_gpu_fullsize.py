"""GPU: the headline grid (c5t: 440 x 64 x 48 voxels, 4.2 M DOFs) checked
through properties that do not depend on size (the oracle cannot assemble
this grid in test time):

* FP64 stencil: K and M symmetric (x.Ky = y.Kx) and linear, x.Kx > 0 and
  x.Mx > 0 on free vectors; the FP32 stencil within FP32 rounding of FP64.
* Static solve (multigrid-preconditioned block PCG, five load cases): the
  recursively updated residual meets rtol = 1e-10 (SciPy's cg stop test, as
  the reference), the true FP64 residual ||F - K u|| / ||F|| is within CG's
  attainable accuracy at this contrast (E_min 1e-9; bar 1e-8, measured 1e-9),
  and the compliance F.u equals the sum of the element energies from the
  adjoint kernel to the same floor (they differ by u.r).
* Modal solve (six modes): every pair's true residual ||K x - lam M x|| below
  the tolerance, X^T M X = I, eigenvalues ascending and positive.
* Multigrid preconditioner, both coarse-level forms: symmetric
  (r1.B r2 = r2.B r1 to FP32 rounding: bar 1e-5 of r1.B r1 rediscretised,
  5e-5 Galerkin, measured 2e-5, its coarse operators being far worse
  conditioned) and positive, as PCG and LOBPCG require.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c5t(gpu):
    from vox8.fem import FEM3D, Material
    from vox8.workloads import E_PA, NU, RHO_MAT, WORKLOADS, load_cases, seeded_density
    wl = WORKLOADS["c5t"]
    geom = wl.geometry()
    fem = FEM3D(geom, Material(E=E_PA, nu=NU, rho=RHO_MAT), precision="mixed")
    rho = torch.tensor(seeded_density(geom, wl.base, wl.amp, wl.seed), device=gpu)
    F = torch.tensor(load_cases(geom, wl.loads), device=gpu)
    free = torch.ones(geom.n_dofs, dtype=torch.bool, device=gpu)
    free[torch.as_tensor(fem.fixed_dofs, device=gpu)] = False
    return wl, geom, fem, rho, F, free


def _apply(fem, op, X, prec=None):
    from vox8 import _lib
    X = X.contiguous()
    Y = torch.empty_like(X)
    p = prec if prec is not None else (_lib.PREC_FP32 if X.dtype == torch.float32 else _lib.PREC_FP64)
    fem._call("vox8_apply", op, p, _lib.ptr(X), _lib.ptr(Y) if op == _lib.OP_K else None,
              _lib.ptr(Y) if op == _lib.OP_M else None, X.shape[0], X.shape[1])
    return Y


def test_fullsize_stencil_properties(c5t):
    from vox8 import _lib
    wl, geom, fem, rho, F, free = c5t
    fem._set_density(rho)
    g = torch.Generator(device=rho.device).manual_seed(5)
    X = torch.randn(3, geom.n_dofs, dtype=torch.float64, device=rho.device, generator=g) * free
    for op in (_lib.OP_K, _lib.OP_M):
        Y = _apply(fem, op, X)
        xy, yx = float(X[0] @ Y[1]), float(X[1] @ Y[0])
        assert abs(xy - yx) <= 1e-12 * float(Y[0].norm() * X[1].norm())
        lin = _apply(fem, op, (2.0 * X[0] - 3.0 * X[2]).reshape(1, -1))[0]
        assert float((lin - (2.0 * Y[0] - 3.0 * Y[2])).norm()) <= 1e-12 * float(lin.norm())
        assert float(X[0] @ Y[0]) > 0.0
    Yk = _apply(fem, _lib.OP_K, X)
    Y32 = _apply(fem, _lib.OP_K, X.float()).double()
    assert float((Y32 - Yk).norm() / Yk.norm()) < 1e-5


def test_fullsize_static_solve_true_residual_and_energy(c5t):
    from vox8 import _lib
    wl, geom, fem, rho, F, free = c5t
    U, info = fem.solve_block(rho, F, rtol=wl.pcg_rtol, maxiter=5000)
    assert info["info"] == [0] * F.shape[0]
    assert max(info["relres"]) < wl.pcg_rtol
    R = F * free - _apply(fem, _lib.OP_K, U) * free
    for j in range(F.shape[0]):
        rel = float(R[j].norm() / (F[j] * free).norm())
        assert rel < 1e-8, (j, rel)
    s = fem.sensitivities(rho, U=U, weights=np.ones(F.shape[0]))
    for j in range(F.shape[0]):
        fu = float(F[j] @ U[j])
        # u.Ku - F.u = -u.r: the same rounding floor as the true residual
        assert abs(s["comp"][j] / fu - 1) < 1e-8, (j, s["comp"][j], fu)


def test_fullsize_modal_residuals_and_orthonormality(c5t):
    from vox8 import _lib
    wl, geom, fem, rho, F, free = c5t
    k = wl.n_modes
    lam, X, info = fem.modal_block(rho, k, 0.0, np.array([], dtype=np.int64), tol=wl.modal_tol,
                                   maxiter=2000)
    assert info["info"] == 0, info
    order = np.argsort(lam)
    assert np.all(lam > 0) and np.all(np.diff(lam[order]) >= 0)
    KX = _apply(fem, _lib.OP_K, X)
    MX = _apply(fem, _lib.OP_M, X)
    lam_t = torch.tensor(lam, device=X.device)[:, None]
    res = ((KX - lam_t * MX) * free).norm(dim=1).cpu().numpy()
    assert np.all(res < 2 * wl.modal_tol), res   # the final Rayleigh-Ritz may move them slightly
    G = (X @ MX.T).cpu().numpy()
    assert np.abs(G - np.eye(k)).max() < 1e-8


@pytest.mark.parametrize("galerkin,bar", [(False, 1e-5), (True, 5e-5)])
def test_fullsize_multigrid_symmetric_positive(c5t, galerkin, bar):
    from vox8 import _lib
    wl, geom, fem, rho, F, free = c5t
    fem.set_mg_galerkin(galerkin)
    fem._set_density(rho)
    g = torch.Generator(device=rho.device).manual_seed(9)
    R = torch.randn(2, geom.n_dofs, dtype=torch.float64, device=rho.device, generator=g) * free
    Z = torch.empty_like(R)
    fem._call("vox8_precond_apply", _lib.ptr(R), _lib.ptr(Z), 2)
    fem.set_mg_galerkin(True)
    a, b = float(R[0] @ Z[1]), float(R[1] @ Z[0])
    assert abs(a - b) <= bar * float(R[0] @ Z[0]), (a, b)
    assert float(R[0] @ Z[0]) > 0 and float(R[1] @ Z[1]) > 0
