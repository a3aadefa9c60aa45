"""GPU: the C5 configuration as specified, at full size.

440x64x48 voxels at 0.5 mm (4,213,755 DOFs), seeded density 0.08 +- 0.02
filtered with the r = 1.5 cone, the five load cases solved together at rtol
1e-10 (FP64 PCG preconditioned by the FP32 multigrid), 12-mode LOBPCG at tol
1e-6 with the FP64 Rayleigh-Ritz, and the modes ranked by their z-direction
effective mass (device M products and FP64 reductions,
sims/modal_participation.py:37-94).

The oracle cannot assemble this grid in test time, so each output is checked
against the plain FP64 path (the reference's Jacobi preconditioner) on the
same density and by properties that fix it:

* per load case: compliance equal to the plain FP64 PCG's (rtol 1e-10)
  within 1e-6, and the FP64 true residual below 1e-8 of ||F||;
* per mode: absolute residual ||K x - lam M x|| < tol recomputed with FP64
  stencil products, X^T M X = I to 1e-8, eigenvalue equal to the Ritz value
  of an independent plain FP64 LOBPCG (cold start, different block) within
  1e-6, MAC >= 0.9999 against it;
* z effective mass: every fraction in [0, 1], their sum <= 1, the ranking
  orders them, and the fractions equal those from the plain modes within 1e-6;
* sensitivities: compliance and mode-1 eigenvalue sensitivities from the two
  paths' solutions within 1e-6 / 1e-5.
"""
import numpy as np
import pytest
import torch

from helpers import rel

pytestmark = pytest.mark.gpu


def _mac(a, b):
    return float((a @ b) ** 2 / ((a @ a) * (b @ b)))


def test_c5_twelve_modes_ranked_by_z_mass(gpu):
    from vox8 import _lib
    from vox8.fem import FEM3D, Material
    from vox8.modal_participation import effective_mass_device, rank_by_effective_mass
    from vox8.topopt import _elem_to_grid, _grid_to_elem, cone_kernel, filter_field
    from vox8.workloads import E_PA, NU, RHO_MAT, WORKLOADS, load_cases, seeded_density
    wl = WORKLOADS["c5"]
    assert wl.n_modes == 12 and wl.pcg_rtol == 1e-10 and wl.modal_tol == 1e-6
    geom = wl.geometry()
    assert geom.n_dofs == 4213755
    x = seeded_density(geom, wl.base, wl.amp, wl.seed)
    xg = _elem_to_grid(x, geom.nx, geom.ny, geom.nz)
    rho_np = _grid_to_elem(filter_field(xg, np.ones(xg.shape, dtype=bool), cone_kernel(wl.r_min_vox)))
    rho = torch.tensor(rho_np, device=gpu)
    F = torch.tensor(load_cases(geom, wl.loads), device=gpu)
    mat = Material(E=E_PA, nu=NU, rho=RHO_MAT)
    n, k = geom.n_dofs, wl.n_modes
    out = {}
    for prec in ("mixed", "fp64"):
        fem = FEM3D(geom, mat, precision=prec)
        U, info = fem.solve_block(rho, F, rtol=wl.pcg_rtol, maxiter=100000)
        assert info["info"] == [0] * 5, (prec, info)
        maxiter = 400 if prec == "mixed" else 6000
        lam, X, li = fem.modal_block(rho, k, 0.0, [], tol=wl.modal_tol, maxiter=maxiter,
                                     seed=0 if prec == "mixed" else 7)
        order = np.argsort(lam)
        lam = lam[order]
        X = X[torch.as_tensor(order, device=gpu)].contiguous()
        frac, L, g, total = effective_mass_device(fem, X, 2)
        s = fem.sensitivities(rho, U=U, weights=np.full(5, 0.2), phi=X[0], lam=float(lam[0]))
        out[prec] = dict(fem=fem, U=U, lam=lam, X=X, li=li, frac=frac, g=g, total=total,
                         comp=s["comp"], dc=s["dc"].cpu().numpy(), dlam=s["dlam"].cpu().numpy())
        print(prec, "pcg", info["iters"], "lobpcg", li)
    m = out["mixed"]
    fem, U, X, lam = m["fem"], m["U"], m["X"], m["lam"]
    assert m["li"]["info"] == 0
    # static: true FP64 residuals and compliance against the plain path
    KU = torch.empty_like(U)
    fem._call("vox8_apply", _lib.OP_K, _lib.PREC_FP64, _lib.ptr(U), _lib.ptr(KU), None, 5, n)
    free = torch.zeros(n, dtype=torch.float64, device=gpu)
    free[torch.as_tensor(fem.free_dofs, device=gpu)] = 1.0
    for j in range(5):
        r = (F[j] * free - KU[j]).norm() / (F[j] * free).norm()
        assert float(r) < 1e-8, (j, float(r))
    comp_err = np.abs(m["comp"] / out["fp64"]["comp"] - 1)
    assert np.all(comp_err < 1e-6), comp_err
    # modal: residuals, M-orthonormality, eigenvalues and shapes against the plain path
    KX, MX = torch.empty_like(X), torch.empty_like(X)
    fem._call("vox8_apply", _lib.OP_KM, _lib.PREC_FP64, _lib.ptr(X), _lib.ptr(KX), _lib.ptr(MX), k, n)
    res = (KX - torch.tensor(lam, device=gpu)[:, None] * MX).norm(dim=1).cpu().numpy()
    assert np.all(res < wl.modal_tol), res
    G = (X @ MX.T).cpu().numpy()
    assert np.max(np.abs(G - np.eye(k))) < 1e-8
    lam_err = np.abs(lam / out["fp64"]["lam"] - 1)
    assert np.all(lam_err < 1e-6), lam_err
    Xp = out["fp64"]["X"].cpu().numpy()
    Xm = X.cpu().numpy()
    macs = np.array([_mac(Xm[j], Xp[j]) for j in range(k)])
    assert np.all(macs >= 0.9999), macs
    # effective mass in z and its ranking
    frac = m["frac"]
    assert np.all(frac >= 0) and np.all(frac <= 1) and frac.sum() <= 1 + 1e-12
    assert np.allclose(m["g"], 1.0, atol=1e-8)
    rank = rank_by_effective_mass(frac)
    assert sorted(rank) == list(range(k)) and np.all(np.diff(frac[rank]) <= 0)
    assert np.max(np.abs(frac - out["fp64"]["frac"])) < 1e-6
    # sensitivities from the two paths' solutions
    assert rel(m["dc"], out["fp64"]["dc"]) < 1e-6
    assert rel(m["dlam"], out["fp64"]["dlam"]) < 1e-5
    f_mixed = np.sqrt(lam) / (2 * np.pi)
    print("c5 per-output errors against the plain FP64 path:")
    print("  compliance rel. error per load case", np.array2string(comp_err, precision=2))
    print("  true residual per load case <= 1e-8 of ||F||")
    print("  f (Hz)", np.array2string(f_mixed, precision=4))
    print("  eigenvalue rel. error per mode", np.array2string(lam_err, precision=2))
    print("  MAC per mode", np.array2string(macs, precision=10))
    print("  residual per mode", np.array2string(res, precision=2))
    print("  z fraction abs. error per mode", np.array2string(np.abs(frac - out["fp64"]["frac"]), precision=2))
    print("  z fractions", np.round(frac, 4), "rank", rank)
    print("  sensitivities rel. error: dc", rel(m["dc"], out["fp64"]["dc"]), "dlam1",
          rel(m["dlam"], out["fp64"]["dlam"]))
    for o in out.values():
        o["fem"].close()


def test_c5_simp_iteration_records_ranking(gpu):
    """One SIMP iteration of the c5 loop (12 tracked modes, effective-mass
    ranking on): the history carries the 12 fractions and their ranking."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    from bench import make_optimizer
    wl, geom, fem, opt, F, x0 = make_optimizer("c5", "mixed")
    opt.step()
    h = opt.history
    assert len(h["eff_mass"]) == 1 and len(h["eff_mass"][0]) == 12
    em = np.array(h["eff_mass"][0])
    assert np.all(np.diff(em[h["mode_rank"][0]]) <= 0)
    assert h["lobpcg_iters"][0] < 400
    opt.close(fem=True)
