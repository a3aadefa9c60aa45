"""c1 workload (110 x 16 x 12 voxels at 2 mm, 73,593 DOF) against the FP64
oracle's direct answers (tests/golden/c1_oracle.npz, written by
tests/golden/make_c1_golden.py with spsolve and eigsh shift-invert).

Bars: compliance per load case within 1e-6 relative; the 6 eigenvalues
within 1e-6 relative (frequencies within 5e-7); mode shapes MAC >= 0.9999 on
the sampled DOFs; compliance / eigenvalue sensitivities within 1e-5.
Both solver paths are checked: Jacobi FP64 ("fp64", the reference's
preconditioner) and multigrid-preconditioned FP64 Krylov ("mixed").
"""
import numpy as np
import pytest

from conftest import GOLDEN, mac
from helpers import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    from vox8.fem import FEM3D, Material
    from vox8.workloads import ALL_LC, E_PA, NU, RHO_MAT, WORKLOADS, load_cases, seeded_density
    wl = WORKLOADS["c1"]
    g = wl.geometry()
    fem = FEM3D(g, Material(E=E_PA, nu=NU, rho=RHO_MAT))
    x = seeded_density(g, wl.base, wl.amp, wl.seed)
    return fem, g, x, load_cases(g, ALL_LC), np.load(GOLDEN / "c1_oracle.npz")


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_c1_static(gpu, c1, prec):
    fem, g, x, F, gold = c1
    U, info = fem.solve_block(x, F, rtol=1e-8, maxiter=20000, precision=prec)
    assert info["info"] == [0] * 5
    comp = (U.cpu().numpy() * F).sum(axis=1)
    assert np.max(np.abs(comp / gold["compliance"] - 1)) < 1e-6
    assert rel(U[0].cpu().numpy()[gold["sample_dofs"]], gold["u_lc1_sample"]) < 1e-5
    if prec == "mixed":
        assert max(info["iters"]) < 200      # vs ~1700 with Jacobi


def test_c1_modal_mg(gpu, c1):
    fem, g, x, F, gold = c1
    lam, X, info = fem.modal_block(x, 6, tip_mass_kg=0.0, tol=1e-6, maxiter=3000, precision="mixed")
    assert np.max(np.abs(lam / gold["lam"] - 1)) < 1e-6
    Xs = X.cpu().numpy()[:, gold["sample_dofs"]]
    for j in range(6):
        assert mac(Xs[j], gold["modes_sample"][:, j]) >= 0.9999
    assert info["info"] == 0, info


def test_c1_sensitivities(gpu, c1):
    fem, g, x, F, gold = c1
    U, _ = fem.solve_block(x, F[:1], rtol=1e-10, maxiter=20000, precision="mixed")
    s = fem.sensitivities(x, U=U, weights=[1.0])
    assert rel(s["ce"].cpu().numpy(), gold["ce_lc1"]) < 1e-5
    lam, X, _ = fem.modal_block(x, 6, tip_mass_kg=0.0, tol=1e-6, maxiter=3000, precision="mixed")
    d = fem.eigenvalue_sensitivity(x, lam[0], X[0].cpu().numpy(), tip_mass_kg=0.0)
    assert rel(d, gold["dlam_mode1"]) < 1e-5
