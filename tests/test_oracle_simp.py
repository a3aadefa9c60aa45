"""Oracle SIMP loop (oracle/topopt.py) against the reference's own run.

tests/golden/simp_tipbox.npz was written by tests/golden/make_simp_golden.py,
which runs src/topopt.py's SIMPOptimizer (with converged inner solves) on the
tip-loaded box of helpers.tipbox_problem: 8 iterations, two weighted load
cases, tip mass, an active frequency constraint (augmented-Lagrangian term,
MAC mode tracking, mu ramp at iteration 5) and the OC bisection.
"""
import numpy as np

from conftest import GOLDEN
from helpers import (TIPBOX_MODAL_MAXITER, TIPBOX_MODAL_TOL, TIPBOX_SOLVE_RTOL, rel,
                     tipbox_problem)

HIST = ("compliance", "mass", "change", "vol_frac", "omega_1_Hz", "mu_freq", "freq_violation")


def run_oracle():
    from oracle.fem import FEM3D, Material
    from oracle.topopt import SIMPOptimizer
    geom, loads, opts, mat = tipbox_problem()
    fem = FEM3D(geom, Material(**mat))
    opt = SIMPOptimizer(geom, fem, loads, solve_rtol=TIPBOX_SOLVE_RTOL, modal_tol=TIPBOX_MODAL_TOL,
                        modal_maxiter=TIPBOX_MODAL_MAXITER, **opts)
    return fem, opt.run()


def test_oracle_simp_matches_reference_run():
    gold = np.load(GOLDEN / "simp_tipbox.npz")
    fem, res = run_oracle()
    assert np.array_equal(fem.elem_grid_idx, gold["elem_grid_idx"])
    assert np.array_equal(fem.free_dofs, gold["free_dofs"])
    h = res["history"]
    for k in HIST:
        a, b = np.asarray(h[k], dtype=float), gold[f"hist_{k}"]
        assert a.shape == b.shape, k
        assert rel(a, b) < 1e-7, (k, rel(a, b))
    assert rel(res["rho_active"], gold["rho_active"]) < 1e-7
    assert rel(np.stack(res["u_per_lc"]), gold["u_per_lc"]) < 1e-7
