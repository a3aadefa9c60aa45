"""Write tests/golden/simp_tipbox.npz with the reference's own SIMP loop.

The SIMP loop (OC bisection, augmented-Lagrangian frequency term, MAC mode
tracking, mu ramp, change test) is pinned otherwise only by invariants, so
its golden vectors come from the reference itself: this script imports
src/fem.py and src/topopt.py read-only from /root/reference, runs
SIMPOptimizer.run (src/topopt.py:217-427) at its own defaults on the small
tip-loaded box of tests/helpers.tipbox_problem, and saves the history, the
final physical densities and displacements.  The loop does not forward solver
tolerances, so FEM3D.solve and FEM3D.modal_analysis are wrapped to pass tight
ones (TIPBOX_*): every inner solve is converged and the fixture is independent
of solver iterates.  It runs in the development container only; the committed
.npz is the fixture.
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, "/root/reference")

from helpers import (TIPBOX_MODAL_MAXITER, TIPBOX_MODAL_TOL, TIPBOX_SOLVE_RTOL,  # noqa: E402
                     tipbox_problem)
from src.fem import FEM3D, Material  # noqa: E402
from src.topopt import SIMPOptimizer  # noqa: E402


def main():
    solve0, modal0 = FEM3D.solve, FEM3D.modal_analysis

    def solve(self, rho_active, F, method="cg", x0=None):
        return self.apply_BCs_and_solve(self.assemble_K(rho_active), F, method=method, x0=x0,
                                        rtol=TIPBOX_SOLVE_RTOL)

    def modal(self, *a, **k):
        k.update(tol=TIPBOX_MODAL_TOL, maxiter=TIPBOX_MODAL_MAXITER)
        return modal0(self, *a, **k)

    FEM3D.solve, FEM3D.modal_analysis = solve, modal
    geom, loads, opts, mat = tipbox_problem()
    fem = FEM3D(geom, Material(**mat))
    res = SIMPOptimizer(geom, fem, loads, **opts).run(verbose=False)
    h = res["history"]
    out = {f"hist_{k}": np.asarray(v, dtype=float) for k, v in h.items()}
    out["rho_active"] = res["rho_active"]
    out["u_per_lc"] = np.stack(res["u_per_lc"])
    out["elem_grid_idx"] = fem.elem_grid_idx
    out["free_dofs"] = fem.free_dofs
    np.savez_compressed(HERE / "simp_tipbox.npz", **out)
    for k, v in h.items():
        print(k, np.round(np.asarray(v, dtype=float), 6))


if __name__ == "__main__":
    main()
