"""Write tests/golden/c1_oracle.npz: FP64 oracle answers for the c1 cantilever.

c1: 110 x 16 x 12 voxels at 2 mm, E 1.7 GPa, nu 0.4, rho 1010, density
clip(0.3 + U(-0.05, 0.05), 1e-3, 1) (numpy default_rng(0)), unfiltered.
Answers are computed with the oracle's assembled CSR matrices and direct /
shift-invert SciPy solvers (exact to ~1e-12), so they pin the converged
solution every iterative solver must reach:
  compliance per load case LC1..LC5 (spsolve), lam/f of the 6 lowest modes
  (eigsh shift-invert), mode shapes sampled on every 7th free DOF, and the
  LC1 element compliance and mode-1 eigenvalue sensitivity.
Run: python tests/golden/make_c1_golden.py   (about a minute)
"""
import sys
from pathlib import Path

import numpy as np
import scipy.sparse.linalg as spla

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.fem import FEM3D, Material  # noqa: E402
from oracle.geometry import Cantilever  # noqa: E402
from oracle import workloads as ow  # noqa: E402


def main():
    g = Cantilever(110, 16, 12, 2e-3)
    fem = FEM3D(g, Material(E=1.7e9, nu=0.4, rho=1010.0))
    rho = ow.seeded_density(g.n_elements, 0.3, 0.05, 0)
    K = fem.assemble_K(rho)
    M = fem.assemble_M(rho)
    free = fem.free_dofs
    Kff = K[free][:, free].tocsc()
    Mff = M[free][:, free].tocsc()
    lu = spla.splu(Kff)
    names = ("LC1", "LC2", "LC3", "LC4", "LC5")
    comp, U = [], []
    for n in names:
        F = ow.load_vector(g, n)
        u = np.zeros(g.n_dofs)
        u[free] = lu.solve(F[free])
        U.append(u)
        comp.append(F @ u)
    lam, V = spla.eigsh(Kff, k=6, M=Mff, sigma=1.0, which="LM", tol=1e-12)
    o = np.argsort(lam)
    lam, V = lam[o], V[:, o]
    modes = np.zeros((g.n_dofs, 6))
    modes[free] = V
    sample = free[::7]
    ce = fem.elem_compliance(U[0])
    dlam = fem.eigenvalue_sensitivity(rho, lam[0], modes[:, 0], tip_mass_kg=0.0)
    out = Path(__file__).resolve().parent / "c1_oracle.npz"
    np.savez_compressed(out, compliance=np.array(comp), lam=lam,
                        freq_Hz=np.sqrt(lam) / (2 * np.pi), sample_dofs=sample,
                        modes_sample=modes[sample], ce_lc1=ce, dlam_mode1=dlam,
                        u_lc1_sample=U[0][sample])
    print("wrote", out, "f =", np.sqrt(lam) / (2 * np.pi), "c =", comp)


if __name__ == "__main__":
    main()
