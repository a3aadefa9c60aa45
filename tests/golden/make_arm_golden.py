"""Write tests/golden/arm_modal_exact.json: converged modes of the reference's
production arm model (restated in oracle/geometry.py) at rho = 1.

Material E 1.7 GPa, nu 0.39, rho 930 kg/m^3, 60 g tip mass lumped over every
node returned by motor_mount_top_nodes() (src/fem.py:383-398, 426-432).
The oracle assembles K and M and scipy eigsh (shift-invert, sigma = 1.0, the
reference's fallback call src/fem.py:457-461, with tol 1e-10) gives the exact
6 lowest eigenpairs.  Takes ~25 minutes (SuperLU factorisation).
"""
import json
import sys
from pathlib import Path

import numpy as np
import scipy.sparse.linalg as spla

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.fem import FEM3D, Material, tip_nodes_of  # noqa: E402
from oracle.geometry import ArmGeometry3D  # noqa: E402


def main():
    g = ArmGeometry3D()
    fem = FEM3D(g, Material(E=1.7e9, nu=0.39, rho=930.0))
    rho = np.ones(fem.n_active)
    K = fem.assemble_K(rho)
    M = fem.assemble_M(rho, 0.060, tip_nodes_of(g))
    fr = fem.free_dofs
    lam = spla.eigsh(K[fr][:, fr].tocsc(), k=6, M=M[fr][:, fr].tocsc(), sigma=1.0, which="LM",
                     mode="normal", tol=1e-10, return_eigenvectors=False)
    f = np.sqrt(np.sort(lam)) / (2 * np.pi)
    out = Path(__file__).resolve().parent / "arm_modal_exact.json"
    out.write_text(json.dumps({"_what": __doc__.strip().splitlines()[0], "exact_Hz": f.tolist()},
                              indent=1))
    print(f)


if __name__ == "__main__":
    main()
