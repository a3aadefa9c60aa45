"""Oracle: hex8 element matrices (test infrastructure only, see oracle/__init__.py).

Restates the reference's element kernels:

* ``LOCAL_NODES`` / 2x2x2 Gauss rule      -- src/fem.py:29-46
* ``shape_funcs_and_grads``               -- src/fem.py:49-59
* ``constitutive_D`` (Voigt xx,yy,zz,xy,yz,xz) -- src/fem.py:62-72
* ``constitutive_D_orthotropic``          -- src/fem.py:75-96
* ``unit_element_K``  (sum_gp B^T D B detJ)  -- src/fem.py:99-132
* ``unit_element_M``  (consistent mass)      -- src/fem.py:135-152
* ``unit_element_B_at_centre``               -- src/fem.py:155-176
"""
from __future__ import annotations

import numpy as np

# Local corner signs, bottom face (zeta=-1) counter-clockwise then top face
# (src/fem.py:37-46; the same order as mesh3d.element_nodes, src/mesh3d.py:140-146).
CORNERS = np.array([(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                    (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)], dtype=float)

_G = 1.0 / np.sqrt(3.0)
# 2x2x2 Gauss points, unit weights (src/fem.py:29-33)
GAUSS_POINTS = np.array([(a, b, c) for a in (-_G, _G) for b in (-_G, _G) for c in (-_G, _G)])


def shape_and_grad(p):
    """Trilinear N (8,) and dN/dxi (8,3) at local point p (src/fem.py:49-59)."""
    xi = np.asarray(p, dtype=float)
    f = 1.0 + CORNERS * xi[None, :]                    # (8,3) factors per axis
    N = 0.125 * f.prod(axis=1)
    dN = np.empty((8, 3))
    dN[:, 0] = 0.125 * CORNERS[:, 0] * f[:, 1] * f[:, 2]
    dN[:, 1] = 0.125 * CORNERS[:, 1] * f[:, 0] * f[:, 2]
    dN[:, 2] = 0.125 * CORNERS[:, 2] * f[:, 0] * f[:, 1]
    return N, dN


def iso_D(E: float, nu: float) -> np.ndarray:
    """Isotropic Voigt stiffness (src/fem.py:62-72)."""
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[[0, 1, 2], [0, 1, 2]] += 2.0 * mu
    D[[3, 4, 5], [3, 4, 5]] = mu
    return D


def ortho_D(Ex, Ey, Ez, nu_xy, nu_yz, nu_xz, Gxy, Gyz, Gxz) -> np.ndarray:
    """Orthotropic Voigt stiffness as the inverse compliance (src/fem.py:75-96)."""
    S = np.diag([1 / Ex, 1 / Ey, 1 / Ez, 1 / Gxy, 1 / Gyz, 1 / Gxz])
    S[0, 1] = S[1, 0] = -nu_xy / Ex
    S[1, 2] = S[2, 1] = -nu_yz / Ey
    S[0, 2] = S[2, 0] = -nu_xz / Ex
    return np.linalg.inv(S)


def strain_matrix(dN_dx: np.ndarray) -> np.ndarray:
    """6x24 B for physical gradients dN_dx (8,3); rows as src/fem.py:118-130."""
    B = np.zeros((6, 24))
    for n in range(8):
        gx, gy, gz = dN_dx[n]
        c = 3 * n
        B[0, c] = gx
        B[1, c + 1] = gy
        B[2, c + 2] = gz
        B[3, c], B[3, c + 1] = gy, gx
        B[4, c + 1], B[4, c + 2] = gz, gy
        B[5, c], B[5, c + 2] = gz, gx
    return B


def element_K(E: float, nu: float, dx: float, dy: float, dz: float, D=None) -> np.ndarray:
    """24x24 box-element stiffness, 2x2x2 Gauss (src/fem.py:99-132)."""
    if D is None:
        D = iso_D(E, nu)
    h = np.array([dx, dy, dz]) / 2.0
    detJ = h.prod()
    K = np.zeros((24, 24))
    for gp in GAUSS_POINTS:
        _, dN = shape_and_grad(gp)
        B = strain_matrix(dN / h[None, :])
        K += B.T @ D @ B * detJ
    return K


def element_M(rho: float, dx: float, dy: float, dz: float) -> np.ndarray:
    """24x24 consistent mass, 2x2x2 Gauss (src/fem.py:135-152)."""
    detJ = dx * dy * dz / 8.0
    M = np.zeros((24, 24))
    for gp in GAUSS_POINTS:
        N, _ = shape_and_grad(gp)
        M += np.kron(np.outer(N, N), np.eye(3)) * (rho * detJ)
    return M


def element_B_centre(dx: float, dy: float, dz: float) -> np.ndarray:
    """B at xi=eta=zeta=0 (src/fem.py:155-176)."""
    h = np.array([dx, dy, dz]) / 2.0
    _, dN = shape_and_grad((0.0, 0.0, 0.0))
    return strain_matrix(dN / h[None, :])
