"""Hex8 box-element matrices by exact integration in the monomial basis.

The trilinear field on a box element is written in the monomial basis
m = [1, xi, eta, zeta, xi*eta, eta*zeta, xi*zeta, xi*eta*zeta]; each shape
function is N_i = (1/8) sum_k H[i,k] m_k with H[i,k] the monomial evaluated at
corner i.  Stiffness and consistent mass then follow from exact integrals of
monomial products over [-1,1]^3, so no quadrature loop is needed.  For box
elements this equals the reference's 2x2x2 Gauss rule exactly, which
integrates these polynomials without error:

* ``constitutive_D``            -- src/fem.py:62-72
* ``unit_element_K``            -- src/fem.py:99-132
* ``unit_element_M``            -- src/fem.py:135-152
* ``unit_element_B_at_centre``  -- src/fem.py:155-176

Local corner order and DOF order [n0x, n0y, n0z, n1x, ...] follow the
reference (src/fem.py:37-46, src/mesh3d.py:140-146).
"""
from __future__ import annotations

import numpy as np

# corner signs in reference local order
CORNER_SIGNS = np.array([(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                         (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)], dtype=float)
# monomial exponents (a, b, c) for xi^a eta^b zeta^c
MONO = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
                 (1, 1, 0), (0, 1, 1), (1, 0, 1), (1, 1, 1)], dtype=int)
# Voigt index of the symmetric pair (p, r): xx yy zz xy yz xz
VOIGT = np.array([[0, 3, 5], [3, 1, 4], [5, 4, 2]])


def _H():
    """H[i, k] = monomial k evaluated at corner i."""
    s = CORNER_SIGNS
    return np.prod(np.where(MONO[None, :, :] == 1, s[:, None, :], 1.0), axis=2)


def _mono_integral(e):
    """Integral of xi^a eta^b zeta^c over [-1,1]^3 for exponent triple(s) e."""
    e = np.asarray(e)
    per = np.where(e % 2 == 0, 2.0 / (e + 1.0), 0.0)
    return np.prod(per, axis=-1)


def _grad_terms():
    """d m_k / d xi_r as (coefficient, exponents) arrays of shape (8, 3)."""
    coef = np.zeros((8, 3))
    expo = np.zeros((8, 3, 3), dtype=int)
    for k, e in enumerate(MONO):
        for r in range(3):
            if e[r] == 1:
                coef[k, r] = 1.0
                ee = e.copy()
                ee[r] = 0
                expo[k, r] = ee
    return coef, expo


def constitutive_D(E: float, nu: float) -> np.ndarray:
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] = lam + 2.0 * mu
    D[np.arange(3, 6), np.arange(3, 6)] = mu
    return D


def gradient_gram(dx: float, dy: float, dz: float) -> np.ndarray:
    """G[i, r, j, s] = integral over the element of dN_i/dx_r * dN_j/dx_s."""
    H = _H() / 8.0
    coef, expo = _grad_terms()
    h = np.array([dx, dy, dz]) / 2.0
    detJ = h.prod()
    # monomial-space integrals P[k, r, l, s] = int d_r m_k d_s m_l dxi
    P = np.zeros((8, 3, 8, 3))
    for k in range(8):
        for r in range(3):
            if coef[k, r] == 0.0:
                continue
            for l in range(8):
                for s in range(3):
                    if coef[l, s] == 0.0:
                        continue
                    P[k, r, l, s] = _mono_integral(expo[k, r] + expo[l, s])
    P /= (h[None, :, None, None] * h[None, None, None, :])
    return detJ * np.einsum("ik,krls,jl->irjs", H, P, H)


def unit_element_K(E: float, nu: float, dx: float, dy: float, dz: float,
                   D: np.ndarray = None) -> np.ndarray:
    """24x24 stiffness K[3i+p, 3j+q] = sum_{r,s} C[p,r,q,s] int dN_i/dx_r dN_j/dx_s."""
    if D is None:
        D = constitutive_D(E, nu)
    G = gradient_gram(dx, dy, dz)
    C4 = D[VOIGT[:, :, None, None], VOIGT[None, None, :, :]]      # C[p,r,q,s]
    K = np.einsum("prqs,irjs->ipjq", C4, G)
    return K.reshape(24, 24)


def unit_element_M(rho_material: float, dx: float, dy: float, dz: float) -> np.ndarray:
    """24x24 consistent mass rho int N_i N_j (identity over the 3 components)."""
    H = _H() / 8.0
    Q = _mono_integral(MONO[:, None, :] + MONO[None, :, :])       # int m_k m_l
    NN = (dx * dy * dz / 8.0) * (H @ Q @ H.T)
    return np.kron(NN, np.eye(3)) * rho_material


def unit_element_B_at_centre(dx: float, dy: float, dz: float) -> np.ndarray:
    """6x24 strain-displacement matrix at the element centre."""
    H = _H() / 8.0
    h = np.array([dx, dy, dz]) / 2.0
    g = H[:, 1:4] / h[None, :]            # only the linear monomials survive at 0
    B = np.zeros((6, 24))
    # row Voigt(p, r), column 3i+p holds dN_i/dx_r: normal strains once, the
    # engineering shears from both (p, r) and (r, p)
    for i in range(8):
        for p in range(3):
            for r in range(3):
                B[VOIGT[p, r], 3 * i + p] = g[i, r]
    return B
