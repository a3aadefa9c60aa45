"""Pin the oracle's element matrices (oracle/element.py) and the product's
(vox8/element.py) against exact rational integration and invariants.

The two implementations are independent (2x2x2 Gauss loop vs. exact monomial
integrals); both are checked against a sympy exact integral of B^T D B and
N N^T over the box, and against rigid-body / mass invariants.
"""
import numpy as np
import pytest
import sympy as sp

from oracle.element import element_K, element_M, iso_D, ortho_D
from vox8.element import unit_element_K, unit_element_M


def _sympy_exact(dx, dy, dz, D):
    x, y, z = sp.symbols("x y z", real=True)
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    N = [((x / dx) if a else (1 - x / dx)) * ((y / dy) if b else (1 - y / dy)) *
         ((z / dz) if c else (1 - z / dz)) for a, b, c in corners]
    B = sp.zeros(6, 24)
    for n in range(8):
        gx, gy, gz = sp.diff(N[n], x), sp.diff(N[n], y), sp.diff(N[n], z)
        c = 3 * n
        B[0, c], B[1, c + 1], B[2, c + 2] = gx, gy, gz
        B[3, c], B[3, c + 1] = gy, gx
        B[4, c + 1], B[4, c + 2] = gz, gy
        B[5, c], B[5, c + 2] = gz, gx
    integrand = B.T * sp.Matrix(D) * B
    K = integrand.applyfunc(lambda e: sp.integrate(sp.expand(e), (x, 0, dx), (y, 0, dy), (z, 0, dz)))
    NN = sp.Matrix(8, 8, lambda i, j: sp.integrate(sp.expand(N[i] * N[j]), (x, 0, dx), (y, 0, dy), (z, 0, dz)))
    return np.array(K, dtype=float), np.array(NN, dtype=float)


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 1, 1)])
def test_exact_integration(dims):
    dx, dy, dz = (sp.Rational(d) for d in dims)
    E, nu = 1.0, sp.Rational(2, 5)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = [[lam + 2 * mu, lam, lam, 0, 0, 0], [lam, lam + 2 * mu, lam, 0, 0, 0],
         [lam, lam, lam + 2 * mu, 0, 0, 0], [0, 0, 0, mu, 0, 0], [0, 0, 0, 0, mu, 0],
         [0, 0, 0, 0, 0, mu]]
    Kx, NNx = _sympy_exact(dx, dy, dz, D)
    fd = [float(v) for v in dims]
    for K in (element_K(1.0, 0.4, *fd), unit_element_K(1.0, 0.4, *fd)):
        assert np.abs(K - Kx).max() < 1e-14 * np.abs(Kx).max()
    Mx = np.kron(NNx, np.eye(3)) * 7.0
    for M in (element_M(7.0, *fd), unit_element_M(7.0, *fd)):
        assert np.abs(M - Mx).max() < 1e-14 * np.abs(Mx).max()


def test_invariants():
    h = 1e-3
    K = element_K(1.7e9, 0.4, h, h, h)
    M = element_M(1010.0, h, h, h)
    assert np.allclose(K, K.T, rtol=0, atol=1e-9 * np.abs(K).max())
    w = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(w) < 1e-9 * w.max()) == 6          # rigid-body null space
    assert w.min() > -1e-9 * w.max()
    corners = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]) * h
    rig = []
    for c in range(3):
        t = np.zeros(24); t[c::3] = 1; rig.append(t)
    for a, b in ((0, 1), (1, 2), (0, 2)):           # small rotations
        r = np.zeros(24)
        r[a::3] = -corners[:, b]; r[b::3] = corners[:, a]
        rig.append(r)
    for r in rig:
        assert np.abs(K @ r).max() < 1e-8 * np.abs(K).max() * h
    np.testing.assert_allclose(M.sum(axis=1), 1010.0 * h ** 3 / 8 * np.ones(24), rtol=1e-14)
    # cube: K scales linearly with E and h
    np.testing.assert_allclose(element_K(3.4e9, 0.4, 2 * h, 2 * h, 2 * h), 4 * K, rtol=1e-13,
                               atol=1e-6)
    # anisotropic D gives a symmetric PSD element with the same null space
    Do = ortho_D(2.0e9, 1.2e9, 1.0e9, 0.35, 0.3, 0.33, 0.6e9, 0.4e9, 0.5e9)
    Ko = unit_element_K(0, 0, h, 2 * h, h, D=Do)
    np.testing.assert_allclose(Ko, element_K(0, 0, h, 2 * h, h, D=Do), rtol=0, atol=1e-12 * np.abs(Ko).max())
    assert np.abs(iso_D(1.0, 0.25) - iso_D(1.0, 0.25).T).max() == 0
