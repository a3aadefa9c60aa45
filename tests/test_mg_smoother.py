"""CPU: the multigrid smoother of tests/mg_reference.py (the device's form:
Chebyshev steps in the three-term recurrence on the iterates alone).

Two independent checks of VCycle.smooth:
* the closed form: on a diagonal operator D^-1 K = diag(lambda), nu steps from
  x_0 = 0 for b = K e leave the error e - x_nu = p_nu(lambda) e with
  p_nu(lambda) = T_nu((theta - lambda) / delta) / T_nu(theta / delta), the
  scaled Chebyshev polynomial of the first kind on [lam_max/alpha, lam_max];
* the residual-direction form of the same semi-iteration (r_k, d_k kept as
  vectors: x += d, r -= K d, d = rho_k rho_{k-1} d + 2 rho_k / delta D^-1 r),
  on a voxel operator from a non-zero start, equal to rounding.
A wrong coefficient, a dropped momentum term or a swapped iterate fails both.
"""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

from mg_reference import VCycle, assemble


def _smoother(lam_max, alpha):
    v = VCycle.__new__(VCycle)
    v.alpha = alpha
    return v


@pytest.mark.parametrize("nu", [1, 2, 3, 4, 8])
@pytest.mark.parametrize("alpha", [20.0, 30.0])
def test_error_polynomial_closed_form(nu, alpha):
    lam_max = 2.3
    lam = np.linspace(1e-3, lam_max, 41)
    e = np.random.default_rng(nu).standard_normal(lam.size)
    K = np.diag(lam * 1.7)                     # D = 1.7 I, so D^-1 K = diag(lam)
    L = {"lam_max": lam_max, "dinv": np.full(lam.size, 1.0 / 1.7), "K": K}
    x = _smoother(lam_max, alpha).smooth(L, np.zeros(lam.size), K @ e, nu)
    lmin = lam_max / alpha
    theta, delta = 0.5 * (lam_max + lmin), 0.5 * (lam_max - lmin)
    Tn = npcheb.Chebyshev.basis(nu)
    p = Tn((theta - lam) / delta) / Tn(theta / delta)
    np.testing.assert_allclose(e - x, p * e, rtol=1e-10, atol=1e-12 * np.abs(e).max())
    inside = lam >= lmin                       # damped on the interval: |p| <= 1 / T_nu(sigma)
    assert np.all(np.abs(p[inside]) <= 1.0 / Tn(theta / delta) + 1e-12)


def _residual_form(K, dinv, lam_max, alpha, x, b, nu):
    lmin = lam_max / alpha
    theta, delta = 0.5 * (lam_max + lmin), 0.5 * (lam_max - lmin)
    sigma = theta / delta
    r = b - K @ x
    d = dinv * r / theta
    rho = 1.0 / sigma
    for _ in range(nu):
        rho1 = 1.0 / (2 * sigma - rho)
        x = x + d
        r = r - K @ d
        d = rho1 * rho * d + 2 * rho1 / delta * dinv * r
        rho = rho1
    return x


@pytest.mark.parametrize("nu", [1, 2, 4])
def test_equals_residual_direction_form(nu):
    from vox8.element import constitutive_D, unit_element_K
    nx, ny, nz = 10, 4, 3
    rng = np.random.default_rng(7)
    f = 1e-9 + (1 - 1e-9) * np.clip(rng.random(nx * ny * nz), 1e-3, 1.0) ** 3
    Ke = unit_element_K(1.0, 0.3, 1.0, 1.0, 1.0, D=constitutive_D(1.0, 0.3))
    fixed = np.zeros((nz + 1, ny + 1, nx + 1), np.uint8)
    fixed[:, :, 0] = 7
    K, free = assemble(nx, ny, nz, f, Ke, fixed.ravel())
    d = K.diagonal()
    dinv = np.where(free, 1.0 / np.where(d != 0, d, 1.0), 0.0)
    lam_max = 1.1 * float(np.max(np.linalg.eigvalsh((K.toarray() * np.sqrt(dinv)[:, None]) * np.sqrt(dinv)[None, :])))
    b = np.where(free, rng.standard_normal(K.shape[0]), 0.0)
    x0 = np.where(free, rng.standard_normal(K.shape[0]), 0.0)
    L = {"lam_max": lam_max, "dinv": dinv, "K": K}
    for start in (np.zeros_like(b), x0):
        got = _smoother(lam_max, 20.0).smooth(L, start, b, nu)
        want = _residual_form(K, dinv, lam_max, 20.0, start, b, nu)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())
