"""CPU: the multigrid smoother's lambda_max(D^-1 K) estimate.

The Chebyshev smoother damps [lam_max/alpha, lam_max]; an estimate below the
true top of the spectrum amplifies the modes above it and can make the
preconditioner indefinite.  The device runs a fixed-length Lanczos in the D
inner product from a fixed start vector (tests/mg_reference.py mirrors it);
here its Ritz value is checked against the exact eigenvalue of the symmetric
form D^-1/2 K D^-1/2 on voxel operators with SIMP densities, void/solid
contrast and holes: within 3 % below, never above (Ritz values lie inside
the spectrum), so the 10 % margin covers the spectrum.
"""
import numpy as np
import pytest

from mg_reference import assemble, exact_lam_max, lanczos_lam_max


def _operator(dims, kind, hole, seed):
    from vox8.element import constitutive_D, unit_element_K
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    ne = nx * ny * nz
    if kind == "solid":
        rho = np.ones(ne)
    elif kind == "simp":
        rho = np.clip(rng.random(ne) ** 2, 1e-3, 1.0)
    else:
        rho = np.where(rng.random(ne) < 0.5, 1.0, 1e-3)
    f = 1e-9 + (1 - 1e-9) * rho ** 3
    f[rng.random(ne) < hole] = 0.0
    Ke = unit_element_K(1.0, 0.3, 1.0, 1.0, 1.0, D=constitutive_D(1.0, 0.3))
    fixed = np.zeros((nz + 1, ny + 1, nx + 1), np.uint8)
    fixed[:, :, 0] = 7
    K, free = assemble(nx, ny, nz, f, Ke, fixed.ravel())
    d = K.diagonal()
    dinv = np.where(free, 1.0 / np.where(d != 0, d, 1.0), 0.0)
    return K, dinv


@pytest.mark.parametrize("dims,kind,hole", [((16, 6, 5), "solid", 0.0), ((24, 8, 6), "simp", 0.0),
                                            ((24, 8, 6), "bimat", 0.1), ((30, 8, 8), "simp", 0.05)])
def test_lanczos_estimate_against_exact(dims, kind, hole):
    K, dinv = _operator(dims, kind, hole, seed=sum(dims))
    lam = exact_lam_max(K, dinv)
    theta = lanczos_lam_max(K, dinv)
    assert 0.97 * lam <= theta <= lam * (1 + 1e-6), (theta, lam)


def test_lanczos_exhausts_small_spaces():
    """Fewer free DOFs than steps: the recurrence stops at the invariant
    subspace and the Ritz value is the exact top eigenvalue."""
    K, dinv = _operator((1, 1, 1), "simp", 0.0, seed=3)
    assert int(np.count_nonzero(dinv)) < 20
    assert abs(lanczos_lam_max(K, dinv) / exact_lam_max(K, dinv) - 1) < 1e-5
