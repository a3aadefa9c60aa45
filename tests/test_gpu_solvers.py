"""GPU parity: Jacobi-PCG and LOBPCG against the oracle (SciPy cg / spsolve /
dense eigh / lobpcg).

Bars (the parity targets): compliance and eigenfrequencies within 1e-6
relative, mode shapes MAC >= 0.9999.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from conftest import mac
from helpers import MaskedBox, random_density, rel

pytestmark = pytest.mark.gpu


def _pair(geom, mat=None):
    from oracle.fem import FEM3D as OFEM, Material as OMat
    from vox8.fem import FEM3D, Material
    m = mat or dict(E=1.7e9, nu=0.4, rho=1010.0)
    return OFEM(geom, OMat(**m)), FEM3D(geom, Material(**m))


def _tip_load(geom, fz=-10.0):
    F = np.zeros(geom.n_dofs)
    iy, iz = np.meshgrid(np.arange(geom.ny + 1), np.arange(geom.nz + 1), indexing="ij")
    t = geom.node_index(geom.nx, iy, iz).ravel()
    F[3 * t + 2] = fz / len(t)
    return F


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_pcg_matches_direct(gpu, prec):
    geom = MaskedBox(10, 4, 4, hole_frac=0.1, seed=11)
    ofem, vfem = _pair(geom)
    rho = random_density(ofem.n_active, 2, 0.2, 1.0)
    F = _tip_load(geom)
    K = ofem.assemble_K(rho)
    u_ref = ofem.apply_BCs_and_solve(K, F, method="lu")
    U, info = vfem.solve_block(rho, F[None, :], rtol=1e-10, precision=prec)
    u = U.cpu().numpy()[0]
    assert info["info"] == [0]
    assert rel(u, u_ref) < 1e-8
    c_ref, c = F @ u_ref, F @ u
    assert abs(c / c_ref - 1) < 1e-9


def test_pcg_scipy_iteration_count(gpu):
    """Same iteration count as scipy cg at the same rtol (same algorithm, FP64).
    The inner products are summed in a different order than scipy's, which on
    this ~160-iteration ill-conditioned case moves the count by a couple of
    iterations; a different algorithm (or a missing preconditioner) would be
    far off."""
    geom = MaskedBox(8, 3, 3, hole_frac=0.0, seed=1)
    ofem, vfem = _pair(geom)
    rho = random_density(ofem.n_active, 4, 0.3, 1.0)
    F = _tip_load(geom)
    K = ofem.assemble_K(rho)
    free = ofem.free_dofs
    Kff = K[free][:, free]
    d = Kff.diagonal()
    its = []
    spla.cg(Kff, F[free], M=spla.aslinearoperator(np.diag(1 / d)), rtol=1e-8, maxiter=5000,
            callback=lambda xk: its.append(1))
    U, info = vfem.solve_block(rho, F[None, :], rtol=1e-8)
    assert abs(info["iters"][0] - len(its)) <= max(2, len(its) // 50), (info["iters"], len(its))


def test_pcg_multi_rhs_and_warm_start(gpu):
    geom = MaskedBox(9, 4, 3, hole_frac=0.05, seed=3)
    ofem, vfem = _pair(geom)
    rho = random_density(ofem.n_active, 5, 0.2, 1.0)
    rng = np.random.default_rng(0)
    F = rng.standard_normal((3, geom.n_dofs))
    K = ofem.assemble_K(rho)
    U, info = vfem.solve_block(rho, F, rtol=1e-10)
    for j in range(3):
        ref = ofem.apply_BCs_and_solve(K, F[j], method="lu")
        assert rel(U[j].cpu().numpy(), ref) < 1e-8
    # warm start from the answer: converges at once
    U2, info2 = vfem.solve_block(rho, F, x0=U, rtol=1e-8)
    assert max(info2["iters"]) <= 2
    # zero right-hand side returns zeros immediately (scipy: bnrm2 == 0)
    Z, infoz = vfem.solve_block(rho, np.zeros((1, geom.n_dofs)), x0=U[:1], rtol=1e-8)
    zmax = float(Z.abs().max())
    assert infoz["info"] == [0] and zmax == 0.0, zmax


def test_pcg_maxiter_reports_info(gpu):
    geom = MaskedBox(10, 4, 4, seed=8)
    ofem, vfem = _pair(geom)
    rho = random_density(ofem.n_active, 2)
    U, info = vfem.solve_block(rho, _tip_load(geom)[None, :], rtol=1e-12, maxiter=3)
    assert info["info"] == [3] and info["iters"] == [3]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_lobpcg_maxiter_reports_info(gpu, prec):
    """LOBPCG stopped by maxiter: info = iterations done (scipy's convention,
    src/fem.py:450-454 then falls back to eigsh; vox8 warns and returns the
    last Ritz pairs, which are still M-orthonormal and above the true
    eigenvalues)."""
    import warnings
    import scipy.linalg as sla
    geom = MaskedBox(10, 4, 4, seed=8)
    ofem, vfem = _pair(geom)
    vfem.precision = prec
    rho = random_density(ofem.n_active, 2)
    lam, X, info = vfem.modal_block(rho, 3, 0.0, [], tol=1e-12, maxiter=2)
    assert info["info"] == 2 and info["iters"] == 2 and info["max_res"] > 1e-12
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        vfem.modal_analysis(rho, n_modes=3, tip_mass_kg=0.0, tol=1e-12, maxiter=2)
    assert any(issubclass(x.category, RuntimeWarning) for x in w)
    free = ofem.free_dofs
    exact = sla.eigh(ofem.assemble_K(rho)[free][:, free].toarray(),
                     ofem.assemble_M(rho)[free][:, free].toarray(), eigvals_only=True,
                     subset_by_index=[0, 2])
    assert np.all(np.sort(lam) >= exact * (1 - 1e-12))


def test_non_finite_density_is_rejected(gpu):
    """NaN or inf densities from the host are an error before any GPU work
    (the reference's LOBPCG would fail and fall back to eigsh, which fails
    as well)."""
    geom = MaskedBox(6, 3, 3, seed=8)
    ofem, vfem = _pair(geom)
    rho = random_density(ofem.n_active, 2)
    rho[3] = np.nan
    with pytest.raises(ValueError):
        vfem.modal_analysis(rho, n_modes=2, tip_mass_kg=0.0)
    rho[3] = np.inf
    with pytest.raises(ValueError):
        vfem.solve_block(rho, _tip_load(geom)[None, :])


@pytest.mark.parametrize("k", [1, 4, 6])
def test_lobpcg_matches_dense_eigh(gpu, k):
    geom = MaskedBox(8, 3, 3, hole_frac=0.1, seed=21)
    ofem, vfem = _pair(geom)
    rho = random_density(ofem.n_active, 6, 0.3, 1.0)
    K = ofem.assemble_K(rho)
    M = ofem.assemble_M(rho)
    free = ofem.free_dofs
    Kff = K[free][:, free].toarray()
    Mff = M[free][:, free].toarray()
    lam_ref, V = sla.eigh(Kff, Mff, subset_by_index=[0, k - 1])
    # absolute residual tol (scipy semantics): 1e-9 of ||K|| ~ E h, well above the
    # FP64 rounding floor of K x for this size
    tol = 1e-9 * 1.7e9 * geom.dx
    f, modes, lam = vfem.modal_analysis(rho, n_modes=k, tip_mass_kg=0.0, tol=tol, maxiter=3000)
    assert vfem.last_info["lobpcg"]["info"] == 0
    assert np.max(np.abs(lam / lam_ref - 1)) < 1e-6
    for j in range(k):
        assert mac(modes[free, j], V[:, j]) > 0.9999
    # M-orthonormal
    G = modes[free].T @ Mff @ modes[free]
    assert np.abs(G - np.eye(k)).max() < 1e-8


@pytest.mark.parametrize("guard,prec", [(1, "fp64"), (2, "fp64"), (2, "mixed")])
def test_lobpcg_guard_columns_same_modes(gpu, guard, prec):
    """Guard columns (vox8_set_lobpcg_guard) are iterated but not returned:
    the k wanted pairs still match dense eigh, cold and warm started, and -1
    restores the default.  vox8_lobpcg_ritz reports the whole block: k + guard
    ascending values, the first k the returned eigenvalues, the guard ones
    upper bounds of the next eigenvalues (Rayleigh-Ritz)."""
    from vox8.fem import FEM3D, Material
    geom = MaskedBox(8, 3, 3, hole_frac=0.1, seed=21)
    ofem, _ = _pair(geom)
    vfem = FEM3D(geom, Material(E=1.7e9, nu=0.4, rho=1010.0), precision=prec)
    rho = random_density(ofem.n_active, 6, 0.3, 1.0)
    free = ofem.free_dofs
    Kff = ofem.assemble_K(rho)[free][:, free].toarray()
    Mff = ofem.assemble_M(rho)[free][:, free].toarray()
    k = 4
    lam_all = sla.eigh(Kff, Mff, eigvals_only=True, subset_by_index=[0, k + guard - 1])
    lam_ref, V = sla.eigh(Kff, Mff, subset_by_index=[0, k - 1])
    tol = 1e-9 * 1.7e9 * geom.dx
    assert vfem.lobpcg_ritz().size == 0
    vfem.set_lobpcg_guard(guard)
    X0 = None
    for _ in range(2):                      # cold, then warm (reusing the guard columns)
        lam, X, info = vfem.modal_block(rho, k, 0.0, np.array([], dtype=np.int64), X0=X0, tol=tol,
                                        maxiter=3000)
        assert info["info"] == 0 and X.shape[0] == k
        order = np.argsort(lam)
        assert np.max(np.abs(lam[order] / lam_ref - 1)) < 1e-6
        Xn = X.cpu().numpy()[order]
        for j in range(k):
            assert mac(Xn[j][free], V[:, j]) > 0.9999
        ritz = vfem.lobpcg_ritz()
        assert ritz.size == k + guard and np.all(np.diff(ritz) >= 0)
        assert np.allclose(ritz[:k], lam[order], rtol=1e-12, atol=0)
        assert np.all(ritz[k:] >= lam_all[k:] * (1 - 1e-9))
        X0 = X
    vfem.set_lobpcg_guard(-1)
    with pytest.raises(Exception):
        vfem.set_lobpcg_guard(3)


def test_lobpcg_tip_mass_and_warm_start(gpu):
    geom = MaskedBox(10, 3, 3, hole_frac=0.0, seed=2)
    ofem, vfem = _pair(geom)
    rho = random_density(ofem.n_active, 7, 0.3, 1.0)
    tip = np.array([geom.node_index(10, iy, iz) for iy in range(4) for iz in range(4)])
    K = ofem.assemble_K(rho)
    M = ofem.assemble_M(rho, tip_mass_kg=1e-6, tip_mass_nodes=tip)
    free = ofem.free_dofs
    lam_ref, V = sla.eigh(K[free][:, free].toarray(), M[free][:, free].toarray(),
                          subset_by_index=[0, 3])
    tol = 1e-8 * 1.7e9 * geom.dx
    f, modes, lam = vfem.modal_analysis(rho, 4, 1e-6, tip, tol=tol, maxiter=6000)
    assert vfem.last_info["lobpcg"]["info"] == 0
    assert np.max(np.abs(lam / lam_ref - 1)) < 1e-6
    first = vfem.last_info["lobpcg"]["iters"]
    f2, modes2, lam2 = vfem.modal_analysis(rho, 4, 1e-6, tip, X0_free=modes[free], tol=tol)
    assert vfem.last_info["lobpcg"]["iters"] <= max(3, first // 10)
    assert np.max(np.abs(lam2 / lam_ref - 1)) < 1e-6
