"""Pins of the oracle's FEM3D, filter, OC and effective-mass restatements
against brute force, SciPy/LAPACK and closed forms (CPU only).

Each check is chosen so that a plausible slip in the restatement (a wrong
local-node order or DOF index, a dropped tip-lump term, a wrong penalisation
branch, a sign) fails it; none retypes the oracle's own formula.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from helpers import MaskedBox, TipBox, random_density, rel

MAT = dict(E=1.7e9, nu=0.4, rho=1010.0)


def _fem(geom):
    from oracle.fem import FEM3D, Material
    return FEM3D(geom, Material(**MAT))


def _brute(geom, fem, factor, Ke):
    """Dense element loop with element nodes from the geometry itself."""
    masks = geom.build_masks()["in_arm"]
    n = geom.n_dofs
    A = np.zeros((n, n))
    e = 0
    for ez in range(geom.nz):
        for ey in range(geom.ny):
            for ex in range(geom.nx):
                if not masks[ex, ey, ez]:
                    continue
                nodes = geom.element_nodes(ex, ey, ez)
                dofs = (3 * np.asarray(nodes)[:, None] + np.arange(3)).ravel()
                A[np.ix_(dofs, dofs)] += factor[e] * Ke
                e += 1
    assert e == fem.n_active
    return A


def test_assembly_against_element_loop():
    geom = MaskedBox(4, 3, 2, hole_frac=0.2, seed=2)
    fem = _fem(geom)
    rho = random_density(fem.n_active, 3, 0.02, 1.0)          # both Pedersen branches
    K = fem.assemble_K(rho).toarray()
    fK = 1e-9 + (1 - 1e-9) * rho ** 3
    assert rel(K, _brute(geom, fem, fK, fem.Ke_solid)) < 1e-14
    fM = np.where(rho >= 0.1, rho, 0.1 * (rho / 0.1) ** 6)
    tip = np.array([geom.node_index(4, 0, 0), geom.node_index(4, 3, 2)])
    M = fem.assemble_M(rho, tip_mass_kg=3e-4, tip_mass_nodes=tip).toarray()
    Mb = _brute(geom, fem, fM, fem.Me_solid)
    for nd in tip:
        for c in range(3):
            Mb[3 * nd + c, 3 * nd + c] += 3e-4 / len(tip)
    assert rel(M, Mb) < 1e-14
    # total mass of the consistent matrix: sum over one translation
    r = np.zeros(geom.n_dofs)
    r[0::3] = 1.0
    V = geom.dx * geom.dy * geom.dz
    assert abs(r @ M @ r - (MAT["rho"] * V * fM.sum() + 3e-4)) < 1e-12 * (r @ M @ r)


def test_rigid_modes_symmetry_energy():
    geom = MaskedBox(5, 3, 3, hole_frac=0.1, seed=4, clamp=False)
    fem = _fem(geom)
    rho = random_density(fem.n_active, 5)
    K = fem.assemble_K(rho)
    xyz = np.array([[i % (geom.nx + 1), (i // (geom.nx + 1)) % (geom.ny + 1),
                     i // ((geom.nx + 1) * (geom.ny + 1))] for i in range(geom.n_nodes)], float) * 1e-3
    scale = abs(K).max() * 1e-3
    for t in range(3):                                          # translations
        u = np.zeros(geom.n_dofs)
        u[t::3] = 1.0
        assert np.abs(K @ u).max() < 1e-9 * scale
    for a, b in ((0, 1), (1, 2), (0, 2)):                        # infinitesimal rotations
        u = np.zeros(geom.n_dofs)
        u[a::3], u[b::3] = -xyz[:, b], xyz[:, a]
        assert np.abs(K @ u).max() < 1e-9 * scale * 1e3
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(geom.n_dofs), rng.standard_normal(geom.n_dofs)
    assert abs(x @ (K @ y) - y @ (K @ x)) < 1e-12 * abs(x @ (K @ y))
    ce = fem.elem_compliance(x)
    fK = 1e-9 + (1 - 1e-9) * rho ** 3
    assert abs(x @ (K @ x) - fK @ ce) < 1e-12 * abs(x @ (K @ x))


def test_cg_against_direct_solve():
    geom = MaskedBox(8, 3, 3, hole_frac=0.1, seed=6)
    fem = _fem(geom)
    rho = random_density(fem.n_active, 7, 0.2, 1.0)
    F = np.zeros(geom.n_dofs)
    F[3 * geom.node_index(8, 3, 3) + 2] = -1.0
    K = fem.assemble_K(rho)
    info = []
    u = fem.apply_BCs_and_solve(K, F, rtol=1e-12, info_out=info)
    assert info == [0]
    free = fem.free_dofs
    ud = spla.spsolve(K[free][:, free].tocsc(), F[free])
    assert rel(u[free], ud) < 1e-9
    assert np.all(u[fem.fixed_dofs] == 0.0)
    res = K[free][:, free] @ u[free] - F[free]
    assert np.linalg.norm(res) <= 2e-12 * np.linalg.norm(F[free])   # recurrence vs true residual


def test_cantilever_tip_deflection_closed_form():
    """Solid cantilever 40x4x4 mm under a 1 N tip load: Timoshenko deflection
    within 10 % (full-integration hex8 is stiffer than the beam theory)."""
    from oracle.fem import FEM3D, Material
    from vox8.workloads import Cantilever, load_cases
    geom = Cantilever(40, 4, 4, 1e-3)
    fem = FEM3D(geom, Material(E=1.7e9, nu=0.4, rho=1010.0))
    F = load_cases(geom, ("LC1",))[0] / 10.0                      # -1 N in z
    u = fem.apply_BCs_and_solve(fem.assemble_K(np.ones(fem.n_active)), F, rtol=1e-12)
    tip = np.array([geom.node_index(40, iy, iz) for iy in range(5) for iz in range(5)])
    w = -u[3 * tip + 2].mean()
    E, nu, L, b = 1.7e9, 0.4, 40e-3, 4e-3
    I, A, G = b ** 4 / 12, b * b, E / (2 * (1 + nu))
    w_t = L ** 3 / (3 * E * I) + L / (5 / 6 * G * A)
    assert 0.90 < w / w_t < 1.0


def test_lobpcg_against_dense_eigh():
    geom = TipBox(6, 2, 2, hole_frac=0.0, seed=1)
    fem = _fem(geom)
    rho = random_density(fem.n_active, 2, 0.05, 1.0)           # includes m(rho) < rho
    tip = geom.motor_mount_top_nodes()
    f, modes, lam = fem.modal_analysis(rho, 4, tip_mass_kg=1e-4, tip_mass_nodes=tip, tol=1e-7,
                                       maxiter=3000)
    K = fem.assemble_K(rho)
    M = fem.assemble_M(rho, 1e-4, tip)
    free = fem.free_dofs
    ev, V = sla.eigh(K[free][:, free].toarray(), M[free][:, free].toarray())
    assert rel(lam, ev[:4]) < 1e-9
    Mf = M[free][:, free]
    G = modes[free].T @ (Mf @ modes[free])
    assert np.abs(G - np.eye(4)).max() < 1e-8
    for k in range(4):
        mac = (modes[free, k] @ (Mf @ V[:, k])) ** 2
        assert mac > 1 - 1e-8


def test_sensitivities_against_central_differences():
    """Compliance and eigenvalue derivatives by exact re-solves (dense)."""
    geom = TipBox(5, 2, 2, hole_frac=0.0, seed=1)
    fem = _fem(geom)
    rho = random_density(fem.n_active, 4, 0.04, 0.9)
    tip = geom.motor_mount_top_nodes()
    free = fem.free_dofs
    F = np.zeros(geom.n_dofs)
    F[3 * tip + 2] = -1.0

    def comp(r):
        K = fem.assemble_K(r)[free][:, free].toarray()
        return F[free] @ np.linalg.solve(K, F[free])

    def lam1(r):
        K = fem.assemble_K(r)[free][:, free].toarray()
        M = fem.assemble_M(r, 2e-4, tip)[free][:, free].toarray()
        ev, V = sla.eigh(K, M)
        return ev[0], V[:, 0]

    K = fem.assemble_K(rho)
    u = fem.apply_BCs_and_solve(K, F, rtol=1e-14)
    dc = -3 * (1 - 1e-9) * rho ** 2 * fem.elem_compliance(u)
    l0, v0 = lam1(rho)
    v = np.zeros(geom.n_dofs)
    v[free] = v0
    dl = fem.eigenvalue_sensitivity(rho, l0, v, tip_mass_kg=2e-4, tip_mass_nodes=tip)
    h = 1e-4                                                    # truncation ~1e-8, roundoff ~1e-8
    for e in range(0, fem.n_active, 3):
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h
        rm[e] -= h
        fd = (comp(rp) - comp(rm)) / (2 * h)
        assert abs(fd - dc[e]) < 1e-5 * abs(dc).max()
        fl = (lam1(rp)[0] - lam1(rm)[0]) / (2 * h)
        assert abs(fl - dl[e]) < 1e-5 * abs(dl).max()


def test_filter_against_brute_force():
    from oracle.topopt import cone_kernel, filter_field
    k = cone_kernel(1.5)
    assert np.count_nonzero(k) == 19
    assert abs(k.sum() - (1.5 + 6 * 0.5 + 12 * (1.5 - np.sqrt(2)))) < 1e-14
    rng = np.random.default_rng(3)
    f = rng.random((6, 5, 4))
    mask = rng.random((6, 5, 4)) > 0.2
    out = filter_field(f, mask, k)
    ref = np.zeros_like(f)
    for i, j, l in zip(*np.nonzero(mask)):
        num = den = 0.0
        for a in range(-1, 2):
            for b in range(-1, 2):
                for c in range(-1, 2):
                    x, y, z = i + a, j + b, l + c
                    if 0 <= x < 6 and 0 <= y < 5 and 0 <= z < 4 and mask[x, y, z]:
                        w = max(0.0, 1.5 - np.sqrt(a * a + b * b + c * c))
                        num += w * f[x, y, z]
                        den += w
        ref[i, j, l] = num / den
    assert np.abs(out - ref).max() < 1e-14
    full = filter_field(np.ones((7, 7, 7)), np.ones((7, 7, 7), bool), k)
    assert np.abs(full - 1).max() < 1e-14


def test_oc_update_invariants():
    from oracle.topopt import oc_update
    rng = np.random.default_rng(5)
    x = rng.uniform(0.01, 1.0, 500)
    dc = -rng.uniform(0.1, 10.0, 500)
    V_e, target, move = 1.0, 0.35 * 500, 0.2                    # multiplier inside [1e-9, 1e9]
    xn = oc_update(x, dc, V_e, target, move)
    assert xn.min() >= 0.001 and xn.max() <= 1.0
    assert np.all(xn <= x + move + 1e-15) and np.all(xn >= np.maximum(0.001, x - move) - 1e-15)
    assert abs(xn.sum() * V_e / target - 1) < 2e-3               # bisection tolerance 1e-4 on lambda
    # the update is monotone in |dc| at fixed lambda: larger sensitivity, larger density
    i, j = np.argmax(np.abs(dc) / x ** -2), np.argmin(np.abs(dc) / x ** -2)
    assert xn[i] / x[i] >= xn[j] / x[j] - 1e-12


@pytest.mark.parametrize("which", ["complete", "single"])
def test_effective_mass(which):
    from oracle.modal import effective_mass_voxel
    geom = TipBox(4, 2, 2, hole_frac=0.0, seed=1)
    fem = _fem(geom)
    rho = random_density(fem.n_active, 8, 0.3, 1.0)
    free = fem.free_dofs
    tip = geom.motor_mount_top_nodes()
    K = fem.assemble_K(rho)[free][:, free].toarray()
    M = fem.assemble_M(rho, 1e-4, tip)[free][:, free].toarray()
    ev, V = sla.eigh(K, M)
    modes = np.zeros((geom.n_dofs, len(free)))
    modes[free] = V
    if which == "complete":
        eff = effective_mass_voxel(fem, geom, rho, modes, 1e-4, tip_nodes=tip)
        assert abs(eff.sum() - 1) < 1e-10
    else:
        # one mode equal to the rigid translation itself carries all the mass
        r = np.zeros(geom.n_dofs)
        r[2::3] = 1.0
        r[fem.fixed_dofs] = 0.0
        eff = effective_mass_voxel(fem, geom, rho, (3.7 * r)[:, None], 1e-4, tip_nodes=tip)
        assert abs(eff[0] - 1) < 1e-13


def test_stress_affine_fields_and_invariants_closed_form():
    """Centre stresses of affine displacement fields (reproduced exactly by
    trilinear elements) equal rho D eps; von Mises and the largest principal
    stress of known tensors (uniaxial, pure shear, a rotated diagonal)."""
    geom = MaskedBox(4, 3, 2, h=2e-3, hole_frac=0.2, seed=9)
    fem = _fem(geom)
    E, nu = MAT["E"], MAT["nu"]
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    idx = np.arange(geom.n_nodes)
    X = np.stack([(idx % (geom.nx + 1)) * geom.dx, ((idx // (geom.nx + 1)) % (geom.ny + 1)) * geom.dy,
                  (idx // ((geom.nx + 1) * (geom.ny + 1))) * geom.dz], axis=1)
    rho = random_density(fem.n_active, 3, 0.2, 1.0)
    eps0, gam = 1e-4, 3e-4
    u = np.zeros((geom.n_nodes, 3))
    u[:, 0] = eps0 * X[:, 0] + gam * X[:, 1]          # eps_xx = eps0, engineering gamma_xy = gam
    s = fem.compute_stress(u.ravel(), rho)
    expect = np.array([(lam + 2 * mu) * eps0, lam * eps0, lam * eps0, mu * gam, 0.0, 0.0])
    assert rel(s, rho[:, None] * expect[None, :]) < 1e-12
    # invariants of known tensors
    sig = 7.5e6
    T = np.array([[sig, 0, 0, 0, 0, 0], [0, 0, 0, sig, 0, 0], [-sig, -sig, -sig, 0, 0, 0]])
    vm = fem.von_mises(T)
    assert vm == pytest.approx([sig, np.sqrt(3) * sig, 0.0], abs=1e-6)
    R = sla.expm(np.array([[0, -0.3, 0.2], [0.3, 0, -0.5], [-0.2, 0.5, 0]]))   # a rotation
    A = R @ np.diag([1.0, 2.0, 3.0]) @ R.T
    row = np.array([[A[0, 0], A[1, 1], A[2, 2], A[0, 1], A[1, 2], A[0, 2]]])
    assert fem.principal_stress_max(row)[0] == pytest.approx(3.0, rel=1e-13)
