"""GPU multigrid preconditioner (vox8_precond_apply) against the NumPy
restatement of the same V-cycle (tests/mg_reference.py), plus the properties
PCG relies on: symmetry and positive definiteness.

The reference uses the device's lambda_max estimates (vox8_mg_info) so both
apply the identical polynomial smoother; agreement is then limited by the
FP32 arithmetic of the device V-cycle (bar 2e-4 relative).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from helpers import MaskedBox, random_density, rel

pytestmark = pytest.mark.gpu


def _setup(shape, seed, hole):
    """A rediscretised hierarchy (the restatement's default form)."""
    from vox8 import _lib
    from vox8.fem import FEM3D, Material
    geom = MaskedBox(*shape, hole_frac=hole, seed=seed)
    fem = FEM3D(geom, Material(E=1.7e9, nu=0.4, rho=1010.0))
    fem.set_mg_galerkin(False)
    rho = random_density(fem.n_active, seed + 1, 0.2, 1.0)
    fem._set_density(rho)
    return geom, fem, rho, _lib


@pytest.mark.parametrize("shape,hole,gamma,wmin", [((24, 6, 5), 0.0, 2, 0), ((33, 9, 7), 0.1, 2, 0),
                                                   ((33, 9, 7), 0.1, 1, 0), ((70, 18, 14), 0.05, 2, 0),
                                                   ((70, 18, 14), 0.05, 2, 10000)])
@pytest.mark.parametrize("nuc", [None, (0, 1), (3, 1)])
def test_vcycle_matches_reference(gpu, shape, hole, gamma, wmin, nuc):
    """nuc: (steps, from level) set through vox8_set_mg_coarse_smoothing;
    None keeps the library default (4 steps from level 2)."""
    from mg_reference import VCycle
    geom, fem, rho, _lib = _setup(shape, 3, hole)
    fem._call("vox8_set_mg_options", 2, 20.0, 700, gamma, wmin)
    if nuc is not None:
        fem._call("vox8_set_mg_coarse_smoothing", *nuc)
    nu_c, nu_l = nuc if nuc is not None else (4, 2)
    nl = C.c_int()
    dims = (C.c_int * 48)()
    lam = (C.c_double * 16)()
    fem._call("vox8_mg_info", C.byref(nl), dims, lam, 16)
    f = np.zeros(geom.n_elements)
    f[fem.active_elems] = fem.E_MIN_FRAC + (1 - fem.E_MIN_FRAC) * rho ** 3
    V = VCycle((geom.nx, geom.ny, geom.nz), f, fem.Ke_solid, fem.node_fixed,
               active=fem.elem_active_grid, alpha=20.0, gamma=gamma, w_min_dofs=wmin,
               nu_coarse=nu_c, nu_coarse_level=nu_l)
    assert [L["dims"] for L in V.levels] == [tuple(dims[3 * i:3 * i + 3]) for i in range(nl.value)]
    for l, L in enumerate(V.levels[:-1]):
        L["lam_max"] = lam[l]
    r = np.random.default_rng(0).standard_normal((2, geom.n_dofs))
    r[:, fem.fixed_dofs] = 0
    rd = torch.tensor(r, device=gpu)
    z = torch.empty_like(rd)
    fem._call("vox8_precond_apply", _lib.ptr(rd), _lib.ptr(z), 2)
    zz = z.cpu().numpy()
    for j in range(2):
        assert rel(zz[j], V(r[j])) < 2e-4
    a, b = float(r[0] @ zz[1]), float(r[1] @ zz[0])
    assert abs(a - b) <= 1e-5 * abs(float(r[0] @ zz[0]))
    assert float(r[0] @ zz[0]) > 0 and float(r[1] @ zz[1]) > 0


@pytest.mark.parametrize("shape,hole,gamma,coarse", [((24, 6, 5), 0.0, 2, 700), ((33, 9, 7), 0.1, 2, 700),
                                                    ((33, 9, 7), 0.1, 1, 2000), ((70, 18, 14), 0.05, 2, 700),
                                                    ((45, 11, 9), 0.1, 2, 2000)])
@pytest.mark.parametrize("lane", [False, True])
def test_galerkin_cycle_matches_reference(gpu, monkeypatch, shape, hole, gamma, coarse, lane):
    """Galerkin coarse levels (vox8_set_mg_galerkin): the device cycle against
    the NumPy restatement with A_{l+1} = P^T A_l P formed by sparse matrix
    products, on densities down to 0.01 (contrast 1e6) with holes and odd
    sizes; FP32 bar 2e-5 relative, symmetric and positive.  lane=True applies
    every coarse level with the one-node-per-lane kernel (entry-major blocks),
    which the library otherwise keeps for levels of 65536 nodes and more."""
    from mg_reference import VCycle
    if lane:
        monkeypatch.setenv("VOX8_MG_LANE_NODES", "1")
    geom = MaskedBox(*shape, hole_frac=hole, seed=3)
    from vox8.fem import FEM3D, Material
    fem = FEM3D(geom, Material(E=1.7e9, nu=0.4, rho=1010.0))
    rho = random_density(fem.n_active, 4, 0.01, 1.0)
    fem.set_mg_galerkin(True)
    fem._set_density(rho)
    fem._call("vox8_set_mg_options", 2, 20.0, coarse, gamma, 0)
    from vox8 import _lib
    nl = C.c_int()
    dims = (C.c_int * 48)()
    lam = (C.c_double * 16)()
    fem._call("vox8_mg_info", C.byref(nl), dims, lam, 16)
    f = np.zeros(geom.n_elements)
    f[fem.active_elems] = fem.E_MIN_FRAC + (1 - fem.E_MIN_FRAC) * rho ** 3
    V = VCycle((geom.nx, geom.ny, geom.nz), f, fem.Ke_solid, fem.node_fixed,
               active=fem.elem_active_grid, alpha=20.0, gamma=gamma, w_min_dofs=0,
               coarse_dofs=coarse, galerkin=True, nu_coarse=4, nu_coarse_level=2)
    assert [L["dims"] for L in V.levels] == [tuple(dims[3 * i:3 * i + 3]) for i in range(nl.value)]
    for l, L in enumerate(V.levels[:-1]):
        assert abs(lam[l] / L["lam_max"] - 1) < 0.03     # Lanczos estimate vs exact (both x1.1)
        L["lam_max"] = lam[l]
    r = np.random.default_rng(0).standard_normal((2, geom.n_dofs))
    r[:, fem.fixed_dofs] = 0
    rd = torch.tensor(r, device=gpu)
    z = torch.empty_like(rd)
    fem._call("vox8_precond_apply", _lib.ptr(rd), _lib.ptr(z), 2)
    zz = z.cpu().numpy()
    for j in range(2):
        assert rel(zz[j], V(r[j])) < 2e-5
    a, b = float(r[0] @ zz[1]), float(r[1] @ zz[0])
    assert abs(a - b) <= 1e-5 * abs(float(r[0] @ zz[0]))
    assert float(r[0] @ zz[0]) > 0 and float(r[1] @ zz[1]) > 0
    fem.close()


def test_vcycle_contracts_energy(gpu):
    """One V-cycle as a stationary iteration reduces the energy-norm error."""
    geom, fem, rho, _lib = _setup((40, 8, 6), 5, 0.05)
    K = fem.assemble_K(rho)
    e = np.random.default_rng(1).standard_normal(geom.n_dofs)
    e[fem.fixed_dofs] = 0
    en = [e @ (K @ e)]
    for _ in range(3):
        Ke = torch.tensor(K @ e, device=gpu).reshape(1, -1).contiguous()
        z = torch.empty_like(Ke)
        fem._call("vox8_precond_apply", _lib.ptr(Ke), _lib.ptr(z), 1)
        e = e - z.cpu().numpy()[0]
        en.append(e @ (K @ e))
    assert all(b < 0.7 * a for a, b in zip(en, en[1:])), en


def test_cheb_step_matches_fp64_products(gpu):
    """vox8_cheb_step (FP32, 3 columns) against FP64 K products and diagonal:
    xs += d, r -= K d, d' = c1 d + c2 D^-1 r.  FP32 bar 1e-5 relative."""
    from vox8 import _lib
    geom, fem, rho, _ = _setup((70, 10, 9), 5, 0.1)
    n = geom.n_dofs
    rng = np.random.default_rng(4)
    free = torch.zeros(n, dtype=torch.float64, device=gpu)
    free[torch.as_tensor(fem.free_dofs, device=gpu)] = 1.0
    d64 = torch.tensor(rng.standard_normal((3, n)), device=gpu) * free
    r64 = torch.tensor(rng.standard_normal((3, n)), device=gpu) * free * 1e6
    x64 = torch.tensor(rng.standard_normal((3, n)), device=gpu)
    d, r, xs = d64.float().contiguous(), r64.float().contiguous(), x64.float().contiguous()
    r0, xs0 = r.double(), xs.double()                  # inputs (r and xs are updated in place)
    dn = torch.empty_like(d)
    c1, c2 = 0.7, 3e-7
    fem._call("vox8_cheb_step", _lib.ptr(d), _lib.ptr(r), _lib.ptr(xs), _lib.ptr(dn), c1, c2, 3)
    Kd = fem.assemble_K(rho) @ d.double().T          # (n, 3), zero rows at fixed DOFs
    diag = torch.as_tensor(fem.assemble_K(rho).diagonal(), device=gpu)
    dinv = torch.where(free > 0, 1.0 / torch.where(diag != 0, diag, torch.ones_like(diag)),
                       torch.zeros_like(diag))
    r_ref = r0 - Kd.T
    assert rel(xs.double().cpu().numpy(), (xs0 + d.double()).cpu().numpy()) < 1e-7
    assert rel(r.double().cpu().numpy(), r_ref.cpu().numpy()) < 1e-5
    dn_ref = c1 * d.double() + c2 * dinv * r_ref
    assert rel(dn.double().cpu().numpy(), dn_ref.cpu().numpy()) < 1e-5


@pytest.mark.parametrize("with_prev", [True, False])
def test_cheb_xstep_matches_fp64_products(gpu, with_prev):
    """vox8_cheb_xstep (FP32, 3 columns) against FP64 K products and diagonal:
    x' = x + c1 (x - x_prev) + c2 D^-1 (b - K x), zero at fixed DOFs (without
    x_prev the c1 term is dropped).  FP32 bar 1e-5 relative."""
    from vox8 import _lib
    geom, fem, rho, _ = _setup((70, 10, 9), 5, 0.1)
    n = geom.n_dofs
    rng = np.random.default_rng(8)
    free = torch.zeros(n, dtype=torch.float64, device=gpu)
    free[torch.as_tensor(fem.free_dofs, device=gpu)] = 1.0
    x = (torch.tensor(rng.standard_normal((3, n)), device=gpu) * free).float().contiguous()
    xp = (torch.tensor(rng.standard_normal((3, n)), device=gpu) * free).float().contiguous()
    b = (torch.tensor(rng.standard_normal((3, n)), device=gpu) * free * 1e6).float().contiguous()
    xn = torch.full_like(x, 7.0)
    c1, c2 = 0.7, 3e-7
    fem._call("vox8_cheb_xstep", _lib.ptr(x), _lib.ptr(xp) if with_prev else None, _lib.ptr(b),
              _lib.ptr(xn), c1, c2, 3)
    K = fem.assemble_K(rho)
    Kx = (K @ x.double().T).T                         # zero rows at fixed DOFs
    diag = torch.as_tensor(K.diagonal(), device=gpu)
    dinv = torch.where(free > 0, 1.0 / torch.where(diag != 0, diag, torch.ones_like(diag)),
                       torch.zeros_like(diag))
    xd = x.double()
    mom = c1 * (xd - xp.double()) if with_prev else 0.0
    want = (xd + mom + c2 * dinv * (b.double() - Kx)) * free
    assert rel(xn.double().cpu().numpy(), want.cpu().numpy()) < 1e-5
    assert torch.all(xn.double()[:, free == 0] == 0)


def test_graph_replays_match_eager(gpu):
    """A multigrid cycle replayed from its captured HIP graph gives the same
    result, bitwise, as launching its kernels one by one: on every replay,
    with other work and synchronisations in between, and after a density
    change (the captured cycle is invalidated and captured again)."""
    geom, fem, rho, _lib = _setup((70, 18, 14), 7, 0.05)
    rng = np.random.default_rng(2)
    r = rng.standard_normal((3, geom.n_dofs))
    r[:, fem.fixed_dofs] = 0
    rd = torch.tensor(r, device=gpu)

    def apply():
        z = torch.full_like(rd, 7.0)
        fem._call("vox8_precond_apply", _lib.ptr(rd), _lib.ptr(z), 3)
        return z

    for dens in (rho, random_density(fem.n_active, 11, 0.2, 1.0)):
        fem._set_density(dens)
        fem._call("vox8_set_mg_graphs", 0)
        ref = apply().cpu().numpy()
        fem._call("vox8_set_mg_graphs", 1)
        for rep in range(4):
            z = apply()
            # eager work between replays: a stencil product and host syncs
            x = torch.randn(1, geom.n_dofs, dtype=torch.float64, device=gpu)
            y = torch.empty_like(x)
            fem._call("vox8_apply", _lib.OP_K, _lib.PREC_FP64, _lib.ptr(x), _lib.ptr(y), None, 1,
                      geom.n_dofs)
            torch.cuda.synchronize()
            assert np.array_equal(z.cpu().numpy(), ref), rep


def _mg_lams(fem):
    nl = C.c_int()
    lam = (C.c_double * 16)()
    fem._call("vox8_mg_info", C.byref(nl), None, lam, 16)
    return [lam[l] for l in range(nl.value)]


def test_lam_max_covers_spectrum_and_is_history_free(gpu):
    """Device lambda_max(D^-1 K) per smoothed level: 1.1 x a Ritz value within
    3 % below the exact top eigenvalue (so at least lambda_max itself), and a
    function of the current density alone: the same density after another one
    gives bitwise the same estimates (no state carried between updates, none
    read from uninitialised work memory)."""
    from mg_reference import VCycle, exact_lam_max
    # NaN left in freed device memory, which the context's work vectors may
    # reuse: the estimate must not read a work vector before writing it
    junk = torch.full((64 << 20,), float("nan"), dtype=torch.float64, device=gpu)
    del junk
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    geom, fem, rho, _lib = _setup((33, 9, 7), 11, 0.1)
    lam_a = _mg_lams(fem)
    f = np.zeros(geom.n_elements)
    f[fem.active_elems] = fem.E_MIN_FRAC + (1 - fem.E_MIN_FRAC) * rho ** 3
    V = VCycle((geom.nx, geom.ny, geom.nz), f, fem.Ke_solid, fem.node_fixed,
               active=fem.elem_active_grid, coarse_dofs=2000)
    assert len(V.levels) == len(lam_a)
    for l, L in enumerate(V.levels[:-1]):
        exact = exact_lam_max(L["K"], L["dinv"])
        assert 0.97 * 1.1 * exact <= lam_a[l] <= 1.1 * exact * (1 + 1e-4), (l, lam_a[l], exact)
    rho_b = random_density(fem.n_active, 99, 0.05, 1.0)
    fem._set_density(rho_b)
    lam_b = _mg_lams(fem)
    assert lam_b[0] != lam_a[0]
    fem._set_density(rho.copy())
    assert _mg_lams(fem) == lam_a
