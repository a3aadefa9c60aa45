"""GPU parity: adjoint sensitivities, cone filter and OC update vs the oracle.

Sensitivity bar 1e-5 relative (the parity target); the kernel computes the same
closed forms, so agreement is checked at 1e-12.  The FD check perturbs the
physical density of sampled elements and differences stationary functionals
(compliance 2F.u - u.K.u, Rayleigh quotient), as the adjoint formulas are the
derivative of those.
"""
import numpy as np
import pytest
import torch

from helpers import MaskedBox, random_density, rel

pytestmark = pytest.mark.gpu


def _pair(geom):
    from oracle.fem import FEM3D as OFEM, Material as OMat
    from vox8.fem import FEM3D, Material
    return OFEM(geom, OMat(E=1.7e9, nu=0.4, rho=1010.0)), FEM3D(geom, Material(E=1.7e9, nu=0.4, rho=1010.0))


def test_elem_compliance_and_eigen_sensitivity(gpu):
    geom = MaskedBox(7, 4, 3, hole_frac=0.1, seed=4)
    ofem, vfem = _pair(geom)
    rho = random_density(ofem.n_active, 3, 0.02, 1.0)   # spans the Pedersen kink at 0.1
    u = np.random.default_rng(1).standard_normal(geom.n_dofs)
    assert rel(vfem.elem_compliance(u, rho), ofem.elem_compliance(u, rho)) < 1e-12
    v = np.random.default_rng(2).standard_normal(geom.n_dofs)
    v[ofem.fixed_dofs] = 0
    tip = np.array([geom.node_index(7, 0, 0), geom.node_index(7, 4, 3)])
    a = vfem.eigenvalue_sensitivity(rho, 1234.5, v, tip_mass_kg=2e-5, tip_mass_nodes=tip)
    b = ofem.eigenvalue_sensitivity(rho, 1234.5, v, tip_mass_kg=2e-5, tip_mass_nodes=tip)
    assert rel(a, b) < 1e-12


def test_compliance_adjoint_vs_central_fd(gpu):
    geom = MaskedBox(12, 4, 4, hole_frac=0.0, seed=1)
    ofem, vfem = _pair(geom)
    rho = random_density(ofem.n_active, 9, 0.3, 0.9)
    F = np.zeros(geom.n_dofs)
    iy, iz = np.meshgrid(np.arange(5), np.arange(5), indexing="ij")
    t = geom.node_index(12, iy, iz).ravel()
    F[3 * t + 2] = -10.0 / len(t)
    U, _ = vfem.solve_block(rho, F[None], rtol=1e-13, maxiter=20000)
    s = vfem.sensitivities(rho, U=U, weights=[1.0])
    dc = s["dc"].cpu().numpy()
    assert abs(s["comp"][0] / (F @ U[0].cpu().numpy()) - 1) < 1e-10
    h = 1e-4
    for e in np.random.default_rng(0).choice(ofem.n_active, 6, replace=False):
        cs = []
        for sgn in (1, -1):
            r = rho.copy()
            r[e] += sgn * h
            Ut, _ = vfem.solve_block(r, F[None], x0=U, rtol=1e-13, maxiter=20000)
            u = Ut[0].cpu().numpy()
            Ku = vfem.assemble_K(r) @ u
            cs.append(2 * F @ u - u @ Ku)
        fd = (cs[0] - cs[1]) / (2 * h)
        assert abs(fd - dc[e]) <= 1e-5 * abs(dc[e]) + 1e-12 * np.abs(dc).max()


def test_filter_matches_oracle(gpu):
    from oracle.topopt import cone_kernel, filter_field as ofilter
    from vox8.topopt import filter_field
    rng = np.random.default_rng(0)
    for shape, r in (((9, 5, 4), 1.5), ((6, 7, 5), 2.3), ((5, 4, 3), 1.0)):
        f = rng.random(shape)
        m = rng.random(shape) > 0.2
        k = cone_kernel(r)
        assert rel(filter_field(f, m, k), ofilter(f, m, k)) < 1e-14


def test_oc_update_matches_oracle(gpu):
    from oracle.topopt import oc_update
    from vox8.fem import FEM3D, Material
    from vox8.workloads import Cantilever
    geom = Cantilever(20, 6, 5, 1e-3)
    vfem = FEM3D(geom, Material())
    n = geom.n_elements
    rng = np.random.default_rng(3)
    x = rng.uniform(0.05, 1.0, n)
    dc = -rng.uniform(1e-3, 1.0, n)
    design = rng.random(n) > 0.1
    Ve = 1e-9
    target = 0.3 * design.sum() * Ve
    ref = oc_update(x[design], dc[design], Ve, target, 0.2)
    xd = torch.tensor(x, device=gpu)
    dcd = torch.tensor(dc, device=gpu)
    md = torch.tensor(design.astype(np.uint8), device=gpu)
    out = torch.empty_like(xd)
    from vox8._lib import ptr
    vfem._call("vox8_oc_update", ptr(md), ptr(xd), ptr(dcd), Ve, target, 0.2, ptr(out), None)
    o = out.cpu().numpy()
    assert rel(o[design], ref) < 1e-12
    np.testing.assert_array_equal(o[~design], x[~design])
