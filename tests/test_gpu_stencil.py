"""GPU parity: matrix-free K.u / M.u / diag against the oracle's assembled CSR.

Oracle: oracle/fem.py (restating src/fem.py:311-398).  The GPU path is
vox8_apply / vox8_diagonal through the C ABI (via vox8.fem.VoxelOperator).
Bar: FP64 relative error <= 1e-13 per product; FP32 <= 1e-5.
"""
import numpy as np
import pytest
import torch

from helpers import MaskedBox, random_density, rel

pytestmark = pytest.mark.gpu


def _pair(geom, rho_seed=1, mat=None):
    from oracle.fem import FEM3D as OFEM, Material as OMat
    from vox8.fem import FEM3D, Material
    m = mat or dict(E=1.7e9, nu=0.4, rho=1010.0)
    ofem = OFEM(geom, OMat(**m))
    vfem = FEM3D(geom, Material(**m))
    rho = random_density(ofem.n_active, rho_seed)
    return ofem, vfem, rho


@pytest.mark.parametrize("shape", [(3, 2, 2), (7, 4, 3), (12, 5, 6)])
def test_K_M_products_fp64(gpu, shape):
    geom = MaskedBox(*shape, seed=sum(shape))
    ofem, vfem, rho = _pair(geom)
    K = ofem.assemble_K(rho)
    M = ofem.assemble_M(rho)
    free = ofem.free_dofs
    np.testing.assert_array_equal(free, vfem.free_dofs)
    x = np.random.default_rng(3).standard_normal(geom.n_dofs)
    xf = np.zeros_like(x)
    xf[free] = x[free]
    yK_ref = np.zeros_like(x)
    yK_ref[free] = (K @ xf)[free]
    yM_ref = np.zeros_like(x)
    yM_ref[free] = (M @ xf)[free]
    yK = vfem.assemble_K(rho) @ x
    yM = vfem.assemble_M(rho) @ x
    assert rel(yK, yK_ref) < 1e-13
    assert rel(yM, yM_ref) < 1e-13
    dK = vfem.assemble_K(rho).diagonal()
    dref = np.zeros_like(x)
    dref[free] = K.diagonal()[free]
    assert rel(dK, dref) < 1e-14


def test_multi_rhs_and_tip_mass(gpu):
    geom = MaskedBox(6, 3, 4, seed=5)
    ofem, vfem, rho = _pair(geom)
    tip = np.array([geom.node_index(6, iy, 4) for iy in range(4)])
    M = ofem.assemble_M(rho, tip_mass_kg=0.05, tip_mass_nodes=tip)
    X = np.random.default_rng(4).standard_normal((geom.n_dofs, 3))
    free = ofem.free_dofs
    Xf = np.zeros_like(X)
    Xf[free] = X[free]
    ref = np.zeros_like(X)
    ref[free] = (M @ Xf)[free]
    out = vfem.assemble_M(rho, 0.05, tip) @ X
    assert rel(out, ref) < 1e-13


def test_fp32_stencil(gpu):
    from vox8 import _lib
    geom = MaskedBox(9, 5, 4, seed=2)
    ofem, vfem, rho = _pair(geom)
    K = ofem.assemble_K(rho)
    free = ofem.free_dofs
    x = np.random.default_rng(3).standard_normal(geom.n_dofs)
    x[ofem.fixed_dofs] = 0.0
    ref = np.zeros_like(x)
    ref[free] = (K @ x)[free]
    vfem._set_density(rho)
    xd = torch.tensor(x, dtype=torch.float32, device=gpu)
    y = torch.empty_like(xd)
    vfem._call("vox8_apply", _lib.OP_K, _lib.PREC_FP32, _lib.ptr(xd), _lib.ptr(y), None, 1,
               geom.n_dofs)
    assert rel(y.cpu().numpy(), ref) < 1e-5


def test_mutated_Ke_is_used(gpu):
    """Callers overwrite FEM3D.Ke_solid (anisotropy study, sims/run_all.py:298-303)."""
    from oracle.element import element_K, ortho_D
    geom = MaskedBox(5, 3, 3, seed=7)
    ofem, vfem, rho = _pair(geom)
    D = ortho_D(2.0e9, 1.2e9, 1.0e9, 0.35, 0.3, 0.33, 0.6e9, 0.4e9, 0.5e9)
    Ke = element_K(0, 0, geom.dx, geom.dy, geom.dz, D=D)
    ofem.Ke_solid = Ke
    ofem.Ke_flat = Ke.ravel()
    vfem.Ke_solid = Ke.copy()
    K = ofem.assemble_K(rho)
    x = np.random.default_rng(9).standard_normal(geom.n_dofs)
    x[ofem.fixed_dofs] = 0
    ref = np.zeros_like(x)
    ref[ofem.free_dofs] = (K @ x)[ofem.free_dofs]
    assert rel(vfem.assemble_K(rho) @ x, ref) < 1e-13


def _general_D():
    """A full (triclinic) elasticity matrix: SPD with shear-normal coupling."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((6, 6))
    return (A @ A.T + 6 * np.eye(6)) * 1e8


@pytest.mark.parametrize("kind", ["iso", "iso_box", "ortho", "general"])
def test_fast_stencil_patterns_vs_dense(gpu, kind):
    """The structured kernel (each pattern) against the dense 24x24 kernel and the oracle."""
    from oracle.element import element_K, ortho_D
    geom = MaskedBox(70, 11, 9, seed=3)            # spans several tiles in x, y and z
    if kind == "iso_box":                          # isotropic, non-cubic voxel
        geom.dy, geom.dz = 1.3e-3, 0.8e-3
    ofem, vfem, rho = _pair(geom)
    if kind in ("ortho", "general"):
        D = ortho_D(2.0e9, 1.2e9, 1.0e9, 0.35, 0.3, 0.33, 0.6e9, 0.4e9, 0.5e9) if kind == "ortho" else _general_D()
        Ke = element_K(0, 0, geom.dx, geom.dy, geom.dz, D=D)
        ofem.Ke_solid, ofem.Ke_flat = Ke, Ke.ravel()
        vfem.Ke_solid = Ke.copy()
    pat = vfem.set_stencil("auto")
    assert pat == {"iso": 2, "iso_box": 0, "ortho": 0, "general": 1}[kind]
    X = np.random.default_rng(7).standard_normal((geom.n_dofs, 3))
    yk = vfem.assemble_K(rho) @ X
    ym = vfem.assemble_M(rho) @ X
    assert vfem.set_stencil("dense") == -1
    yk_d = vfem.assemble_K(rho) @ X
    ym_d = vfem.assemble_M(rho) @ X
    vfem.set_stencil("auto")
    assert rel(yk, yk_d) < 1e-13 and rel(ym, ym_d) < 1e-13
    K, M = ofem.assemble_K(rho), ofem.assemble_M(rho)
    free = ofem.free_dofs
    Xf = np.zeros_like(X)
    Xf[free] = X[free]
    ref = np.zeros_like(X)
    ref[free] = (K @ Xf)[free]
    assert rel(yk, ref) < 1e-13
