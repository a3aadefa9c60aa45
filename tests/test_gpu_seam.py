"""GPU: the reference's seam, enable_gpu_solver(force, fem_cls, precision).

A reference-shaped class stands in for src/fem.py's FEM3D: a subclass of the
oracle's FEM3D (its restatement), patched through vox8 and compared with an
unpatched instance of the oracle class on the same inputs.  Every call is
made the way the reference's callers make it:

* ``FEM3D.solve`` (the reference's assemble_K + apply_BCs_and_solve pair,
  src/fem.py:558-561), and ``K.nnz`` / ``K[free][:, free]`` on the patched
  ``assemble_K`` result (src/fem.py:675, 436-437);
* ``modal_analysis``, ``eigenvalue_sensitivity``, ``elem_compliance``,
  ``compute_stress`` + ``von_mises``;
* ``effective_mass_voxel`` of sims/modal_participation.py:37-57 (oracle
  restatement), which slices ``assemble_M(...)[fd][:, fd]``;
* three iterations of the SIMP loop of src/topopt.py (oracle restatement)
  driving the patched instance.

Bars: 1e-6 relative on compliance, frequencies, sensitivities, effective
masses, stresses and the loop history; the materialised CSR equals the
oracle's assembled matrix to 1e-14.
"""
import numpy as np
import pytest

from helpers import TIPBOX_MODAL_MAXITER, TIPBOX_MODAL_TOL, TIPBOX_SOLVE_RTOL, rel, tipbox_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["mixed", "fp64"])
def patched(gpu, request):
    from oracle.fem import FEM3D as OFEM
    from vox8.fem_gpu import disable_gpu_solver, enable_gpu_solver

    class RefFEM3D(OFEM):
        """Stand-in for the reference's FEM3D class (patched in place)."""

    assert enable_gpu_solver(fem_cls=RefFEM3D, precision=request.param)
    yield RefFEM3D, request.param
    disable_gpu_solver(RefFEM3D)


def _problem():
    from oracle.fem import FEM3D as OFEM, Material as OMat
    geom, loads, opts, mat = tipbox_problem()
    rho = np.random.default_rng(4).uniform(0.15, 1.0, OFEM(geom, OMat(**mat)).n_active)
    return geom, loads, opts, OMat(**mat), rho


def test_seam_solve_modal_sensitivities(patched):
    from oracle.fem import FEM3D as OFEM
    from oracle.modal import effective_mass_voxel
    from vox8.fem import VoxelOperator
    cls, prec = patched
    geom, loads, opts, mat, rho = _problem()
    ref, pat = OFEM(geom, mat), cls(geom, mat)
    d, m = geom.load_dofs_for_force((0.0, 0.0, -1.0))
    F = np.zeros(geom.n_dofs)
    F[d] += m
    # static: the reference's solve() chain and the assembled-matrix surface
    K = pat.assemble_K(rho)
    assert isinstance(K, VoxelOperator)
    Kref = ref.assemble_K(rho)
    assert K.nnz == Kref.nnz and K.shape == Kref.shape
    fd = pat.free_dofs
    assert rel(K[fd][:, fd].toarray(), Kref[fd][:, fd].toarray()) < 1e-14
    u = pat.apply_BCs_and_solve(K, F, rtol=1e-12)
    u_ref = ref.apply_BCs_and_solve(Kref, F, method="lu")
    assert isinstance(u, np.ndarray) and u.shape == (geom.n_dofs,)
    assert abs(F @ u / (F @ u_ref) - 1) < 1e-6
    assert rel(pat.solve(rho, F), u_ref) < 1e-3          # the reference's solve(): rtol 1e-5
    # modal
    tip = geom.motor_mount_top_nodes()
    f, modes, lam = pat.modal_analysis(rho, n_modes=4, tip_mass_kg=2e-3, tol=1e-7, maxiter=4000)
    f_ref, modes_ref, lam_ref = ref.modal_analysis(rho, n_modes=4, tip_mass_kg=2e-3, tol=1e-7,
                                                   maxiter=4000)
    assert modes.shape == (geom.n_dofs, 4)
    assert np.max(np.abs(f / f_ref - 1)) < 1e-6, (f, f_ref)
    # sensitivities and element energies
    dl = pat.eigenvalue_sensitivity(rho, lam[0], modes[:, 0], tip_mass_kg=2e-3)
    dl_ref = ref.eigenvalue_sensitivity(rho, lam[0], modes[:, 0], tip_mass_kg=2e-3)
    assert rel(dl, dl_ref) < 1e-6
    ce, ce_ref = pat.elem_compliance(u_ref, rho), ref.elem_compliance(u_ref, rho)
    assert rel(ce, ce_ref) < 1e-10
    vm = pat.von_mises(pat.compute_stress(u_ref, rho))
    vm_ref = ref.von_mises(ref.compute_stress(u_ref, rho))
    assert rel(vm, vm_ref) < 1e-10
    # effective mass through the sliced mass matrix (tip lump included)
    M = pat.assemble_M(rho, tip_mass_kg=2e-3, tip_mass_nodes=tip)
    assert rel(M.toarray(), ref.assemble_M(rho, 2e-3, tip).toarray()) < 1e-14
    em = effective_mass_voxel(pat, geom, rho, modes, 2e-3)
    em_ref = effective_mass_voxel(ref, geom, rho, modes_ref, 2e-3)
    assert np.max(np.abs(em - em_ref)) < 1e-6, (em, em_ref)
    # the patch is per class: the oracle class itself is untouched
    assert not isinstance(ref.assemble_K(rho), VoxelOperator)


def test_seam_drives_simp_loop(patched):
    from oracle.fem import FEM3D as OFEM
    from oracle.topopt import SIMPOptimizer
    cls, prec = patched
    geom, loads, opts, mat, _ = _problem()
    opts = dict(opts, max_iter=3)
    kw = dict(solve_rtol=TIPBOX_SOLVE_RTOL, modal_tol=TIPBOX_MODAL_TOL,
              modal_maxiter=TIPBOX_MODAL_MAXITER)
    a = SIMPOptimizer(geom, cls(geom, mat), loads, **opts, **kw).run()
    b = SIMPOptimizer(geom, OFEM(geom, mat), loads, **opts, **kw).run()
    for k in ("compliance", "mass", "change", "vol_frac", "omega_1_Hz", "mu_freq"):
        assert rel(np.asarray(a["history"][k]), np.asarray(b["history"][k])) < 1e-6, k
    assert rel(a["rho_active"], b["rho_active"]) < 1e-6


def test_seam_switches_precision_and_restores(gpu):
    from oracle.fem import FEM3D as OFEM
    from vox8.fem_gpu import disable_gpu_solver, enable_gpu_solver

    class RefFEM3D(OFEM):
        pass

    orig = RefFEM3D.assemble_K
    assert enable_gpu_solver(fem_cls=RefFEM3D, precision="fp64")
    geom, loads, opts, mat, rho = _problem()
    inst = RefFEM3D(geom, mat)
    inst.assemble_K(rho)
    assert inst._vox8_shadow.precision == "fp64"
    assert enable_gpu_solver(fem_cls=RefFEM3D, precision="mixed")
    inst.assemble_K(rho)
    assert inst._vox8_shadow.precision == "mixed"
    disable_gpu_solver(RefFEM3D)
    assert RefFEM3D.assemble_K is orig
    with pytest.raises(ValueError):
        enable_gpu_solver(fem_cls=RefFEM3D, precision="fp32")
