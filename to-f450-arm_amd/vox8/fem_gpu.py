"""Drop-in seam: route a reference-style FEM3D class through vox8.

The reference's seam is ``enable_gpu_solver(force=False) -> bool``
(src/fem_gpu.py:176-228), which monkey-patches ``FEM3D.apply_BCs_and_solve``
and ``FEM3D.modal_analysis`` so every caller uses the GPU without code
changes.  ``enable_gpu_solver(force, fem_cls, precision)`` does the same for
any class with the reference FEM3D surface (``geom``, ``mat``, ``Ke_solid``,
``Me_solid``, ``D_solid``, ``B_centre`` and the SIMP/mass class constants)
and also moves ``assemble_K``/``assemble_M`` (matrix-free operators that
materialise the reference's CSR when a caller indexes or slices them),
``eigenvalue_sensitivity``, ``elem_compliance`` and the stress
post-processing onto the device.  ``FEM3D.solve`` needs no patch: it calls
the patched ``assemble_K`` and ``apply_BCs_and_solve``.

Each instance of the patched class is shadowed by a vox8 FEM3D built from
the same geometry and material, with the chosen precision ("mixed": FP64
Krylov solvers with the FP32 multigrid preconditioner; "fp64": the
reference's Jacobi preconditioner).  Ke/Me/D and the class constants are
taken from the instance at call time, so callers that overwrite
``Ke_solid`` (sims/run_parallel.py:303-307) keep working.

Unlike the reference there is no CPU fallback after the patch: a GPU error
raises vox8.Vox8Error.  ``disable_gpu_solver(fem_cls)`` restores the class.
"""
from __future__ import annotations

import ctypes as C

from . import _lib

__all__ = ["gpu_available", "enable_gpu_solver", "disable_gpu_solver"]


def gpu_available() -> bool:
    """True when libvox8 loads and a HIP device is visible (src/fem_gpu.py:42-43)."""
    try:
        lib = _lib.load()
    except Exception:
        return False
    n = C.c_int(0)
    return lib.vox8_device_count(C.byref(n)) == 0 and n.value > 0


_PATCHED = ("assemble_K", "assemble_M", "apply_BCs_and_solve", "modal_analysis",
            "eigenvalue_sensitivity", "elem_compliance", "compute_stress", "von_mises",
            "principal_stress_max")


def _shadow(self):
    """The vox8 FEM3D behind a patched instance, synchronised with its attributes."""
    from .fem import FEM3D
    prec = type(self)._vox8_precision
    sh = self.__dict__.get("_vox8_shadow")
    if sh is None or sh.precision != prec:
        if sh is not None:
            sh.close()
        sh = FEM3D(self.geom, self.mat, precision=prec)
        self.__dict__["_vox8_shadow"] = sh
    for name in ("SIMP_PENALTY", "E_MIN_FRAC", "MASS_THRESHOLD", "MASS_PENALTY_Q"):
        setattr(sh, name, getattr(self, name, getattr(sh, name)))
    sh.Ke_solid = self.Ke_solid
    sh.Me_solid = self.Me_solid
    sh.D_solid = getattr(self, "D_solid", sh.D_solid)
    sh.B_centre = getattr(self, "B_centre", sh.B_centre)
    return sh


def enable_gpu_solver(force: bool = False, fem_cls=None, precision: str = "mixed") -> bool:
    """Patch ``fem_cls`` to run on vox8 (mirrors src/fem_gpu.py:176-228).

    Returns True when the patch was applied, False when no GPU is visible and
    ``force`` is not set.  ``fem_cls=None`` means vox8.fem.FEM3D, which is
    native already (nothing to patch).  Calling it again on a patched class
    switches its precision.
    """
    if precision not in ("mixed", "fp64"):
        raise ValueError("precision must be 'mixed' or 'fp64'")
    if not (gpu_available() or force):
        return False
    if fem_cls is None:
        return True
    fem_cls._vox8_precision = precision
    if fem_cls.__dict__.get("_vox8_patched", False):
        return True

    def assemble_K(self, rho_active):
        return _shadow(self).assemble_K(rho_active)

    def assemble_M(self, rho_active, tip_mass_kg=0.0, tip_mass_nodes=None):
        return _shadow(self).assemble_M(rho_active, tip_mass_kg, tip_mass_nodes)

    def apply_BCs_and_solve(self, K, F, method="cg", x0=None, rtol=1e-5):
        return _shadow(self).apply_BCs_and_solve(K, F, method=method, x0=x0, rtol=rtol)

    def modal_analysis(self, rho_active, n_modes=6, tip_mass_kg=0.060, tip_mass_nodes=None,
                       X0_free=None, tol=1e-4, maxiter=500):
        return _shadow(self).modal_analysis(rho_active, n_modes, tip_mass_kg, tip_mass_nodes,
                                            X0_free, tol, maxiter)

    def eigenvalue_sensitivity(self, rho_active, omega_sq_k, v_k, tip_mass_kg=0.060,
                               tip_mass_nodes=None):
        return _shadow(self).eigenvalue_sensitivity(rho_active, omega_sq_k, v_k, tip_mass_kg,
                                                    tip_mass_nodes)

    def elem_compliance(self, u, rho_active=None):
        return _shadow(self).elem_compliance(u, rho_active)

    def compute_stress(self, u, rho_active):
        return _shadow(self).compute_stress(u, rho_active)

    def von_mises(self, stress):
        from .fem import FEM3D
        return FEM3D.von_mises(stress)

    def principal_stress_max(self, stress):
        from .fem import FEM3D
        return FEM3D.principal_stress_max(stress)

    fns = dict(assemble_K=assemble_K, assemble_M=assemble_M,
               apply_BCs_and_solve=apply_BCs_and_solve, modal_analysis=modal_analysis,
               eigenvalue_sensitivity=eigenvalue_sensitivity, elem_compliance=elem_compliance,
               compute_stress=compute_stress, von_mises=von_mises,
               principal_stress_max=principal_stress_max)
    originals = {}
    for name in _PATCHED:
        originals[name] = fem_cls.__dict__.get(name)
        setattr(fem_cls, name, fns[name])
    fem_cls._vox8_originals = originals
    fem_cls._vox8_patched = True
    return True


def disable_gpu_solver(fem_cls) -> None:
    """Undo enable_gpu_solver on ``fem_cls``: the class's own methods come
    back; shadow contexts of its instances are closed when they are next
    collected."""
    if not fem_cls.__dict__.get("_vox8_patched", False):
        return
    for name, fn in fem_cls._vox8_originals.items():
        if fn is None:
            delattr(fem_cls, name)
        else:
            setattr(fem_cls, name, fn)
    for attr in ("_vox8_originals", "_vox8_patched", "_vox8_precision"):
        delattr(fem_cls, attr)
