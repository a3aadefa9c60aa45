"""Drop-in seam: route a reference-style FEM3D class through vox8.

The reference's seam is ``enable_gpu_solver(force=False) -> bool``
(src/fem_gpu.py:176-228), which monkey-patches ``FEM3D.apply_BCs_and_solve``
and ``FEM3D.modal_analysis`` so every caller uses the GPU without code
changes.  ``enable_gpu_solver(force, fem_cls)`` does the same for any class
with the reference FEM3D surface (``geom``, ``mat``, ``Ke_solid``,
``Me_solid`` and the SIMP/mass class constants), and also moves
``assemble_K``/``assemble_M`` (to matrix-free operators),
``eigenvalue_sensitivity`` and ``elem_compliance`` onto the device.  The
patched class keeps its own attributes; a vox8 FEM3D built from the same
geometry and material shadows each instance and takes Ke/Me/constants from
the instance at call time.

Unlike the reference there is no CPU fallback after the patch: a GPU error
raises vox8.Vox8Error.
"""
from __future__ import annotations

import ctypes as C

from . import _lib

__all__ = ["gpu_available", "enable_gpu_solver"]


def gpu_available() -> bool:
    """True when libvox8 loads and a HIP device is visible (src/fem_gpu.py:42-43)."""
    try:
        lib = _lib.load()
    except Exception:
        return False
    n = C.c_int(0)
    return lib.vox8_device_count(C.byref(n)) == 0 and n.value > 0


_PATCHED = ("assemble_K", "assemble_M", "apply_BCs_and_solve", "modal_analysis",
            "eigenvalue_sensitivity", "elem_compliance")


def _shadow(self):
    from .fem import FEM3D
    sh = getattr(self, "_vox8_shadow", None)
    if sh is None:
        sh = FEM3D(self.geom, self.mat)
        self._vox8_shadow = sh
    for name in ("SIMP_PENALTY", "E_MIN_FRAC", "MASS_THRESHOLD", "MASS_PENALTY_Q"):
        setattr(sh, name, getattr(self, name, getattr(sh, name)))
    sh.Ke_solid = self.Ke_solid
    sh.Me_solid = self.Me_solid
    sh.D_solid = getattr(self, "D_solid", sh.D_solid)
    return sh


def enable_gpu_solver(force: bool = False, fem_cls=None) -> bool:
    """Patch ``fem_cls`` (default: vox8.fem.FEM3D, already native) to run on vox8.

    Returns True when the patch was applied, False when no GPU is visible and
    ``force`` is not set (mirrors src/fem_gpu.py:176-185).
    """
    if not (gpu_available() or force):
        return False
    if fem_cls is None:
        from .fem import FEM3D as fem_cls  # noqa: N813 - native class, nothing to patch
        return True
    if getattr(fem_cls, "_vox8_patched", False):
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

    for name, fn in (("assemble_K", assemble_K), ("assemble_M", assemble_M),
                     ("apply_BCs_and_solve", apply_BCs_and_solve),
                     ("modal_analysis", modal_analysis),
                     ("eigenvalue_sensitivity", eigenvalue_sensitivity),
                     ("elem_compliance", elem_compliance)):
        setattr(fem_cls, "_vox8_orig_" + name, getattr(fem_cls, name, None))
        setattr(fem_cls, name, fn)
    fem_cls._vox8_patched = True
    return True
