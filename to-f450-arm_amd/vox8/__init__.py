"""vox8 — MI355X-native voxel hex8 FEM hot path (stiffness / mass products,
Jacobi-PCG, LOBPCG, adjoint sensitivities, density filter, SIMP loop).

Python surface mirrors the reference's src/fem.py, src/fem_gpu.py and
src/topopt.py; all numerics run in libvox8.so (HIP, gfx950) through ctypes.
"""
from . import _lib
from ._lib import Vox8Error

__all__ = ["Vox8Error", "lib_path", "FEM3D", "Material", "SIMPOptimizer", "filter_field",
           "cone_kernel", "enable_gpu_solver", "gpu_available"]


def lib_path():
    return str(_lib.LIB_PATH)


def __getattr__(name):
    # torch-dependent modules load lazily so `import vox8` stays cheap
    if name in ("FEM3D", "Material"):
        from . import fem
        return getattr(fem, name)
    if name in ("SIMPOptimizer", "filter_field", "cone_kernel"):
        from . import topopt
        return getattr(topopt, name)
    if name in ("enable_gpu_solver", "gpu_available"):
        from . import fem_gpu
        return getattr(fem_gpu, name)
    raise AttributeError(name)
