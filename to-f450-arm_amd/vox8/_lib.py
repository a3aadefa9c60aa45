"""ctypes binding of libvox8.so (C ABI in include/vox8.h).

The library is loaded from this package directory.  There is no fallback:
if it is missing or no GPU is visible, every GPU entry point raises.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libvox8.so"

OP_K, OP_M, OP_KM = 1, 2, 3
PREC_FP64, PREC_MIXED, PREC_FP32 = 0, 1, 2

_vp = C.c_void_p
_i32 = C.c_int
_i64 = C.c_int64
_f64 = C.c_double
_pi32 = C.POINTER(C.c_int)
_pf64 = C.POINTER(C.c_double)

class TailArgs(C.Structure):
    """struct vox8_tail_args (include/vox8.h), field for field."""
    _fields_ = [("d_X", _vp), ("h_lam", _pf64), ("d_vprev", _vp), ("d_rho", _vp),
                ("d_dlam", _vp), ("d_dlam_f", _vp), ("d_sens_f", _vp), ("d_design", _vp),
                ("d_x", _vp), ("d_xnew", _vp), ("d_check", _vp), ("d_phys_prev", _vp),
                ("ldx", _i64), ("w_target", _f64), ("mu", _f64), ("p", _f64), ("e_min", _f64),
                ("mthr", _f64), ("mq", _f64), ("V_e", _f64), ("target_volume", _f64),
                ("move", _f64), ("k", _i32), ("have_prev", _i32)]


TAIL_STATE = 20

# name -> argtypes (all return int status, except vox8_last_error)
SIGNATURES = {
    "vox8_version": [],
    "vox8_device_count": [_pi32],
    "vox8_create": [C.POINTER(_vp), _i32, _i32, _i32, _vp],
    "vox8_destroy": [_vp],
    "vox8_set_stream": [_vp, _vp],
    "vox8_set_element_matrices": [_vp, _pf64, _pf64],
    "vox8_set_stencil": [_vp, _i32, _pi32],
    "vox8_set_topology": [_vp, _vp, _vp],
    "vox8_set_tip_mass": [_vp, C.POINTER(C.c_int64), _i32, _f64],
    "vox8_set_density": [_vp, _vp, _f64, _f64, _f64, _f64],
    "vox8_apply": [_vp, _i32, _i32, _vp, _vp, _vp, _i32, _i64],
    "vox8_diagonal": [_vp, _i32, _vp],
    "vox8_pcg": [_vp, _vp, _vp, _i32, _i64, _i32, _f64, _i32, _i32, _pi32, _pi32, _pf64],
    "vox8_lobpcg": [_vp, _i32, _f64, _i32, _i32, _vp, _vp, _pf64, _pi32, _pi32, _pf64,
                    C.c_uint64],
    "vox8_set_lobpcg_guard": [_vp, _i32],
    "vox8_lobpcg_ritz": [_vp, _pf64, _i32, _pi32],
    "vox8_precond_apply": [_vp, _vp, _vp, _i32],
    "vox8_cheb_step": [_vp, _vp, _vp, _vp, _vp, _f64, _f64, _i32],
    "vox8_cheb_xstep": [_vp, _vp, _vp, _vp, _vp, _f64, _f64, _i32],
    "vox8_mg_info": [_vp, _pi32, _pi32, _pf64, _i32],
    "vox8_set_mg_options": [_vp, _i32, _f64, _i32, _i32, C.c_int64],
    "vox8_set_mg_graphs": [_vp, _i32],
    "vox8_set_mg_galerkin": [_vp, _i32],
    "vox8_set_mg_coarse_smoothing": [_vp, _i32, _i32],
    "vox8_share_multigrid": [_vp, _vp],
    "vox8_prepare_multigrid": [_vp, _i32],
    "vox8_stress": [_vp, _vp, _vp, _pf64, _pf64, _vp],
    "vox8_stress_invariants": [_vp, C.c_int64, _vp, _vp, _vp],
    "vox8_fatigue": [_vp, C.c_int64, _vp, _vp, _f64, _pf64, _pf64, _f64, _f64, _f64,
                     _vp, _vp, _vp, _vp, _vp, _vp],
    "vox8_sensitivities": [_vp, _vp, _i32, _i64, _pf64, _vp, _f64, _vp, _f64, _f64, _f64,
                           _f64, _vp, _vp, _vp, _pf64, _pf64],
    "vox8_filter_setup": [_vp, _f64, _vp],
    "vox8_filter": [_vp, _vp, _vp],
    "vox8_filter_adjoint": [_vp, _vp, _vp],
    "vox8_oc_update": [_vp, _vp, _vp, _vp, _f64, _f64, _f64, _vp, _pf64],
    "vox8_stats": [_vp, _i64, _vp, _vp, _vp, _pf64],
    "vox8_block_dots": [_vp, _i64, _i32, _vp, _i64, _vp, _pf64],
    "vox8_block_dots_noalloc": [_vp, _i64, _i32, _vp, _i64, _vp, _vp],
    "vox8_axpby": [_vp, _i64, _vp, _f64, _vp, _f64, _vp],
    "vox8_compose_density": [_vp, _vp, _vp, _vp, _vp],
    "vox8_simp_tail": [_vp, C.POINTER(TailArgs), _pf64],
    "vox8_set_tail_graph": [_vp, _i32],
    "vox8_effective_mass": [_vp, _i32, _vp, _i64, _i32, _pf64, _pf64, _pf64, _pf64],
    "vox8_trial_mfma_apply": [_vp, _vp, _vp, _i32],
}


class Vox8Error(RuntimeError):
    """A vox8 C-ABI call returned an error status."""


_lib = None


def load():
    """Load libvox8.so once and attach the ctypes signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise Vox8Error(f"{LIB_PATH} is not built; run vox8.build.build() (or __graft_entry__.build())")
    lib = C.CDLL(str(LIB_PATH))
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = C.c_int
    lib.vox8_last_error.argtypes = []
    lib.vox8_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def call(name, *args):
    """Call a vox8 entry point and raise Vox8Error on a non-zero status."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != 0:
        msg = lib.vox8_last_error().decode(errors="replace")
        raise Vox8Error(f"{name} failed (status {st}): {msg}")
    return st


_pending = []
_pending_lock = threading.Lock()


def _drain():
    with _pending_lock:
        ctxs = list(_pending)
        _pending.clear()
    lib = load()
    for c in ctxs:
        lib.vox8_destroy(c)


def create(nx, ny, nz, stream):
    """New context for an nx x ny x nz grid bound to `stream` (a HIP stream
    handle as int, 0 = default stream).  Contexts released by the garbage
    collector since the last call are destroyed first, on this thread."""
    _drain()
    ctx = C.c_void_p()
    call("vox8_create", C.byref(ctx), int(nx), int(ny), int(nz), C.c_void_p(stream))
    return ctx


def destroy(ctx):
    """Destroy a context now (and any released ones)."""
    _drain()
    if ctx is not None and ctx.value:
        load().vox8_destroy(ctx)


def release(ctx):
    """Hand a context over for destruction at the next create()/destroy():
    what a finaliser calls, so no library call is made from the collector."""
    if ctx is not None and ctx.value:
        with _pending_lock:
            _pending.append(C.c_void_p(ctx.value))


def ptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())


def f64p(arr):
    return arr.ctypes.data_as(_pf64)
