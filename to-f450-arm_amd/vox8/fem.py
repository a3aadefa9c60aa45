"""GPU-native FEM3D with the reference's call surface (src/fem.py:182-625).

Drop-in differences, all deliberate:

* ``assemble_K`` / ``assemble_M`` return a matrix-free ``VoxelOperator``;
  ``apply_BCs_and_solve`` accepts that operator (an assembled scipy matrix
  raises TypeError).  The solvers never assemble.  Callers that index, slice
  or count the matrix (``M[fd][:, fd]``, ``K.nnz``, ``tocsr()``) get the
  reference's CSR, assembled on the host on first such use.
* Every product, solve and sensitivity runs in libvox8.so on the GPU; there
  is no CPU fallback.  Non-convergence is reported as a RuntimeWarning and the
  best iterate is returned (the reference silently switches to spsolve /
  eigsh, src/fem.py:548-550, 455-461).
* Inputs may be NumPy arrays (results come back as NumPy, like the reference)
  or CUDA torch tensors (results stay on the device).
"""
from __future__ import annotations

import ctypes as C
import math
import warnings
import weakref
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from ._lib import call, ptr
from .element import (constitutive_D, unit_element_B_at_centre, unit_element_K,
                      unit_element_M)

__all__ = ["Material", "FEM3D", "VoxelOperator", "constitutive_D", "unit_element_K",
           "unit_element_M", "unit_element_B_at_centre"]


@dataclass
class Material:
    """src/fem.py:182-188."""
    E: float = 1700e6
    nu: float = 0.39
    sigma_y: float = 38e6
    sigma_u: float = 48e6
    rho: float = 930.0


_PREC = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}


def _device():
    if not torch.cuda.is_available():
        raise _lib.Vox8Error("vox8 needs a visible GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _tip_nodes_of(geom):
    """Tip-mass node lookup of src/fem.py:426-432."""
    if hasattr(geom, "motor_mount_top_nodes"):
        return np.asarray(geom.motor_mount_top_nodes(), dtype=np.int64)
    if hasattr(geom, "_motor_top_boundary_nodes"):
        return np.asarray(geom._motor_top_boundary_nodes(), dtype=np.int64)
    return np.array([], dtype=np.int64)


def _stress_to_dev(stress):
    """(n, 6) device float64 tensor of Voigt rows, converter back to the caller's type."""
    if isinstance(stress, torch.Tensor):
        t = stress if stress.is_cuda else stress.to("cuda")
        return t.to(torch.float64).reshape(-1, 6).contiguous(), (lambda y: y)
    arr = np.ascontiguousarray(np.asarray(stress, dtype=np.float64).reshape(-1, 6))
    return torch.from_numpy(arr).to("cuda"), (lambda y: y.cpu().numpy())


class VoxelOperator:
    """Matrix-free K(rho) or M(rho) on the full DOF vector (what assemble_* return).

    ``A @ x`` applies the free-DOF operator (x read as zero at fixed DOFs, zero
    rows there), i.e. the embedding of K_ff / M_ff used by the solvers; it
    equals the assembled product on every vector that is zero at the fixed
    DOFs.  Indexing, ``nnz``, ``tocsr()`` and ``toarray()`` use the assembled
    matrix of src/fem.py:311-325 / 362-398 (all DOFs, M with the tip lump),
    built on the host once per operator for callers off the solver path such
    as sims/modal_participation.py:44-49.
    """

    def __init__(self, fem, kind, rho, tip_mass_kg=0.0, tip_nodes=None):
        self.fem, self.kind, self.rho = fem, kind, rho
        self.tip_mass_kg = tip_mass_kg
        self.tip_nodes = np.array([], dtype=np.int64) if tip_nodes is None else np.asarray(tip_nodes, dtype=np.int64)
        n = fem.geom.n_dofs
        self.shape = (n, n)
        self.dtype = np.dtype(np.float64)
        self._csr = None

    def tocsr(self):
        if self._csr is None:
            self._csr = self.fem._host_csr(self.kind, self.rho, self.tip_mass_kg, self.tip_nodes)
        return self._csr

    def toarray(self):
        return self.tocsr().toarray()

    @property
    def nnz(self):
        return self.tocsr().nnz

    def __getitem__(self, key):
        return self.tocsr()[key]

    def _prepare(self):
        self.fem._set_density(self.rho)
        if self.kind == "M":
            self.fem._set_tip(self.tip_mass_kg, self.tip_nodes)

    def __matmul__(self, x):
        self._prepare()
        xt, back = self.fem._to_dev(x)
        X = xt.reshape(-1, self.shape[0]) if xt.dim() == 1 else xt.T.contiguous()
        Y = torch.empty_like(X)
        op = _lib.OP_K if self.kind == "K" else _lib.OP_M
        self.fem._call("vox8_apply", op, _lib.PREC_FP64, ptr(X),
                       ptr(Y) if op == _lib.OP_K else None, ptr(Y) if op == _lib.OP_M else None,
                       X.shape[0], X.shape[1])
        Y = Y.reshape(-1) if xt.dim() == 1 else Y.T
        return back(Y)

    def dot(self, x):
        return self @ x

    def diagonal(self):
        self._prepare()
        d = torch.empty(self.shape[0], dtype=torch.float64, device=self.fem.device)
        self.fem._call("vox8_diagonal", 1 if self.kind == "K" else 2, ptr(d))
        return d.cpu().numpy()


class FEM3D:
    """Linear-elastic voxel FEM on the GPU (reference: src/fem.py:191-625)."""

    SIMP_PENALTY = 3.0
    E_MIN_FRAC = 1e-9
    MASS_THRESHOLD = 0.1
    MASS_PENALTY_Q = 6.0

    def __init__(self, geom, mat: Material, precision: str = "fp64"):
        self.geom, self.mat = geom, mat
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        self.precision = precision
        self.D_solid = constitutive_D(mat.E, mat.nu)
        self.Ke_solid = unit_element_K(mat.E, mat.nu, geom.dx, geom.dy, geom.dz, D=self.D_solid)
        self.B_centre = unit_element_B_at_centre(geom.dx, geom.dy, geom.dz)
        self.Me_solid = unit_element_M(mat.rho, geom.dx, geom.dy, geom.dz)
        nx, ny, nz = geom.nx, geom.ny, geom.nz
        in_arm = np.asarray(geom.build_masks()["in_arm"], dtype=bool)
        # element grid e = ex + nx(ey + ny ez): the reference's active order
        # (src/fem.py:244-263) is this order restricted to in_arm
        self.elem_active_grid = np.ascontiguousarray(in_arm.transpose(2, 1, 0)).reshape(-1)
        self.active_elems = np.nonzero(self.elem_active_grid)[0].astype(np.int64)
        self.n_active = len(self.active_elems)
        ex = self.active_elems % nx
        ey = (self.active_elems // nx) % ny
        ez = self.active_elems // (nx * ny)
        self.elem_grid_idx = np.stack([ex, ey, ez], axis=1)
        # used nodes -> ghosts are fixed (src/fem.py:292-305)
        nxp, nyp, nzp = nx + 1, ny + 1, nz + 1
        act3 = self.elem_active_grid.reshape(nz, ny, nx)
        used = np.zeros((nzp, nyp, nxp), dtype=bool)
        for oz in (0, 1):
            for oy in (0, 1):
                for ox in (0, 1):
                    used[oz:oz + nz, oy:oy + ny, ox:ox + nx] |= act3
        fixed_nodes = np.asarray(geom.fixed_nodes(), dtype=np.int64)
        fixed = ~used.reshape(-1)
        fixed[fixed_nodes] = True
        self.node_fixed = np.where(fixed, 7, 0).astype(np.uint8)
        fn = np.nonzero(fixed)[0]
        self.fixed_dofs = np.sort(np.concatenate([3 * fn, 3 * fn + 1, 3 * fn + 2]))
        free = np.ones(geom.n_dofs, dtype=bool)
        free[self.fixed_dofs] = False
        self.free_dofs = np.nonzero(free)[0]
        self.rho_active = np.ones(self.n_active)
        self._elem_node_ids = None
        self.mg_galerkin = True
        # device context
        self.device = _device()
        self._stream = torch.cuda.current_stream(self.device).cuda_stream
        self._ctx = _lib.create(nx, ny, nz, self._stream)
        self._Ke_dev = None
        self._Me_dev = None
        self._sync_element_matrices()
        call("vox8_set_topology", self._ctx, self.elem_active_grid.astype(np.uint8).ctypes.data_as(C.c_void_p),
             self.node_fixed.ctypes.data_as(C.c_void_p))
        self._rho_grid = torch.zeros(self.geom.n_elements, dtype=torch.float64, device=self.device)
        self._active_idx_dev = torch.as_tensor(self.active_elems, device=self.device)
        self._full_box = self.n_active == self.geom.n_elements
        self._density_key = None
        self._tip_key = None
        self.last_info = {}
        self._mg_src = None                    # share_multigrid source
        self._mg_followers = weakref.WeakSet()

    # ------------------------------------------------------------------ plumbing
    def close(self):
        """Release the device context (its work space and captured graphs).
        The object is unusable afterwards.  Contexts still open when they are
        garbage-collected are handed to ``_lib.release`` instead, which
        destroys them on the next ``close()`` / context creation of the
        process, never from the collector's thread."""
        ctx = getattr(self, "_ctx", None)
        for f in list(getattr(self, "_mg_followers", ())):
            f.share_multigrid(None)            # no context keeps reading this one's hierarchy
        if getattr(self, "_mg_src", None) is not None:
            self._mg_src._mg_followers.discard(self)
            self._mg_src = None
        self._ctx = None
        if ctx is not None and ctx.value:
            _lib.destroy(ctx)

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value:
            self._ctx = None
            try:
                _lib.release(ctx)
            except (AttributeError, TypeError):
                # interpreter shutdown: module globals are already cleared
                # and the process is ending, so the context goes with it
                pass

    def _call(self, name, *args):
        if self._ctx is None:
            raise _lib.Vox8Error("FEM3D context is closed")
        s = torch.cuda.current_stream(self.device).cuda_stream
        if s != self._stream:
            self._stream = s
            call("vox8_set_stream", self._ctx, C.c_void_p(s))
        self._sync_element_matrices()
        return call(name, self._ctx, *args)

    def _sync_element_matrices(self):
        """Re-upload Ke/Me when a caller has replaced or mutated them."""
        Ke = np.ascontiguousarray(self.Ke_solid, dtype=np.float64)
        Me = np.ascontiguousarray(self.Me_solid, dtype=np.float64)
        if (self._Ke_dev is not None and np.array_equal(Ke, self._Ke_dev)
                and np.array_equal(Me, self._Me_dev)):
            return
        if Ke.shape != (24, 24) or Me.shape != (24, 24):
            raise ValueError("Ke_solid and Me_solid must be 24x24")
        call("vox8_set_element_matrices", self._ctx, _lib.f64p(Ke), _lib.f64p(Me))
        self._Ke_dev, self._Me_dev = Ke.copy(), Me.copy()
        self.Ke_flat = self.Ke_solid.ravel()
        self.Me_flat = self.Me_solid.ravel()
        self._density_key = None

    def _to_dev(self, x):
        """(device float64 tensor, converter back to the caller's type)."""
        if isinstance(x, torch.Tensor):
            return x.to(self.device, torch.float64).contiguous(), (lambda y: y)
        arr = np.asarray(x, dtype=np.float64)
        return (torch.from_numpy(np.ascontiguousarray(arr)).to(self.device),
                (lambda y: y.cpu().numpy()))

    def grid_from_active(self, rho_active):
        """Active-order densities -> element-grid device tensor."""
        r, _ = self._to_dev(rho_active)
        if r.numel() != self.n_active:
            raise ValueError(f"expected {self.n_active} active densities, got {r.numel()}")
        if self._full_box:
            return r
        g = torch.zeros(self.geom.n_elements, dtype=torch.float64, device=self.device)
        g[self._active_idx_dev] = r
        return g

    def active_from_grid(self, g):
        return g if self._full_box else g[self._active_idx_dev]

    def _set_density(self, rho_active):
        """Upload densities and form the element factors on the device.

        Device tensors are always re-read (kernels may have written them in
        place without torch noticing) and taken as given; a NumPy array is
        checked for NaN/inf (ValueError), and one equal to the last one set is
        skipped, which saves the host-to-device copy.
        """
        params = (self.SIMP_PENALTY, self.E_MIN_FRAC, self.MASS_THRESHOLD, self.MASS_PENALTY_Q)
        if not isinstance(rho_active, torch.Tensor):
            rho_active = np.ascontiguousarray(rho_active, dtype=np.float64)
            if not np.all(np.isfinite(rho_active)):
                raise ValueError("densities must be finite (NaN or inf given)")
            last = self._density_key
            if (last is not None and last[1] == params and last[0].shape == rho_active.shape
                    and np.array_equal(last[0], rho_active)):
                return self._rho_grid
        g = self.grid_from_active(rho_active)
        self._rho_grid = g
        self._call("vox8_set_density", ptr(g), *params)
        self._density_key = ((rho_active.copy(), params)
                             if isinstance(rho_active, np.ndarray) else None)
        return g

    def set_stencil(self, kind: str = "auto") -> int:
        """Select the stencil ("auto" or "dense"); returns the kernel pattern in
        use (-1 dense 24x24, 0 orthotropic Hadamard blocks, 1 general blocks)."""
        pat = C.c_int()
        self._call("vox8_set_stencil", {"auto": 0, "dense": 1}[kind], C.byref(pat))
        return pat.value

    def _set_tip(self, tip_mass_kg, tip_nodes):
        nodes = np.ascontiguousarray(np.asarray(tip_nodes, dtype=np.int64))
        key = (float(tip_mass_kg), nodes.tobytes())
        if key == self._tip_key:
            return
        n = len(nodes) if tip_mass_kg > 0.0 else 0
        self._call("vox8_set_tip_mass", nodes.ctypes.data_as(C.POINTER(C.c_int64)), n,
                   float(tip_mass_kg))
        self._tip_key = key

    # ------------------------------------------------------------------ index tables
    @property
    def elem_node_ids(self):
        """(n_active, 8) global node ids in the reference local order."""
        if self._elem_node_ids is None:
            nxp, nyp = self.geom.nx + 1, self.geom.ny + 1
            off = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                            (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)])
            g = self.elem_grid_idx
            ix = g[:, 0:1] + off[None, :, 0]
            iy = g[:, 1:2] + off[None, :, 1]
            iz = g[:, 2:3] + off[None, :, 2]
            self._elem_node_ids = ((iz * nyp + iy) * nxp + ix).astype(np.int64)
        return self._elem_node_ids

    @property
    def elem_dof_ids(self):
        n = self.elem_node_ids
        return (3 * n[:, :, None] + np.arange(3)[None, None, :]).reshape(len(n), 24)

    # ------------------------------------------------------------------ assembly surface
    def assemble_K(self, rho_active):
        return VoxelOperator(self, "K", rho_active)

    def assemble_M(self, rho_active, tip_mass_kg: float = 0.0, tip_mass_nodes=None):
        return VoxelOperator(self, "M", rho_active, tip_mass_kg,
                             tip_mass_nodes if tip_mass_nodes is not None else [])

    def _host_csr(self, kind, rho_active, tip_mass_kg=0.0, tip_nodes=()):
        """Assembled K(rho) or M(rho) over all DOFs as scipy CSR (duplicates
        summed, src/fem.py:311-325, 362-381) with the tip lump of
        src/fem.py:383-398 on M.  Host-side, for off-solver callers only."""
        rho = (rho_active.detach().cpu().numpy() if isinstance(rho_active, torch.Tensor)
               else np.asarray(rho_active, dtype=np.float64))
        if kind == "K":
            f = self.E_MIN_FRAC + (1.0 - self.E_MIN_FRAC) * rho ** self.SIMP_PENALTY
            Ae = np.asarray(self.Ke_solid, dtype=np.float64)
        else:
            f = self._mass_simp_factor(rho)
            Ae = np.asarray(self.Me_solid, dtype=np.float64)
        d = self.elem_dof_ids
        n = self.geom.n_dofs
        rows = np.repeat(d, 24, axis=1).ravel()
        cols = np.tile(d, (1, 24)).ravel()
        vals = (f[:, None] * Ae.ravel()[None, :]).ravel()
        A = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
        nodes = np.asarray(tip_nodes, dtype=np.int64)
        if kind == "M" and tip_mass_kg > 0.0 and len(nodes):
            dofs = np.concatenate([3 * nodes, 3 * nodes + 1, 3 * nodes + 2])
            lump = sp.coo_matrix((np.full(len(dofs), tip_mass_kg / len(nodes)), (dofs, dofs)),
                                 shape=(n, n)).tocsr()
            A = (A + lump).tocsr()
        return A

    def _mass_simp_factor(self, rho_active):
        t, q = self.MASS_THRESHOLD, self.MASS_PENALTY_Q
        rho_active = np.asarray(rho_active, dtype=float)
        return np.where(rho_active >= t, rho_active, t * (np.maximum(rho_active, 0.0) / t) ** q)

    # ------------------------------------------------------------------ static solve
    def solve_block(self, rho_active, F, x0=None, rtol=1e-8, maxiter=5000, precision=None):
        """All load cases at once: F (n_lc, n_dofs) -> U (n_lc, n_dofs) on the device.

        Each column is an independent Jacobi-PCG with scipy cg semantics; columns
        share every kernel launch.  Returns (U, info dict).
        """
        self._set_density(rho_active)
        Ft, _ = self._to_dev(F)
        Ft = Ft.reshape(-1, self.geom.n_dofs).contiguous()
        nrhs = Ft.shape[0]
        if x0 is not None:
            U, _ = self._to_dev(x0)
            U = U.reshape(nrhs, self.geom.n_dofs).clone()
        else:
            U = torch.zeros_like(Ft)
        prec = _PREC[precision or self.precision]
        iters = (C.c_int * nrhs)()
        info = (C.c_int * nrhs)()
        rel = (C.c_double * nrhs)()
        self._call("vox8_pcg", ptr(Ft), ptr(U), nrhs, self.geom.n_dofs, int(x0 is not None),
                   float(rtol), int(maxiter), prec, iters, info, rel)
        out = {"iters": list(iters), "info": list(info), "relres": list(rel)}
        self.last_info["pcg"] = out
        return U, out

    def apply_BCs_and_solve(self, K, F, method: str = "cg", x0=None, rtol: float = 1e-5):
        """K_ff u_f = F_f with u = 0 on fixed DOFs (src/fem.py:529-555)."""
        if not isinstance(K, VoxelOperator) or K.kind != "K":
            raise TypeError("vox8 FEM3D is matrix-free: pass the operator returned by assemble_K")
        if method not in ("cg", "lu"):
            raise ValueError(f"unknown method {method!r}")
        if method == "lu":           # no direct factorisation on the device: converge PCG tightly
            rtol = min(rtol, 1e-12)
        _, back = self._to_dev(F)
        U, info = self.solve_block(K.rho, F, x0=x0, rtol=rtol, maxiter=5000)
        if info["info"][0] != 0:
            U, info = self.solve_block(K.rho, F, x0=U, rtol=rtol, maxiter=100000)
            if info["info"][0] != 0:
                warnings.warn(f"vox8 PCG did not reach rtol={rtol} (relres {info['relres'][0]:.3e})",
                              RuntimeWarning, stacklevel=2)
        return back(U.reshape(-1))

    def solve(self, rho_active, F, method: str = "cg", x0=None):
        """src/fem.py:558-561 (rtol is not forwarded: 1e-5)."""
        return self.apply_BCs_and_solve(self.assemble_K(rho_active), F, method=method, x0=x0)

    # ------------------------------------------------------------------ modal
    def twin(self):
        """A second context for the same model (geometry, material, precision,
        element matrices, SIMP constants) on the caller's current stream, e.g.
        to run the modal solve concurrently with the static one."""
        t = FEM3D(self.geom, self.mat, precision=self.precision)
        if not self.mg_galerkin:
            t.set_mg_galerkin(False)
        t.Ke_solid = np.array(self.Ke_solid, copy=True)
        t.Me_solid = np.array(self.Me_solid, copy=True)
        for k in ("SIMP_PENALTY", "E_MIN_FRAC", "MASS_THRESHOLD", "MASS_PENALTY_Q"):
            if k in self.__dict__:
                setattr(t, k, getattr(self, k))
        return t

    def share_multigrid(self, src: "FEM3D | None"):
        """Take the per-density multigrid operators of ``src`` (a context of the
        same model) by a device copy instead of computing them
        (vox8_share_multigrid).  The caller gives both the same densities and
        runs ``src.prepare_multigrid`` before each solve of this context; a
        stale source only weakens the preconditioner, never the answer.
        ``None`` ends the sharing."""
        if self._ctx is None:
            return
        call("vox8_share_multigrid", self._ctx, None if src is None else src._ctx)
        if self._mg_src is not None:
            self._mg_src._mg_followers.discard(self)
        self._mg_src = src
        if src is not None:
            src._mg_followers.add(self)

    def prepare_multigrid(self, rho_active, ncols: int = 1):
        """Set the density and bring the multigrid hierarchy up to date for it
        (with work space for ``ncols`` columns); returns when the device work
        is done."""
        self._set_density(rho_active)
        self._call("vox8_prepare_multigrid", int(ncols))

    def set_mg_galerkin(self, enable: bool = True):
        """Coarse levels of this context's multigrid preconditioner: Galerkin
        P^T A P (True, the default) or rediscretised (False: cheaper
        coarse-level passes, but more Krylov iterations at high density
        contrast, and at some SIMP states the cycle stops helping and PCG
        stalls)."""
        self._call("vox8_set_mg_galerkin", int(bool(enable)))
        self.mg_galerkin = bool(enable)

    def set_lobpcg_guard(self, guard: int):
        """Guard columns of the modal solve (0..2; -1 = the default): iterated
        with the wanted modes, not returned; they speed up the last wanted mode
        when the next eigenvalue is close to it."""
        self._call("vox8_set_lobpcg_guard", int(guard))

    def lobpcg_ritz(self):
        """Ritz values of the last modal solve's whole block, ascending: the
        wanted eigenvalues, then the guard columns' (numpy; empty before the
        first solve)."""
        out = (C.c_double * 12)()
        n = C.c_int()
        self._call("vox8_lobpcg_ritz", out, 12, C.byref(n))
        return np.array(out[:n.value])

    def modal_block(self, rho_active, n_modes=6, tip_mass_kg=0.060, tip_mass_nodes=None,
                    X0=None, tol=1e-4, maxiter=500, seed=0, precision=None):
        """Device-resident modal solve: returns (lam (k,) numpy, X (k, n_dofs) torch, info)."""
        if tip_mass_nodes is None:
            tip_mass_nodes = _tip_nodes_of(self.geom)
        self._set_density(rho_active)
        self._set_tip(tip_mass_kg, tip_mass_nodes)
        k = int(n_modes)
        n = self.geom.n_dofs
        X = torch.empty(k, n, dtype=torch.float64, device=self.device)
        lam = (C.c_double * k)()
        iters, info, mres = C.c_int(), C.c_int(), C.c_double()
        x0p = None
        if X0 is not None:
            X0d, _ = self._to_dev(X0)
            x0p = ptr(X0d.reshape(k, n).contiguous())
            X0keep = X0d
        prec = _PREC[precision or self.precision]
        self._call("vox8_lobpcg", k, float(tol), int(maxiter), prec, x0p, ptr(X), lam,
                   C.byref(iters), C.byref(info), C.byref(mres), int(seed))
        out = {"iters": iters.value, "info": info.value, "max_res": mres.value}
        self.last_info["lobpcg"] = out
        return np.array(lam[:]), X, out

    def modal_analysis(self, rho_active, n_modes: int = 6, tip_mass_kg: float = 0.060,
                       tip_mass_nodes=None, X0_free=None, tol: float = 1e-4, maxiter: int = 500):
        """(frequencies_Hz, mode_shapes (n_dofs, k), omega_sq) as src/fem.py:400-470."""
        n, k = self.geom.n_dofs, int(n_modes)
        nf = len(self.free_dofs)
        X0 = None
        to_numpy = not isinstance(rho_active, torch.Tensor)
        if X0_free is not None and tuple(X0_free.shape) == (nf, k):
            Xf, _ = self._to_dev(X0_free)
            X0 = torch.zeros(k, n, dtype=torch.float64, device=self.device)
            X0[:, torch.as_tensor(self.free_dofs, device=self.device)] = Xf.T
        lam, X, info = self.modal_block(rho_active, k, tip_mass_kg, tip_mass_nodes, X0, tol, maxiter)
        if info["info"] != 0:
            warnings.warn(f"vox8 LOBPCG stopped after {info['iters']} iterations with max residual "
                          f"{info['max_res']:.3e} > tol {tol}", RuntimeWarning, stacklevel=2)
        order = np.argsort(lam)
        lam = np.maximum(lam[order], 0.0)
        X = X[torch.as_tensor(order, device=self.device)]
        freqs = np.sqrt(lam) / (2.0 * np.pi)
        modes = X.T.contiguous()
        return freqs, (modes.cpu().numpy() if to_numpy else modes), lam

    # ------------------------------------------------------------------ adjoint
    def sensitivities(self, rho_active, U=None, weights=None, phi=None, lam=0.0):
        """Device adjoint kernel: returns dict with ce, dc, comp (per LC), dlam, vmv.

        U: (n_lc, n_dofs) device block; phi: (n_dofs,) mode.  Element arrays are
        in element-grid order (equal to active order for a full box).
        """
        g = self._set_density(rho_active)
        ne = self.geom.n_elements
        nrhs = 0 if U is None else U.reshape(-1, self.geom.n_dofs).shape[0]
        w = np.ones(max(nrhs, 1)) if weights is None else np.asarray(weights, dtype=float)
        out = {}
        ce = torch.empty(ne, dtype=torch.float64, device=self.device) if nrhs else None
        dc = torch.empty(ne, dtype=torch.float64, device=self.device) if nrhs else None
        dl = torch.empty(ne, dtype=torch.float64, device=self.device) if phi is not None else None
        comp = (C.c_double * max(nrhs, 1))()
        vmv = C.c_double()
        Ud = U.reshape(nrhs, -1).contiguous() if nrhs else None
        ph = phi.reshape(-1).contiguous() if phi is not None else None
        self._call("vox8_sensitivities", ptr(Ud), nrhs, self.geom.n_dofs,
                   _lib.f64p(np.ascontiguousarray(w)), ptr(ph), float(lam), ptr(g),
                   self.SIMP_PENALTY, self.E_MIN_FRAC, self.MASS_THRESHOLD, self.MASS_PENALTY_Q,
                   ptr(ce), ptr(dc), ptr(dl), comp, C.byref(vmv))
        out.update(ce=ce, dc=dc, dlam=dl, comp=np.array(comp[:nrhs]), vmv=vmv.value)
        return out

    def eigenvalue_sensitivity(self, rho_active, omega_sq_k: float, v_k, tip_mass_kg: float = 0.060,
                               tip_mass_nodes=None):
        """d(lam)/d(rho_e) per active element (src/fem.py:472-526)."""
        if tip_mass_nodes is None:
            tip_mass_nodes = _tip_nodes_of(self.geom)
        self._set_tip(tip_mass_kg, tip_mass_nodes)
        v, back = self._to_dev(v_k)
        s = self.sensitivities(rho_active, phi=v, lam=float(omega_sq_k))
        return back(self.active_from_grid(s["dlam"]))

    def elem_compliance(self, u, rho_active=None):
        """u_e^T Ke u_e per active element (src/fem.py:605-615)."""
        ud, back = self._to_dev(u)
        rho = self.rho_active if rho_active is None else rho_active
        s = self.sensitivities(rho, U=ud.reshape(1, -1), weights=[1.0])
        return back(self.active_from_grid(s["ce"]))

    # ------------------------------------------------------------------ post-processing
    def compute_stress(self, u, rho_active):
        """Centre-point Voigt stress per active element, linear in rho
        (src/fem.py:564-579): rho_e D (B u_e), shape (n_active, 6), on the
        device (vox8_stress); numpy in, numpy out."""
        ud, back = self._to_dev(u)
        rg = self.grid_from_active(rho_active)
        out = torch.empty((self.geom.n_elements, 6), dtype=torch.float64, device=self.device)
        B = np.ascontiguousarray(self.B_centre, dtype=np.float64)
        D = np.ascontiguousarray(self.D_solid, dtype=np.float64)
        self._call("vox8_stress", ptr(ud.reshape(-1)), ptr(rg), _lib.f64p(B), _lib.f64p(D), ptr(out))
        return back(self.active_from_grid(out))

    @staticmethod
    def _invariants(stress):
        st, back = _stress_to_dev(stress)
        n = st.shape[0]
        vm = torch.empty(n, dtype=torch.float64, device=st.device)
        pm = torch.empty(n, dtype=torch.float64, device=st.device)
        call("vox8_stress_invariants", C.c_void_p(torch.cuda.current_stream(st.device).cuda_stream), n,
             ptr(st), ptr(vm), ptr(pm))
        return back(vm), back(pm)

    @staticmethod
    def von_mises(stress):
        """Von Mises equivalent stress of Voigt rows (src/fem.py:582-590), on
        the device (vox8_stress_invariants)."""
        return FEM3D._invariants(stress)[0]

    @staticmethod
    def principal_stress_max(stress):
        """Largest principal stress of Voigt rows (src/fem.py:593-602): closed
        form of the symmetric 3x3 eigenvalue problem, on the device."""
        return FEM3D._invariants(stress)[1]

    def total_mass(self, rho_active) -> float:
        V_e = self.geom.dx * self.geom.dy * self.geom.dz
        r = rho_active.sum().item() if isinstance(rho_active, torch.Tensor) else float(np.sum(rho_active))
        return float(r * V_e * self.mat.rho)

    def total_solid_mass(self) -> float:
        return self.total_mass(np.ones(self.n_active))
