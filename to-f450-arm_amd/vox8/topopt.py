"""Device-resident SIMP optimiser with the reference's surface (src/topopt.py:32-427).

Every per-element array of the loop (design variables, physical densities,
sensitivities, filtered fields) lives on the GPU in element-grid order
e = ex + nx (ey + ny ez); the host sees only scalars (compliance, volume,
change, eigenvalues, MAC) that drive the loop control.  Kernels:
cone filter, adjoint sensitivities, OC bisection and the reductions are
libvox8 entry points; the static solve is one batched PCG over all load cases
and the modal solve is LOBPCG warm-started from the previous iteration.  The
two solves depend only on the physical density, so the modal one runs in a
second vox8 context on its own stream and host thread, concurrently with the
static one: both spend much of their time in latency-bound coarse multigrid
levels that leave most of the GPU idle.  Every kernel is deterministic and
the two contexts share no state, so the results are bitwise those of running
the solves one after the other.
"""
from __future__ import annotations

import time
from concurrent.futures import Future, ThreadPoolExecutor

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr
from .fem import FEM3D, _tip_nodes_of
from .modal_participation import effective_mass_device, rank_by_effective_mass

import ctypes as C

__all__ = ["cone_kernel", "filter_field", "SIMPOptimizer"]


def cone_kernel(r_min_vox: float) -> np.ndarray:
    """w(d) = max(0, r - d) on the (2R+1)^3 box, R = ceil(r) (src/topopt.py:32-41)."""
    R = int(np.ceil(r_min_vox))
    c = np.arange(-R, R + 1, dtype=float)
    d = np.sqrt(c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2)
    return np.maximum(0.0, r_min_vox - d)


def _grid_to_elem(a):
    """(nx, ny, nz) C-order array -> element-grid order (x fastest)."""
    return np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0)).reshape(-1)


def _elem_to_grid(v, nx, ny, nz):
    return np.ascontiguousarray(np.asarray(v).reshape(nz, ny, nx).transpose(2, 1, 0))


class _GridContext:
    """A bare device context for a grid shape (the filter needs no element data)."""

    _cache = {}

    def __init__(self, nx, ny, nz):
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.ctx = _lib.create(nx, ny, nz, torch.cuda.current_stream().cuda_stream)

    def close(self):
        ctx, self.ctx = getattr(self, "ctx", None), None
        _lib.destroy(ctx)

    def __del__(self):
        if getattr(self, "ctx", None) is not None:
            _lib.release(self.ctx)
            self.ctx = None

    def _call(self, name, *args):
        call("vox8_set_stream", self.ctx, C.c_void_p(torch.cuda.current_stream().cuda_stream))
        return call(name, self.ctx, *args)

    @classmethod
    def get(cls, shape):
        if shape not in cls._cache:
            cls._cache[shape] = cls(*shape)
        return cls._cache[shape]


def filter_field(field_grid, mask_grid, kernel, fem: FEM3D = None):
    """Masked normalised cone filter on the GPU (src/topopt.py:44-58).

    Same signature as the reference; ``kernel`` must be a cone kernel (its
    radius is the centre weight).  Arrays are (nx, ny, nz) as in the reference.
    """
    k = np.asarray(kernel)
    r = float(k[tuple(s // 2 for s in k.shape)])
    nx, ny, nz = np.asarray(field_grid).shape
    g = fem if fem is not None else _GridContext.get((nx, ny, nz))
    dev = g.device
    mask = torch.as_tensor(_grid_to_elem(np.asarray(mask_grid, dtype=np.uint8)), device=dev)
    g._call("vox8_filter_setup", float(r), ptr(mask))
    f = torch.as_tensor(_grid_to_elem(np.asarray(field_grid, dtype=np.float64)), device=dev)
    out = torch.empty_like(f)
    g._call("vox8_filter", ptr(f), ptr(out))
    return _elem_to_grid(out.cpu().numpy(), nx, ny, nz)


class SIMPOptimizer:
    """src/topopt.py:64-427 on the device.

    Extra keyword arguments (defaults reproduce the reference):
    F_list (n_lc load vectors, instead of geometry load dicts), solve_rtol (None =
    the reference's effective 1e-5, src/fem.py:558-561), modal_tol / modal_maxiter
    (reference loop defaults 1e-4 / 500), solve_maxiter (5000), concurrent_solves
    (True: the modal solve overlaps the static one in a second context),
    device_tail (True: mode tracking, penalty, OC and bookkeeping in one device
    call with one host read; False: the same steps driven from the host),
    modal_guard (guard columns of the modal solve: 0..2 (default 0), or "auto":
    two guard columns or none, whichever has cost less per solve (LOBPCG
    iterations x block width).  The first solve runs without them and the
    second with; afterwards the setting not in use is tried for one solve,
    every 2 iterations at first and up to every GUARD_RETRY while the trials
    confirm the choice.  A trial is compared with the solve just before it
    (nearly the same design), and the choice follows the median of the last
    GUARD_SAMPLES such cost ratios.  They pay
    when the last tracked eigenvalues have close neighbours or the block is
    too small for the cluster it tracks, and cost a third more per iteration
    otherwise),
    effective_mass_direction (None,
    or 0/1/2: after each modal solve the tracked modes' effective-mass
    fractions in that translation direction, sims/modal_participation.py:37-57,
    are computed on the device and the history gains "eff_mass" (per mode, in
    frequency order) and "mode_rank" (mode indices by decreasing fraction)).
    The multigrid coarse levels follow ``fem.set_mg_galerkin`` in both
    contexts.
    """

    GUARD_RETRY = 16
    GUARD_SAMPLES = 5

    def __init__(self, geom, fem: FEM3D, load_cases, vol_frac=0.30, r_min_mm=2.0, move=0.2,
                 max_iter=100, tol_change=0.01, weights=None, min_freq_Hz=None, tip_mass_kg=0.060,
                 n_modes_track=4, mu_freq_init=0.10, mu_freq_max=50.0, mu_freq_ramp_every=5,
                 mu_freq_ramp_factor=1.5, F_list=None, solve_rtol=None, modal_tol=1e-4,
                 modal_maxiter=500, solve_maxiter=5000, concurrent_solves=True, device_tail=True,
                 modal_guard=0, effective_mass_direction=None):
        self.geom, self.fem, self.load_cases = geom, fem, load_cases
        self.vol_frac = vol_frac
        self.r_min_vox = max(1.0, r_min_mm / (geom.dx * 1e3))
        self.move, self.max_iter, self.tol_change = move, max_iter, tol_change
        self.min_freq_Hz, self.tip_mass_kg = min_freq_Hz, tip_mass_kg
        self.n_modes_track = max(1, int(n_modes_track))
        self.mu_freq, self.mu_freq_max = mu_freq_init, mu_freq_max
        self.mu_freq_ramp_every = max(1, int(mu_freq_ramp_every))
        self.mu_freq_ramp_factor = mu_freq_ramp_factor
        self.solve_rtol = 1e-5 if solve_rtol is None else solve_rtol
        self.solve_maxiter = solve_maxiter
        self.modal_tol, self.modal_maxiter = modal_tol, modal_maxiter
        self.concurrent_solves = bool(concurrent_solves)
        self.device_tail = bool(device_tail)
        if modal_guard != "auto" and int(modal_guard) not in (0, 1, 2):
            raise ValueError("modal_guard must be 'auto' or 0..2")
        self.modal_guard = modal_guard
        if effective_mass_direction not in (None, 0, 1, 2):
            raise ValueError("effective_mass_direction must be None, 0, 1 or 2")
        self.eff_dir = effective_mass_direction
        self._guard = 0 if modal_guard == "auto" else int(modal_guard)
        self._guard_set = None
        self._guard_prev = None        # "auto": (guard, cost) of the last modal solve
        self._guard_probing = False    # "auto": the next solve tries the other setting
        self._guard_next_probe = 1     # "auto": iteration after which to try it
        self._guard_interval = 2       # "auto": iterations between trials
        self._guard_ratios = []        # "auto": log cost ratios (guard / none) of recent trials
        self._widths = set()           # block widths the modal solve has run with
        self._fem_modal = None
        if F_list is None:
            F_list = []
            for lc in load_cases:
                F = np.zeros(geom.n_dofs)
                if lc.get("F") is not None:
                    d, m = geom.load_dofs_for_force(lc["F"])
                    F[d] += m
                if lc.get("Mz") is not None:
                    d, m = geom.load_dofs_for_moment_Mz(lc["Mz"])
                    F[d] += m
                F_list.append(F)
        n_lc = len(F_list)
        w = np.asarray(weights if weights is not None else [1.0 / n_lc] * n_lc, dtype=float)
        self.weights = w / w.sum()
        masks = geom.build_masks()
        self.in_arm, self.design, self.keep_solid = masks["in_arm"], masks["design"], masks["keep_solid"]
        dev = fem.device
        self._design = torch.as_tensor(_grid_to_elem(self.design.astype(np.uint8)), device=dev)
        self._in_arm = torch.as_tensor(_grid_to_elem(self.in_arm.astype(np.uint8)), device=dev)
        self.active_idx = fem.elem_grid_idx
        ex, ey, ez = self.active_idx.T
        self.is_design_active = self.design[ex, ey, ez]
        self.n_design = int(self.is_design_active.sum())
        self.kernel = cone_kernel(self.r_min_vox)
        self.design_mask_grid = self.design.astype(bool)
        self.F = torch.as_tensor(np.stack([np.asarray(f, dtype=np.float64) for f in F_list]),
                                 device=dev).contiguous()
        self.F_list = F_list
        self.V_e = geom.dx * geom.dy * geom.dz
        self.design_volume = self.n_design * self.V_e
        self.target_volume = vol_frac * self.design_volume
        self.timings = {}

    # -------------------------------------------------------------------------
    def _filter(self, src, dst):
        self.fem._call("vox8_filter", ptr(src), ptr(dst))

    def _stats(self, a, b=None, mask=None):
        out = (C.c_double * 4)()
        self.fem._call("vox8_stats", a.numel(), ptr(mask), ptr(a), ptr(b), out)
        return out[0], out[1], out[2], out[3]

    def start(self, rho_init=None):
        """Initialise the loop state (step() then advances one iteration)."""
        fem, geom = self.fem, self.geom
        dev = fem.device
        fem._call("vox8_filter_setup", float(self.r_min_vox), ptr(self._design))
        if rho_init is None:
            rho_active = np.where(self.is_design_active, self.vol_frac, 1.0)
        else:
            rho_active = np.asarray(rho_init, dtype=float)
        self.x = fem.grid_from_active(torch.as_tensor(rho_active, device=dev)).clone()
        self.phys_prev = self.x.clone()
        ne = geom.n_elements
        self._filt = torch.empty(ne, dtype=torch.float64, device=dev)
        self._phys = torch.empty(ne, dtype=torch.float64, device=dev)
        self._sens_f = torch.empty(ne, dtype=torch.float64, device=dev)
        self._dlam_f = torch.empty(ne, dtype=torch.float64, device=dev)
        self._dlam = torch.zeros(ne, dtype=torch.float64, device=dev)
        self._vprev = torch.zeros(geom.n_dofs, dtype=torch.float64, device=dev)
        self._xnew = torch.empty(ne, dtype=torch.float64, device=dev)
        self._check = torch.empty(ne, dtype=torch.float64, device=dev)
        self.U = None
        self.X = None
        self.v_prev = None
        self.modal = self.min_freq_Hz is not None
        if self.modal and self.concurrent_solves and self._fem_modal is None:
            # the modal solve runs in a second context on its own stream and
            # host thread, concurrently with the static solve (independent
            # given the density; their latency-bound coarse-grid phases leave
            # most of the GPU idle)
            self._modal_stream = torch.cuda.Stream(device=dev)
            with torch.cuda.stream(self._modal_stream):
                self._fem_modal = fem.twin()
            if fem.precision == "mixed":
                # both solves use one density per iteration: the modal context
                # copies the static context's multigrid operators (prepared
                # before the hand-over in step()) instead of building its own
                self._fem_modal.share_multigrid(fem)
            self._pool = ThreadPoolExecutor(max_workers=1)
            self._overlap = False
            # one pair of events for the whole run: the solves' hand-over
            # creates and destroys nothing while another thread may hold an
            # open graph capture
            self._ev_ready = torch.cuda.Event()
            self._ev_done = torch.cuda.Event()
        self.tip = _tip_nodes_of(geom) if self.modal else np.array([], dtype=np.int64)
        self.w_target = (2.0 * np.pi * self.min_freq_Hz) ** 2 if self.modal else None
        self.history = {k: [] for k in ("compliance", "mass", "change", "vol_frac", "omega_1_Hz",
                                        "mu_freq", "freq_violation", "pcg_iters", "lobpcg_iters",
                                        "time_s", "eff_mass", "mode_rank", "lobpcg_guard")}
        self.it = 0
        return self

    def step(self):
        """One SIMP iteration (src/topopt.py:267-412); returns the change."""
        fem = self.fem
        t0 = time.perf_counter()
        self.it += 1
        it = self.it
        self._filter(self.x, self._filt)
        fem._call("vox8_compose_density", ptr(self._design), ptr(self._in_arm), ptr(self._filt),
                  ptr(self._phys))
        rho_phys = fem.active_from_grid(self._phys)
        self.rho_phys = rho_phys
        modal_job = None
        if self.modal:
            mfem = self._fem_modal if self._fem_modal is not None else fem
            if self._guard_set != (id(mfem), self._guard):
                mfem.set_lobpcg_guard(self._guard)
                self._guard_set = (id(mfem), self._guard)
        if self.modal and self._fem_modal is not None:
            # Stream hand-over without record_stream: the modal stream waits
            # for everything this thread queued so far (the previous tail's
            # reads of the modes included), and this thread's stream waits
            # for the modal solve before reading its outputs.  Tensors of
            # either stream are then safe to free from any thread, and a free
            # never records an event on another thread's stream (which could
            # land inside a graph capture that thread has open).
            ready, done = self._ev_ready, self._ev_done
            if self._fem_modal._mg_src is fem:
                fem.prepare_multigrid(rho_phys, ncols=self.F.shape[0])
            ready.record()
            X0 = self.X

            def run_modal():
                torch.cuda.set_device(fem.device)
                with torch.cuda.stream(self._modal_stream):
                    self._modal_stream.wait_event(ready)
                    res = self._fem_modal.modal_block(rho_phys, self.n_modes_track, self.tip_mass_kg,
                                                      self.tip, X0=X0, tol=self.modal_tol,
                                                      maxiter=self.modal_maxiter)
                    if self.eff_dir is not None:
                        res = res + (effective_mass_device(self._fem_modal, res[1], self.eff_dir)[0],)
                    done.record()
                return res

            if self._overlap:
                modal_job = self._pool.submit(run_modal)
            else:
                # first step: the solves run one after the other, so the
                # stencil launch-shape autotuner times each new grid and
                # width on an otherwise idle GPU; later steps overlap
                modal_job = Future()
                modal_job.set_result(run_modal())
                self._overlap = True
        self.U, pinfo = fem.solve_block(rho_phys, self.F, x0=self.U, rtol=self.solve_rtol,
                                        maxiter=self.solve_maxiter)
        s = fem.sensitivities(rho_phys, U=self.U, weights=self.weights)
        c_total = float(np.dot(self.weights, s["comp"]))
        self._filter(s["dc"], self._sens_f)
        lob_iters = 0
        lam = X = None
        if self.modal:
            em = None
            if modal_job is not None:
                res = modal_job.result()
                lam, X, linfo = res[:3]
                em = res[3] if len(res) > 3 else None
                torch.cuda.current_stream().wait_event(self._ev_done)
            else:
                lam, X, linfo = fem.modal_block(rho_phys, self.n_modes_track, self.tip_mass_kg,
                                                self.tip, X0=self.X, tol=self.modal_tol,
                                                maxiter=self.modal_maxiter)
                if self.eff_dir is not None:
                    em = effective_mass_device(fem, X, self.eff_dir)[0]
            lob_iters = linfo["iters"]
            self.history["lobpcg_guard"].append(self._guard)
            if self.modal_guard == "auto":
                self._auto_guard(lob_iters, it)
            order = np.argsort(lam)
            lam = np.maximum(lam[order], 0.0)
            if not np.array_equal(order, np.arange(len(order))):
                X = X[torch.as_tensor(order, device=fem.device)].contiguous()
            self.X = X
            if em is not None:
                em = em[order]
                self.history["eff_mass"].append([float(v) for v in em])
                self.history["mode_rank"].append([int(v) for v in rank_by_effective_mass(em)])
        if self.device_tail:
            change, vol = self._device_tail(lam, X, it)
        else:
            change, vol = self._host_tail(lam, X, rho_phys, it)
        self.x, self._xnew = self._xnew, self.x
        self.phys_prev, self._check = self._check, self.phys_prev
        m = vol * self.V_e
        h = self.history
        h["compliance"].append(c_total)
        h["mass"].append(m)
        h["change"].append(change)
        h["vol_frac"].append(m / self.design_volume)
        h["pcg_iters"].append(list(pinfo["iters"]))
        h["lobpcg_iters"].append(lob_iters)
        h["time_s"].append(time.perf_counter() - t0)
        return change

    def _auto_guard(self, lob_iters, it):
        """The "auto" guard policy (class docstring): after a trial of the
        other setting, record the cost ratio of the two adjacent solves and
        choose the setting the recorded ratios favour."""
        g, k = self._guard, self.n_modes_track
        if k + 2 > 12:
            return                                  # no room for guard columns
        cost = lob_iters * (k + g) / k
        prev = self._guard_prev
        self._guard_prev = (g, cost)
        self._widths.add(k + g)
        nxt = g
        if self._guard_probing:
            # the trial ran right after a solve with the other setting:
            # adjacent designs, so the two iteration counts compare
            self._guard_probing = False
            c2, c0 = (cost, prev[1]) if g == 2 else (prev[1], cost)
            self._guard_ratios.append(np.log(c2 / c0))
            del self._guard_ratios[:-self.GUARD_SAMPLES]
            # the median of the recent ratios: one trial next to a slow
            # solve (a mode swap) does not flip the choice
            nxt = 2 if np.median(self._guard_ratios) < 0 else 0
            same = nxt == prev[0]
            self._guard_interval = min(2 * self._guard_interval, self.GUARD_RETRY) if same else 2
            self._guard_next_probe = it + self._guard_interval
        elif it >= self._guard_next_probe:
            nxt = 2 - g
            self._guard_probing = True
        if nxt != g:
            self._guard = nxt
            if k + nxt not in self._widths:
                # a block width not run before: that solve runs alone, so the
                # launch shapes of the width are tuned on an idle GPU
                self._overlap = False

    def close(self, fem: bool = False):
        """Release the modal context, its worker thread and stream, in that
        order (the thread is joined before its context is destroyed).  Streams
        share a few hardware queues per process, so an optimiser kept alive
        after its run can take the queues another one needs to overlap its
        solves.  The optimiser keeps working afterwards, with the two solves
        one after the other in the main context; ``fem=True`` also closes the
        main FEM3D context (the optimiser is then unusable)."""
        pool = getattr(self, "_pool", None)
        if pool is not None:
            pool.shutdown(wait=True)
        self._pool = None
        mf = getattr(self, "_fem_modal", None)
        if mf is not None:
            torch.cuda.synchronize(mf.device)
            mf.close()
        self._fem_modal = None
        self._modal_stream = None
        if fem:
            self.fem.close()

    def _track(self, lam1, viol, it):
        """History and mu ramp of the frequency constraint (src/topopt.py:363-373)."""
        self.history["omega_1_Hz"].append(float(np.sqrt(lam1) / (2.0 * np.pi)))
        self.history["mu_freq"].append(self.mu_freq)
        self.history["freq_violation"].append(viol)
        if it % self.mu_freq_ramp_every == 0 and viol > 0.0 and self.mu_freq < self.mu_freq_max:
            self.mu_freq = min(self.mu_freq_max, self.mu_freq * self.mu_freq_ramp_factor)

    def _device_tail(self, lam, X, it):
        """Mode tracking, eigenvalue-sensitivity penalty, OC update, filter and
        bookkeeping in one device call (vox8_simp_tail): one host read per
        iteration.  Returns (change, design volume sum)."""
        fem = self.fem
        a = _lib.TailArgs()
        lam_c = None
        if self.modal:
            k = X.shape[0]
            lam_c = (C.c_double * k)(*[float(v) for v in lam])
            fem._set_tip(self.tip_mass_kg, self.tip)
            a.d_X, a.ldx, a.h_lam = ptr(X), X.shape[1], C.cast(lam_c, C.POINTER(C.c_double))
            a.d_vprev, a.have_prev = ptr(self._vprev), int(self.v_prev is not None)
            a.w_target, a.mu = self.w_target, self.mu_freq
            a.d_rho, a.d_dlam, a.d_dlam_f = ptr(fem._rho_grid), ptr(self._dlam), ptr(self._dlam_f)
            a.p, a.e_min = fem.SIMP_PENALTY, fem.E_MIN_FRAC
            a.mthr, a.mq = fem.MASS_THRESHOLD, fem.MASS_PENALTY_Q
            a.k = k
        a.d_sens_f, a.d_design, a.d_x = ptr(self._sens_f), ptr(self._design), ptr(self.x)
        a.d_xnew, a.d_check, a.d_phys_prev = ptr(self._xnew), ptr(self._check), ptr(self.phys_prev)
        a.V_e, a.target_volume, a.move = self.V_e, self.target_volume, self.move
        out = (C.c_double * _lib.TAIL_STATE)()
        fem._call("vox8_simp_tail", C.byref(a), out)
        if self.modal:
            self.v_prev = self._vprev
            self._track(out[1], out[2], it)
        else:
            for key in ("omega_1_Hz", "mu_freq", "freq_violation"):
                self.history[key].append(float("nan"))
        return out[13], out[15]

    def _host_tail(self, lam, X, rho_phys, it):
        """The same steps driven from the host (one read-back per reduction),
        the parity reference of _device_tail."""
        fem = self.fem
        if self.modal:
            if self.v_prev is None:
                k = 0
            else:
                k_modes = X.shape[0]
                d = (C.c_double * (2 * k_modes + 1))()
                fem._call("vox8_block_dots", X.shape[1], k_modes, ptr(X), X.shape[1],
                          ptr(self.v_prev), d)
                num = np.array(d[:k_modes])
                a2 = np.array(d[k_modes:2 * k_modes])
                mac = num ** 2 / np.maximum(a2 * d[2 * k_modes], 1e-24)
                k = int(np.argmax(mac))
            lam1 = float(lam[k])
            v1 = X[k].clone()
            self.v_prev = v1
            viol = max(0.0, 1.0 - lam1 / self.w_target)
            if viol > 0.0:
                fem._set_tip(self.tip_mass_kg, self.tip)
                sm = fem.sensitivities(rho_phys, phi=v1, lam=lam1)
                self._filter(sm["dlam"], self._dlam_f)
                _, sc_abs, _, cnt = self._stats(self._sens_f, mask=self._design)
                _, sm_abs, _, _ = self._stats(self._dlam_f, mask=self._design)
                s_c, s_m = sc_abs / cnt, sm_abs / cnt
                a = -self.mu_freq * viol * (s_c / max(s_m, 1e-30))
                fem._call("vox8_axpby", self._sens_f.numel(), None, a, ptr(self._dlam_f), 1.0,
                          ptr(self._sens_f))
            self._track(lam1, viol, it)
        else:
            for key in ("omega_1_Hz", "mu_freq", "freq_violation"):
                self.history[key].append(float("nan"))
        fem._call("vox8_oc_update", ptr(self._design), ptr(self.x), ptr(self._sens_f), self.V_e,
                  self.target_volume, self.move, ptr(self._xnew), None)
        self._filter(self._xnew, self._check)
        _, _, change, _ = self._stats(self._check, self.phys_prev)
        vol, _, _, _ = self._stats(self._xnew, mask=self._design)
        return change, vol

    # ------------------------------------------------------------------ checkpoint
    def state(self):
        """The loop state after the last step, as host arrays: design
        variables and the previous filtered design, the warm starts of both
        solves (U, X), the tracked mode, mu, the iteration count and the
        guard setting.  The reference keeps the same warm starts in memory
        (src/topopt.py:237, 246, 288-290, 323); ``load_state`` resumes from it,
        in this optimiser or another one on the same problem (e.g. the plain
        FP64 path from the multigrid path's state)."""
        st = {"x": self.x.cpu().numpy(), "phys_prev": self.phys_prev.cpu().numpy(),
              "it": np.int64(self.it), "mu_freq": np.float64(self.mu_freq),
              "guard": np.int64(self._guard)}
        if self.U is not None:
            st["U"] = self.U.cpu().numpy()
        if self.X is not None:
            st["X"] = self.X.cpu().numpy()
        if self.v_prev is not None:
            st["v_prev"] = self.v_prev.cpu().numpy()
        return st

    def load_state(self, st, overlap=False):
        """Resume from ``state()`` (or ``np.load`` of ``save``); call after
        start().  overlap=True lets the next step overlap the two solves at
        once (otherwise, as after start(), the first step runs them one after
        the other so the multigrid launch shapes are tuned on an idle GPU)."""
        dev = self.fem.device
        if overlap and getattr(self, "_fem_modal", None) is not None:
            self._overlap = True
        self.x = torch.as_tensor(np.asarray(st["x"]), device=dev).clone()
        self.phys_prev = torch.as_tensor(np.asarray(st["phys_prev"]), device=dev).clone()
        self.it = int(st["it"])
        self.mu_freq = float(st["mu_freq"])
        if self.modal_guard == "auto":
            self._guard = int(st["guard"])
        self.U = torch.as_tensor(np.asarray(st["U"]), device=dev).clone() if "U" in st else None
        self.X = torch.as_tensor(np.asarray(st["X"]), device=dev).clone() if "X" in st else None
        if "v_prev" in st:
            self._vprev.copy_(torch.as_tensor(np.asarray(st["v_prev"]), device=dev))
            self.v_prev = self._vprev
        else:
            self.v_prev = None
        return self

    def save(self, path):
        """Checkpoint the loop state to an .npz file (no pickled objects)."""
        np.savez(path, **self.state())

    def load(self, path):
        """Resume from a checkpoint written by ``save``."""
        with np.load(path, allow_pickle=False) as f:
            return self.load_state({k: f[k] for k in f.files})

    def result(self):
        fem, g = self.fem, self.geom
        self._filter(self.x, self._filt)
        fem._call("vox8_compose_density", ptr(self._design), ptr(self._in_arm), ptr(self._filt),
                  ptr(self._phys))
        full = self._phys.cpu().numpy()
        rho_active = full[fem.active_elems]
        u = [] if self.U is None else [self.U[i].cpu().numpy() for i in range(self.U.shape[0])]
        return {"rho_active": rho_active, "rho_grid": _elem_to_grid(full, g.nx, g.ny, g.nz),
                "history": self.history, "u_per_lc": u}

    def run(self, verbose: bool = True, rho_init=None):
        self.start(rho_init)
        for it in range(1, self.max_iter + 1):
            change = self.step()
            if verbose:
                h = self.history
                extra = f"  w1={h['omega_1_Hz'][-1]:6.1f}Hz  mu={self.mu_freq:.2e}" if self.modal else ""
                print(f"  iter {it:3d}  c={h['compliance'][-1]:.4e}  Vf={h['vol_frac'][-1]:.3f}  "
                      f"change={change:.4f}  [{h['time_s'][-1]:.3f}s]{extra}")
            if change < self.tol_change and it > 10:
                break
        return self.result()
