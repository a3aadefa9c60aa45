"""Benchmark: SIMP iterations/s of the voxel hex8 FEM hot path on one MI355X.

One step = one full SIMP iteration of the workload (default c5t: 440x64x48
voxels at 0.5 mm, 4,213,755 DOF, 5 load cases, 6 modes): density filter, the
5-load-case static solve as one block PCG (rtol 1e-10, warm-started),
compliance adjoint, 6-mode LOBPCG (tol 1e-6, warm-started), mode-1
eigenvalue adjoint (the frequency target is above f1, so the constraint is
active every iteration), sensitivity filters, OC update.  Inputs are resident
in HBM before the timed region.

Prints one JSON line.  Extra objects:
  roofline      the dominant kernel (FP32 Chebyshev smoothing step of the
                multigrid preconditioner on the fine grid, three-term form on
                the iterates, 6 columns), timed
                with HIP events on the solver stream; algorithmic bytes per
                launch in "alg_bytes_per_launch"; "traffic" from the committed
                rocprofv3 FETCH_SIZE/WRITE_SIZE summary when present
  kernels       the same measurement for the other hot kernels
  solvers       the static (PCG) and modal (LOBPCG) solves alone on the final
                design, cold-started: wall time, iterations, PCG iterations/s
  plain_fp64    one SIMP iteration of the same loop on the plain FP64 path
                (the reference's Jacobi preconditioner in PCG and LOBPCG),
                resumed from the multigrid path's loop state at the start of
                the timed window, beside the multigrid path's own time for
                that iteration
  secondary     the other workloads given with --also (default c2, c3, c5 --
                12 modes ranked by z effective mass, at most 5 timed steps --
                and c4, the 16 batched designs run concurrently and one by
                one, at most 5 timed steps, in a child process)

Other modes: --trajectory N (per-iteration times of both paths over N SIMP
iterations), --kernels-only (kernel timings for counter runs), --c4-only.
  cpu_baseline  the SciPy oracle on this host (bounded z-slab sample of the
                same grid, scaled by DOF count and by the plain path's
                iteration counts, i.e. the reference's algorithm)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "to-f450-arm_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

HBM_PEAK_GBS = 8000.0
DOMINANT = "cheb_x_fp32"


def make_optimizer(wl_name, precision, modal_maxiter=2000, galerkin=True, modal_guard="auto"):
    from vox8.fem import FEM3D, Material
    from vox8.topopt import SIMPOptimizer
    from vox8.workloads import E_PA, NU, RHO_MAT, WORKLOADS, load_cases, seeded_density

    wl = WORKLOADS[wl_name]
    geom = wl.geometry()
    fem = FEM3D(geom, Material(E=E_PA, nu=NU, rho=RHO_MAT), precision=precision)
    if not galerkin:
        fem.set_mg_galerkin(False)
    F = load_cases(geom, wl.loads)
    x0 = seeded_density(geom, wl.base, wl.amp, wl.seed)
    # frequency target above the initial f1 keeps the eigen-adjoint active
    opt = SIMPOptimizer(geom, fem, load_cases=[{}] * len(F), vol_frac=wl.base,
                        r_min_mm=wl.r_min_vox * geom.dx * 1e3, max_iter=10 ** 6,
                        min_freq_Hz=1e4, tip_mass_kg=0.0, n_modes_track=wl.n_modes,
                        F_list=list(F), solve_rtol=wl.pcg_rtol, modal_tol=wl.modal_tol,
                        modal_maxiter=modal_maxiter, solve_maxiter=40000,
                        effective_mass_direction=wl.eff_dir, modal_guard=modal_guard)
    opt.start(rho_init=x0)
    return wl, geom, fem, opt, F, x0


def time_steps(opt, warmup, steps, snapshot=None):
    """Time `steps` SIMP iterations after `warmup` untimed ones; with a
    `snapshot` dict, the loop state at the start of the timed window is
    stored in it (host copies, for the plain-path comparison)."""
    import torch
    for _ in range(warmup):
        opt.step()
        print(f"[bench] warmup step {opt.it} done", file=sys.stderr, flush=True)
    if snapshot is not None:
        snapshot.update(opt.state())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        opt.step()
        print(f"[bench] step {opt.it} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _event_time(fem, fn, reps):
    """Average time of fn() on the solver's stream (HIP events on that stream)."""
    import torch
    for _ in range(3):
        fn()
    st = torch.cuda.current_stream()          # fem._call binds the library to this stream
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def kernel_rooflines(fem, geom, rho, nrhs=6, reps=30):
    """HIP-event timings of the hot kernels at the workload's size and density."""
    import torch
    from vox8 import _lib
    n, ne, nn = geom.n_dofs, geom.n_elements, geom.n_nodes
    fem._set_density(rho)
    out = {}
    # The FP32 launch shape is tuned when a multigrid cycle of this width is
    # first captured; run one so the timings below use the tuned shape.
    if fem.precision == "mixed":
        rz = torch.zeros(nrhs, n, dtype=torch.float64, device=fem.device)
        fem._call("vox8_precond_apply", _lib.ptr(rz), _lib.ptr(rz), nrhs)
    # FP32 Chebyshev smoothing step of the multigrid's fine level (three-term
    # form on the iterates), nrhs columns
    d = torch.randn(nrhs, n, dtype=torch.float32, device=fem.device)
    r = torch.randn_like(d)
    xs = torch.randn_like(d)
    dn = torch.empty_like(d)
    t = _event_time(fem, lambda: fem._call("vox8_cheb_xstep", _lib.ptr(d), _lib.ptr(r), _lib.ptr(xs),
                                           _lib.ptr(dn), 0.5, 1e-3, nrhs), reps)
    alg = nrhs * 16 * n + 4 * n + 4 * ne + nn
    out["cheb_x_fp32"] = dict(kernel=f"k_apply_fast<float,CUBE,K,EPI_X> (FP32 Chebyshev step, {nrhs} columns)",
                              t=t, alg=alg, per_unit="16 B per DOF and column (read x_k, x_k-1, b; write "
                                                     "x_k+1), + 4 B/DOF D^-1, 4 B/element factor, 1 B/node mask")
    # the same step in the residual-direction form (vox8_cheb_step)
    t = _event_time(fem, lambda: fem._call("vox8_cheb_step", _lib.ptr(d), _lib.ptr(r), _lib.ptr(xs),
                                           _lib.ptr(dn), 0.5, 1e-3, nrhs), reps)
    alg = nrhs * 24 * n + 4 * n + 4 * ne + nn
    out["cheb_fp32"] = dict(kernel=f"k_apply_fast<float,CUBE,K,EPI_CHEB> (FP32 Chebyshev step, residual "
                                   f"form, {nrhs} columns)",
                            t=t, alg=alg, per_unit="24 B per DOF and column (read d, r, xs; write r, xs, d'), "
                                                   "+ 4 B/DOF D^-1, 4 B/element factor, 1 B/node mask")
    # FP64 fused K and M products (LOBPCG), nrhs columns
    x = torch.randn(nrhs, n, dtype=torch.float64, device=fem.device)
    yk, ym = torch.empty_like(x), torch.empty_like(x)
    t = _event_time(fem, lambda: fem._call("vox8_apply", _lib.OP_KM, _lib.PREC_FP64, _lib.ptr(x),
                                           _lib.ptr(yk), _lib.ptr(ym), nrhs, n), reps)
    alg = nrhs * 24 * n + 16 * ne + nn
    out["km_fp64"] = dict(kernel=f"k_apply_fast<double,CUBE,K+M> (FP64 K and M products, {nrhs} columns)",
                          t=t, alg=alg, per_unit="24 B per DOF and column (read x; write Kx, Mx), "
                                                 "+ 16 B/element factors, 1 B/node mask")
    # FP64 K product, one column (PCG matvec shape)
    x1, y1 = x[0].contiguous(), yk[0].contiguous()
    t = _event_time(fem, lambda: fem._call("vox8_apply", _lib.OP_K, _lib.PREC_FP64, _lib.ptr(x1),
                                           _lib.ptr(y1), None, 1, n), reps)
    alg = 16 * n + 8 * ne + nn
    out["k_fp64"] = dict(kernel="k_apply_fast<double,CUBE,K> (FP64 K product, 1 column)", t=t, alg=alg,
                         per_unit="16 B per DOF (read x; write Kx), + 8 B/element factor, 1 B/node mask")
    # FP64 K product, 5 columns: the block PCG's matvec width over the five
    # load cases (this launch is the unfused form; the PCG's own also sums
    # p.Kp per workgroup, at the same z chunk)
    x5, y5 = x[:5].contiguous(), yk[:5].contiguous()
    t = _event_time(fem, lambda: fem._call("vox8_apply", _lib.OP_K, _lib.PREC_FP64, _lib.ptr(x5),
                                           _lib.ptr(y5), None, 5, n), reps)
    alg = 5 * 16 * n + 8 * ne + nn
    out["k_fp64_pcg5"] = dict(kernel="k_apply_fast<double,CUBE,K> (FP64 K product, 5 columns: the PCG shape)",
                              t=t, alg=alg, per_unit="16 B per DOF and column (read x; write Kx), "
                                                     "+ 8 B/element factor, 1 B/node mask")
    # FP32 K product, nrhs columns: the structured stencil against the MFMA
    # trial of the 24x24 element contraction (same inputs, same output)
    x32, y32 = d, dn
    t = _event_time(fem, lambda: fem._call("vox8_apply", _lib.OP_K, _lib.PREC_FP32, _lib.ptr(x32),
                                           _lib.ptr(y32), None, nrhs, n), reps)
    alg = nrhs * 8 * n + 4 * ne + nn
    out["k_fp32"] = dict(kernel=f"k_apply_fast<float,CUBE,K> (FP32 K product, {nrhs} columns)", t=t,
                         alg=alg, per_unit="8 B per DOF and column (read x; write Kx), "
                                           "+ 4 B/element factor, 1 B/node mask")
    t = _event_time(fem, lambda: fem._call("vox8_trial_mfma_apply", _lib.ptr(x32), _lib.ptr(y32), nrhs),
                    reps)
    out["k_fp32_mfma_trial"] = dict(kernel=f"k_apply_mfma_trial (FP32 K product on MFMA, {nrhs} columns)",
                                    t=t, alg=alg, per_unit="as k_fp32")
    # attainable streaming rate on this part: a 1 GiB device-to-device copy
    # (read + write bytes), timed the same way
    src = torch.ones(1 << 27, dtype=torch.float64, device=fem.device)
    dst = torch.empty_like(src)
    tcopy = _event_time(fem, lambda: dst.copy_(src), 5)
    copy_gbs = 2 * src.numel() * 8 / tcopy / 1e9
    del src, dst
    res = {}
    for k, v in out.items():
        ach = v["alg"] / v["t"] / 1e9
        res[k] = {"kernel": v["kernel"], "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": load_traffic(k, geom),
                  "avg_launch_us": round(v["t"] * 1e6, 2), "alg_bytes_per_launch": int(v["alg"]),
                  "alg_bytes_per_unit": v["per_unit"], "copy_GBs": round(copy_gbs, 1),
                  "frac_of_copy": round(ach / copy_gbs, 4)}
    return res


def summarize_trace(path, top=40, name_filter=""):
    """Per-(kernel, grid) table of a rocprofv3 kernel trace (CSV): rocprofv3's
    --stats groups by kernel name, which merges a kernel's launches on every
    multigrid level and at every tuning candidate; split by grid size, one
    kernel at one problem size can be compared with the HIP-event timings."""
    import csv
    from collections import defaultdict
    agg = defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"]
            if name_filter and name_filter not in name:
                continue
            grid = r.get("Grid_Size") or r.get("Grid_Size_X") or ""
            agg[(name, grid)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    total = sum(sum(v) for v in agg.values())
    rows = sorted(agg.items(), key=lambda kv: -sum(kv[1]))
    out = [f"Total GPU kernel time: {total / 1e3:.1f} ms over {sum(len(v) for v in agg.values())} launches.",
           "", "| total ms | calls | mean us | min us | max us | grid | kernel |",
           "|---:|---:|---:|---:|---:|---:|---|"]
    for (name, grid), t in rows[:top]:
        short = name.split("(")[0].replace("|", "\\|")
        out.append(f"| {sum(t) / 1e3:.2f} | {len(t)} | {sum(t) / len(t):.1f} | {min(t):.1f} | {max(t):.1f} "
                   f"| {grid} | `{short}` |")
    return "\n".join(out) + "\n"


def load_traffic(name, geom):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if it
    covers this kernel at this grid."""
    p = ROOT / "profiles" / "pmc_traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text()).get(f"{name}@{geom.nx}x{geom.ny}x{geom.nz}")
        return None if d is None else d.get("traffic_bytes")
    except (ValueError, OSError):
        return None


def cpu_baseline(wl, geom, pcg_iters, lob_iters, basis, slab_nz=6, budget_s=25.0):
    """SciPy oracle on this host, timed on a z-slab of the workload's grid
    (same voxel size, all nx x ny, slab_nz layers), scaled by the DOF ratio."""
    import warnings
    import numpy as np
    import scipy.sparse.linalg as spla
    from oracle.fem import FEM3D as OFEM, Material as OMat
    from oracle.topopt import cone_kernel, filter_field, oc_update
    from vox8.workloads import E_PA, NU, RHO_MAT, Cantilever, load_cases, seeded_density

    cores = len(os.sched_getaffinity(0))
    nz = min(slab_nz, geom.nz)
    sg = Cantilever(geom.nx, geom.ny, nz, geom.dx)
    scale = geom.n_dofs / sg.n_dofs
    of = OFEM(sg, OMat(E=E_PA, nu=NU, rho=RHO_MAT))
    x0 = seeded_density(sg, wl.base, wl.amp, wl.seed)
    F = load_cases(sg, wl.loads)
    g3 = lambda v: v.reshape(sg.nz, sg.ny, sg.nx).transpose(2, 1, 0)
    mask = np.ones((sg.nx, sg.ny, sg.nz), dtype=bool)
    ker = cone_kernel(wl.r_min_vox)
    t = time.perf_counter()
    phys = filter_field(g3(x0), mask, ker)
    t_filter = time.perf_counter() - t
    rho = phys.transpose(2, 1, 0).reshape(-1)
    t = time.perf_counter()
    K = of.assemble_K(rho)
    M = of.assemble_M(rho)
    free = of.free_dofs
    Kff = K[free][:, free].tocsr()
    Mff = M[free][:, free].tocsr()
    t_asm = time.perf_counter() - t
    d = Kff.diagonal()
    Minv = spla.LinearOperator(Kff.shape, matvec=lambda v: v / d, dtype=float)
    n_cg = 60
    t = time.perf_counter()
    spla.cg(Kff, F[0][free], M=Minv, rtol=1e-30, maxiter=n_cg)
    t_cg = (time.perf_counter() - t) / n_cg
    X0 = np.random.default_rng(0).standard_normal((len(free), wl.n_modes))
    Tinv = spla.LinearOperator(Kff.shape, matvec=lambda v: v / d, matmat=lambda V: V / d[:, None],
                               dtype=float)
    n_lob = 10
    t = time.perf_counter()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        spla.lobpcg(Kff, X0, B=Mff, M=Tinv, tol=1e-30, largest=False, maxiter=n_lob)
    t_lob = (time.perf_counter() - t) / n_lob
    u = np.random.default_rng(1).standard_normal(sg.n_dofs)
    t = time.perf_counter()
    ce = of.elem_compliance(u)
    of.eigenvalue_sensitivity(rho, 1.0, u, tip_mass_kg=0.0)
    t_sens = time.perf_counter() - t
    t = time.perf_counter()
    oc_update(x0, -np.abs(ce) - 1.0, sg.dx ** 3, wl.base * x0.size * sg.dx ** 3, 0.2)
    t_oc = time.perf_counter() - t
    n_lc = len(wl.loads)
    per_iter = scale * (2 * t_asm + sum(pcg_iters) * t_cg + lob_iters * t_lob + n_lc * t_sens
                        + 4 * t_filter + t_oc)
    return {"value": round(1.0 / per_iter, 8), "unit": "SIMP iters/s", "cores": cores, "kind": "port",
            "sample": (f"SciPy oracle (reference algorithm: Jacobi CG, Jacobi LOBPCG) on a "
                       f"{sg.nx}x{sg.ny}x{sg.nz} z-slab of the {wl.name} grid ({sg.n_dofs} DOF; "
                       f"times x{scale:.2f} for {geom.n_dofs} DOF): COO->CSR assembly {t_asm:.2f}s, "
                       f"{n_cg} CG iters ({t_cg * 1e3:.1f} ms/iter), {n_lob} LOBPCG iters "
                       f"({t_lob * 1e3:.1f} ms/iter), einsum sensitivities {t_sens:.2f}s, ndimage filter "
                       f"{t_filter:.2f}s x4, OC {t_oc:.2f}s; per SIMP iteration "
                       f"{sum(pcg_iters)} CG iters over {n_lc} load cases and {lob_iters} LOBPCG iters "
                       f"({basis})"),
            "seconds_per_iter": round(per_iter, 2)}


def run_workload(name, precision, steps, warmup, snapshot=None, galerkin=True, modal_guard="auto"):
    import numpy as np
    wl, geom, fem, opt, F, x0 = make_optimizer(name, precision, galerkin=galerkin,
                                               modal_guard=modal_guard)
    print(f"[bench] {wl.name}: {geom.nx}x{geom.ny}x{geom.nz}, {geom.n_dofs} DOF, {F.shape[0]} LC, "
          f"{wl.n_modes} modes ({precision}); warmup {warmup}", file=sys.stderr, flush=True)
    dt = time_steps(opt, warmup, steps, snapshot)
    h = opt.history
    pcg_it = float(np.mean([np.mean(v) for v in h["pcg_iters"][-steps:]]))
    lob_it = float(np.mean(h["lobpcg_iters"][-steps:]))
    print(f"[bench] {wl.name}: {steps} steps in {dt:.3f}s; pcg {pcg_it:.0f} it/LC, lobpcg {lob_it:.0f} it",
          file=sys.stderr, flush=True)
    return wl, geom, fem, opt, dt, pcg_it, lob_it


def solver_timings(opt):
    """The static and modal solves alone (no overlap), cold-started (U = 0, a
    random LOBPCG block) on the optimiser's final design: wall time,
    iterations, and PCG iterations/s (one block iteration advances every load
    case, so column iterations/s = iterations/s x load cases).  Inside the
    SIMP loop both solves are warm-started from the previous iteration's
    solutions; their mean iteration counts are in the config."""
    import torch
    fem, rho = opt.fem, opt.rho_phys
    mfem = opt._fem_modal if opt._fem_modal is not None else fem
    mst = opt._modal_stream if opt._fem_modal is not None else torch.cuda.current_stream()

    def clock(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    dt, (_, info) = clock(lambda: fem.solve_block(rho, opt.F, rtol=opt.solve_rtol,
                                                  maxiter=opt.solve_maxiter))
    it, nlc = max(info["iters"]), opt.F.shape[0]
    out = {"pcg": {"s": round(dt, 4), "iters": it, "iters_per_s": round(it / dt, 1),
                   "column_iters_per_s": round(it * nlc / dt, 1), "load_cases": nlc,
                   "rtol": opt.solve_rtol, "start": "cold"}}
    if opt.modal:
        with torch.cuda.stream(mst):
            dt, (_, _, info) = clock(lambda: mfem.modal_block(
                rho, opt.n_modes_track, opt.tip_mass_kg, opt.tip, tol=opt.modal_tol,
                maxiter=opt.modal_maxiter))
        out["lobpcg"] = {"s": round(dt, 4), "iters": info["iters"], "modes": opt.n_modes_track,
                         "tol": opt.modal_tol, "start": "cold"}
    return out


def trajectory(name, precision, n_iter, modal_maxiter):
    """n_iter SIMP iterations from the seeded density, each timed on its own
    (the first includes the launch-shape tuning of the multigrid path)."""
    import torch
    wl, geom, fem, opt, F, x0 = make_optimizer(name, precision, modal_maxiter=modal_maxiter,
                                               modal_guard="auto" if precision == "mixed" else 0)
    secs = []
    for i in range(n_iter):
        secs.append(time_steps(opt, 0, 1))
        h = opt.history
        print(f"[bench] {name} {precision} iteration {i + 1}: {secs[-1]:.3f}s, pcg {h['pcg_iters'][-1]}, "
              f"lobpcg {h['lobpcg_iters'][-1]}, f1 {h['omega_1_Hz'][-1]:.6f} Hz", file=sys.stderr, flush=True)
    h = opt.history
    out = {"workload": name, "precision": precision, "seconds": [round(t, 4) for t in secs],
           "pcg_iters": h["pcg_iters"], "lobpcg_iters": h["lobpcg_iters"],
           "lobpcg_maxiter": modal_maxiter, "compliance": h["compliance"], "omega_1_Hz": h["omega_1_Hz"]}
    _release(opt, fem)
    del opt, fem
    torch.cuda.synchronize()
    return out


def _release(*objs):
    """Free the device state of finished optimizers (their contexts, streams
    and worker threads) before the next workload runs."""
    import gc
    import torch
    for o in objs:
        close = getattr(o, "close", None)
        if close is not None:
            close()
    gc.collect()
    torch.cuda.synchronize()


C4_MAX_STEPS, C4_MAX_WARMUP = 5, 1
# secondary workloads measured on fewer iterations than the headline one (steps, warmup)
SECONDARY_CAPS = {"c5": (5, 2)}


def bench_c4(precision, steps, warmup):
    """The 16 uneven designs: concurrent (one context, stream and thread per
    design) against one after another; design-iterations per second.  Each
    phase closes its optimisers (worker threads joined, then contexts
    destroyed) before the next starts."""
    from vox8.batch import close_designs, make_designs, step_designs
    from vox8.workloads import c4_designs
    import torch
    out = {"unit": "design SIMP iters/s", "designs": 16, "steps": steps, "warmup": warmup}
    for mode, conc in (("concurrent", True), ("sequential", False)):
        ds = make_designs(c4_designs(), precision, concurrent=conc)
        step_designs(ds, warmup, concurrent=conc)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step_designs(ds, steps, concurrent=conc)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[mode] = round(16 * steps / dt, 3)
        out[mode + "_mean_pcg_iters"] = round(float(sum(sum(d.opt.history["pcg_iters"][-1])
                                                        for d in ds)) / 16, 1)
        out[mode + "_mean_lobpcg_iters"] = round(float(sum(d.opt.history["lobpcg_iters"][-1]
                                                           for d in ds)) / 16, 1)
        print(f"[bench] c4 {mode}: {16 * steps} design-iterations in {dt:.3f}s", file=sys.stderr, flush=True)
        close_designs(ds)
        del ds
    return out


def bench_c4_child(precision, steps, warmup):
    """bench_c4 in a child process (a fresh stream pool: the 32 streams of the
    batch would otherwise shift the stream-to-queue mapping of the process).
    At most C4_MAX_WARMUP + C4_MAX_STEPS iterations per design and phase, so
    the measurement fits the bench whatever --steps is; the time limit scales
    with them.  The child prints a Python traceback of every thread if it
    crashes (faulthandler); on a time-out its output so far is kept."""
    import subprocess
    steps, warmup = min(steps, C4_MAX_STEPS), min(warmup, C4_MAX_WARMUP)
    timeout_s = 180 + 40 * (steps + warmup)
    cmd = [sys.executable, "-u", "-X", "faulthandler", str(Path(__file__).resolve()), "--c4-only",
           "--precision", precision, "--steps", str(steps), "--warmup", str(warmup)]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s, cwd=str(ROOT))
    except subprocess.TimeoutExpired as e:
        err = e.stderr.decode(errors="replace") if isinstance(e.stderr, bytes) else (e.stderr or "")
        sys.stderr.write(err[-4000:])
        return {"error": f"child timed out after {timeout_s}s", "stderr_tail": err[-1500:]}
    sys.stderr.write(r.stderr[-4000:])
    sys.stderr.flush()
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"child exited with status {r.returncode}", "stderr_tail": r.stderr[-1500:]}
    out = json.loads(lines[-1])
    out["process"] = "child"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5t")
    ap.add_argument("--also", default="c2,c3,c5,c4",
                    help="comma-separated secondary workloads ('' for none; c4 = the 16 batched designs)")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--plain-maxiter", type=int, default=2000,
                    help="LOBPCG iteration cap of the plain FP64 (Jacobi) comparison path")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--modal-guard", default="auto", choices=["auto", "0", "1", "2"],
                    help="LOBPCG guard columns of the SIMP loop's modal solve (auto: two or none, "
                         "whichever costs less; SIMPOptimizer modal_guard)")
    ap.add_argument("--coarse", default="galerkin", choices=["galerkin", "rediscretised"],
                    help="coarse levels of the multigrid preconditioner")
    ap.add_argument("--no-kernels", action="store_true",
                    help="skip the per-kernel HIP-event timings and the cold-start solver timings "
                         "(for profiles of the timed window alone)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--trajectory", type=int, default=0, metavar="N",
                    help="instead of the bench line: N SIMP iterations of the workload on the "
                         "multigrid path and on the plain FP64 path from the same seeded "
                         "density, one JSON line per path with per-iteration times")
    ap.add_argument("--summarize-trace", default=None, metavar="KERNEL_TRACE_CSV",
                    help="print a per-(kernel, grid) table of a rocprofv3 kernel trace and exit "
                         "(with --top N and --trace-filter SUBSTRING)")
    ap.add_argument("--top", type=int, default=40)
    ap.add_argument("--trace-filter", default="")
    ap.add_argument("--c4-only", action="store_true",
                    help="only the batched-designs measurement, as one JSON object (the main "
                         "bench runs it this way, in a child process)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the per-kernel timings (and a 1 GiB device copy as a counter "
                         "calibration), for rocprofv3 --pmc runs")
    args = ap.parse_args()
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace, args.top, args.trace_filter), end="")
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)

    if args.c4_only:
        print(json.dumps(bench_c4(args.precision, args.steps, args.warmup)), flush=True)
        return

    if args.trajectory > 0:
        for prec in ("mixed", "fp64"):
            print(json.dumps(trajectory(args.workload, prec, args.trajectory, args.plain_maxiter),
                             default=lambda v: v.item()), flush=True)
        return

    if args.kernels_only:
        wl, geom, fem, opt, F, x0 = make_optimizer(args.workload, args.precision)
        src = torch.ones(1 << 27, dtype=torch.float64, device=fem.device)      # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        roofs = kernel_rooflines(fem, geom, torch.as_tensor(x0, device=fem.device), reps=5)
        print(json.dumps({"workload": wl.name, "copy_bytes_each_way": src.numel() * 8,
                          "kernels": roofs}), flush=True)
        return

    snap = {}
    wl, geom, fem, opt, dt, pcg_it, lob_it = run_workload(args.workload, args.precision, args.steps,
                                                          args.warmup, snap, args.coarse == "galerkin",
                                                          args.modal_guard)
    value = args.steps / dt
    # the multigrid path's own time for the first timed iteration (the state
    # the plain path is timed from below)
    first_timed = {"ms": round(opt.history["time_s"][args.warmup] * 1e3, 2),
                   "pcg_iters": opt.history["pcg_iters"][args.warmup],
                   "lobpcg_iters": opt.history["lobpcg_iters"][args.warmup]}
    solvers = None if args.no_kernels else solver_timings(opt)
    print(f"[bench] solves alone: {json.dumps(solvers)}", file=sys.stderr, flush=True)
    h = opt.history
    for i in range(args.warmup, len(h["time_s"])):
        print(f"[bench] iteration {i + 1}: {h['time_s'][i] * 1e3:.1f} ms, pcg {h['pcg_iters'][i]}, "
              f"lobpcg {h['lobpcg_iters'][i]}"
              + (f" (guard {h['lobpcg_guard'][i]})" if i < len(h["lobpcg_guard"]) else ""),
              file=sys.stderr, flush=True)
    roofs = {} if args.no_kernels else kernel_rooflines(fem, geom, opt.rho_phys)
    # Each workload's contexts are released before the next one starts: every
    # stream alive in the process shares the few hardware queues HIP maps
    # streams onto, and a context kept alive for nothing can make the next
    # workload's static and modal streams share a queue (and serialise).
    _release(opt, fem)
    del opt, fem
    secondary = {}
    also = [s for s in args.also.split(",") if s and s != args.workload]
    run_c4 = "c4" in also
    if run_c4:
        also.remove("c4")
    for name in also:
        cap_s, cap_w = SECONDARY_CAPS.get(name, (args.steps, args.warmup))
        ss, sw = min(args.steps, cap_s), min(args.warmup, cap_w)
        swl, sgeom, _, sopt, sdt, spcg, slob = run_workload(name, args.precision, ss, sw,
                                                            galerkin=args.coarse == "galerkin")
        secondary[name] = {"value": round(ss / sdt, 4), "unit": "SIMP iters/s",
                           "ms_per_step": round(sdt / ss * 1e3, 2), "steps": ss, "warmup": sw,
                           "n_dofs": sgeom.n_dofs, "load_cases": len(swl.loads), "modes": swl.n_modes,
                           "mean_pcg_iters": round(spcg, 1), "mean_lobpcg_iters": round(slob, 1)}
        if swl.eff_dir is not None:
            h = sopt.history
            secondary[name]["effective_mass_direction"] = "xyz"[swl.eff_dir]
            secondary[name]["last_eff_mass"] = [round(v, 6) for v in h["eff_mass"][-1]]
            secondary[name]["last_mode_rank"] = h["mode_rank"][-1]
        _release(sopt)
        del sopt
    # The plain path runs after c2 and c3: every optimiser takes a fresh pooled
    # stream for its modal solve, and streams share a few hardware queues in
    # turn, so the fourth optimiser of the process can get its modal stream on
    # the static stream's queue (measured: c3 12.5 instead of 17.4 it/s).
    plain = None
    if not args.no_plain and args.precision != "fp64":
        # one iteration of the plain path from the multigrid path's state at
        # the start of the timed window (same design, same warm starts)
        _, _, _, popt, _, _ = make_optimizer(args.workload, "fp64", modal_maxiter=args.plain_maxiter,
                                             modal_guard=0)
        popt.load_state(snap, overlap=True)
        pdt = time_steps(popt, 0, 1)
        ph = popt.history
        lob = ph["lobpcg_iters"][-1]
        plain = {"value": round(1 / pdt, 4), "unit": "SIMP iters/s", "ms_per_step": round(pdt * 1e3, 2),
                 "steps": 1, "path": "FP64 Jacobi-PCG + Jacobi-LOBPCG (reference preconditioner)",
                 "state": f"the multigrid path's loop state after {args.warmup} iterations "
                          "(iteration {} of the timed window's design)".format(args.warmup + 1),
                 "pcg_iters": ph["pcg_iters"][-1], "lobpcg_iters": lob,
                 "lobpcg_maxiter": args.plain_maxiter,
                 "lobpcg_converged": bool(lob < args.plain_maxiter),
                 "multigrid_same_iteration": first_timed,
                 "speedup_same_iteration": round(pdt * 1e3 / first_timed["ms"], 2),
                 "speedup_of_value": round(value * pdt, 2)}
        print(f"[bench] plain FP64 path: {pdt:.3f}s for iteration {args.warmup + 1} "
              f"(multigrid path {first_timed['ms'] / 1e3:.3f}s)", file=sys.stderr, flush=True)
        _release(popt)
        del popt
    if run_c4:
        # last: its designs take 32 pooled streams (a design stream and a modal
        # stream each), which would shift the stream-to-queue mapping of any
        # workload measured after it
        secondary["c4"] = bench_c4_child(args.precision, args.steps, args.warmup)
    line = {
        "metric": "SIMP iters/s (solve+6 modes+adjoint) & matvec HBM GB/s vs plain FP64 path",
        "value": round(value, 4), "unit": "SIMP iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 Krylov / f32 multigrid",
        "data": "synthetic seeded voxel cantilever",
        "config": {"workload": f"{wl.name}: SIMP iteration on {geom.nx}x{geom.ny}x{geom.nz} voxels @ "
                               f"{geom.dx * 1e3:g} mm, {len(wl.loads)} load cases, {wl.n_modes} modes",
                   "n_dofs": geom.n_dofs, "pcg_rtol": wl.pcg_rtol, "lobpcg_tol": wl.modal_tol,
                   "mean_pcg_iters_per_lc": round(pcg_it, 1), "mean_lobpcg_iters": round(lob_it, 1),
                   "precision": args.precision,
                   "multigrid_coarse_levels": args.coarse, "lobpcg_guard": args.modal_guard},
        "roofline": roofs.get(DOMINANT),
        "kernels": {k: v for k, v in roofs.items() if k != DOMINANT},
        "solvers": solvers,
        "plain_fp64": plain,
        "secondary": secondary,
    }
    if not args.no_cpu_baseline:
        if plain is not None:
            pi = plain["pcg_iters"]
            li = plain["lobpcg_iters"]
            basis = f"iteration counts of the plain FP64 GPU path, LOBPCG capped at {args.plain_maxiter}"
        else:
            pi = [int(round(pcg_it))] * len(wl.loads)
            li = int(round(lob_it))
            basis = "iteration counts of the multigrid path (plain path not run)"
        line["cpu_baseline"] = cpu_baseline(wl, geom, pi, li, basis)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
