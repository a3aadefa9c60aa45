"""Benchmark: SIMP iterations/s of the voxel hex8 FEM hot path on one MI355X.

One step = one full SIMP iteration of the workload (default c2: 220x32x24
voxels at 1 mm, 546,975 DOF): density filter, Jacobi-PCG static solve (rtol
1e-8, warm-started), compliance adjoint, 6-mode LOBPCG (tol 1e-6,
warm-started), mode-1 eigenvalue adjoint (the frequency constraint is set
above f1 so it is active every iteration), sensitivity filters, OC update.
Inputs are resident in HBM before the timed region.

Prints one JSON line (rank 0).  Extra objects: "roofline" for the stencil
K.u (HIP-event timing of the dominant kernel, algorithmic bytes
6*8*N_nodes + 8*N_elem per FP64 launch), "plain_fp64" (the same loop on the
plain FP64 path) and "cpu_baseline" (the SciPy oracle on this host, bounded
sample, extrapolated per SIMP iteration with the GPU run's iteration counts).
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


def make_optimizer(wl_name, precision, modal_maxiter=2000):
    import numpy as np
    from vox8.fem import FEM3D, Material
    from vox8.topopt import SIMPOptimizer
    from vox8.workloads import E_PA, NU, RHO_MAT, WORKLOADS, load_cases, seeded_density

    wl = WORKLOADS[wl_name]
    geom = wl.geometry()
    fem = FEM3D(geom, Material(E=E_PA, nu=NU, rho=RHO_MAT), precision=precision)
    F = load_cases(geom, wl.loads)
    x0 = seeded_density(geom, wl.base, wl.amp, wl.seed)
    # frequency target above the initial f1 keeps the eigen-adjoint active
    opt = SIMPOptimizer(geom, fem, load_cases=[{}] * len(F), vol_frac=wl.base,
                        r_min_mm=wl.r_min_vox * geom.dx * 1e3, max_iter=10 ** 6,
                        min_freq_Hz=1e4, tip_mass_kg=0.0, n_modes_track=wl.n_modes,
                        F_list=list(F), solve_rtol=wl.pcg_rtol, modal_tol=wl.modal_tol,
                        modal_maxiter=modal_maxiter, solve_maxiter=20000)
    opt.start(rho_init=x0)
    return wl, geom, fem, opt, F, x0


def time_steps(opt, warmup, steps):
    import torch
    for _ in range(warmup):
        opt.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        opt.step()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def stencil_roofline(fem, geom, reps=50):
    """HIP-event timing of the FP64 single-RHS K.u launch on the solver stream."""
    import torch
    from vox8 import _lib
    n = geom.n_dofs
    x = torch.randn(n, dtype=torch.float64, device=fem.device)
    y = torch.empty_like(x)
    rho = torch.full((geom.n_elements,), 0.3, dtype=torch.float64, device=fem.device)
    fem._set_density(rho)
    args = (_lib.OP_K, _lib.PREC_FP64, _lib.ptr(x), _lib.ptr(y), None, 1, n)
    for _ in range(5):
        fem._call("vox8_apply", *args)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        fem._call("vox8_apply", *args)
    e1.record(st)
    e1.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / reps
    alg_bytes = 6 * 8 * geom.n_nodes + 8 * geom.n_elements
    achieved = alg_bytes / t / 1e9
    return {"kernel": "k_apply_dense<double,K> (K.u, FP64, 1 RHS)", "bound": "hbm",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(),
            "avg_launch_us": round(t * 1e6, 2), "alg_bytes_per_launch": alg_bytes}


def load_traffic():
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if present."""
    p = ROOT / "profiles" / "stencil_Ku_fp64_c2_pmc.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("traffic_bytes")
        except Exception:
            return None
    return None


def cpu_baseline(wl, geom, F, x0, pcg_iters, lob_iters, budget_s=25.0):
    """SciPy oracle on this host: per-phase timings on a bounded sample, combined
    into seconds per SIMP iteration with the GPU run's mean iteration counts."""
    import numpy as np
    import scipy.sparse.linalg as spla
    from oracle.fem import FEM3D as OFEM, Material as OMat
    from oracle.topopt import cone_kernel, filter_field, oc_update
    from vox8.workloads import E_PA, NU, RHO_MAT

    cores = len(os.sched_getaffinity(0))
    of = OFEM(geom, OMat(E=E_PA, nu=NU, rho=RHO_MAT))
    g3 = lambda v: v.reshape(geom.nz, geom.ny, geom.nx).transpose(2, 1, 0)
    mask = np.ones((geom.nx, geom.ny, geom.nz), dtype=bool)
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
    n_cg = 0
    t = time.perf_counter()
    spla.cg(Kff, F[0][free], M=Minv, rtol=1e-30, maxiter=40)
    n_cg = 40
    t_cg = (time.perf_counter() - t) / n_cg
    X0 = np.random.default_rng(0).standard_normal((len(free), wl.n_modes))
    Tinv = spla.LinearOperator(Kff.shape, matvec=lambda v: v / d, matmat=lambda V: V / d[:, None],
                               dtype=float)
    n_lob = 8
    t = time.perf_counter()
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        spla.lobpcg(Kff, X0, B=Mff, M=Tinv, tol=1e-30, largest=False, maxiter=n_lob)
    t_lob = (time.perf_counter() - t) / n_lob
    u = np.zeros(geom.n_dofs)
    t = time.perf_counter()
    ce = of.elem_compliance(u)
    of.eigenvalue_sensitivity(rho, 1.0, u, tip_mass_kg=0.0)
    t_sens = time.perf_counter() - t
    t = time.perf_counter()
    oc_update(x0, -np.abs(ce) - 1.0, geom.dx ** 3, 0.3 * x0.size * geom.dx ** 3, 0.2)
    t_oc = time.perf_counter() - t
    n_lc = F.shape[0]
    per_iter = (2 * t_asm + n_lc * pcg_iters * t_cg + lob_iters * t_lob + t_sens
                + 4 * t_filter + t_oc)
    return {"value": round(1.0 / per_iter, 6), "unit": "SIMP iters/s", "cores": cores,
            "kind": "port",
            "sample": (f"SciPy oracle on {wl.name}: K/M COO->CSR assembly {t_asm:.2f}s, "
                       f"{n_cg} CG iters ({t_cg * 1e3:.1f} ms/iter, CSR SpMV single-threaded), "
                       f"{n_lob} LOBPCG iters ({t_lob * 1e3:.1f} ms/iter), einsum sensitivities "
                       f"{t_sens:.2f}s, ndimage filter {t_filter:.2f}s x4, OC {t_oc:.2f}s; "
                       f"extrapolated with GPU mean {pcg_iters:.0f} CG iters x {n_lc} LC and "
                       f"{lob_iters:.0f} LOBPCG iters per SIMP iteration"),
            "seconds_per_iter": round(per_iter, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--plain-maxiter", type=int, default=2000,
                    help="LOBPCG iteration cap of the plain FP64 (Jacobi) comparison path")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    torch.cuda.set_device(0)

    wl, geom, fem, opt, F, x0 = make_optimizer(args.workload, args.precision)
    print(f"[bench] {wl.name}: {geom.nx}x{geom.ny}x{geom.nz}, {geom.n_dofs} DOF, "
          f"{F.shape[0]} LC, {wl.n_modes} modes; warmup {args.warmup}", file=sys.stderr, flush=True)
    dt = time_steps(opt, args.warmup, args.steps)
    h = opt.history
    pcg_it = float(np.mean([np.mean(v) for v in h["pcg_iters"][-args.steps:]]))
    lob_it = float(np.mean(h["lobpcg_iters"][-args.steps:]))
    print(f"[bench] timed {args.steps} steps in {dt:.3f}s; pcg {pcg_it:.0f} it, lobpcg {lob_it:.0f} it",
          file=sys.stderr, flush=True)
    value = args.steps / dt
    plain = None
    if not args.no_plain and args.precision != "fp64":
        _, _, _, popt, _, _ = make_optimizer(args.workload, "fp64", modal_maxiter=args.plain_maxiter)
        pdt = time_steps(popt, 0, 1)
        ph = popt.history
        plain = {"value": round(1 / pdt, 4), "unit": "SIMP iters/s", "ms_per_step": round(pdt * 1e3, 2),
                 "steps": 1, "path": "FP64 Jacobi-PCG + Jacobi-LOBPCG (reference preconditioner)",
                 "pcg_iters": ph["pcg_iters"][-1], "lobpcg_iters": ph["lobpcg_iters"][-1],
                 "lobpcg_maxiter": args.plain_maxiter}
        print(f"[bench] plain FP64 path: {pdt:.3f}s per SIMP iteration", file=sys.stderr, flush=True)
    roof = stencil_roofline(fem, geom)
    line = {
        "metric": "SIMP iters/s (solve+6 modes+adjoint) & matvec HBM GB/s vs plain FP64 path",
        "value": round(value, 4), "unit": "SIMP iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-stencil/f64-refinement",
        "data": "synthetic seeded voxel cantilever",
        "config": {"workload": f"{wl.name} SIMP iteration {geom.nx}x{geom.ny}x{geom.nz} @ {geom.dx*1e3:g} mm, "
                               f"{F.shape[0]} load case(s), {wl.n_modes} modes",
                   "n_dofs": geom.n_dofs, "pcg_rtol": wl.pcg_rtol, "lobpcg_tol": wl.modal_tol,
                   "mean_pcg_iters": round(pcg_it, 1), "mean_lobpcg_iters": round(lob_it, 1),
                   "precision": args.precision},
        "roofline": roof,
        "plain_fp64": plain,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(wl, geom, F, x0, pcg_it, lob_it)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
