"""GPU: the block multi-load-case configuration (c3) and the batched designs
configuration (c4).

c3: 220x32x24 at 1 mm, LC1-LC5 solved together as one block PCG (rtol 1e-8)
against each load case solved alone: per-load-case compliance within 1e-6
(the parity bar), displacements within 1e-6, and the weighted summed
compliance gradient equal to the weighted sum of the single-case gradients.

c4: the 16 uneven designs (nx 80..220, V_f 0.08..0.50), two SIMP iterations
each, run concurrently (one context, stream and thread per design) against
the same designs run one after another: the histories must agree bitwise-close
(deterministic kernels, independent state).  Each design's iteration-1
compliance and first frequency must match a plain FP64 solve of that design
alone (Jacobi PCG to rtol 1e-10, Jacobi LOBPCG to tol 1e-6, cold start) within
1e-6 (sims/run_vf_gpu.py:133-141 runs each V_f as its own job).
"""
import numpy as np
import pytest
import torch

from helpers import rel

pytestmark = pytest.mark.gpu


def test_c3_block_pcg_matches_single_solves(gpu):
    from vox8.fem import FEM3D, Material
    from vox8.workloads import E_PA, NU, RHO_MAT, WORKLOADS, load_cases, seeded_density
    wl = WORKLOADS["c3"]
    geom = wl.geometry()
    fem = FEM3D(geom, Material(E=E_PA, nu=NU, rho=RHO_MAT), precision="mixed")
    F = load_cases(geom, wl.loads)
    assert F.shape == (5, geom.n_dofs)
    rho = torch.tensor(seeded_density(geom, wl.base, wl.amp, wl.seed), device=gpu)
    Ft = torch.tensor(F, device=gpu)
    U, info = fem.solve_block(rho, Ft, rtol=wl.pcg_rtol)
    assert info["info"] == [0] * 5
    w = np.array([0.3, 0.2, 0.2, 0.2, 0.1])
    s = fem.sensitivities(rho, U=U, weights=w)
    dc_sum = torch.zeros_like(s["dc"])
    for j in range(5):
        Uj, ij = fem.solve_block(rho, Ft[j:j + 1], rtol=wl.pcg_rtol)
        assert ij["info"] == [0]
        cj = float(Ft[j] @ Uj[0])
        assert abs(cj / float(Ft[j] @ U[j]) - 1) < 1e-6
        assert abs(s["comp"][j] / cj - 1) < 1e-6
        assert rel(U[j].cpu().numpy(), Uj[0].cpu().numpy()) < 1e-6
        dc_sum += w[j] * fem.sensitivities(rho, U=Uj, weights=[1.0])["dc"]
    assert rel(s["dc"].cpu().numpy(), dc_sum.cpu().numpy()) < 1e-6


def _first_physical_density(wl):
    from oracle.topopt import cone_kernel, filter_field
    from vox8.workloads import seeded_density
    g = wl.geometry()
    x = seeded_density(g, wl.base, wl.amp, wl.seed)
    grid = x.reshape(g.nz, g.ny, g.nx).transpose(2, 1, 0)
    phys = filter_field(grid, np.ones(grid.shape, dtype=bool), cone_kernel(wl.r_min_vox))
    return g, np.ascontiguousarray(phys.transpose(2, 1, 0)).reshape(-1)


def test_c4_batched_designs_match_sequential_and_plain(gpu):
    from vox8.batch import close_designs, histories, run_designs
    from vox8.fem import FEM3D, Material
    from vox8.workloads import E_PA, NU, RHO_MAT, c4_designs, load_cases
    designs = c4_designs()
    assert len(designs) == 16 and designs[0].nx == 80 and designs[-1].nx == 220
    conc = run_designs(designs, steps=2, concurrent=True)
    (cc, fc) = histories(conc)
    xc = [d.opt.x.cpu().numpy() for d in conc]
    close_designs(conc)
    seq = run_designs(designs, steps=2, concurrent=False)
    (cs, fs) = histories(seq)
    xs = [d.opt.x.cpu().numpy() for d in seq]
    close_designs(seq)
    assert cc.shape == (16, 2)
    assert rel(cc, cs) < 1e-12 and rel(fc, fs) < 1e-12
    for a, b in zip(xc, xs):
        assert rel(a, b) < 1e-12
    # each design against a plain FP64 solve of that design alone
    mat = Material(E=E_PA, nu=NU, rho=RHO_MAT)
    for d, wl in enumerate(designs):
        geom, rho = _first_physical_density(wl)
        F = load_cases(geom, wl.loads)
        fem = FEM3D(geom, mat, precision="fp64")
        U, info = fem.solve_block(rho, F, rtol=1e-10, maxiter=200000)
        assert info["info"] == [0]
        c = float(F[0] @ U[0].cpu().numpy())
        lam, _, li = fem.modal_block(rho, wl.n_modes, 0.0, [], tol=1e-6, maxiter=40000)
        assert li["info"] == 0, (d, li)
        f1 = float(np.sqrt(np.min(lam)) / (2 * np.pi))
        fem.close()
        print(f"design {d}: c {cc[d, 0]:.10e} vs {c:.10e}, f1 {fc[d, 0]:.9f} vs {f1:.9f}")
        assert abs(cc[d, 0] / c - 1) < 1e-6, (d, cc[d, 0], c)
        assert abs(fc[d, 0] / f1 - 1) < 1e-6, (d, fc[d, 0], f1)
