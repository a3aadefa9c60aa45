"""GPU SIMP loop (vox8.topopt.SIMPOptimizer) against the reference's own run.

Same fixture as tests/test_oracle_simp.py (tests/golden/simp_tipbox.npz,
written by the reference's SIMPOptimizer with converged inner solves): eight
iterations with two weighted load cases, tip mass, an active frequency
constraint, MAC mode tracking, the mu ramp and the OC bisection, all on the
device.  Bar: compliance and frequencies within 1e-6 relative; the rest of
the history and the final densities within 1e-6.
"""
import numpy as np
import pytest

from conftest import GOLDEN
from helpers import (TIPBOX_MODAL_MAXITER, TIPBOX_MODAL_TOL, TIPBOX_SOLVE_RTOL, rel,
                     tipbox_problem)

pytestmark = pytest.mark.gpu

HIST = ("compliance", "mass", "change", "vol_frac", "omega_1_Hz", "mu_freq", "freq_violation")


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_simp_loop_matches_reference_run(gpu, prec):
    from vox8.fem import FEM3D, Material
    from vox8.topopt import SIMPOptimizer
    gold = np.load(GOLDEN / "simp_tipbox.npz")
    geom, loads, opts, mat = tipbox_problem()
    fem = FEM3D(geom, Material(**mat), precision=prec)
    res = SIMPOptimizer(geom, fem, loads, solve_rtol=TIPBOX_SOLVE_RTOL, modal_tol=TIPBOX_MODAL_TOL,
                        modal_maxiter=TIPBOX_MODAL_MAXITER, **opts).run(verbose=False)
    h = res["history"]
    errs = {k: rel(np.asarray(h[k], dtype=float), gold[f"hist_{k}"]) for k in HIST}
    errs["rho"] = rel(res["rho_active"], gold["rho_active"])
    errs["u"] = rel(np.stack(res["u_per_lc"]), gold["u_per_lc"])
    print(prec, {k: f"{v:.1e}" for k, v in errs.items()})
    assert all(v < 1e-6 for v in errs.values()), errs


def test_concurrent_modal_solve_matches_sequential(gpu):
    """The modal solve in a second context on its own stream and thread,
    overlapping the static solve, reproduces the history of running the two
    solves one after the other in one context (deterministic kernels, no
    shared state; launch shapes are tuned per grid, so FP32 multigrid rounding
    may differ at the 1e-12 level)."""
    from vox8.fem import FEM3D, Material
    from vox8.topopt import SIMPOptimizer
    geom, loads, opts, mat = tipbox_problem()
    runs = {}
    for conc, gal in ((False, True), (True, True), (True, False)):
        fem = FEM3D(geom, Material(**mat), precision="mixed")
        fem.set_mg_galerkin(gal)
        runs[conc, gal] = SIMPOptimizer(geom, fem, loads, solve_rtol=TIPBOX_SOLVE_RTOL,
                                        modal_tol=TIPBOX_MODAL_TOL,
                                        modal_maxiter=TIPBOX_MODAL_MAXITER, concurrent_solves=conc,
                                        **opts).run(verbose=False)
    a, b = runs[False, True]["history"], runs[True, True]["history"]
    for k in HIST + ("pcg_iters", "lobpcg_iters"):
        assert rel(np.asarray(b[k], dtype=float), np.asarray(a[k], dtype=float)) < 1e-10, k
    assert rel(runs[True, True]["rho_active"], runs[False, True]["rho_active"]) < 1e-10
    # rediscretised coarse levels: another preconditioner, the same converged
    # answers (the parity bar)
    c = runs[True, False]["history"]
    for k in HIST:
        assert rel(np.asarray(c[k], dtype=float), np.asarray(a[k], dtype=float)) < 1e-6, k


@pytest.mark.parametrize("case", ["active", "never_binding", "no_modal"])
def test_device_tail_matches_host_tail(gpu, case):
    """The device-resident loop tail (vox8_simp_tail: MAC tracking, eigenvalue
    penalty, OC, filter, bookkeeping, one host read) against the same steps
    driven from the host: identical histories and designs.  Cases: frequency
    constraint active (penalty applied), a target below every frequency (the
    gated branch never runs) and no modal constraint."""
    from vox8.fem import FEM3D, Material
    from vox8.topopt import SIMPOptimizer
    geom, loads, opts, mat = tipbox_problem()
    opts = dict(opts)
    if case == "never_binding":
        opts["min_freq_Hz"] = 1.0
    elif case == "no_modal":
        opts["min_freq_Hz"] = None
    runs = {}
    for dev_tail in (False, True):
        fem = FEM3D(geom, Material(**mat), precision="mixed")
        runs[dev_tail] = SIMPOptimizer(geom, fem, loads, solve_rtol=TIPBOX_SOLVE_RTOL,
                                       modal_tol=TIPBOX_MODAL_TOL, modal_maxiter=TIPBOX_MODAL_MAXITER,
                                       device_tail=dev_tail, **opts).run(verbose=False)
    a, b = runs[False]["history"], runs[True]["history"]
    if case == "active":
        assert max(a["freq_violation"]) > 0.0
    if case == "never_binding":
        assert max(a["freq_violation"]) == 0.0
    for k in HIST:
        x, y = np.asarray(a[k], dtype=float), np.asarray(b[k], dtype=float)
        if case == "no_modal" and k in ("omega_1_Hz", "mu_freq", "freq_violation"):
            assert np.all(np.isnan(x)) and np.all(np.isnan(y))
            continue
        assert rel(y, x) < 1e-12, (k, x, y)
    assert rel(runs[True]["rho_active"], runs[False]["rho_active"]) < 1e-12


@pytest.mark.parametrize("case", ["active", "no_modal"])
def test_tail_graph_replays_match_direct_launches(gpu, case):
    """vox8_simp_tail replayed from captured HIP graphs (required: a refused
    capture fails the run) against the same launches issued one by one:
    bitwise equal histories and designs over the whole run (the graphs are
    replayed with new eigenvalues, penalty coefficients and mode blocks, and
    the ping-pong design buffers select between cached graphs)."""
    from vox8.fem import FEM3D, Material
    from vox8.topopt import SIMPOptimizer
    geom, loads, opts, mat = tipbox_problem()
    opts = dict(opts)
    if case == "no_modal":
        opts["min_freq_Hz"] = None
    runs = {}
    for mode in (0, 2):
        fem = FEM3D(geom, Material(**mat), precision="mixed")
        fem._call("vox8_set_tail_graph", mode)
        runs[mode] = SIMPOptimizer(geom, fem, loads, solve_rtol=TIPBOX_SOLVE_RTOL,
                                   modal_tol=TIPBOX_MODAL_TOL, modal_maxiter=TIPBOX_MODAL_MAXITER,
                                   device_tail=True, **opts).run(verbose=False)
    a, b = runs[0]["history"], runs[2]["history"]
    assert len(a["compliance"]) >= 4
    for k in HIST:
        x, y = np.asarray(a[k], dtype=float), np.asarray(b[k], dtype=float)
        assert np.array_equal(x, y, equal_nan=True), (k, x, y)
    assert np.array_equal(runs[2]["rho_active"], runs[0]["rho_active"])


def test_auto_guard_columns_keep_the_history(gpu):
    """The "auto" guard policy (two guard columns or none, whichever has cost
    less per solve; here also a variant that tries the other setting every
    iteration)
    changes how the modes are reached, not what they are: the history equals
    the run without guard columns to the golden test's bar (1e-6), and the
    policy runs both settings."""
    from vox8.fem import FEM3D, Material
    from vox8.topopt import SIMPOptimizer

    class Alternating(SIMPOptimizer):
        GUARD_RETRY = 1
        GUARD_CONFIRM = 1

    geom, loads, opts, mat = tipbox_problem()
    runs = {}
    for name, cls, guard in (("none", SIMPOptimizer, 0), ("auto", SIMPOptimizer, "auto"),
                             ("alt", Alternating, "auto")):
        fem = FEM3D(geom, Material(**mat), precision="mixed")
        opt = cls(geom, fem, loads, solve_rtol=TIPBOX_SOLVE_RTOL, modal_tol=TIPBOX_MODAL_TOL,
                  modal_maxiter=TIPBOX_MODAL_MAXITER, modal_guard=guard, **opts)
        runs[name] = opt.run(verbose=False)
        if guard == "auto":
            assert set(runs[name]["history"]["lobpcg_guard"]) == {0, 2}, name
    for name in ("auto", "alt"):
        a, b = runs["none"]["history"], runs[name]["history"]
        for k in HIST:
            assert rel(np.asarray(b[k], dtype=float), np.asarray(a[k], dtype=float)) < 1e-6, (name, k)



def test_checkpoint_resume_continues_the_run(gpu, tmp_path):
    """save() after three iterations, load() into a fresh optimiser, two more
    iterations: the same history as five uninterrupted iterations (the warm
    starts, tracked mode and mu travel with the state)."""
    from vox8.fem import FEM3D, Material
    from vox8.topopt import SIMPOptimizer
    geom, loads, opts, mat = tipbox_problem()
    kw = dict(solve_rtol=TIPBOX_SOLVE_RTOL, modal_tol=TIPBOX_MODAL_TOL,
              modal_maxiter=TIPBOX_MODAL_MAXITER, concurrent_solves=False, **opts)
    a = SIMPOptimizer(geom, FEM3D(geom, Material(**mat), precision="mixed"), loads, **kw).start()
    for _ in range(5):
        a.step()
    b = SIMPOptimizer(geom, FEM3D(geom, Material(**mat), precision="mixed"), loads, **kw).start()
    for _ in range(3):
        b.step()
    path = tmp_path / "state.npz"
    b.save(path)
    c = SIMPOptimizer(geom, FEM3D(geom, Material(**mat), precision="mixed"), loads, **kw).start()
    c.load(path)
    assert c.it == 3
    for _ in range(2):
        c.step()
    for k in ("compliance", "change", "omega_1_Hz", "mu_freq"):
        assert rel(np.asarray(c.history[k]), np.asarray(a.history[k][3:])) < 1e-6, k
    assert rel(c.x.cpu().numpy(), a.x.cpu().numpy()) < 1e-6
