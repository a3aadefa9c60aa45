"""GPU: run results in the reference's pickle schema (vox8.results).

On the tip-box problem with its two load cases: the device static analysis
against the oracle (spsolve displacements, the oracle's stress, von Mises and
principal stress of the same displacements, sf_yield, nodal displacement
norms), the fatigue fields against the oracle's, the record's keys and values
(including every key sims/consolidate_v6.py reads), and the written pickles:
file name, schema, and only NumPy / builtin types inside (they load without
vox8 installed).
"""
import pickle

import numpy as np
import pytest

from helpers import rel, tipbox_problem

pytestmark = pytest.mark.gpu

SIGMA_Y, SIGMA_U = 38e6, 48e6
NAMES = ["LC1_hover", "LC2_maneuver"]


def _globals(path):
    """Modules the pickle's objects come from (recorded while loading it)."""
    mods = set()

    class Rec(pickle.Unpickler):
        def find_class(self, module, name):
            mods.add(module)
            return super().find_class(module, name)

    with open(path, "rb") as f:
        Rec(f).load()
    return mods


def test_static_fatigue_record_and_pickles(gpu, tmp_path):
    from oracle.fatigue import fatigue_life_field as o_fatigue
    from oracle.fem import FEM3D as OFEM, Material as OMat
    from vox8.fem import FEM3D, Material
    from vox8.results import (fatigue_analysis, modal_results, result_path, run_record, save_baseline,
                              save_run, static_analysis)
    geom, loads, opts, mat = tipbox_problem()
    m = Material(E=mat["E"], nu=mat["nu"], rho=mat["rho"], sigma_y=SIGMA_Y, sigma_u=SIGMA_U)
    fem = FEM3D(geom, m, precision="mixed")
    ofem = OFEM(geom, OMat(E=mat["E"], nu=mat["nu"], rho=mat["rho"]))
    F = []
    for lc in loads:
        f = np.zeros(geom.n_dofs)
        dofs, mags = geom.load_dofs_for_force(lc["F"])
        f[dofs] += mags
        F.append(f)
    rho = np.random.default_rng(3).uniform(0.2, 1.0, fem.n_active)
    st = static_analysis(fem, F, rho, NAMES, SIGMA_Y, rtol=1e-10)
    assert list(st) == NAMES
    for name, f in zip(NAMES, F):
        s = st[name]
        assert set(s) == {"u", "sigma", "vm", "sigma_max", "disp_node", "sf_yield", "solve_time"}
        u_ref = ofem.apply_BCs_and_solve(ofem.assemble_K(rho), f, method="direct")
        assert rel(s["u"], u_ref) < 1e-7
        sig = ofem.compute_stress(s["u"], rho)
        assert rel(s["sigma"], sig) < 1e-12
        assert rel(s["vm"], ofem.von_mises(sig)) < 1e-12
        assert rel(s["sigma_max"], ofem.principal_stress_max(sig)) < 1e-10
        assert np.allclose(s["sf_yield"], SIGMA_Y / np.maximum(s["vm"], 1e-3), rtol=1e-14)
        assert np.allclose(s["disp_node"], np.linalg.norm(s["u"].reshape(-1, 3), axis=1))
        assert all(isinstance(s[k], np.ndarray) for k in ("u", "sigma", "vm", "sigma_max"))
    fat = fatigue_analysis(st, SIGMA_U)
    ofat = o_fatigue(st[NAMES[0]]["vm"], st[NAMES[1]]["vm"], SIGMA_U)
    for k in ("sigma_mean", "sigma_alt", "sigma_alt_eq", "life", "damage", "factor_of_safety"):
        assert rel(fat[k], ofat[k]) < 1e-12, k
    freqs, modes, lam = fem.modal_analysis(rho, 3, tip_mass_kg=2e-3)
    modal = modal_results(freqs, modes, lam, 2e-3)
    rec = run_record(300.0, 0.3, fem.total_mass(rho) * 1e3, freqs, st, fat, n_iter=8,
                     wall_time_min=0.1)
    for k in ("omega_target_Hz", "Vf", "mass_g", "omega_1_Hz", "omega_2_Hz", "omega_3_Hz",
              "constraint_satisfied", "sigma_max_LC1_MPa", "sigma_max_LC2_MPa", "sf_yield_LC2",
              "fs_fatigue", "n_iter", "wall_time_min", "gpu_used"):
        assert k in rec, k
    assert rec["sigma_max_LC2_MPa"] == pytest.approx(st[NAMES[1]]["vm"].max() / 1e6)
    assert rec["fs_fatigue"] == pytest.approx(fat["factor_of_safety"].min())
    assert rec["constraint_satisfied"] == bool(freqs[0] >= 300.0)
    hist = {"compliance": [1.0, 0.9], "omega_1_Hz": [float(freqs[0])] * 2, "pcg_iters": [[3, 4], [2, 2]]}
    p = save_run(str(tmp_path), 0.3, 300.0, rho, st, fat, modal, hist, rec)
    assert p == result_path(str(tmp_path), 0.3, 300.0) and p.endswith("freq_vf30_omega0300.pkl")
    pb = save_baseline(str(tmp_path), st, fat, modal, m, geom, fem.n_active)
    for path in (p, pb):
        mods = _globals(path)
        assert mods <= {"numpy", "numpy.core.multiarray", "numpy._core.multiarray",
                        "numpy.core.numeric", "numpy._core.numeric", "builtins"}, mods
    d = pickle.load(open(p, "rb"))
    assert set(d) == {"rho", "static", "fatigue", "modal", "history", "record"}
    assert np.array_equal(d["rho"], rho)
    assert d["modal"]["mode_shapes"].shape == (geom.n_dofs, 3)
    assert d["record"]["Vf"] == 0.3
    b = pickle.load(open(pb, "rb"))
    assert set(b) == {"static", "fatigue", "modal", "material", "mesh_info"}
    assert b["material"]["sigma_y"] == SIGMA_Y
    assert b["mesh_info"] == dict(nx=geom.nx, ny=geom.ny, nz=geom.nz, n_elements=geom.n_elements,
                                  n_active_elements=fem.n_active, n_nodes=geom.n_nodes,
                                  n_dofs=geom.n_dofs)
