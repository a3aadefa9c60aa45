"""Run results in the reference's pickle schema (sims/run_freq_gpu.py:95-260).

The reference's runner saves, per (V_f, omega_target) run,
``freq_vf{VF:02d}_omega{OMEGA:04d}.pkl`` = dict(rho, static, fatigue, modal,
history, record) and once ``raw_baseline_modal.pkl`` = dict(static, fatigue,
modal, material, mesh_info); its consolidation and figure scripts
(sims/consolidate_v6.py:53-80 and others) read them back.  The functions here
produce the same dictionaries from the device path, so those scripts consume
vox8 runs unchanged:

* ``static_analysis``: per load case u, sigma (Voigt rows), vm, sigma_max
  (largest principal), disp_node, sf_yield = sigma_y / max(vm, 1e-3),
  solve_time.  All load cases are solved as one block PCG (the reference
  solves them one after another); stresses and invariants on the device.
* ``fatigue_analysis``: the Goodman / S-N fields of the first two load cases.
* ``modal_results``, ``run_record``, ``save_run``, ``save_baseline``.

Every value stored is a NumPy array or a Python scalar/str/dict/list, so the
files load without vox8 installed (the reference pickles its Material object;
here ``material`` is the equivalent plain dict, which none of the reference's
readers uses).
"""
from __future__ import annotations

import dataclasses
import os
import pickle
import time

import numpy as np
import torch

from .fatigue import fatigue_life_field

__all__ = ["static_analysis", "fatigue_analysis", "modal_results", "run_record",
           "result_path", "save_run", "save_baseline"]


def _np(x):
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def static_analysis(fem, F_list, rho_active, lc_names, sigma_y, rtol=1e-5, maxiter=5000):
    """sims/run_freq_gpu.py:95-116 on the device: dict load-case name ->
    dict(u, sigma, vm, sigma_max, disp_node, sf_yield, solve_time), NumPy
    arrays.  rtol 1e-5 is the reference's effective cg tolerance
    (src/fem.py:558-561); solve_time is the block solve's time split evenly
    over the load cases."""
    names = list(lc_names)
    F = np.stack([np.asarray(f, dtype=float) for f in F_list])
    if F.shape[0] != len(names):
        raise ValueError(f"{F.shape[0]} load vectors for {len(names)} names")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    U, info = fem.solve_block(rho_active, F, rtol=rtol, maxiter=maxiter)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / len(names)
    if any(i != 0 for i in info["info"]):
        raise RuntimeError(f"static solve did not converge: {info}")
    out = {}
    for j, name in enumerate(names):
        u = U[j]
        sigma = fem.compute_stress(u, rho_active)
        vm = fem.von_mises(sigma)
        smax = fem.principal_stress_max(sigma)
        vm_np = _np(vm)
        out[name] = dict(u=_np(u), sigma=_np(sigma), vm=vm_np, sigma_max=_np(smax),
                         disp_node=np.linalg.norm(_np(u).reshape(-1, 3), axis=1),
                         sf_yield=sigma_y / np.maximum(vm_np, 1e-3), solve_time=dt)
    return out


def fatigue_analysis(static, sigma_u, lc_pair=None):
    """Goodman / S-N fields from the von Mises stresses of two load cases
    (sims/run_freq_gpu.py:119-122: hover and manoeuvre, i.e. the first two)."""
    a, b = lc_pair if lc_pair is not None else list(static)[:2]
    return {k: _np(v) for k, v in fatigue_life_field(static[a]["vm"], static[b]["vm"],
                                                     sigma_u).items()}


def modal_results(freqs_Hz, mode_shapes, omega_sq, tip_mass_kg):
    """The ``modal`` entry: frequencies, (n_dofs, k) shapes, eigenvalues, tip mass."""
    return dict(frequencies_Hz=_np(freqs_Hz), mode_shapes=_np(mode_shapes),
                omega_sq=_np(omega_sq), tip_mass_kg=float(tip_mass_kg))


def _tag(name):
    """Record key of a load case: its name up to the first underscore
    ("LC3_landing" -> "LC3")."""
    return str(name).split("_")[0]


def run_record(omega_target_Hz, vf, mass_g, freqs_Hz, static, fatigue, n_iter, wall_time_min,
               gpu_used=True):
    """The ``record`` entry (sims/run_freq_gpu.py:201-220): sigma_max_<LC>_MPa
    (peak von Mises) and sf_yield_<LC> (minimum yield safety factor) for every
    load case, the reference's keys among them."""
    f = [float(x) for x in _np(freqs_Hz)]
    rec = dict(omega_target_Hz=float(omega_target_Hz), Vf=float(vf), mass_g=float(mass_g),
               omega_1_Hz=f[0], omega_2_Hz=f[1], omega_3_Hz=f[2],
               constraint_satisfied=bool(f[0] >= omega_target_Hz))
    for name, s in static.items():
        rec[f"sigma_max_{_tag(name)}_MPa"] = float(np.max(s["vm"]) / 1e6)
    for name, s in static.items():
        rec[f"sf_yield_{_tag(name)}"] = float(np.min(s["sf_yield"]))
    rec.update(fs_fatigue=float(np.min(fatigue["factor_of_safety"])), n_iter=int(n_iter),
               wall_time_min=float(wall_time_min), gpu_used=bool(gpu_used))
    return rec


def result_path(results_dir, vf, omega_target_Hz):
    """``freq_vf{VF:02d}_omega{OMEGA:04d}.pkl`` (sims/run_freq_gpu.py:225-228)."""
    return os.path.join(results_dir, f"freq_vf{int(vf * 100):02d}_omega{int(omega_target_Hz):04d}.pkl")


def _history(h):
    return {k: (list(v) if isinstance(v, (list, tuple)) else v) for k, v in h.items()}


def save_run(results_dir, vf, omega_target_Hz, rho, static, fatigue, modal, history, record):
    """Write one run (sims/run_freq_gpu.py:229-240); returns the path."""
    os.makedirs(results_dir, exist_ok=True)
    path = result_path(results_dir, vf, omega_target_Hz)
    with open(path, "wb") as f:
        pickle.dump(dict(rho=_np(rho), static=static, fatigue=fatigue, modal=modal,
                         history=_history(history), record=record), f)
    return path


def save_baseline(results_dir, static, fatigue, modal, material, geom, n_active):
    """Write ``raw_baseline_modal.pkl`` (sims/run_freq_gpu.py:242-258); returns
    the path."""
    os.makedirs(results_dir, exist_ok=True)
    path = os.path.join(results_dir, "raw_baseline_modal.pkl")
    mat = dataclasses.asdict(material) if dataclasses.is_dataclass(material) else dict(material)
    mesh = dict(nx=int(geom.nx), ny=int(geom.ny), nz=int(geom.nz),
                n_elements=int(geom.n_elements), n_active_elements=int(n_active),
                n_nodes=int(geom.n_nodes), n_dofs=int(geom.n_dofs))
    with open(path, "wb") as f:
        pickle.dump(dict(static=static, fatigue=fatigue, modal=modal, material=mat,
                         mesh_info=mesh), f)
    return path
