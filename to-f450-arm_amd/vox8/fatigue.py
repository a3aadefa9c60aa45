"""Goodman / S-N fatigue post-processing (src/fatigue.py:20-121) on the device.

`fatigue_life_field` runs the `vox8_fatigue` kernel (one thread per element:
Goodman mean-stress correction and log-log S-N interpolation); `sn_life` and
`equivalent_alt_stress_goodman` are the same kernel on single fields.  Inputs
may be numpy arrays, lists or device tensors; numpy in, numpy out.
`sn_alt_stress_at` (a scalar of the S-N table) stays on the host.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import call, f64p, ptr

# PA12 SLS S-N curve at R = -1 (src/fatigue.py:20-21)
SN_CYCLES = np.array([1e2, 1e3, 1e4, 1e5, 1e6, 5e6, 1e7], dtype=float)
SN_ALT_MPA = np.array([45.0, 38.0, 28.0, 20.0, 13.0, 12.0, 12.0], dtype=float)


def _interp_table():
    """Stress -> life table in log-log space with strictly increasing stress
    abscissae (repeated plateau values separated by 1e-12)."""
    xs = np.log10(SN_ALT_MPA[::-1]).copy()
    ys = np.log10(SN_CYCLES[::-1])
    for i in range(1, len(xs)):
        if xs[i] <= xs[i - 1]:
            xs[i] = xs[i - 1] + 1e-12
    return np.ascontiguousarray(xs), np.ascontiguousarray(ys)


def sn_alt_stress_at(N):
    xs = np.log10(SN_CYCLES)
    ys = np.log10(SN_ALT_MPA)
    return float(10.0 ** np.interp(np.log10(N), xs, ys, left=ys[0], right=ys[-1]))


def _dev(x):
    if isinstance(x, torch.Tensor):
        t = x if x.is_cuda else x.to("cuda")
        return t.to(torch.float64).reshape(-1).contiguous(), (lambda y: y)
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    return torch.from_numpy(a).to("cuda"), (lambda y: y.cpu().numpy())


def _run(s1, s2, sigma_uts_Pa, want):
    a, back = _dev(s1)
    b, _ = _dev(s2)
    if a.numel() != b.numel():
        raise ValueError("stress fields differ in length")
    n = a.numel()
    out = {k: torch.empty(n, dtype=torch.float64, device=a.device) for k in want}
    xs, ys = _interp_table()
    st = C.c_void_p(torch.cuda.current_stream(a.device).cuda_stream)
    call("vox8_fatigue", st, n, ptr(a), ptr(b), float(sigma_uts_Pa), f64p(xs), f64p(ys),
         float(SN_ALT_MPA[-1]), float(SN_ALT_MPA[0]), sn_alt_stress_at(1e6),
         *[ptr(out.get(k)) for k in ("sigma_mean", "sigma_alt", "sigma_alt_eq", "life", "damage",
                                     "factor_of_safety")])
    return {k: back(v) for k, v in out.items()}


def sn_life(sigma_alt_eq_Pa):
    """Cycles to failure at an equivalent alternating stress (src/fatigue.py:24-41)."""
    s, back = _dev(sigma_alt_eq_Pa)
    s = s.clamp(min=0.0)            # non-positive stresses: below the endurance limit
    return back(_dev_field(s, -s, 1e300, "life"))


def equivalent_alt_stress_goodman(sigma_alt, sigma_mean, sigma_uts_Pa):
    """Goodman-corrected alternating stress (src/fatigue.py:44-60), sigma_alt >= 0."""
    a, back = _dev(sigma_alt)
    m, _ = _dev(sigma_mean)
    return back(_dev_field(m + a, m - a, sigma_uts_Pa, "sigma_alt_eq"))


def _dev_field(s1, s2, uts, key):
    return _run(s1, s2, uts, (key,))[key]


def fatigue_life_field(stress_lc1, stress_lc2, sigma_uts_Pa):
    """Per-element fatigue from the two load-case stress fields (src/fatigue.py:63-121)."""
    return _run(stress_lc1, stress_lc2, sigma_uts_Pa,
                ("sigma_mean", "sigma_alt", "sigma_alt_eq", "life", "damage", "factor_of_safety"))
