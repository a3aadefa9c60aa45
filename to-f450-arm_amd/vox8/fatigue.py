"""Goodman / S-N fatigue post-processing (src/fatigue.py:20-121), unchanged.

Element-wise scalar arithmetic on von Mises fields; kept on the host because
it is a few vector operations on already-reduced data.
"""
from __future__ import annotations

import numpy as np

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
    return xs, ys


def sn_life(sigma_alt_eq_Pa):
    s = np.asarray(sigma_alt_eq_Pa, dtype=float) / 1e6
    xs, ys = _interp_table()
    log_s = np.log10(np.where(s > 0, s, 1e-10))
    life = 10.0 ** np.interp(log_s, xs, ys, left=ys[0], right=ys[-1])
    life = np.where(s < SN_ALT_MPA[-1] * 0.99, 1e9, life)        # below endurance
    return np.where(s > SN_ALT_MPA[0] * 1.5, 1.0, life)           # gross overload


def equivalent_alt_stress_goodman(sigma_alt, sigma_mean, sigma_uts_Pa):
    return sigma_alt / np.maximum(1.0 - np.maximum(sigma_mean, 0.0) / sigma_uts_Pa, 1e-6)


def sn_alt_stress_at(N):
    xs = np.log10(SN_CYCLES)
    ys = np.log10(SN_ALT_MPA)
    return float(10.0 ** np.interp(np.log10(N), xs, ys, left=ys[0], right=ys[-1]))


def fatigue_life_field(stress_lc1, stress_lc2, sigma_uts_Pa):
    s1 = np.asarray(stress_lc1, dtype=float)
    s2 = np.asarray(stress_lc2, dtype=float)
    peak, valley = np.maximum(s1, s2), np.minimum(s1, s2)
    mean, alt = 0.5 * (peak + valley), 0.5 * (peak - valley)
    eq = equivalent_alt_stress_goodman(alt, mean, sigma_uts_Pa)
    life = sn_life(eq)
    return {"sigma_mean": mean, "sigma_alt": alt, "sigma_alt_eq": eq, "life": life,
            "damage": 1.0 / np.maximum(life, 1.0),
            "factor_of_safety": sn_alt_stress_at(1e6) * 1e6 / np.maximum(eq, 1e-6)}
