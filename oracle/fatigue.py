"""Oracle: Goodman mean-stress correction and PA12 S-N life (test infrastructure only).

Restates src/fatigue.py:20-121 (S-N table :20-21, sn_life :24-60,
equivalent_alt_stress_goodman :63-75, fatigue_life_field :78-112,
sn_alt_stress_at :115-121).
"""
from __future__ import annotations

import numpy as np

SN_CYCLES = np.array([1e2, 1e3, 1e4, 1e5, 1e6, 5e6, 1e7])
SN_ALT_MPA = np.array([45.0, 38.0, 28.0, 20.0, 13.0, 12.0, 12.0])


def sn_life(sigma_alt_eq_Pa):
    s = np.asarray(sigma_alt_eq_Pa, dtype=float) / 1e6
    ls = np.log10(SN_ALT_MPA)[::-1].copy()
    ln = np.log10(SN_CYCLES)[::-1]
    for i in range(1, len(ls)):          # strictly increasing abscissae (ties nudged by 1e-12)
        if ls[i] <= ls[i - 1]:
            ls[i] = ls[i - 1] + 1e-12
    logn = np.interp(np.log10(np.where(s > 0, s, 1e-10)), ls, ln, left=ln[0], right=ln[-1])
    life = 10.0 ** logn
    life = np.where(s < SN_ALT_MPA[-1] * 0.99, 1e9, life)
    return np.where(s > SN_ALT_MPA[0] * 1.5, 1.0, life)


def equivalent_alt_stress_goodman(sigma_alt, sigma_mean, sigma_uts_Pa):
    denom = np.maximum(1.0 - np.maximum(sigma_mean, 0.0) / sigma_uts_Pa, 1e-6)
    return sigma_alt / denom


def sn_alt_stress_at(N):
    return float(10.0 ** np.interp(np.log10(N), np.log10(SN_CYCLES), np.log10(SN_ALT_MPA),
                                   left=np.log10(SN_ALT_MPA)[0], right=np.log10(SN_ALT_MPA)[-1]))


def fatigue_life_field(stress_lc1, stress_lc2, sigma_uts_Pa):
    s1 = np.asarray(stress_lc1, dtype=float)
    s2 = np.asarray(stress_lc2, dtype=float)
    hi, lo = np.maximum(s1, s2), np.minimum(s1, s2)
    sm, sa = 0.5 * (hi + lo), 0.5 * (hi - lo)
    eq = equivalent_alt_stress_goodman(sa, sm, sigma_uts_Pa)
    life = sn_life(eq)
    return {"sigma_mean": sm, "sigma_alt": sa, "sigma_alt_eq": eq, "life": life,
            "damage": 1.0 / np.maximum(life, 1.0),
            "factor_of_safety": sn_alt_stress_at(1e6) * 1e6 / np.maximum(eq, 1e-6)}
