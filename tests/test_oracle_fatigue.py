"""Oracle fatigue (oracle/fatigue.py) against the reference's committed OEM
known answer (notes/oem_baseline.json) and the hand-worked cases of
src/fatigue.py:124-141 (the device implementation is checked against the same
cases in tests/test_gpu_post.py)."""
import json

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import fatigue as of


def _impls():
    return (of,)


@pytest.mark.parametrize("impl", _impls())
def test_oem_known_answer(impl):
    d = json.loads((GOLDEN / "reference_notes.json").read_text())["oem_baseline"]
    r = impl.fatigue_life_field([d["max_vm_MPa_LC1"] * 1e6], [d["max_vm_MPa_LC2"] * 1e6],
                                d["sigma_u_MPa"] * 1e6)
    assert r["factor_of_safety"][0] == pytest.approx(d["sf_fatigue"], rel=1e-14)
    assert r["life"][0] == d["N_cycles"]


@pytest.mark.parametrize("impl", _impls())
def test_hand_cases(impl):
    cases = [((15e6, 15e6), 21.818182e6, 5.513165e4), ((6e6, 6e6), 6.857143e6, 1e9),
             ((40e6, 5e6), 44.651163e6, 111.1802)]
    for (alt, mean), eq, life in cases:
        e = impl.equivalent_alt_stress_goodman(np.array([alt]), np.array([mean]), 48e6)
        assert e[0] == pytest.approx(eq, rel=1e-7)
        assert impl.sn_life(e)[0] == pytest.approx(life, rel=1e-6)
    assert impl.sn_alt_stress_at(1e6) == pytest.approx(13.0, rel=1e-14)
    # closed form inside a table segment: log-log line between (1e4, 28) and (1e5, 20)
    s = 24.0
    n = 10 ** (4 + (np.log10(s) - np.log10(28)) / (np.log10(20) - np.log10(28)))
    assert impl.sn_life(np.array([s * 1e6]))[0] == pytest.approx(n, rel=1e-12)
