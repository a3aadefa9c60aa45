"""GPU: element stress recovery and stress invariants (vox8_stress,
vox8_stress_invariants) against the oracle's restatement of
FEM3D.compute_stress / von_mises / principal_stress_max (src/fem.py:564-602).
FP64 throughout: bar 1e-12 relative for stresses and von Mises; the closed-form
principal stress against LAPACK eigvalsh within 1e-9 of the stress scale."""
import numpy as np
import pytest
import torch

from helpers import MaskedBox, random_density, rel

pytestmark = pytest.mark.gpu

MAT = dict(E=1.7e9, nu=0.4, rho=1010.0)


def _pair(geom):
    from oracle.fem import FEM3D as OFEM, Material as OMat
    from vox8.fem import FEM3D, Material
    return OFEM(geom, OMat(**MAT)), FEM3D(geom, Material(**MAT))


@pytest.mark.parametrize("shape,hole", [((9, 4, 3), 0.2), ((30, 7, 5), 0.05)])
def test_stress_matches_oracle(gpu, shape, hole):
    geom = MaskedBox(*shape, hole_frac=hole, seed=4)
    ofem, vfem = _pair(geom)
    rng = np.random.default_rng(1)
    u = rng.standard_normal(geom.n_dofs) * 1e-5
    rho = random_density(ofem.n_active, 2, 0.05, 1.0)
    ref = ofem.compute_stress(u, rho)
    got = vfem.compute_stress(u, rho)                       # numpy in, numpy out
    assert got.shape == ref.shape == (ofem.n_active, 6)
    assert rel(got, ref) < 1e-12
    # device tensors in: device tensor out
    gt = vfem.compute_stress(torch.tensor(u, device=gpu), torch.tensor(rho, device=gpu))
    assert gt.is_cuda and rel(gt.cpu().numpy(), ref) < 1e-12
    vm = vfem.von_mises(got)
    assert rel(vm, ofem.von_mises(ref)) < 1e-12
    pm = vfem.principal_stress_max(got)
    scale = np.abs(ref).max()
    assert np.abs(pm - ofem.principal_stress_max(ref)).max() < 1e-9 * scale


def test_invariants_special_tensors(gpu):
    """Repeated and zero eigenvalues, pure shear, hydrostatic: the closed form
    against eigvalsh."""
    from vox8.fem import FEM3D
    rng = np.random.default_rng(5)
    rows = [[1, 1, 1, 0, 0, 0], [0, 0, 0, 0, 0, 0], [2, 2, -1, 0, 0, 0], [0, 0, 0, 1, 0, 0],
            [0, 0, 0, 0, 1, 1], [-3, -3, -3, 0, 0, 0], [5, 5, 5, 1e-9, 0, 0]]
    S = np.vstack([np.array(rows, dtype=float), rng.standard_normal((200, 6))]) * 1e7
    T = np.empty((S.shape[0], 3, 3))
    T[:, 0, 0], T[:, 1, 1], T[:, 2, 2] = S[:, 0], S[:, 1], S[:, 2]
    T[:, 0, 1] = T[:, 1, 0] = S[:, 3]
    T[:, 1, 2] = T[:, 2, 1] = S[:, 4]
    T[:, 0, 2] = T[:, 2, 0] = S[:, 5]
    lmax = np.linalg.eigvalsh(T).max(axis=1)
    pm = FEM3D.principal_stress_max(S)
    assert np.abs(pm - lmax).max() < 1e-9 * np.abs(S).max()
    vm = FEM3D.von_mises(S)
    s = S
    vm_ref = np.sqrt(np.maximum(0.5 * ((s[:, 0] - s[:, 1]) ** 2 + (s[:, 1] - s[:, 2]) ** 2
                                       + (s[:, 2] - s[:, 0]) ** 2) + 3 * (s[:, 3:] ** 2).sum(1), 0))
    assert rel(vm, vm_ref) < 1e-13


def test_fatigue_known_answers_and_oracle(gpu):
    """vox8_fatigue: the reference's OEM known answer and hand cases
    (src/fatigue.py:124-141), and random stress fields against the oracle."""
    import json
    from conftest import GOLDEN
    from oracle import fatigue as of
    from vox8 import fatigue as vf
    d = json.loads((GOLDEN / "reference_notes.json").read_text())["oem_baseline"]
    r = vf.fatigue_life_field([d["max_vm_MPa_LC1"] * 1e6], [d["max_vm_MPa_LC2"] * 1e6], d["sigma_u_MPa"] * 1e6)
    assert r["factor_of_safety"][0] == pytest.approx(d["sf_fatigue"], rel=1e-14)
    assert r["life"][0] == d["N_cycles"]
    for (alt, mean), eq, life in [((15e6, 15e6), 21.818182e6, 5.513165e4), ((6e6, 6e6), 6.857143e6, 1e9),
                                  ((40e6, 5e6), 44.651163e6, 111.1802)]:
        e = vf.equivalent_alt_stress_goodman(np.array([alt]), np.array([mean]), 48e6)
        assert e[0] == pytest.approx(eq, rel=1e-7)
        assert vf.sn_life(e)[0] == pytest.approx(life, rel=1e-6)
    rng = np.random.default_rng(3)
    s1 = np.abs(rng.standard_normal(5000)) * 2e7
    s2 = s1 * rng.uniform(-1, 1, 5000)
    s1[:3] = [0.0, 11.8e6, 70e6]          # zero, below endurance, gross overload
    a, b = vf.fatigue_life_field(s1, s2, 48e6), of.fatigue_life_field(s1, s2, 48e6)
    for k in b:
        assert rel(a[k], b[k]) < 1e-12, k
    t = vf.fatigue_life_field(torch.tensor(s1, device=gpu), torch.tensor(s2, device=gpu), 48e6)
    assert t["life"].is_cuda and rel(t["life"].cpu().numpy(), b["life"]) < 1e-12
