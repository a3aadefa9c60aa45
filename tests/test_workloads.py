"""Workload generator (vox8.workloads) vs the oracle's independent restatement,
plus closed-form checks of the load cases (resultant force and moment)."""
import numpy as np
import pytest

from oracle import workloads as ow
from oracle.geometry import Cantilever as OCant
from vox8.workloads import ALL_LC, WORKLOADS, Cantilever, c4_designs, load_cases, seeded_density


@pytest.mark.parametrize("shape", [(110, 16, 12, 2e-3), (7, 4, 5, 1e-3)])
def test_loads_match_oracle_and_statics(shape):
    g, og = Cantilever(*shape), OCant(*shape)
    F = load_cases(g, ALL_LC)
    for j, name in enumerate(ALL_LC):
        ref = ow.load_vector(og, name)
        np.testing.assert_allclose(F[j], ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    t = g.tip_nodes()
    iy = (t % ((g.nx + 1) * (g.ny + 1))) // (g.nx + 1)
    iz = t // ((g.nx + 1) * (g.ny + 1))
    y = iy * g.dy - 0.5 * g.ny * g.dy
    z = iz * g.dz - 0.5 * g.nz * g.dz
    res = lambda f: (f[0::3].sum(), f[1::3].sum(), f[2::3].sum())
    np.testing.assert_allclose(res(F[0]), (0, 0, -10), atol=1e-12)
    np.testing.assert_allclose(res(F[1]), (0, 0, 0), atol=1e-12)
    Mx = np.sum(y * F[1][3 * t + 2] - z * F[1][3 * t + 1])
    assert abs(Mx - 0.5) < 1e-12
    np.testing.assert_allclose(res(F[2]), (0, 5, -8.660254), atol=1e-12)
    np.testing.assert_allclose(np.hypot(*res(F[2])[1:]), 10.0, rtol=1e-7)
    np.testing.assert_allclose(res(F[3]), (0, 3, 0), atol=1e-12)
    np.testing.assert_allclose(F[4], 0.6 * F[0] + 0.4 * F[1], atol=1e-15)


def test_density_and_counts():
    for wl in WORKLOADS.values():
        g = wl.geometry()
        x = seeded_density(g, wl.base, wl.amp, wl.seed)
        np.testing.assert_array_equal(x, ow.seeded_density(g.n_elements, wl.base, wl.amp, wl.seed))
        assert x.min() >= 1e-3 and x.max() <= 1.0
    # exact sizes quoted for the workloads
    c1, c2, c5 = WORKLOADS["c1"].geometry(), WORKLOADS["c2"].geometry(), WORKLOADS["c5"].geometry()
    assert (c1.n_elements, c1.n_nodes, c1.n_dofs) == (21120, 24531, 73593)
    assert (c2.n_elements, c2.n_nodes, c2.n_dofs) == (168960, 182325, 546975)
    assert (c5.n_elements, c5.n_nodes, c5.n_dofs) == (1351680, 1404585, 4213755)
    assert 3 * len(c2.fixed_nodes()) == 546975 - 544500
    nxs = [d.nx for d in c4_designs()]
    assert nxs == [80, 89, 99, 108, 117, 127, 136, 145, 155, 164, 173, 183, 192, 201, 211, 220]
    assert sum(d.nx * 32 * 24 for d in c4_designs()) == 1843200
