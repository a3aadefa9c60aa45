"""CPU: Abaqus / CalculiX deck export (vox8.export.write_inp).

The deck is parsed back and pinned two ways.  Structure: node coordinates
from the node index, connectivity against the geometry's own
``element_nodes`` for every kept element (Abaqus C3D8 order), the kept body
against scipy.ndimage's component labelling, the material classes (equal
sizes, ordered by stiffness, class E and density the means of their members'
SIMP and Pedersen values, total mass preserved), tip-mass nodes and value,
clamped nodes, the frequency step.  Physics: K and M assembled from nothing but the
parsed deck (oracle element matrices per material) give the natural
frequencies of the in-house model of the same structure (oracle FEM3D with
its SIMP interpolation and tip mass): to rounding for the solid deck of a
fully dense design, to the binning error for the SIMP deck.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from helpers import TipBox

MAT = dict(E=1.7e9, nu=0.39, rho=930.0)


class _Mat:
    E, nu, rho = MAT["E"], MAT["nu"], MAT["rho"]


def parse_inp(path):
    """Minimal reader of the keywords the exporter writes."""
    nodes, elems, elsets, mats, sections = {}, {}, {}, {}, {}
    nsets, mass_elems, mass_value, boundary, step_modes = {}, [], None, [], None
    lines = [l.rstrip("\n") for l in open(path) if not l.startswith("**")]
    i, kw, cur = 0, None, None
    while i < len(lines):
        l = lines[i].strip()
        i += 1
        if l.startswith("*"):
            parts = [p.strip() for p in l[1:].split(",")]
            kw = parts[0].upper()
            opts = dict(p.split("=", 1) for p in parts[1:] if "=" in p)
            cur = opts
            if kw == "MATERIAL":
                mats[opts["NAME"]] = {}
            elif kw == "ELASTIC":
                E, nu = [float(x) for x in lines[i].split(",")]
                mats[list(mats)[-1]].update(E=E, nu=nu)
                i += 1
            elif kw == "DENSITY":
                mats[list(mats)[-1]]["rho"] = float(lines[i])
                i += 1
            elif kw == "SOLID SECTION":
                sections[opts["ELSET"]] = opts["MATERIAL"]
            elif kw == "MASS":
                mass_value = float(lines[i])
                i += 1
            elif kw == "FREQUENCY":
                step_modes = int(lines[i])
                i += 1
            elif kw == "ELSET":
                elsets[opts["ELSET"]] = []
            elif kw == "NSET":
                nsets[opts["NSET"]] = []
            continue
        vals = [v.strip() for v in l.split(",")]
        if kw == "NODE":
            nodes[int(vals[0])] = tuple(float(v) for v in vals[1:])
        elif kw == "ELEMENT" and cur.get("TYPE") == "C3D8":
            elems[int(vals[0])] = [int(v) for v in vals[1:]]
        elif kw == "ELEMENT" and cur.get("TYPE") == "MASS":
            mass_elems.append(int(vals[1]))
        elif kw == "ELSET":
            elsets[cur["ELSET"]] += [int(v) for v in vals]
        elif kw == "NSET":
            nsets[cur["NSET"]] += [int(v) for v in vals]
        elif kw == "BOUNDARY":
            boundary.append(vals)
    return dict(nodes=nodes, elems=elems, elsets=elsets, mats=mats, sections=sections,
                nsets=nsets, mass_elems=mass_elems, mass_value=mass_value, boundary=boundary,
                step_modes=step_modes)


def _design(geom, seed):
    from vox8.export import active_element_grid
    ex, _, _ = active_element_grid(geom)
    return np.random.default_rng(seed).uniform(0.05, 1.0, len(ex))


@pytest.mark.parametrize("mode", ["simp", "solid"])
def test_deck_structure(tmp_path, mode):
    from vox8.export import active_element_grid, simp_E, simp_rho, write_inp
    geom = TipBox(12, 4, 4, h=1e-3, hole_frac=0.08, seed=11)
    rho = _design(geom, 5)
    info = write_inp(str(tmp_path / "d.inp"), geom, rho, mode, _Mat, tip_mass_kg=2e-3, n_modes=6)
    d = parse_inp(info["path"])
    ex, ey, ez = active_element_grid(geom)
    keep = np.ones(len(rho), bool) if mode == "simp" else rho > 0.5
    if mode == "solid":
        from scipy import ndimage
        g = np.zeros((geom.nx, geom.ny, geom.nz), bool)
        g[ex[keep], ey[keep], ez[keep]] = True
        lab, n = ndimage.label(g)
        big = np.argmax(np.bincount(lab.ravel())[1:]) + 1
        keep &= lab[ex, ey, ez] == big
    kept = np.nonzero(keep)[0]
    assert info["elements"] == len(d["elems"]) == len(kept)
    # node numbering: ascending node index; coordinates from the index
    conn = np.array([geom.element_nodes(ex[e], ey[e], ez[e]) for e in kept])
    used = np.unique(conn)
    assert info["nodes"] == len(d["nodes"]) == len(used)
    nx1, ny1 = geom.nx + 1, geom.ny + 1
    for k, nid in enumerate(used, start=1):
        x = np.array([nid % nx1, (nid // nx1) % ny1, nid // (nx1 * ny1)]) * 1e-3
        assert np.allclose(d["nodes"][k], x, rtol=0, atol=1e-12)
    back = {k + 1: nid for k, nid in enumerate(used)}
    for k, e in enumerate(kept, start=1):
        assert [back[v] for v in d["elems"][k]] == list(conn[k - 1])
    # materials per element
    elem_mat = {}
    for name, members in d["elsets"].items():
        if name == "EALL":
            continue
        for m in members:
            elem_mat[m] = d["mats"][d["sections"][name]]
    assert sorted(elem_mat) == list(range(1, len(kept) + 1))
    if mode == "simp":
        r = rho[kept]
        E_el, D_el = simp_E(r, MAT["E"]), simp_rho(r, MAT["rho"])
        members = {}
        for k in range(1, len(kept) + 1):
            members.setdefault(id(elem_mat[k]), []).append(k - 1)
        groups = sorted(members.values(), key=lambda g: E_el[g].min())
        sizes = [len(g) for g in groups]
        assert len(groups) == min(40, len(kept)) and max(sizes) - min(sizes) <= 1
        for a, b in zip(groups, groups[1:]):
            assert E_el[a].max() <= E_el[b].min()
        for g in groups:
            m = elem_mat[g[0] + 1]
            assert m["E"] == pytest.approx(E_el[g].mean(), rel=1e-6)
            assert m["rho"] == pytest.approx(D_el[g].mean(), rel=1e-6)
            assert m["nu"] == MAT["nu"]
        mass = sum(elem_mat[k]["rho"] for k in range(1, len(kept) + 1))
        assert mass == pytest.approx(D_el.sum(), rel=1e-6)
    else:
        assert all(m["E"] == pytest.approx(MAT["E"]) and m["rho"] == pytest.approx(MAT["rho"])
                   for m in elem_mat.values())
    # tip mass, clamp, step
    tip = np.intersect1d(geom.motor_mount_top_nodes(), used)
    assert sorted(back[v] for v in d["nsets"]["NTIP"]) == list(tip)
    assert sorted(back[v] for v in d["mass_elems"]) == list(tip)
    assert d["mass_value"] == pytest.approx(2e-3 / len(tip), rel=1e-6)
    fixed = np.intersect1d(geom.fixed_nodes(), used)
    assert sorted(back[v] for v in d["nsets"]["NFIX"]) == list(fixed)
    assert d["boundary"] == [["NFIX", "1", "3"]]
    assert d["step_modes"] == 6


def _deck_frequencies(d, nu, k):
    """Lowest k natural frequencies of the parsed deck (dense, free DOFs)."""
    from oracle.element import element_K, element_M
    nn = len(d["nodes"])
    X = np.array([d["nodes"][i + 1] for i in range(nn)])
    elem_mat = {m: d["mats"][d["sections"][s]] for s, ms in d["elsets"].items() if s != "EALL"
                for m in ms}
    rows, cols, kv, mv = [], [], [], []
    for e, c in d["elems"].items():
        c = np.array(c) - 1
        h = X[c[6]] - X[c[0]]
        m = elem_mat[e]
        dofs = (3 * c[:, None] + np.arange(3)).ravel()
        rows.append(np.repeat(dofs, 24))
        cols.append(np.tile(dofs, 24))
        kv.append(element_K(m["E"], nu, *h).ravel())
        mv.append(element_M(m["rho"], *h).ravel())
    n = 3 * nn
    K = sp.coo_matrix((np.concatenate(kv), (np.concatenate(rows), np.concatenate(cols))), (n, n)).toarray()
    M = sp.coo_matrix((np.concatenate(mv), (np.concatenate(rows), np.concatenate(cols))), (n, n)).toarray()
    for v in d["mass_elems"]:
        for c in range(3):
            M[3 * (v - 1) + c, 3 * (v - 1) + c] += d["mass_value"]
    fixed = np.array(d["nsets"]["NFIX"]) - 1
    free = np.setdiff1d(np.arange(n), np.concatenate([3 * fixed, 3 * fixed + 1, 3 * fixed + 2]))
    lam = sla.eigh(K[np.ix_(free, free)], M[np.ix_(free, free)], eigvals_only=True,
                   subset_by_index=[0, k - 1])
    return np.sqrt(lam) / (2 * np.pi)


def _inhouse_frequencies(geom, rho, k, tip):
    from oracle.fem import FEM3D, Material
    f = FEM3D(geom, Material(E=MAT["E"], nu=MAT["nu"], rho=MAT["rho"]))
    K = f.assemble_K(rho)
    M = f.assemble_M(rho, tip_mass_kg=tip, tip_mass_nodes=geom.motor_mount_top_nodes())
    fr = f.free_dofs
    lam = sla.eigh(K[fr][:, fr].toarray(), M[fr][:, fr].toarray(), eigvals_only=True,
                   subset_by_index=[0, k - 1])
    return np.sqrt(lam) / (2 * np.pi)


def test_solid_deck_reproduces_inhouse_frequencies(tmp_path):
    """Fully dense design: the solid deck is the in-house model exactly."""
    from vox8.export import write_inp
    geom = TipBox(10, 4, 3, h=1e-3, hole_frac=0.0, seed=2)
    rho = np.ones(geom.n_elements)
    info = write_inp(str(tmp_path / "s.inp"), geom, rho, "solid", _Mat, tip_mass_kg=2e-3)
    f_deck = _deck_frequencies(parse_inp(info["path"]), MAT["nu"], 6)
    f_in = _inhouse_frequencies(geom, rho, 6, 2e-3)
    assert np.allclose(f_deck, f_in, rtol=1e-6), (f_deck, f_in)


def test_simp_deck_frequencies_within_binning_error(tmp_path):
    """Graded design: with fine bins the SIMP deck converges to the in-house
    model; 40 bins (the default) stay within a few percent."""
    from vox8.export import write_inp
    geom = TipBox(10, 4, 3, h=1e-3, hole_frac=0.1, seed=4)
    from vox8.export import active_element_grid
    rho = np.random.default_rng(9).uniform(0.3, 1.0, len(active_element_grid(geom)[0]))
    f_in = _inhouse_frequencies(geom, rho, 6, 2e-3)
    errs = []
    for nb in (40, 2000):
        info = write_inp(str(tmp_path / f"b{nb}.inp"), geom, rho, "simp", _Mat, tip_mass_kg=2e-3,
                         n_bins=nb)
        f_deck = _deck_frequencies(parse_inp(info["path"]), MAT["nu"], 6)
        errs.append(np.max(np.abs(f_deck / f_in - 1)))
    assert errs[0] < 0.03 and errs[1] < 1e-3 and errs[1] < errs[0], errs


def test_largest_body_and_tip_fallback(tmp_path):
    """Two bodies of kept voxels (a larger clamped bar and a detached block):
    the solid deck keeps the larger, face-connected one only; with the top
    layer removed, the tip mass moves to the highest kept node of each
    motor-mount column."""
    from vox8.export import active_element_grid, write_inp
    geom = TipBox(10, 4, 4, h=1e-3, hole_frac=0.0, seed=1)
    ex, ey, ez = active_element_grid(geom)
    rho = np.full(len(ex), 0.01)
    rho[(ey <= 1) & (ez <= 2)] = 1.0                      # bar along x, three layers high
    rho[(ey == 3) & (ez == 3) & (ex >= 6) & (ex <= 8)] = 1.0   # detached block (edge contact only)
    info = write_inp(str(tmp_path / "s.inp"), geom, rho, "solid", _Mat, tip_mass_kg=2e-3)
    assert info["elements"] == int(((ey <= 1) & (ez <= 2)).sum())
    d = parse_inp(info["path"])
    nxy = (geom.nx + 1) * (geom.ny + 1)
    used = np.unique(np.concatenate([geom.element_nodes(ex[e], ey[e], ez[e])
                                     for e in np.nonzero((ey <= 1) & (ez <= 2))[0]]))
    back = {k + 1: nid for k, nid in enumerate(used)}
    tip = sorted(back[v] for v in d["nsets"]["NTIP"])
    mount_cols = np.unique(np.asarray(geom.motor_mount_top_nodes()) % nxy)
    want = [n for n in used if n % nxy in mount_cols and n // nxy == 3]
    assert tip == sorted(want) and len(tip) > 0


def test_rejects_bad_input(tmp_path):
    from vox8.export import write_inp
    geom = TipBox(6, 3, 3, h=1e-3, hole_frac=0.0, seed=1)
    with pytest.raises(ValueError):
        write_inp(str(tmp_path / "x.inp"), geom, np.ones(5), "simp", _Mat)
    with pytest.raises(ValueError):
        write_inp(str(tmp_path / "x.inp"), geom, np.ones(geom.n_elements), "both", _Mat)
    with pytest.raises(ValueError):
        write_inp(str(tmp_path / "x.inp"), geom, np.zeros(geom.n_elements), "solid", _Mat)
