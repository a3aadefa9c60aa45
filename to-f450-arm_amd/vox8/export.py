"""Abaqus / CalculiX input deck of the modal problem.

The deck lets a solver outside this project reproduce the modal result
(the reference writes the same kind of deck, sims/export_abaqus_inp.py:57-185,
and compares its CalculiX frequencies with the in-house ones): plain Abaqus
keywords, so the same file runs in CalculiX and Abaqus/Standard.  Two models:

``simp``   every active element with the SIMP stiffness E (e_min + (1 - e_min)
           rho^p) and the Pedersen mass density of FEM3D.  A deck with one
           material per element is valid but slow for the external solvers, so
           the elements are grouped into ``n_bins`` material classes of equal
           size (at most one element apart) by rank of their stiffness, and a
           class gets the mean stiffness and the mean mass density of its
           members: the deck's total mass is the model's, and with as many
           classes as elements the deck is the model itself;
``solid``  only elements with rho > 0.5, as solid material, restricted to the
           largest body connected through voxel faces (a detached fragment
           would otherwise contribute six rigid-body modes).  The bodies are
           the connected components of the face-adjacency graph of the kept
           voxels.

Deck layout (the file format the reference's tools read): nodes referenced by
a kept element, numbered 1.. in ascending node-index order, coordinates
(ix dx, iy dy, iz dz); C3D8 elements numbered 1.. in active-element order
with the voxel node order (bottom face counter-clockwise, then top), which is
Abaqus' C3D8 order; one ELSET / MATERIAL / SOLID SECTION per class; the tip
mass split equally over the motor-mount nodes as MASS elements (when a design
has no material at the motor-mount nodes, over the highest material node in
each of their (x, y) columns, the surface the motor would rest on); the fixed
nodes clamped in 1-3; a FREQUENCY step for ``n_modes``.

Connectivity and node numbering are vectorised index arithmetic on the voxel
grid: the production arm (61,142 active elements) takes 0.3 s, a 440x64x48
grid (1.35 M elements, 179 MB deck) 6 s.  Runs on the host only; it needs the
geometry, the active-order densities and the material.
"""
from __future__ import annotations

import os

import numpy as np

__all__ = ["simp_E", "simp_rho", "active_element_grid", "write_inp"]

# SIMP constants of FEM3D (src/fem.py:204-205, 331-332); kept here so the deck can be
# written without a device context.
SIMP_PENALTY = 3.0
E_MIN_FRAC = 1e-9
MASS_THRESHOLD = 0.1
MASS_PENALTY_Q = 6.0

_LOCAL = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                   (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], dtype=np.int64)


def simp_E(rho, E, p=SIMP_PENALTY, e_min=E_MIN_FRAC):
    """Effective Young's modulus of the SIMP stiffness (src/fem.py:311-320)."""
    rho = np.asarray(rho, dtype=float)
    return E * (e_min + (1.0 - e_min) * rho ** p)


def simp_rho(rho, rho_solid, thr=MASS_THRESHOLD, q=MASS_PENALTY_Q):
    """Effective mass density with the Pedersen low-density penalty
    (src/fem.py:334-349): rho above thr, thr (rho/thr)^q below."""
    rho = np.asarray(rho, dtype=float)
    m = np.where(rho >= thr, rho, (np.maximum(rho, 0.0) / thr) ** q * thr)
    return rho_solid * m


def active_element_grid(geom):
    """(ex, ey, ez) of the active elements in the reference's active order
    (ez outer, then ey, ex inner; src/fem.py:244-247)."""
    in_arm = np.asarray(geom.build_masks()["in_arm"], dtype=bool)        # (nx, ny, nz)
    e = np.nonzero(np.ascontiguousarray(in_arm.transpose(2, 1, 0)).reshape(-1))[0]
    nx, ny = geom.nx, geom.ny
    return e % nx, (e // nx) % ny, e // (nx * ny)


def _largest_body(geom, ex, ey, ez, keep):
    """keep restricted to the largest face-connected body of kept voxels
    (ties: the body holding the lowest active index)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    idx = np.nonzero(keep)[0]
    if len(idx) < 2:
        return keep
    nx, ny = geom.nx, geom.ny
    lin = ex[idx] + nx * (ey[idx] + ny * ez[idx])           # voxel index of each kept element
    pos = np.full(geom.nx * geom.ny * geom.nz, -1, dtype=np.int64)
    pos[lin] = np.arange(len(idx))
    src, dst = [], []
    for axis, step, lim in ((ex, 1, nx), (ey, nx, ny), (ez, nx * ny, geom.nz)):
        ok = axis[idx] + 1 < lim                            # the +1 neighbour exists
        nb = pos[lin[ok] + step]
        has = nb >= 0
        src.append(np.nonzero(ok)[0][has])
        dst.append(nb[has])
    src, dst = np.concatenate(src), np.concatenate(dst)
    g = coo_matrix((np.ones(len(src)), (src, dst)), shape=(len(idx), len(idx)))
    n, label = connected_components(g, directed=False)
    if n == 1:
        return keep
    size = np.bincount(label, minlength=n)
    big = int(np.flatnonzero(size == size.max())[0])       # lowest label = lowest index on ties
    out = np.zeros_like(keep)
    out[idx[label == big]] = True
    return out


def _tip_nodes(geom, used):
    """Tip-mass nodes: the geometry's motor-mount nodes that the design keeps;
    if it keeps none, the highest kept node in each (x, y) column of the
    motor-mount nodes, where the motor would bear on the material.  Empty for
    a geometry without a motor mount."""
    if not hasattr(geom, "motor_mount_top_nodes"):
        return np.array([], dtype=np.int64)
    mount = np.asarray(geom.motor_mount_top_nodes(), dtype=np.int64)
    tip = np.intersect1d(mount, used)
    if len(tip):
        return tip
    nxy = (geom.nx + 1) * (geom.ny + 1)
    columns = np.unique(mount % nxy)                        # (x, y) columns of the mount
    col_of_used = used % nxy
    sel = np.isin(col_of_used, columns)
    cand, ccol = used[sel], col_of_used[sel]
    if len(cand) == 0:
        raise ValueError("the design keeps no material under the motor mount; cannot place the "
                         "tip mass")
    top = np.zeros(nxy, dtype=np.int64) - 1
    np.maximum.at(top, ccol, cand // nxy)                   # highest layer per column
    return np.sort(cand[cand // nxy == top[ccol]])


def _classes(E_el, D_el, n_bins):
    """Equal-size material classes by stiffness rank: (class of each element,
    class E = mean member E, class density = mean member density)."""
    n = len(E_el)
    k = max(1, min(int(n_bins), n))
    order = np.argsort(E_el, kind="stable")
    cls = np.empty(n, dtype=np.int64)
    cls[order] = (np.arange(n) * k) // n                    # sizes differ by at most one
    cnt = np.bincount(cls, minlength=k)
    return (cls, np.bincount(cls, weights=E_el, minlength=k) / cnt,
            np.bincount(cls, weights=D_el, minlength=k) / cnt)


def _lines16(ids):
    ids = [str(v) for v in ids]
    return "".join(", ".join(ids[i:i + 16]) + "\n" for i in range(0, len(ids), 16))


def write_inp(path, geom, rho_active, mode, mat, tip_mass_kg=0.06, n_modes=6, n_bins=40,
              title="F450 arm modal analysis"):
    """Write the ``mode`` ('simp' or 'solid') deck for active-order densities
    ``rho_active`` to ``path``; returns dict(nodes, elements, fixed, tip_nodes,
    path).  ``mat`` needs E, nu, rho.  tip_mass_kg = 0 writes no tip mass; a
    positive tip mass on a geometry without a motor mount is an error."""
    rho_active = np.asarray(rho_active, dtype=float)
    ex, ey, ez = active_element_grid(geom)
    if rho_active.shape != ex.shape:
        raise ValueError(f"expected {len(ex)} active densities, got {rho_active.shape}")
    if mode == "solid":
        keep = _largest_body(geom, ex, ey, ez, rho_active > 0.5)
    elif mode == "simp":
        keep = np.ones(len(rho_active), dtype=bool)
    else:
        raise ValueError(f"mode must be 'simp' or 'solid', got {mode!r}")
    e_idx = np.nonzero(keep)[0]
    if len(e_idx) == 0:
        raise ValueError("no element kept")
    nx1, ny1 = geom.nx + 1, geom.ny + 1
    conn = ((ex[e_idx, None] + _LOCAL[None, :, 0])
            + nx1 * ((ey[e_idx, None] + _LOCAL[None, :, 1]) + ny1 * (ez[e_idx, None] + _LOCAL[None, :, 2])))
    used = np.unique(conn)
    remap = np.zeros(geom.n_nodes, dtype=np.int64)
    remap[used] = np.arange(1, len(used) + 1)
    iz, iy, ix = used // (nx1 * ny1), (used % (nx1 * ny1)) // nx1, used % nx1
    fixed = np.intersect1d(np.asarray(geom.fixed_nodes(), dtype=np.int64), used)
    tip = _tip_nodes(geom, used) if tip_mass_kg > 0 else np.array([], dtype=np.int64)
    if tip_mass_kg > 0 and len(tip) == 0:
        raise ValueError("tip mass requested but the geometry has no motor-mount nodes")

    if mode == "solid":
        bins = np.zeros(len(e_idx), dtype=np.int64)
        bin_E, bin_D = np.array([mat.E]), np.array([mat.rho])
    else:
        r = rho_active[e_idx]
        bins, bin_E, bin_D = _classes(simp_E(r, mat.E), simp_rho(r, mat.rho), n_bins)

    out = [f"** F450 quadcopter arm - modal verification deck ({mode})\n",
           f"** {len(used)} nodes, {len(e_idx)} C3D8 elements, {3 * len(used)} DoF\n",
           "** Units: m, kg, N, Pa. Generated by vox8.export\n",
           f"*HEADING\n{title}\n", "*NODE\n"]
    out += [f"{n}, {x:.6e}, {y:.6e}, {z:.6e}\n"
            for n, x, y, z in zip(remap[used].tolist(), (ix * geom.dx).tolist(),
                                  (iy * geom.dy).tolist(), (iz * geom.dz).tolist())]
    out.append("*ELEMENT, TYPE=C3D8, ELSET=EALL\n")
    rc = remap[conn]
    out += [f"{k}, {c[0]}, {c[1]}, {c[2]}, {c[3]}, {c[4]}, {c[5]}, {c[6]}, {c[7]}\n"
            for k, c in enumerate(rc.tolist(), start=1)]
    for b in range(len(bin_E)):
        members = np.nonzero(bins == b)[0] + 1
        if len(members) == 0:
            continue
        out.append(f"*ELSET, ELSET=EB{b:03d}\n")
        out.append(_lines16(members.tolist()))
        out.append(f"*MATERIAL, NAME=M{b:03d}\n*ELASTIC\n{bin_E[b]:.6e}, {mat.nu}\n"
                   f"*DENSITY\n{bin_D[b]:.6e}\n*SOLID SECTION, ELSET=EB{b:03d}, MATERIAL=M{b:03d}\n")
    if len(tip):
        out.append("*NSET, NSET=NTIP\n" + _lines16(remap[tip].tolist()))
        out.append("*ELEMENT, TYPE=MASS, ELSET=ETIP\n")
        out += [f"{j}, {n}\n" for j, n in enumerate(remap[tip].tolist(), start=len(e_idx) + 1)]
        out.append(f"*MASS, ELSET=ETIP\n{tip_mass_kg / len(tip):.6e}\n")
    out.append("*NSET, NSET=NFIX\n" + _lines16(remap[fixed].tolist()))
    out.append("*BOUNDARY\nNFIX, 1, 3\n")
    out.append(f"*STEP\n*FREQUENCY, STORAGE=YES\n{n_modes}\n*NODE FILE\nU\n*END STEP\n")
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with open(path, "w") as f:
        f.write("".join(out))
    return dict(nodes=int(len(used)), elements=int(len(e_idx)), fixed=int(len(fixed)),
                tip_nodes=int(len(tip)), path=path)
