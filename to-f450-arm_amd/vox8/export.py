"""Abaqus / CalculiX input deck of the modal problem (sims/export_abaqus_inp.py:57-185).

The deck lets a solver outside this project reproduce the modal result: plain
Abaqus keywords, so the same file runs in CalculiX and Abaqus/Standard.  Two
models, as in the reference:

``simp``   every active element with the SIMP stiffness E (e_min + (1 - e_min)
           rho^p) and the Pedersen mass density of FEM3D baked into
           per-element materials, binned into ``n_bins`` density classes;
``solid``  only elements with rho > 0.5, as solid material, restricted to the
           largest face-connected body (a detached fragment would otherwise
           contribute six rigid-body modes).

Deck layout (same keywords, numbering and order as the reference's): nodes
referenced by a kept element, numbered 1.. in ascending node-index order,
coordinates (ix dx, iy dy, iz dz); C3D8 elements numbered 1.. in active-element
order with the voxel node order (bottom face counter-clockwise, then top),
which is Abaqus' C3D8 order; one ELSET / MATERIAL / SOLID SECTION per
non-empty bin; the tip mass split equally over the motor-mount nodes as MASS
elements; the fixed nodes clamped in 1-3; a FREQUENCY step for ``n_modes``.

The connectivity and node numbering are vectorised index arithmetic on the
voxel grid (the reference loops over elements in Python): the production arm
(61,142 active elements) takes 0.3 s, a 440x64x48 grid (1.35 M elements,
179 MB deck) 6 s.  Runs on the host only; it needs the geometry, the
active-order densities and the material.
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
    from scipy import ndimage
    g = np.zeros((geom.nx, geom.ny, geom.nz), dtype=bool)
    g[ex[keep], ey[keep], ez[keep]] = True
    lab, n = ndimage.label(g)
    if n > 1:
        main = int(np.argmax(np.bincount(lab.ravel())[1:])) + 1
        keep = keep & (lab[ex, ey, ez] == main)
    return keep


def _tip_nodes(geom, used):
    """Motor-mount nodes among the used ones; when the top face is absent, the
    highest used nodes within the motor-pad radius (sims/export_abaqus_inp.py:
    104-127).  Empty for a geometry without a motor mount."""
    if not hasattr(geom, "motor_mount_top_nodes"):
        return np.array([], dtype=np.int64)
    tip = np.intersect1d(np.asarray(geom.motor_mount_top_nodes(), dtype=np.int64), used)
    if len(tip):
        return tip
    nxy = (geom.nx + 1) * (geom.ny + 1)
    ix, iy = used % (geom.nx + 1), (used % nxy) // (geom.nx + 1)
    r = np.hypot(ix * geom.dx - geom.motor_centre[0], iy * geom.dy - geom.motor_centre[1])
    near = used[r <= geom.motor_dia / 2.0 + geom.motor_keep_ring]
    if len(near) == 0:
        raise ValueError("no nodes near the motor mount; cannot place the tip mass")
    zt = near // nxy
    return near[zt >= zt.max() - 1]


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
        edges = np.linspace(r.min(), r.max() + 1e-12, n_bins + 1)
        bins = np.clip(np.digitize(r, edges) - 1, 0, n_bins - 1)
        centres = 0.5 * (edges[:-1] + edges[1:])
        bin_E, bin_D = simp_E(centres, mat.E), simp_rho(centres, mat.rho)

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
