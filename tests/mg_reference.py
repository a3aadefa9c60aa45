"""NumPy restatement of the vox8 multigrid V-cycle (test infrastructure).

Same hierarchy and operators as to-f450-arm_amd/csrc/vox8_mg.hip, written
with assembled sparse matrices so the preconditioner can be checked on the
CPU (symmetry, positivity, contraction) and compared with the device kernel:
levels coarsened 2:1 (ceil), factor = mean of 8 children, element matrix
2^l Ke, trilinear P (masked by fixed DOFs on both sides; a coarse DOF
is fixed when any fine node of its 3x3x3 interpolation stencil has it fixed), R = P^T, Chebyshev
smoothing of D^-1 K on [lam_max/alpha, lam_max] (nu steps, nu_coarse from
level nu_coarse_level on; the three-term recurrence on the iterates, as the
device runs it), dense coarsest solve; cycle
index gamma (1 = V, 2 = W with a fresh residual after each coarse correction).
galerkin=True replaces the rediscretised coarse operators by P^T K P (sparse
matrix products), the device's Galerkin option (vox8_set_mg_galerkin).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

LOCAL = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)])


def assemble(nx, ny, nz, f, Ke, fixed):
    """Free-DOF-embedded K (identity-free: rows/cols of fixed DOFs zero)."""
    NX, NY = nx + 1, ny + 1
    ez, ey, ex = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    nodes = (ex[:, None] + LOCAL[None, :, 0]) + NX * ((ey[:, None] + LOCAL[None, :, 1]) + NY * (ez[:, None] + LOCAL[None, :, 2]))
    dofs = (3 * nodes[:, :, None] + np.arange(3)).reshape(-1, 24)
    rows = np.repeat(dofs, 24, axis=1).ravel()
    cols = np.tile(dofs, (1, 24)).ravel()
    vals = (f[:, None] * Ke.ravel()[None, :]).ravel()
    n = 3 * NX * NY * (nz + 1)
    K = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    free = ~np.repeat(fixed, 3).astype(bool) if fixed.dtype == bool else np.array(
        [not ((fixed[i // 3] >> (i % 3)) & 1) for i in range(n)])
    D = sp.diags(free.astype(float))
    return (D @ K @ D).tocsr(), free


def prolongation(g_f, g_c, free_f, free_c):
    (nxf, nyf, nzf), (nxc, nyc, nzc) = g_f, g_c

    def axis(nf, nc):
        P = np.zeros((nf + 1, nc + 1))
        for i in range(nf + 1):
            if i % 2 == 0:
                P[i, i // 2] = 1.0
            else:
                P[i, (i - 1) // 2] = 0.5
                P[i, (i + 1) // 2] = 0.5
        return sp.csr_matrix(P)
    Pn = sp.kron(axis(nzf, nzc), sp.kron(axis(nyf, nyc), axis(nxf, nxc)))
    P = sp.kron(Pn, sp.eye(3)).tocsr()
    return (sp.diags(free_f.astype(float)) @ P @ sp.diags(free_c.astype(float))).tocsr()


class VCycle:
    def __init__(self, dims, f, Ke, fixed_nodes_mask, active=None, nu=2, alpha=20.0, coarse_dofs=700,
                 gamma=2, w_min_dofs=0, galerkin=False, nu_coarse=0, nu_coarse_level=1):
        self.nu, self.alpha, self.gamma, self.w_min_dofs = nu, alpha, gamma, w_min_dofs
        self.nu_coarse, self.nu_coarse_level = nu_coarse, nu_coarse_level
        nx, ny, nz = dims
        fixed = fixed_nodes_mask.astype(np.uint8) * 7 if fixed_nodes_mask.dtype == bool else fixed_nodes_mask
        act = np.ones(nx * ny * nz, bool) if active is None else active.astype(bool)
        self.levels = []
        scale = 1.0
        Kg = None
        while True:
            K, free = assemble(nx, ny, nz, f, scale * Ke, fixed)
            if Kg is not None:
                K = Kg
            d = K.diagonal()
            dinv = np.where(free, 1.0 / np.where(d != 0, d, 1.0), 0.0)
            self.levels.append(dict(dims=(nx, ny, nz), K=K, free=free, dinv=dinv))
            n = 3 * (nx + 1) * (ny + 1) * (nz + 1)
            if n <= coarse_dofs or min(nx, ny, nz) < 2:
                break
            cx, cy, cz = (nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2
            fc = np.zeros(cx * cy * cz)
            ac = np.zeros(cx * cy * cz, bool)
            for Z in range(cz):
                for Y in range(cy):
                    for X in range(cx):
                        s, a = 0.0, False
                        for k in range(8):
                            x, y, z = 2 * X + (k & 1), 2 * Y + ((k >> 1) & 1), 2 * Z + (k >> 2)
                            if x < nx and y < ny and z < nz:
                                e = x + nx * (y + ny * z)
                                s += f[e]
                                a |= act[e]
                        fc[X + cx * (Y + cy * Z)] = s / 8
                        ac[X + cx * (Y + cy * Z)] = a
            used = np.zeros((cz + 1, cy + 1, cx + 1), bool)
            a3 = ac.reshape(cz, cy, cx)
            for oz in (0, 1):
                for oy in (0, 1):
                    for ox in (0, 1):
                        used[oz:oz + cz, oy:oy + cy, ox:ox + cx] |= a3
            fx3 = fixed.reshape(nz + 1, ny + 1, nx + 1)
            fc_fix = np.zeros((cz + 1, cy + 1, cx + 1), np.uint8)
            for K_ in range(cz + 1):
                for J in range(cy + 1):
                    for I in range(cx + 1):
                        blk = fx3[max(2 * K_ - 1, 0):2 * K_ + 2, max(2 * J - 1, 0):2 * J + 2,
                                  max(2 * I - 1, 0):2 * I + 2]
                        fc_fix[K_, J, I] = np.bitwise_or.reduce(blk.ravel()) if blk.size else 0
            fc_fix[~used] = 7
            fixed_c = fc_fix.ravel()
            _, free_c = assemble(cx, cy, cz, fc, Ke, fixed_c)
            self.levels[-1]["P"] = prolongation((nx, ny, nz), (cx, cy, cz), free, free_c)
            if galerkin:
                Pl = self.levels[-1]["P"]
                Kg = (Pl.T @ self.levels[-1]["K"] @ Pl).tocsr()
            nx, ny, nz, f, fixed, act = cx, cy, cz, fc, fixed_c, ac
            scale *= 2.0
        for L in self.levels[:-1]:
            A = sp.diags(L["dinv"]) @ L["K"]
            L["lam_max"] = 1.1 * float(np.max(np.abs(np.linalg.eigvals(A.toarray())))) if A.shape[0] < 3000 \
                else 1.1 * lanczos_lam_max(L["K"], L["dinv"])
        C = self.levels[-1]
        Kc = C["K"].toarray()
        Kc[~C["free"], :] = 0
        Kc[:, ~C["free"]] = 0
        Kc[~C["free"], ~C["free"]] = 1.0
        C["inv"] = np.linalg.inv(Kc)

    def smooth(self, L, x, b, nu):
        """nu Chebyshev steps of the first kind on [lam_max/alpha, lam_max] in
        the three-term form on the iterates (the device's form):
        x_1 = x_0 + D^-1 (b - K x_0) / theta, then
        x_{k+1} = x_k + rho_k rho_{k-1} (x_k - x_{k-1}) + 2 rho_k / delta D^-1 (b - K x_k)."""
        lmax = L["lam_max"]
        lmin = lmax / self.alpha
        theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
        sigma = theta / delta
        dinv, K = L["dinv"], L["K"]
        xprev, x = x, x + dinv * (b - K @ x) / theta
        rho = 1.0 / sigma
        for _ in range(nu - 1):
            rho1 = 1.0 / (2 * sigma - rho)
            xprev, x = x, x + rho1 * rho * (x - xprev) + 2 * rho1 / delta * dinv * (b - K @ x)
            rho = rho1
        return x

    def __call__(self, b, l=0):
        """Cycle of index gamma: pre-smoothing from zero, gamma coarse-grid
        corrections each on a fresh fine residual (one when the next level is
        the exact coarsest solve or this level has fewer than w_min_dofs
        DOFs), post-smoothing."""
        L = self.levels[l]
        if l == len(self.levels) - 1:
            return np.where(L["free"], L["inv"] @ b, 0.0)
        nu = self.nu_coarse if (self.nu_coarse > 0 and l >= self.nu_coarse_level) else self.nu
        x = self.smooth(L, np.zeros_like(b), b, nu)
        g = self.gamma if (l + 1 < len(self.levels) - 1 and L["K"].shape[0] >= self.w_min_dofs) else 1
        for _ in range(g):
            r = b - L["K"] @ x
            x = x + L["P"] @ self(L["P"].T @ r, l + 1)
        return self.smooth(L, x, b, nu)


def hash_start(n):
    """The device's Lanczos start vector: a fixed hash of the DOF index in [-0.5, 0.5)."""
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = i * np.uint64(0x9E3779B97F4A7C15) + np.uint64(0x632BE59BD9B4E019)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return ((z >> np.uint64(40)).astype(np.float64) / 16777216.0 - 0.5).astype(np.float32)


def lanczos_lam_max(K, dinv, steps=20):
    """Largest Ritz value of D^-1 K after `steps` Lanczos steps in the D inner
    product from hash_start, FP32 vectors and FP64 sums as on the device."""
    dinv = np.asarray(dinv, dtype=np.float32)
    free = dinv > 0
    d = np.where(free, 1.0 / np.where(free, dinv, 1.0), 0.0)
    q = np.where(free, hash_start(K.shape[0]), 0.0).astype(np.float32)
    q = (q / np.sqrt(np.sum(q.astype(np.float64) ** 2 * d))).astype(np.float32)
    qp = np.zeros_like(q)
    beta = 0.0
    a, b = [], []
    for _ in range(steps):
        w = (dinv * (K @ q.astype(np.float64)).astype(np.float32) - np.float32(beta) * qp).astype(np.float32)
        s1 = np.sum(w.astype(np.float64) * q * d)
        s2 = np.sum(w.astype(np.float64) ** 2 * d)
        s3 = np.sum(q.astype(np.float64) ** 2 * d)
        alpha = s1 / s3
        beta = np.sqrt(max(s2 - s1 * s1 / s3, 0.0))
        a.append(alpha)
        b.append(beta)
        if not beta > 1e-7 * max(abs(x) for x in a):
            break
        qp, q = q, ((w - np.float32(alpha) * q) / np.float32(beta)).astype(np.float32)
    m = len(a)
    T = np.diag(a) + np.diag(b[:m - 1], 1) + np.diag(b[:m - 1], -1)
    return float(np.linalg.eigvalsh(T)[-1])


def exact_lam_max(K, dinv):
    """lambda_max(D^-1 K) from the symmetric form D^-1/2 K D^-1/2."""
    s = np.sqrt(np.asarray(dinv, dtype=np.float64))
    S = (sp.diags(s) @ K @ sp.diags(s)).tocsr()
    if S.shape[0] < 3000:
        return float(np.linalg.eigvalsh(S.toarray())[-1])
    import scipy.sparse.linalg as sla
    return float(sla.eigsh(S, k=1, which="LA", tol=1e-10)[0][0])
