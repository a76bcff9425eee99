"""CPU restatement of the geometric multigrid preconditioner (numpy/scipy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
V-cycle (msa_batch_mg_apply) and of multigrid-preconditioned solves, never by
the msolve_amd product.

The reference has no multigrid: its PCG uses Jacobi (SolverPCG.cs:105-139,
SolverPCGSimpleSearchVectorCalculator.cs:18-41) or, for Monte Carlo, the
first sample's LDL^T (MonteCarloSolverPCGDirectMatrixCalculator.cs:38-80).
This module defines the cycle the GPU implements, written independently of
the device code from its description:

* hierarchy: a tensor-product grid mesh whose element counts are even and
  >= 4 in every direction is coarsened by keeping every other grid line;
  a coarse node inherits the constraints of the fine node it sits on; the
  coarse free DOFs are numbered node by node, component by component; a
  coarse element takes the local node order of its low-corner child and the
  mean modulus of its 8 children; coarse matrices are assembled from the
  coarse mesh with the oracle's own element and CSR routines;
* transfer: trilinear interpolation P (a fine node on an odd grid line takes
  the two neighbouring coarse lines, weighted by distance), restriction P^T;
* smoother: two Chebyshev steps preconditioned by D^-1 on [lam/ratio, lam],
  lam = max_i sum_j |K_ij| / K_ii (Gershgorin);
* cycle: pre-smoothing from zero, residual, restriction, coarse correction
  (recursive; the coarsest level solved exactly), prolongation,
  post-smoothing with the same polynomial.

Pins (tests/test_oracle_multigrid.py): with one modulus everywhere the
rediscretized coarse matrix equals the Galerkin product P^T K P (nested
trilinear spaces), which fixes P, the coarse connectivity and orientation;
the two-step smoother's error propagator equals the scaled Chebyshev
polynomial T_2 on the spectrum of D^-1 K; the cycle is symmetric and
positive definite; preconditioned CG converges to the direct solution.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

from . import oracle as O


def _axis(xyz, d):
    c = xyz[:, d]
    s = np.sort(c)
    tol = 1e-9 * max(1e-300, s[-1] - s[0])
    vals = [s[0]]
    for v in s[1:]:
        if v - vals[-1] > tol:
            vals.append(v)
    vals = np.array(vals)
    idx = np.searchsorted(vals, c - tol)
    return vals, idx


class Level:
    """One level: mesh arrays, free-DOF numbering, CSR matrix, D^-1, lam."""

    def __init__(self, xyz, conn, dof, n, E, nu=0.3):
        self.xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
        self.conn = np.asarray(conn, np.int32).reshape(-1, 8)
        self.dof = np.asarray(dof, np.int32).ravel()
        self.n = int(n)
        self.E = np.asarray(E, np.float64)
        ke, bad = O.ke_mesh(self.conn, self.xyz, self.E, nu)
        assert bad == 0
        rp, col, val = O.csr_assemble(self.conn.ravel(), self.dof, self.n, ke)
        self.K = sp.csr_matrix((val, col, rp), shape=(self.n, self.n))
        d = self.K.diagonal()
        self.dinv = 1.0 / d
        self.lam = float((abs(self.K).sum(axis=1).A1 * self.dinv).max())


def coarsen(level: Level, nu=0.3):
    """Coarse Level and the prolongation P (fine free DOFs x coarse free
    DOFs), or None when the mesh does not halve."""
    xyz, conn = level.xyz, level.conn
    g, gi = zip(*[_axis(xyz, d) for d in range(3)])
    nx, ny, nz = (len(v) for v in g)
    ex, ey, ez = nx - 1, ny - 1, nz - 1
    if nx * ny * nz != xyz.shape[0] or ex * ey * ez != conn.shape[0]:
        return None
    if min(ex, ey, ez) < 4 or ex % 2 or ey % 2 or ez % 2:
        return None
    grid = np.full((nz, ny, nx), -1)
    grid[gi[2], gi[1], gi[0]] = np.arange(xyz.shape[0])
    lo = np.stack([gi[d][conn].min(axis=1) for d in range(3)], axis=1)  # [ne, 3]
    cell = np.full((ez, ey, ex), -1)
    cell[lo[:, 2], lo[:, 1], lo[:, 0]] = np.arange(conn.shape[0])
    cx, cy, cz = ex // 2 + 1, ey // 2 + 1, ez // 2 + 1
    cnode = lambda I, J, K: I + cx * (J + cy * K)
    cxyz = np.zeros((cx * cy * cz, 3))
    fixed = np.zeros((cx * cy * cz, 3), bool)
    for K in range(cz):
        for J in range(cy):
            for I in range(cx):
                fn = grid[2 * K, 2 * J, 2 * I]
                cxyz[cnode(I, J, K)] = xyz[fn]
                fixed[cnode(I, J, K)] = level.dof[3 * fn:3 * fn + 3] < 0
    cdof = np.full(3 * cx * cy * cz, -1, np.int32)
    free = ~fixed.ravel()
    cdof[free] = np.arange(free.sum())
    nce = (ex // 2) * (ey // 2) * (ez // 2)
    cconn = np.zeros((nce, 8), np.int32)
    cE = np.zeros(nce)
    for K in range(ez // 2):
        for J in range(ey // 2):
            for I in range(ex // 2):
                ce = I + (ex // 2) * (J + (ey // 2) * K)
                kids = [cell[2 * K + (o >> 2), 2 * J + ((o >> 1) & 1), 2 * I + (o & 1)]
                        for o in range(8)]
                cE[ce] = sum(level.E[k] for k in kids) / 8.0
                e0 = kids[0]
                for a in range(8):
                    n = conn[e0, a]
                    cconn[ce, a] = cnode(I + gi[0][n] - 2 * I, J + gi[1][n] - 2 * J,
                                         K + gi[2][n] - 2 * K)
    coarse = Level(cxyz, cconn, cdof, int(free.sum()), cE, nu)
    # trilinear prolongation by grid-line distance
    def axis_w(d, i):
        if i % 2 == 0:
            return [(i // 2, 1.0)]
        x0, x1, x2 = g[d][i - 1], g[d][i], g[d][i + 1]
        return [((i - 1) // 2, (x2 - x1) / (x2 - x0)), ((i + 1) // 2, (x1 - x0) / (x2 - x0))]
    rows, cols, vals = [], [], []
    for fn in range(xyz.shape[0]):
        for (a, wa) in axis_w(0, gi[0][fn]):
            for (b, wb) in axis_w(1, gi[1][fn]):
                for (c, wc) in axis_w(2, gi[2][fn]):
                    cn = cnode(a, b, c)
                    for d in range(3):
                        r, q = level.dof[3 * fn + d], cdof[3 * cn + d]
                        if r >= 0 and q >= 0:
                            rows.append(r)
                            cols.append(q)
                            vals.append(wa * wb * wc)
    P = sp.csr_matrix((vals, (rows, cols)), shape=(level.n, coarse.n))
    return coarse, P


def hierarchy(xyz, conn, dof, n, E, nu=0.3, max_levels=0):
    """[Level 0, Level 1, ...] and [P_0, P_1, ...] (P_l: level l+1 -> l)."""
    levels = [Level(xyz, conn, dof, n, E, nu)]
    Ps = []
    while max_levels <= 0 or len(Ps) < max_levels:
        nxt = coarsen(levels[-1], nu)
        if nxt is None:
            break
        levels.append(nxt[0])
        Ps.append(nxt[1])
    return levels, Ps


def cheb_coef(lam, ratio):
    hi, lo = lam, lam / ratio
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    sigma = theta / delta
    rho0 = 1.0 / sigma
    rho1 = 1.0 / (2.0 * sigma - rho0)
    return 1.0 / theta, rho1 * rho0, 2.0 * rho1 / delta


def smooth(L: Level, b, x, ratio):
    """Two Chebyshev-Jacobi steps from x (None: zero)."""
    ith, c1, c2 = cheb_coef(L.lam, ratio)
    r = b.copy() if x is None else b - L.K @ x
    x0 = np.zeros_like(b) if x is None else x
    d1 = ith * L.dinv * r
    x1 = x0 + d1
    r1 = b - L.K @ x1
    return x1 + c1 * d1 + c2 * L.dinv * r1


def vcycle(levels, Ps, b, ratio=12.0, l=0, chol=None):
    """z = B b.  chol: cached Cholesky factor of the coarsest matrix."""
    L = levels[l]
    if l == len(levels) - 1:
        c = chol if chol is not None else sla.cho_factor(L.K.toarray(), lower=True)
        return sla.cho_solve(c, b)
    x = smooth(L, b, None, ratio)
    r = b - L.K @ x
    xc = vcycle(levels, Ps, Ps[l].T @ r, ratio, l + 1, chol)
    x = x + Ps[l] @ xc
    return smooth(L, b, x, ratio)


def mg_pcg(levels, Ps, f, tol=1e-10, max_it=500, ratio=12.0):
    """PCG (SolverPCG.cs:105-139 recurrence, stopping |r|/|f| < tol) with
    the V-cycle as preconditioner; returns (x, iterations)."""
    chol = sla.cho_factor(levels[-1].K.toarray(), lower=True)
    K = levels[0].K
    x = np.zeros_like(f)
    r = f.copy()
    z = vcycle(levels, Ps, r, ratio, 0, chol)
    p = z.copy()
    rz = r @ z
    fn = np.linalg.norm(f)
    for it in range(1, max_it + 1):
        q = K @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(r) / fn < tol:
            return x, it
        z = vcycle(levels, Ps, r, ratio, 0, chol)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, max_it
