"""Pins the oracle's zero-energy modes (SkylineMatrix2D.Factorize's zems /
zemCols, SkylineMatrix2D.cs:63-158) by what the mathematics fixes, on a
floating Hexa8 box (no supports: six rigid-body modes):

* six modes at any tolerance between 1e-6 and 1e-1 (the singular pivots
  are rounding-sized, the others of the order of the modulus);
* each mode is in the null space (K z = 0 to rounding) and is 1 at its own
  column and 0 at the other modes' columns, which fixes it uniquely:
  Z = N N_S^-1 for any basis N of the null space (here the six smallest
  eigenvectors of the dense K) and S the mode columns;
* the modes span the rigid-body motions (three translations, three
  rotations about the coordinate axes);
* with one node clamped (three translations removed) three modes remain.
A dropped term, a wrong sign or a wrong index in the vector's formation or
its back-substitution fails the closed form."""
import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd.model import box_mesh
from oracle import oracle as O


def floating(nex, ney, nez, seed, clamp_first=False):
    xyz, conn = box_mesh(nex, ney, nez)
    cons = np.zeros((xyz.shape[0], 3), bool)
    if clamp_first:
        cons[0] = True
    dof, nf = msa.enumerate_dofs(cons)
    rng = np.random.default_rng(seed)
    E = 2.1e5 * np.exp(0.3 * rng.standard_normal(conn.shape[0]))
    ke, bad = O.ke_mesh(conn, xyz, E, 0.3)
    assert bad == 0
    ri, data = O.skyline_assemble(conn, dof, nf, ke)
    return xyz, dof, nf, ri, data


def rigid_modes(xyz, dof, nf):
    R = np.zeros((nf, 6))
    for nd in range(xyz.shape[0]):
        x, y, z = xyz[nd]
        rows = [(0, [1, 0, 0]), (1, [0, 1, 0]), (2, [0, 0, 1]), (3, [0, -z, y]),
                (4, [z, 0, -x]), (5, [-y, x, 0])]
        for k, v in rows:
            for p in range(3):
                if dof[3 * nd + p] >= 0:
                    R[dof[3 * nd + p], k] = v[p]
    return R


def angles_cos(A, B):
    qa, _ = np.linalg.qr(A)
    qb, _ = np.linalg.qr(B)
    return np.linalg.svd(qa.T @ qb, compute_uv=False)


@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 2, 3)])
def test_oracle_zero_energy_modes_closed_form(shape):
    xyz, dof, nf, ri, data = floating(*shape, seed=sum(shape))
    K = O.skyline_to_dense(ri, data)
    counts = {tol: O.skyline_factorize_zem(ri, data, tol, 16)[1] for tol in (1e-6, 1e-3, 1e-1)}
    assert set(counts.values()) == {6}, counts
    fac, k, cols, Z = O.skyline_factorize_zem(ri, data, 1e-3, 16)
    assert k == 6 and np.all(np.diff(cols) > 0)
    w, V = np.linalg.eigh(K)
    assert w[5] < 1e-8 * w[-1] < w[6]
    N = V[:, :6]
    Zc = N @ np.linalg.inv(N[cols, :])      # 1 at its own column, 0 at the others
    for j in range(6):
        assert abs(Z[j, cols[j]] - 1.0) == 0.0
        assert np.abs(Z[j, np.delete(cols, j)]).max() == 0.0
        assert np.linalg.norm(K @ Z[j]) <= 1e-10 * np.linalg.norm(K) * np.linalg.norm(Z[j])
        assert np.abs(Z[j] - Zc[:, j]).max() <= 1e-8 * np.abs(Zc[:, j]).max()
    assert angles_cos(Z.T, rigid_modes(xyz, dof, nf)).min() > 1 - 1e-10
    # the factor is that of K with the mode DOFs decoupled (pivot 1, row and
    # column cleared), so SkylineMatrix2D.Solve returns the load value at a
    # mode's column and solves the other rows with the mode DOFs as supports
    f = np.random.default_rng(3).standard_normal(nf)
    x = O.skyline_solve(ri, fac, f)
    assert np.allclose(x[cols], f[cols], rtol=0, atol=1e-12 * np.abs(f).max())
    free = np.setdiff1d(np.arange(nf), cols)
    xf = np.linalg.solve(K[np.ix_(free, free)], f[free])
    assert np.abs(x[free] - xf).max() <= 1e-8 * np.abs(xf).max()


def test_oracle_zero_energy_modes_one_node_clamped():
    xyz, dof, nf, ri, data = floating(3, 3, 2, seed=9, clamp_first=True)
    _, k, cols, Z = O.skyline_factorize_zem(ri, data, 1e-3, 16)
    assert k == 3
    K = O.skyline_to_dense(ri, data)
    for j in range(3):
        assert np.linalg.norm(K @ Z[j]) <= 1e-10 * np.linalg.norm(K) * np.linalg.norm(Z[j])
    # rotations about the clamped node (the origin)
    R = rigid_modes(xyz, dof, nf)[:, 3:]
    assert angles_cos(Z.T, R).min() > 1 - 1e-10
