"""GPU parity of the direct solver's zero-energy modes
(SkylineMatrix2D.Factorize's zems / zemCols, SkylineMatrix2D.cs:63-158) on
floating Hexa8 boxes (no supports, six rigid-body modes), against the
oracle's restatement, which the CPU tests pin by the closed form:

* a cube's band order is the DOF order, so the GPU finds the oracle's mode
  columns and vectors (1e-8 of the largest entry) and the direct solve
  returns the oracle's SkylineMatrix2D.Solve result (1e-9);
* on a 4 x 2 x 3 box (another band order) the modes are in the null space
  and are the closed form N N_S^-1 for the GPU's own columns;
* a sample with more modes than kept gets MSA_BREAKDOWN."""
import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd import _lib as L
from msolve_amd.model import box_mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def floating(shape, ns, seed):
    xyz, conn = box_mesh(*shape)
    dof, nf = msa.enumerate_dofs(np.zeros((xyz.shape[0], 3), bool))
    rng = np.random.default_rng(seed)
    E = 2.1e5 * np.exp(0.3 * rng.standard_normal((ns, conn.shape[0])))
    return xyz, conn, dof, nf, E


def run(ctx, xyz, conn, dof, nf, E, f, max_modes=16):
    mesh = msa.Mesh(ctx, xyz, conn, dof, nf)
    b = msa.Batch(ctx, [mesh], [E.shape[0]], L.MSA_FP64, 0)
    b.set_modulus(0, E)
    b.set_rhs(0, f)
    b.assemble(0.3)
    b.set_solver(L.MSA_SOLVER_DIRECT)
    b.set_zem(1e-3, max_modes)
    b.solve()
    return mesh, b


@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 2, 3)])
def test_zero_energy_modes_match_oracle(ctx, shape):
    print(f"\n[case] zero-energy modes {shape}", flush=True)
    ns = 3
    xyz, conn, dof, nf, E = floating(shape, ns, sum(shape))
    f = np.random.default_rng(5).standard_normal(nf)
    mesh, b = run(ctx, xyz, conn, dof, nf, E, f)
    st = b.stats(0)
    assert (st["status"] == msa.MSA_OK).all(), st
    for s in range(ns):
        ke, bad = O.ke_mesh(conn, xyz, E[s], 0.3)
        ri, data = O.skyline_assemble(conn, dof, nf, ke)
        K = O.skyline_to_dense(ri, data)
        fac, k, ocols, oZ = O.skyline_factorize_zem(ri, data, 1e-3, 16)
        cols, Z = b.zem(0, s)
        print(f"[zem] sample {s}: GPU columns {cols.tolist()} oracle {ocols.tolist()}", flush=True)
        assert k == 6 and len(cols) == 6
        for j in range(6):
            assert np.linalg.norm(K @ Z[j]) <= 1e-10 * np.linalg.norm(K) * np.linalg.norm(Z[j])
        if shape[0] == shape[1] == shape[2]:
            assert np.array_equal(cols, ocols)
            assert np.abs(Z - oZ).max() <= 1e-8 * np.abs(oZ).max()
            x = O.skyline_solve(ri, fac, f)
            u = b.solution(0, s)
            assert np.linalg.norm(u - x) <= 1e-9 * np.linalg.norm(x)
        else:
            w, V = np.linalg.eigh(K)
            N = V[:, :6]
            Zc = (N @ np.linalg.inv(N[cols, :])).T
            assert np.abs(Z - Zc).max() <= 1e-7 * np.abs(Zc).max()
    b.close()
    mesh.close()


def test_more_modes_than_kept_is_breakdown(ctx):
    print("\n[case] zero-energy modes over the limit", flush=True)
    xyz, conn, dof, nf, E = floating((2, 2, 2), 2, 1)
    mesh, b = run(ctx, xyz, conn, dof, nf, E, np.ones(nf), max_modes=4)
    assert (b.stats(0)["status"] == msa.MSA_BREAKDOWN).all()
    with pytest.raises(msa.MsaError):
        b.zem(0, 0)
    b.close()
    mesh.close()
