"""Edge cases of the solve path on the GPU: a zero load (the reference's
SolverPCG returns x = 0 without iterating: its |r|/|b| test is 0/0), the
smallest mesh (one Hexa8 element), ragged lane tiles (a last tile with one
live lane), an empty element batch, and the argument checks of the C ABI
(zero samples, no free DOF, bad lane tile, too many classes)."""
import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd import _lib as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def direct(m, E):
    ke, bad = O.ke_mesh(m.conn, m.xyz, E, 0.3)
    assert bad == 0
    ri, data = O.skyline_assemble(m.conn, m.dof, m.n_free, ke)
    fac, rc = O.skyline_factorize(ri, data, 1e-15)
    assert rc == 0
    return O.skyline_solve(ri, fac, m.f)


@pytest.mark.parametrize("precision,precond", [(L.MSA_FP64, L.MSA_PRECOND_JACOBI),
                                               (L.MSA_MIXED, L.MSA_PRECOND_JACOBI),
                                               (L.MSA_MIXED, L.MSA_PRECOND_MULTIGRID)])
def test_zero_load_returns_zero(ctx, precision, precond):
    print(f"\n[case] zero load precision={precision} precond={precond}", flush=True)
    model = msa.cantilever(4)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [2], precision, 1)
    if precond != L.MSA_PRECOND_JACOBI:
        b.set_preconditioner(precond)
    b.set_modulus(0, np.full((2, model.n_elem), 2.1e5))
    b.set_rhs(0, np.zeros(model.n_free))
    b.assemble(0.3)
    b.solve(tol=1e-10, max_it=100, graph_iters=8)
    st = b.stats(0)
    assert (st["status"] == msa.MSA_OK).all() and (st["iters"] == 0).all(), st
    for s in range(2):
        assert not np.any(b.solution(0, s))
    b.close()
    mesh.close()


def test_one_element_mesh(ctx):
    print("\n[case] one-element mesh", flush=True)
    model = msa.cantilever(1)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    E = np.array([[2.1e5], [3.3e5]])
    for solver in (L.MSA_SOLVER_PCG, L.MSA_SOLVER_DIRECT):
        b = msa.Batch(ctx, [mesh], [2], L.MSA_FP64, 1)
        b.set_solver(solver)
        b.set_modulus(0, E)
        b.set_rhs(0, model.f)
        b.assemble(0.3)
        b.solve(tol=1e-13, max_it=200)
        assert (b.stats(0)["status"] == msa.MSA_OK).all()
        for s in range(2):
            assert rel(b.solution(0, s), direct(model, E[s])) < 1e-12
        b.close()
    mesh.close()


@pytest.mark.parametrize("tile,ns", [(4, 5), (64, 3), (2, 65)])
def test_ragged_lane_tiles(ctx, tile, ns):
    """ns samples in lane tiles of `tile`: the last tile has ns % tile live
    lanes; every sample solves to the oracle's direct solution."""
    print(f"\n[case] ragged tiles tile={tile} samples={ns}", flush=True)
    model = msa.cantilever(3, 2, 2)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    rng = np.random.default_rng(tile + ns)
    E = 2.1e5 * np.exp(0.3 * rng.standard_normal((ns, model.n_elem)))
    b = msa.Batch(ctx, [mesh], [ns], L.MSA_FP64, tile)
    b.set_modulus(0, E)
    b.set_rhs(0, model.f)
    b.assemble(0.3)
    b.solve(tol=1e-12, max_it=2000, graph_iters=8)
    assert (b.stats(0)["status"] == msa.MSA_OK).all()
    for s in (0, ns // 2, ns - 1):
        assert rel(b.solution(0, s), direct(model, E[s])) < 1e-10
    b.close()
    mesh.close()


def test_argument_checks(ctx):
    print("\n[case] argument checks", flush=True)
    ke, st = msa.ke_packed(ctx, np.zeros((0, 24)), np.zeros(0), 0.3)
    assert ke.shape == (0, 300) and st.shape == (0,)
    model = msa.cantilever(2)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    with pytest.raises(msa.MsaError):
        msa.Batch(ctx, [mesh], [0], L.MSA_FP64, 1)
    with pytest.raises(msa.MsaError):
        msa.Batch(ctx, [mesh], [2], L.MSA_FP64, 3)
    with pytest.raises(msa.MsaError):
        msa.Batch(ctx, [mesh] * 33, [1] * 33, L.MSA_FP64, 1)
    cons = np.ones((model.n_nodes, 3), bool)     # every DOF fixed
    dof, nf = msa.enumerate_dofs(cons)
    fixed = msa.Mesh(ctx, model.xyz, model.conn, dof, nf)
    with pytest.raises(msa.MsaError, match="no free DOF"):
        msa.Batch(ctx, [fixed], [1], L.MSA_FP64, 1)
    fixed.close()
    b = msa.Batch(ctx, [mesh], [1], L.MSA_FP64, 1)
    with pytest.raises(msa.MsaError):
        b.solve()                               # not assembled
    with pytest.raises(msa.MsaError):
        b.set_rhs(0, np.zeros(model.n_free + 1))
    b.close()
    mesh.close()
