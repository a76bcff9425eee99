"""GPU parity at the workload's full sizes, through size-independent checks:
the true residual |f - K u|/|f| of the GPU solution evaluated with the
oracle's own CSR operator (assembled on the CPU from the oracle's Ke and the
same modulus field), and exact scaling laws (u(cE) = u(E)/c)."""
import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd import _lib as L
from msolve_amd.stochastic import kl_table, standard_normal
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_residual(cl, E, u):
    ke, bad = O.ke_mesh(cl.conn, cl.xyz, E, 0.3)
    assert bad == 0
    rp, col, val = O.csr_assemble(cl.conn, cl.dof, cl.n, ke)
    return np.linalg.norm(cl.f - O.csr_multiply(rp, col, val, u)) / np.linalg.norm(cl.f)


def test_plain_fp64_64cube_true_residual(ctx):
    """64^3 cantilever (811 200 free DOFs), E = 2.1e5, Jacobi-PCG
    to 1e-10; the oracle's operator confirms the residual."""
    print("\n[case] 64^3 fp64 single", flush=True)
    model = msa.cantilever(64)
    assert model.n_free == 811200
    E = np.full(model.n_elem, 2.1e5)
    solver = msa.SolverPCG(ctx, model, E, 0.3)
    solver.Initialize()
    u = solver.Solve(20000, 1e-10, graph_iters=32)
    assert solver.Status == msa.MSA_OK and solver.RelativeResidual < 1e-10
    cl = O.Cantilever(64)
    res = oracle_residual(cl, E, u)
    assert res < 1.5e-10, res
    solver.batch.close()
    solver.mesh.close()


def test_batched_32cube_kl_true_residuals(ctx):
    """configs[2] shape: 32^3, lognormal KL modulus (COV 0.2, 40 terms), FP64
    batched PCG to 1e-10; three samples checked with the oracle operator and
    the oracle's modulus field."""
    print("\n[case] 32^3 fp64 batch kl", flush=True)
    model = msa.cantilever(32)
    table = kl_table([1.0, 1.0, 1.0], 0.25, 40)
    ns = 24
    xi = standard_normal(20261202, ns, 40)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [ns], L.MSA_FP64, 0)
    b.set_kl(0, table, model.domain, 2.1e5, 0.2)
    b.set_xi(0, xi)
    b.set_rhs(0, model.f)
    b.run(0.3, tol=1e-10, max_it=20000, graph_iters=16)
    st = b.stats(0)
    assert (st["status"] == msa.MSA_OK).all()
    cl = O.Cantilever(32)
    for s in (0, 11, 23):
        E, bad = O.kl_field(cl.conn, cl.xyz, table, xi[s], cl.domain, 2.1e5, 0.2)
        assert bad == 0
        assert np.abs(b.modulus(0, s) - E).max() <= 1e-13 * E.max()
        assert oracle_residual(cl, E, b.solution(0, s)) < 1.5e-10
    b.close()
    mesh.close()


def test_mixed_48cube_scaling_and_residual(ctx):
    """configs[3] largest size: mixed precision to |f-Ku|/|f| < 1e-9; u(2E)
    equals u(E)/2 to the solve accuracy, residual confirmed by the oracle."""
    print("\n[case] 48^3 mixed", flush=True)
    model = msa.cantilever(48)
    rng = np.random.default_rng(5)
    s = np.sqrt(np.log(1 + 0.3 ** 2))
    E = 2.1e5 * np.exp(-0.5 * s * s + s * rng.standard_normal(model.n_elem))
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [2], L.MSA_MIXED, 0)
    b.set_modulus(0, np.stack([E, 2 * E]))
    b.set_rhs(0, model.f)
    b.assemble(0.3)
    b.solve(tol=1e-9, max_it=20000, graph_iters=16)
    st = b.stats(0)
    assert (st["status"] == msa.MSA_OK).all(), st
    u1, u2 = b.solution(0, 0), b.solution(0, 1)
    assert np.linalg.norm(2 * u2 - u1) / np.linalg.norm(u1) < 1e-8
    cl = O.Cantilever(48)
    assert oracle_residual(cl, E, u1) < 1e-9
    b.close()
    mesh.close()
