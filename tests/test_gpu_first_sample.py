"""GPU parity of the first-sample direct preconditioner (the Monte Carlo
calculator of MonteCarloSolverPCGDirectMatrixCalculator.cs:38-80: every
sample's PCG is preconditioned by the LDL^T factorisation of the FIRST
sample's stiffness) against the oracle's restatement of the reference loop
(orc_mc_reference: per sample KL modulus -> Ke -> skyline K, PCG
preconditioned by the skyline LDL^T of the first sample's K).

The same preconditioner gives the same Krylov iteration: iteration counts
agree within one (dot products are summed in a different order), and the
first sample, whose preconditioner is exact, converges in at most two
products.  FP64 displacements agree to 1e-10 at tol 1e-13; a mixed batch
reaches the oracle's displacements to 1e-8 at |f-Ku|/|f| < 1e-10.  8^3 runs
both band sweeps in one launch (1944 rows), 10^3 (3630 rows) per 64-column
block."""
import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd import _lib as L
from msolve_amd.stochastic import kl_table, standard_normal
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("n,precision", [(8, L.MSA_FP64), (10, L.MSA_FP64), (8, L.MSA_MIXED),
                                         (10, L.MSA_MIXED)])
def test_first_sample_preconditioner_matches_reference_loop(ctx, n, precision):
    print(f"\n[case] first-sample preconditioner {n}^3 precision={precision}", flush=True)
    ns, M = 5, 12
    table = kl_table([1.0, 1.0, 1.0], 0.25, M)
    otab = O.kl_table([1.0, 1.0, 1.0], 0.25, M)
    xi = standard_normal(77 + n, ns, M)
    tol = 1e-13 if precision == L.MSA_FP64 else 1e-10
    ref = O.mc_reference(n, otab, xi, cov=0.3, tol=1e-13, keep_u=True, storage="skyline")
    model = msa.cantilever(n)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [ns], precision, 1)
    b.set_preconditioner(L.MSA_PRECOND_FIRST_SAMPLE)
    b.set_kl(0, table, model.domain, 2.1e5, 0.3)
    b.set_xi(0, xi)
    b.set_rhs(0, model.f)
    b.run(0.3, tol=tol, max_it=2000, graph_iters=8)
    st = b.stats(0)
    assert (st["status"] == msa.MSA_OK).all(), st
    errs = [rel(b.solution(0, s), ref["u"][s]) for s in range(ns)]
    print(f"[first-sample] iterations GPU {st['iters'].tolist()} oracle {ref['iters'].tolist()}; "
          f"displacement errors {['%.1e' % e for e in errs]}", flush=True)
    if precision == L.MSA_FP64:
        assert np.abs(st["iters"] - ref["iters"]).max() <= 1
        assert st["iters"][0] <= 2
        assert max(errs) < 1e-10
    else:
        assert max(errs) < 1e-8
    # far fewer iterations than Jacobi on the same samples
    bj = msa.Batch(ctx, [mesh], [ns], precision, 1)
    bj.set_kl(0, table, model.domain, 2.1e5, 0.3)
    bj.set_xi(0, xi)
    bj.set_rhs(0, model.f)
    bj.run(0.3, tol=tol, max_it=5000, graph_iters=8)
    assert st["iters"].max() < bj.stats(0)["iters"].min() / 3
    bj.close()
    b.close()
    mesh.close()


def test_first_sample_then_direct_solver(ctx):
    """The shared factor and the direct solver's per-sample factors use the
    same storage: switching between them re-factors what each needs."""
    print("\n[case] first-sample preconditioner and direct solver", flush=True)
    n, ns, M = 6, 3, 8
    table = kl_table([1.0, 1.0, 1.0], 0.25, M)
    otab = O.kl_table([1.0, 1.0, 1.0], 0.25, M)
    xi = standard_normal(5, ns, M)
    ref = O.mc_reference(n, otab, xi, cov=0.3, tol=1e-13, keep_u=True, storage="skyline")
    model = msa.cantilever(n)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [ns], L.MSA_FP64, 1)
    b.set_preconditioner(L.MSA_PRECOND_FIRST_SAMPLE)
    b.set_kl(0, table, model.domain, 2.1e5, 0.3)
    b.set_xi(0, xi)
    b.set_rhs(0, model.f)
    b.run(0.3, tol=1e-13, max_it=2000, graph_iters=8)
    first = [b.solution(0, s) for s in range(ns)]
    b.set_solver(L.MSA_SOLVER_DIRECT)
    b.direct_solve()
    direct = [b.solution(0, s) for s in range(ns)]
    b.set_solver(L.MSA_SOLVER_PCG)
    b.solve(tol=1e-13, max_it=2000, graph_iters=8)
    again = [b.solution(0, s) for s in range(ns)]
    for s in range(ns):
        assert rel(first[s], ref["u"][s]) < 1e-10
        assert rel(direct[s], ref["u"][s]) < 1e-10
        assert rel(again[s], ref["u"][s]) < 1e-10
    assert b.stats(0)["iters"].max() < 40
    b.close()
    mesh.close()
