"""GPU parity of the direct solver (banded Cholesky on the GPU) against the
oracle's skyline LDL^T direct solve (SolverSkyline semantics) on the same
inputs: 1e-12 relative in FP64, batched over samples and mesh classes, and
as the per-step solver of the Newmark loop."""
import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd import _lib as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def skyline_direct(model, E, f):
    ke, bad = O.ke_mesh(model.conn, model.xyz, E, 0.3)
    assert bad == 0
    n = model.n_free if hasattr(model, "n_free") else model.n
    ri, data = O.skyline_assemble(model.conn, model.dof, n, ke)
    fac, rc = O.skyline_factorize(ri, data, 1e-15)
    assert rc == 0
    return O.skyline_solve(ri, fac, f)


def test_direct_solve_matches_skyline_oracle(ctx):
    print("\n[case] direct solver 3 classes", flush=True)
    models = [msa.cantilever(3), msa.cantilever(6), msa.cantilever(5, 4, 3)]
    meshes = [msa.Mesh(ctx, m.xyz, m.conn, m.dof, m.n_free) for m in models]
    rng = np.random.default_rng(30)
    ns = [2, 3, 2]
    s = np.sqrt(np.log(1.09))
    E = [2.1e5 * np.exp(-0.5 * s * s + s * rng.standard_normal((k, m.n_elem)))
         for k, m in zip(ns, models)]
    b = msa.Batch(ctx, meshes, ns, L.MSA_FP64, 0)
    for c, m in enumerate(models):
        b.set_modulus(c, E[c])
        b.set_rhs(c, m.f)
        info = b.direct_info(c)
        assert info["n"] == m.n_free and 0 < info["bandwidth"] < m.n_free
    b.set_solver(msa.MSA_SOLVER_DIRECT)
    b.assemble(0.3)
    b.solve()
    for c, m in enumerate(models):
        st = b.stats(c)
        assert (st["status"] == msa.MSA_OK).all(), st
        assert (st["relres"] < 1e-13).all(), st["relres"]
        for k in range(ns[c]):
            assert rel(b.solution(c, k), skyline_direct(m, E[c][k], m.f)) < 1e-12
    b.close()
    for m in meshes:
        m.close()


def test_direct_solver_rejects_mixed_and_flags_indefinite(ctx):
    model = msa.cantilever(3)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    bm = msa.Batch(ctx, [mesh], [1], L.MSA_MIXED, 0)
    with pytest.raises(msa.MsaError):
        bm.set_solver(msa.MSA_SOLVER_DIRECT)
    bm.close()
    b = msa.Batch(ctx, [mesh], [2], L.MSA_FP64, 0)
    E = np.full((2, model.n_elem), 2.1e5)
    E[1, :] *= -1.0                     # negative definite: non-positive pivots
    b.set_modulus(0, E)
    b.set_rhs(0, model.f)
    b.assemble(0.3)
    b.direct_solve()
    st = b.stats(0)
    assert st["status"][0] == msa.MSA_OK and st["status"][1] == msa.MSA_BREAKDOWN
    b.close()
    mesh.close()


def test_newmark_with_direct_solver(ctx):
    print("\n[case] newmark direct 4x4x2", flush=True)
    model = msa.cantilever(4, 4, 2, lz=0.5)
    om = O.Cantilever(4, 4, 2, lz=0.5)
    rng = np.random.default_rng(31)
    E = 2.1e5 * np.exp(0.2 * rng.standard_normal((2, model.n_elem)))
    nsteps, dt = 30, 2e-3
    acc = np.array([O.sinusoidal_acc(i, 1.0, 0.05, dt, nsteps * dt) for i in range(nsteps)])
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [2], L.MSA_FP64, 0)
    b.set_modulus(0, E)
    b.set_dynamics(50.0, 30.0, 2e-4, 0.25, 0.5, dt)
    b.set_probe(0, model.n_free - 1)
    b.set_solver(msa.MSA_SOLVER_DIRECT)
    b.assemble(0.3)
    b.newmark(acc, acc_dir=2)
    for k in range(2):
        h_ref, u_ref, _, _ = O.newmark_skyline(om, E[k], rho=50.0, alphaR=30.0, betaR=2e-4, dt=dt,
                                               nsteps=nsteps, acc_dir=2, acc_mag=1.0,
                                               acc_period=0.05, acc_tlen=nsteps * dt,
                                               probe=model.n_free - 1)
        assert rel(b.history(0, k), h_ref) < 1e-11
        assert rel(b.dynamic_state(0, k)[0], u_ref) < 1e-11
    b.close()
    mesh.close()


def test_solver_skyline_mirror_config1(ctx):
    """The plain CPU-runnable configuration (10^3 cantilever, E = 2.1e5): the
    GPU direct solve equals the oracle's skyline LDL^T solve."""
    model = msa.cantilever(10)
    solver = msa.SolverSkyline(ctx, model, np.full(model.n_elem, 2.1e5))
    solver.Initialize()
    u = solver.Solve()
    assert solver.Status == msa.MSA_OK and solver.RelativeResidual < 1e-12
    assert rel(u, skyline_direct(model, np.full(model.n_elem, 2.1e5), model.f)) < 1e-12
    solver.batch.close()
    solver.mesh.close()
