"""GPU parity of every workload configuration at its stated size.

  C2  64^3 cantilever, one FP64 Jacobi-PCG solve
  C3  256 lognormal-KL samples of 32^3, FP64 batched Jacobi-PCG
  C4  the 1024-sample uneven mixed-precision batch, multigrid-preconditioned
      PCG with FP64 residual replacement (matrix-free smoothing on 16^3-48^3)
  C5  batched Newmark dynamics, 48x48x24 block, 64 samples, 500 steps

Displacements are compared with the CPU oracle's own tightly converged
solves (tests/golden/fullsize_golden.npz, written by make_fullsize_golden.py
beside it): the loaded face's DOFs exactly, and the whole vector through 64
seeded Gaussian projections, whose mean squared mismatch is an unbiased
estimate of |u_gpu - u_oracle|^2 with 64 degrees of freedom (an estimate at
half the bar bounds the true error below the bar except with probability
~1e-10).  Residuals are evaluated with the oracle's CSR operator assembled on
the CPU from the oracle's element matrices, and the Newmark step with the
oracle's stiffness, consistent mass and load.

Bars: FP64 displacements 1e-10 relative (both sides converged to 1e-14; the
attainable FP64 agreement of independent solvers here is ~1e-14); mixed
precision 1e-8 (the workload's bar), oracle residuals at the solve tolerance.
"""
import os

import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd import _lib as L
from msolve_amd.stochastic import kl_table, standard_normal
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fullsize_golden.npz"))
SEED_MC = 20261202
SEED_C3 = SEED_MC + 3


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def face_dofs(nex, dof, n_nodes):
    nodes = np.nonzero(np.arange(n_nodes) % (nex + 1) == nex)[0]
    d = np.asarray(dof).reshape(-1, 3)[nodes].ravel()
    return d[d >= 0]


def against_fixture(key, model, u):
    """(face error, projected estimate of the whole-vector error), both
    relative to the oracle's displacement."""
    face = u[face_dofs(model.nex, model.dof, model.n_nodes)]
    e_face = rel(face, GOLD[key + "_face"])
    g = np.random.Generator(np.random.Philox(int(GOLD[key + "_seed"][0])))
    p = np.array([g.standard_normal(u.size) @ u for _ in range(GOLD[key + "_proj"].size)])
    e_proj = float(np.sqrt(np.mean((p - GOLD[key + "_proj"]) ** 2)) / GOLD[key + "_norm"][0])
    return e_face, e_proj


def oracle_residual(model, E, u):
    ke, bad = O.ke_mesh(model.conn, model.xyz, E, 0.3)
    assert bad == 0
    rp, col, val = O.csr_assemble(model.conn, model.dof, model.n_free, ke)
    return float(np.linalg.norm(model.f - O.csr_multiply(rp, col, val, u)) /
                 np.linalg.norm(model.f))


def test_c2_64cube_fp64_matches_oracle(ctx):
    """64^3 (811 200 DOFs), E 2.1e5: FP64 Jacobi-PCG to 1e-10, residual by
    the oracle operator; then to 1e-14 against the oracle's own PCG."""
    print("\n[case] C2 64^3 fp64", flush=True)
    model = msa.cantilever(64)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [1], L.MSA_FP64, 1)
    E = np.full(model.n_elem, 2.1e5)
    b.set_modulus(0, E)
    b.set_rhs(0, model.f)
    b.run(0.3, tol=1e-10, max_it=50000, graph_iters=32)
    st = b.stats(0)
    assert st["status"][0] == msa.MSA_OK and st["relres"][0] < 1e-10
    res = oracle_residual(model, E, b.solution(0, 0))
    b.solve(tol=1e-14, max_it=50000, graph_iters=32)
    assert b.stats(0)["status"][0] == msa.MSA_OK
    e_face, e_proj = against_fixture("c2", model, b.solution(0, 0))
    print(f"[C2] oracle residual at 1e-10: {res:.2e}; at 1e-14 vs oracle PCG: face {e_face:.2e}, "
          f"whole vector (estimate) {e_proj:.2e}", flush=True)
    assert res < 1.5e-10
    assert e_face < 1e-10 and e_proj < 0.5e-10
    b.close()
    mesh.close()


def test_c3_256_samples_32cube_fp64(ctx):
    """All 256 samples of 32^3 (KL COV 0.2) solved in one FP64 batch to
    1e-10; three samples' residuals by the oracle operator with the oracle's
    own modulus field; two samples to 1e-14 against the oracle's loop."""
    print("\n[case] C3 256 x 32^3 fp64", flush=True)
    model = msa.cantilever(32)
    table = kl_table([1.0, 1.0, 1.0], 0.25, 40)
    otab = O.kl_table([1.0, 1.0, 1.0], 0.25, 40)
    ns = 256
    xi = standard_normal(SEED_C3, ns, 40)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [ns], L.MSA_FP64, 1)
    b.set_kl(0, table, model.domain, 2.1e5, 0.2)
    b.set_xi(0, xi)
    b.set_rhs(0, model.f)
    b.run(0.3, tol=1e-10, max_it=20000, graph_iters=16)
    st = b.stats(0)
    assert (st["status"] == msa.MSA_OK).all() and (st["relres"] < 1e-10).all()
    res = {}
    for s in (0, 127, 255):
        E, bad = O.kl_field(model.conn, model.xyz, otab, xi[s], model.domain, 2.1e5, 0.2)
        assert bad == 0
        assert np.abs(b.modulus(0, s) - E).max() <= 1e-13 * E.max()
        res[s] = oracle_residual(model, E, b.solution(0, s))
    b.solve(tol=1e-14, max_it=20000, graph_iters=16)
    assert (b.stats(0)["status"] == msa.MSA_OK).all()
    errs = {s: against_fixture(f"c3_{s}", model, b.solution(0, s)) for s in (0, 255)}
    print(f"[C3] iterations {st['iters'].min()}-{st['iters'].max()}; oracle residuals {res}; "
          f"vs oracle at 1e-14 (face, whole): {errs}", flush=True)
    assert max(res.values()) < 1.5e-10
    for e_face, e_proj in errs.values():
        assert e_face < 1e-10 and e_proj < 0.5e-10
    b.close()
    mesh.close()


def test_c4_1024_uneven_mixed_multigrid(ctx):
    """The headline batch: 1024 samples over n in {16..48} (counts from the
    workload's seed), COV 0.3, mixed precision with the multigrid
    preconditioner (matrix-free smoothing on every class) to |f-Ku|/|f| <
    1e-9.  Each class's first sample against the oracle's loop to 1e-12
    (bar 1e-8), the 48^3 one's residual by the oracle operator."""
    print("\n[case] C4 1024 mixed multigrid", flush=True)
    sizes = [16, 24, 32, 40, 48]
    counts = [int(c) for c in GOLD["counts"]]
    assert sum(counts) == 1024
    models = [msa.cantilever(n) for n in sizes]
    meshes = [msa.Mesh(ctx, m.xyz, m.conn, m.dof, m.n_free) for m in models]
    table = kl_table([1.0, 1.0, 1.0], 0.25, 40)
    xi = standard_normal(SEED_MC, 1024, 40)
    b = msa.Batch(ctx, meshes, counts, L.MSA_MIXED, 1)
    b.set_preconditioner(L.MSA_PRECOND_MULTIGRID)
    b.set_refinement(L.MSA_REFINE_RELIABLE, 2, 1e-3)
    off = []
    o = 0
    for c, m in enumerate(models):
        info = b.mg_info(c)
        assert info["levels"] >= 2 and info["matrix_free"], (sizes[c], info)
        b.set_kl(c, table, m.domain, 2.1e5, 0.3)
        b.set_xi(c, xi[o:o + counts[c]])
        b.set_rhs(c, m.f)
        off.append(o)
        o += counts[c]
    b.run(0.3, tol=1e-9, max_it=20000, graph_iters=16)
    errs, iters = {}, {}
    for c, n in enumerate(sizes):
        st = b.stats(c)
        assert (st["status"] == msa.MSA_OK).all() and (st["relres"] < 1e-9).all(), n
        iters[n] = (int(st["iters"].min()), int(st["iters"].max()))
        errs[n] = against_fixture(f"c4_{n}", models[c], b.solution(c, 0))
    E, bad = O.kl_field(models[-1].conn, models[-1].xyz, O.kl_table([1.0, 1.0, 1.0], 0.25, 40),
                        xi[off[-1]], models[-1].domain, 2.1e5, 0.3)
    res48 = oracle_residual(models[-1], E, b.solution(len(sizes) - 1, 0))
    print(f"[C4] iterations {iters}; first sample of each class vs oracle at 1e-12 "
          f"(face, whole): {errs}; 48^3 oracle residual {res48:.2e}", flush=True)
    for e_face, e_proj in errs.values():
        assert e_face < 1e-8 and e_proj < 0.5e-8
    assert res48 < 1.05e-9
    b.close()
    for m in meshes:
        m.close()


def test_c5_newmark_48x48x24_64_samples_500_steps(ctx):
    """The dynamics workload at size: 64 samples x 500 steps of the
    48x48x24 block (rho 7.85e-9, beta 0.25, gamma 0.5, dt 1e-4, 20 Hz base
    acceleration in Z, KL COV 0.2), mixed multigrid PCG to 1e-8 warm-started
    from the previous displacement.  Checks: every sample converged at every
    step; probe histories against FP64 runs (PCG to 1e-12, bar 1e-7); a
    4-sample batch reproduces the 64-sample histories (samples are
    independent); and the last step against the oracle: the effective
    equation (K + a0 M) u_500 = F_500 + M (a0 u_499 + a2 v_499 + a3 a_499)
    with the oracle's K, consistent M and mass-acceleration load, and the
    Newmark velocity/acceleration updates."""
    print("\n[case] C5 newmark 48x48x24 x64 x500", flush=True)
    ns, nsteps, dt = 64, 500, 1e-4
    model = msa.cantilever(48, 48, 24, lz=0.5)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    table = kl_table(list(model.domain), 0.25, 40)
    xi = standard_normal(SEED_MC, ns, 40)
    acc = msa.sinusoidal_history(0.05, dt, 0.05, 1.0, nsteps)
    probe = model.n_free - 1

    def make(precision, nsamp):
        b = msa.Batch(ctx, [mesh], [nsamp], precision, 1)
        b.set_preconditioner(L.MSA_PRECOND_MULTIGRID)
        if precision == L.MSA_MIXED:
            b.set_refinement(L.MSA_REFINE_RELIABLE, 2, 1e-3)
        b.set_kl(0, table, model.domain, 2.1e5, 0.2)
        b.set_xi(0, xi[:nsamp])
        b.eval_kl()
        b.set_dynamics(7.85e-9, 0.0, 0.0, 0.25, 0.5, dt)
        b.set_probe(0, probe)
        b.assemble(0.3)
        return b

    b = make(L.MSA_MIXED, ns)
    b.newmark(acc, 2, 1e-8, 20000, 16)
    st = b.stats(0)
    assert (st["status"] == msa.MSA_OK).all()
    hist = [b.history(0, s) for s in range(4)]
    its = np.array([b.dynamic_state(0, s)[3] for s in range(ns)]) / nsteps
    b.close()
    b64 = make(L.MSA_FP64, 4)
    b64.newmark(acc, 2, 1e-12, 20000, 16)
    e64 = [rel(hist[s], b64.history(0, s)) for s in range(4)]
    b64.close()
    b4 = make(L.MSA_MIXED, 4)
    b4.newmark(acc[:nsteps - 1], 2, 1e-8, 20000, 16)
    prev = [b4.dynamic_state(0, s)[:3] for s in (0, 3)]
    b4.newmark(acc, 2, 1e-8, 20000, 16)
    e_batch = [rel(b4.history(0, s), hist[s]) for s in range(4)]
    last = [b4.dynamic_state(0, s)[:3] for s in (0, 3)]
    E4 = [b4.modulus(0, s) for s in (0, 3)]
    b4.close()
    mesh.close()
    # the oracle's last step
    a0, a2, a3 = 1 / (0.25 * dt * dt), 1 / (0.25 * dt), 1 / (2 * 0.25) - 1
    a6, a7 = dt * (1 - 0.5), 0.5 * dt
    x8 = model.xyz[model.conn[0]]
    me = O.h8_mass_csharp(x8, 7.85e-9)           # every cell of the grid is the same brick
    Me = O.unpack300(me)
    mrow = Me[:, 2::3].sum(axis=1)               # M_e times the unit Z vector
    edof = model.dof.reshape(-1, 3)[model.conn].reshape(model.n_elem, 24)
    bacc = np.zeros(model.n_free)
    ok = edof >= 0
    np.add.at(bacc, edof[ok], np.broadcast_to(mrow, edof.shape)[ok])
    mall = np.broadcast_to(me, (model.n_elem, 300))
    rpm, colm, valm = O.csr_assemble(model.conn, model.dof, model.n_free, mall)
    ag = O.sinusoidal_acc(nsteps - 1, 1.0, 0.05, dt, nsteps * dt)
    assert abs(ag - acc[nsteps - 1]) <= 1e-15
    res, upd = [], []
    for k in range(2):
        ke, bad = O.ke_mesh(model.conn, model.xyz, E4[k], 0.3)
        assert bad == 0
        rpk, colk, valk = O.csr_assemble(model.conn, model.dof, model.n_free, ke)
        (u0, v0, w0), (u1, v1, w1) = prev[k], last[k]
        rhs = bacc * ag + O.csr_multiply(rpm, colm, valm, a0 * u0 + a2 * v0 + a3 * w0)
        lhs = O.csr_multiply(rpk, colk, valk, u1) + a0 * O.csr_multiply(rpm, colm, valm, u1)
        res.append(rel(lhs, rhs))
        w_ref = a0 * (u1 - u0) - a2 * v0 - a3 * w0
        v_ref = v0 + a6 * w0 + a7 * w_ref
        upd.append(max(rel(w1, w_ref), rel(v1, v_ref)))
    print(f"[C5] iterations per step {its.min():.1f}-{its.max():.1f}; probe history vs FP64 "
          f"{max(e64):.2e}; 4-sample vs 64-sample batch {max(e_batch):.2e}; last step vs oracle: "
          f"residual {max(res):.2e}, updates {max(upd):.2e}", flush=True)
    assert max(e64) < 1e-7
    assert max(e_batch) < 1e-12
    assert max(res) < 1.5e-8
    assert max(upd) < 1e-10
