"""GPU parity of the Newmark dynamics path (consistent mass, Rayleigh
damping, effective stiffness, mass-acceleration history load, warm-started
PCG per step, velocity/acceleration update) against the oracle's skyline
Newmark loop on the same inputs.  Bars: FP64 batches 1e-9 relative on the
probe history and the final displacement (PCG to 1e-12 per step), mixed
batches 1e-7 (FP64 residual replacement to 1e-10 per step)."""
import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd import _lib as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def lognormal(rng, n, cov=0.2):
    s = np.sqrt(np.log(1 + cov * cov))
    return 2.1e5 * np.exp(-0.5 * s * s + s * rng.standard_normal(n))


def run_case(ctx, precision, nex, ney, nez, lz, rho, dt, nsteps, alphaR, betaR, tol, n_samp=2,
             static_load=False, tile=0):
    model = msa.cantilever(nex, ney, nez, lz=lz)
    om = O.Cantilever(nex, ney, nez, lz=lz)
    rng = np.random.default_rng(nex * 100 + nez)
    E = np.stack([lognormal(rng, model.n_elem) for _ in range(n_samp)])
    period = 0.05
    acc = np.array([O.sinusoidal_acc(i, 1.0, period, dt, nsteps * dt) for i in range(nsteps)])
    probe = model.n_free - 1  # Z of the last free node (tip corner)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [n_samp], precision, tile)
    b.set_modulus(0, E)
    if static_load:
        b.set_rhs(0, model.f)
    b.set_dynamics(rho, alphaR, betaR, 0.25, 0.5, dt)
    b.set_probe(0, probe)
    b.assemble(0.3)
    b.newmark(acc, acc_dir=2, tol=tol, max_it=20000, graph_iters=8)
    out = []
    for s in range(n_samp):
        h_ref, u_ref, v_ref, a_ref = O.newmark_skyline(
            om, E[s], rho=rho, alphaR=alphaR, betaR=betaR, dt=dt, nsteps=nsteps,
            f_static=model.f if static_load else None, acc_dir=2, acc_mag=1.0,
            acc_period=period, acc_tlen=nsteps * dt, probe=probe)
        u, v, a, its = b.dynamic_state(0, s)
        out.append(dict(h=b.history(0, s), h_ref=h_ref, u=u, u_ref=u_ref, v=v, v_ref=v_ref,
                        a=a, a_ref=a_ref, iters=its))
    st = b.stats(0)
    b.close()
    mesh.close()
    return out, st


@pytest.mark.parametrize("alphaR,betaR", [(0.0, 0.0), (30.0, 2e-4)])
def test_newmark_fp64_matches_skyline_oracle(ctx, alphaR, betaR):
    print(f"\n[case] newmark fp64 4x4x2 rayleigh=({alphaR},{betaR})", flush=True)
    res, st = run_case(ctx, L.MSA_FP64, 4, 4, 2, 0.5, 50.0, 2e-3, 40, alphaR, betaR, 1e-12)
    assert (st["status"] == msa.MSA_OK).all(), st
    for r in res:
        assert np.abs(r["h_ref"]).max() > 0
        assert rel(r["h"], r["h_ref"]) < 1e-9
        assert rel(r["u"], r["u_ref"]) < 1e-9
        assert rel(r["v"], r["v_ref"]) < 1e-8
        assert rel(r["a"], r["a_ref"]) < 1e-7
        assert r["iters"] > 0


def test_newmark_mixed_matches_skyline_oracle(ctx):
    print("\n[case] newmark mixed 4x4x2", flush=True)
    res, st = run_case(ctx, L.MSA_MIXED, 4, 4, 2, 0.5, 50.0, 2e-3, 40, 30.0, 2e-4, 1e-10)
    assert (st["status"] == msa.MSA_OK).all(), st
    for r in res:
        assert rel(r["h"], r["h_ref"]) < 1e-7
        assert rel(r["u"], r["u_ref"]) < 1e-7
        assert rel(r["v"], r["v_ref"]) < 1e-6


def test_newmark_workload_parameters_small_block(ctx):
    """The dynamics workload's parameters (density 7.85e-9, beta 0.25,
    gamma 0.5, dt 1e-4, 20 Hz base acceleration, COV 0.2) on a 12x12x6
    block with a static tip load as well; mixed precision, tile 1."""
    print("\n[case] newmark workload params 12x12x6 mixed", flush=True)
    res, st = run_case(ctx, L.MSA_MIXED, 12, 12, 6, 0.5, 7.85e-9, 1e-4, 12, 0.0, 0.0, 1e-10,
                       n_samp=3, static_load=True)
    assert (st["status"] == msa.MSA_OK).all(), st
    for r in res:
        assert rel(r["h"], r["h_ref"]) < 1e-7
        assert rel(r["u"], r["u_ref"]) < 1e-7


def test_newmark_warm_start_saves_iterations(ctx):
    """Warm starts from the previous step's displacement: the same history
    (base acceleration only, smooth in time) with fewer summed iterations
    than cold starts from zero."""
    print("\n[case] newmark warm vs cold 8x8x4", flush=True)
    model = msa.cantilever(8, 8, 4, lz=0.5)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    rng = np.random.default_rng(4)
    E = np.stack([lognormal(rng, model.n_elem) for _ in range(2)])
    nsteps = 60
    acc = np.array([O.sinusoidal_acc(i, 1.0, 0.05, 1e-4, 0.05) for i in range(nsteps)])
    res = {}
    for prec in (L.MSA_FP64, L.MSA_MIXED):
        b = msa.Batch(ctx, [mesh], [2], prec, 0)
        b.set_modulus(0, E)
        b.set_dynamics(7.85e-9, 0, 0, 0.25, 0.5, 1e-4)
        b.set_probe(0, model.n_free - 1)
        b.assemble(0.3)
        for mode in (1, 2, 0):
            b.newmark(acc, acc_dir=2, tol=1e-10, max_it=20000, graph_iters=8, warm_start=mode)
            res[prec, mode] = [(b.history(0, s), b.dynamic_state(0, s)[3]) for s in range(2)]
        b.close()
    mesh.close()
    for prec in (L.MSA_FP64, L.MSA_MIXED):
        for s in range(2):
            hc, ic = res[prec, 0][s]
            for mode in (1, 2):
                assert rel(res[prec, mode][s][0], hc) < 1e-8
            iw = res[prec, 1][s][1]
            print(f"[warm start] precision {prec} sample {s}: {iw} / {ic} iterations "
                  f"({100 * (1 - iw / ic):.1f}% saved)", flush=True)
            # the response is quasi-static (a0 M << K), so the previous
            # displacement is nearly parallel to the new one: a warm start
            # removes only the amplitude change of the step from the initial
            # residual (about 1/i at step i of a load growing like i), a few
            # decades of the ten the tolerance asks for
            assert iw <= 0.97 * ic, (prec, iw, ic)


@pytest.mark.parametrize("precision,solver", [(L.MSA_FP64, L.MSA_SOLVER_PCG),
                                              (L.MSA_MIXED, L.MSA_SOLVER_PCG),
                                              (L.MSA_FP64, L.MSA_SOLVER_DIRECT)])
def test_newmark_step_graph_equals_eager(ctx, precision, solver):
    """Captured time steps (one graph per step: the previous update, the
    step's right-hand side and solve start, the first iteration block)
    give the same histories as eager launches, bit for bit, with the
    multigrid preconditioner; one capture set per call."""
    print(f"\n[case] newmark graph vs eager precision={precision} solver={solver}", flush=True)
    model = msa.cantilever(8, 8, 4, lz=0.5)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    rng = np.random.default_rng(9)
    E = np.stack([lognormal(rng, model.n_elem) for _ in range(3)])
    acc = np.array([O.sinusoidal_acc(i, 1.0, 0.05, 1e-4, 0.05) for i in range(25)])
    out = []
    for g in (0, 8):
        b = msa.Batch(ctx, [mesh], [3], precision, 1)
        if solver == L.MSA_SOLVER_DIRECT:
            b.set_solver(solver)
        else:
            b.set_preconditioner(L.MSA_PRECOND_MULTIGRID)
        b.set_modulus(0, E)
        b.set_rhs(0, model.f)
        b.set_dynamics(7.85e-9, 0, 0, 0.25, 0.5, 1e-4)
        b.set_probe(0, model.n_free - 1)
        b.assemble(0.3)
        b.newmark(acc, acc_dir=2, tol=1e-10, max_it=20000, graph_iters=g)
        gi = b.graph_info()
        out.append([(b.history(0, s), b.dynamic_state(0, s)) for s in range(3)])
        assert (b.stats(0)["status"] == msa.MSA_OK).all()
        if g:
            assert gi["captures"] >= 3 and gi["replays"] >= 25, gi
        b.close()
    mesh.close()
    for s in range(3):
        assert np.array_equal(out[0][s][0], out[1][s][0])
        for k in range(3):
            assert np.array_equal(out[0][s][1][k], out[1][s][1][k])


def test_newmark_argument_checks(ctx):
    model = msa.cantilever(2)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [1], L.MSA_FP64, 0)
    with pytest.raises(msa.MsaError):
        b.set_dynamics(1.0, 0, 0, 0.5, 0.25, 1e-3)   # delta < 0.5 (reference throws)
    with pytest.raises(msa.MsaError):
        b.set_dynamics(1.0, 0, 0, 0.2, 0.5, 1e-3)    # alpha below the limit
    with pytest.raises(msa.MsaError):
        b.newmark(np.zeros(3))                       # not in dynamics mode
    b.close()
    mesh.close()
