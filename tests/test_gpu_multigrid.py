"""GPU parity of the multigrid preconditioner (msa_batch_set_preconditioner)
against oracle/multigrid.py: one V-cycle on the same vector (FP64 within
1e-10, FP32 levels within 2e-5), preconditioned solves against the direct
solution (FP64 PCG to 1e-11: within 1e-9; mixed reliable updates to 1e-9:
within 1e-8, the Monte Carlo workload's bar), iteration counts against the
oracle's preconditioned CG, graph replay against eager launches, lane tiles,
a class that does not coarsen (keeps Jacobi), and Newmark steps."""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

import msolve_amd as msa
from msolve_amd import _lib as L
from oracle import multigrid as MG
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def lognormal(rng, n, cov=0.3):
    s = np.sqrt(np.log(1 + cov * cov))
    return 2.1e5 * np.exp(-0.5 * s * s + s * rng.standard_normal(n))


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make(ctx, sizes, nsamp, precision, tile=0, seed=0, mg=True):
    models = [msa.cantilever(n) for n in sizes]
    meshes = [msa.Mesh(ctx, m.xyz, m.conn, m.dof, m.n_free) for m in models]
    b = msa.Batch(ctx, meshes, [nsamp] * len(sizes), precision, tile)
    rng = np.random.default_rng(seed)
    Es = []
    for c, m in enumerate(models):
        E = np.stack([lognormal(rng, m.n_elem) for _ in range(nsamp)])
        b.set_modulus(c, E)
        b.set_rhs(c, m.f)
        Es.append(E)
    b.assemble(0.3)
    if mg:
        b.set_preconditioner(msa.MSA_PRECOND_MULTIGRID)
    return b, models, meshes, Es


def oracle_levels(m, E):
    return MG.hierarchy(m.xyz.reshape(-1, 3), m.conn, m.dof, m.n_free, E)


def with_gpu_interval(levels, b, c, s):
    """The oracle's levels smoothing on the GPU's interval [lam/ratio, lam]
    of class c, sample s (the Lanczos estimate; pinned against the true
    spectrum by test_lanczos_interval_brackets_the_spectrum)."""
    lam = b.mg_lambda(c, s)
    assert len(lam) == len(levels) - 1
    for L_, l in zip(levels[:-1], lam):
        L_.lam = float(l)
    return levels


def close(b, meshes):
    b.close()
    for m in meshes:
        m.close()


def test_hierarchy_sizes_match_oracle(ctx):
    print("\n[case] mg hierarchy sizes 8,10,12,9", flush=True)
    b, models, meshes, Es = make(ctx, [8, 10, 12, 9], 1, L.MSA_FP64)
    for c, m in enumerate(models):
        info = b.mg_info(c)
        levels, _ = oracle_levels(m, Es[c][0])
        assert info["levels"] == len(levels) - 1
        if info["levels"]:
            assert info["coarsest_dofs"] == levels[-1].n
            assert info["coarse_dofs"] == sum(L_.n for L_ in levels[1:])
    assert b.mg_info(3)["levels"] == 0
    close(b, meshes)


@pytest.mark.parametrize("precision,bar", [(L.MSA_FP64, 1e-10), (L.MSA_MIXED, 2e-5)])
def test_vcycle_matches_oracle(ctx, precision, bar):
    print(f"\n[case] mg apply precision={precision}", flush=True)
    b, models, meshes, Es = make(ctx, [8, 10, 12], 2, precision, seed=4)
    rng = np.random.default_rng(9)
    for c, m in enumerate(models):
        levels, Ps = oracle_levels(m, Es[c][1])
        with_gpu_interval(levels, b, c, 1)
        r = rng.standard_normal(m.n_free)
        z = b.mg_apply(c, 1, r)
        zr = MG.vcycle(levels, Ps, r)
        assert rel(z, zr) < bar, (c, rel(z, zr))
    close(b, meshes)


def test_fp64_solves_match_direct_and_oracle_iterations(ctx):
    print("\n[case] mg fp64 solves 8,12,9", flush=True)
    b, models, meshes, Es = make(ctx, [8, 12, 9], 3, L.MSA_FP64, seed=5)
    b.solve(1e-11, 5000)
    for c, m in enumerate(models):
        st = b.stats(c)
        assert (st["status"] == msa.MSA_OK).all(), st
        for s in range(3):
            levels, Ps = oracle_levels(m, Es[c][s])
            ud = spl.spsolve(levels[0].K.tocsc(), m.f)
            assert rel(b.solution(c, s), ud) < 1e-9
            if c < 2:
                with_gpu_interval(levels, b, c, s)
                _, it = MG.mg_pcg(levels, Ps, m.f, tol=1e-11)
                assert abs(int(st["iters"][s]) - it) <= 1, (c, s, st["iters"][s], it)
                assert it < 25
        if c == 2:  # 9^3 does not coarsen: Jacobi
            assert st["iters"].min() > 40
    close(b, meshes)


def test_mixed_reliable_updates_reach_the_workload_bar(ctx):
    print("\n[case] mg mixed 8,12,10", flush=True)
    b, models, meshes, Es = make(ctx, [8, 12, 10], 3, L.MSA_MIXED, seed=6)
    b.solve(1e-9, 5000, graph_iters=8)
    for c, m in enumerate(models):
        st = b.stats(c)
        assert (st["status"] == msa.MSA_OK).all(), st
        assert st["iters"].max() < 40, st["iters"]
        for s in range(3):
            K = oracle_levels(m, Es[c][s])[0][0].K
            ud = spl.spsolve(K.tocsc(), m.f)
            assert rel(b.solution(c, s), ud) < 1e-8
    close(b, meshes)


@pytest.mark.parametrize("precision", [L.MSA_FP64, L.MSA_MIXED])
def test_graph_equals_eager(ctx, precision):
    print(f"\n[case] mg graph vs eager precision={precision}", flush=True)
    out = []
    for g in (0, 8, 20):  # 20: a long first block, then blocks of 2
        b, models, meshes, Es = make(ctx, [8, 12], 2, precision, seed=7)
        b.solve(1e-10, 5000, graph_iters=g)
        out.append(([b.solution(c, s) for c in range(2) for s in range(2)],
                    [b.stats(c)["iters"].copy() for c in range(2)]))
        close(b, meshes)
    for k in (1, 2):
        for u0, u1 in zip(out[0][0], out[k][0]):
            assert np.array_equal(u0, u1)
        for i0, i1 in zip(out[0][1], out[k][1]):
            assert np.array_equal(i0, i1)


def test_lane_tiles(ctx):
    print("\n[case] mg tile 4", flush=True)
    b, models, meshes, Es = make(ctx, [8, 12], 5, L.MSA_FP64, tile=4, seed=8)
    b.solve(1e-11, 5000)
    for c, m in enumerate(models):
        assert (b.stats(c)["status"] == msa.MSA_OK).all()
        for s in range(5):
            K = oracle_levels(m, Es[c][s])[0][0].K
            assert rel(b.solution(c, s), spl.spsolve(K.tocsc(), m.f)) < 1e-9
    close(b, meshes)


def test_restarted_refinement(ctx):
    print("\n[case] mg restarted refinement", flush=True)
    b, models, meshes, Es = make(ctx, [8], 2, L.MSA_MIXED, seed=10)
    b.set_refinement(msa.MSA_REFINE_RESTART)
    b.solve(1e-9, 5000, inner_tol=1e-4, max_outer=20)
    m = models[0]
    for s in range(2):
        K = oracle_levels(m, Es[0][s])[0][0].K
        assert rel(b.solution(0, s), spl.spsolve(K.tocsc(), m.f)) < 1e-8
    close(b, meshes)


def test_switch_back_to_jacobi(ctx):
    print("\n[case] mg then jacobi", flush=True)
    b, models, meshes, Es = make(ctx, [8], 2, L.MSA_FP64, seed=11)
    b.solve(1e-11, 5000)
    it_mg = b.stats(0)["iters"].copy()
    u_mg = b.solution(0, 0)
    b.set_preconditioner(msa.MSA_PRECOND_JACOBI)
    b.solve(1e-11, 5000)
    it_j = b.stats(0)["iters"]
    assert (it_j > 3 * it_mg).all()
    assert rel(b.solution(0, 0), u_mg) < 1e-9
    with pytest.raises(msa.MsaError):
        b.set_preconditioner(msa.MSA_PRECOND_MULTIGRID, max_levels=1)
    with pytest.raises(msa.MsaError):
        b.set_preconditioner(7)
    close(b, meshes)


@pytest.mark.parametrize("precision,bar", [(L.MSA_FP64, 1e-9), (L.MSA_MIXED, 1e-7)])
def test_newmark_with_multigrid_matches_skyline_oracle(ctx, precision, bar):
    print(f"\n[case] mg newmark precision={precision}", flush=True)
    nex, ney, nez, lz = 8, 8, 4, 0.5
    model = msa.cantilever(nex, ney, nez, lz=lz)
    om = O.Cantilever(nex, ney, nez, lz=lz)
    rng = np.random.default_rng(12)
    E = np.stack([lognormal(rng, model.n_elem, 0.2) for _ in range(2)])
    dt, nsteps, period = 2e-3, 20, 0.05
    acc = np.array([O.sinusoidal_acc(i, 1.0, period, dt, nsteps * dt) for i in range(nsteps)])
    probe = model.n_free - 1
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [2], precision, 0)
    b.set_modulus(0, E)
    b.set_dynamics(50.0, 30.0, 2e-4, 0.25, 0.5, dt)
    b.set_probe(0, probe)
    b.assemble(0.3)
    b.set_preconditioner(msa.MSA_PRECOND_MULTIGRID)
    assert b.mg_info(0)["levels"] == 1
    b.newmark(acc, acc_dir=2, tol=1e-12 if precision == L.MSA_FP64 else 1e-10, max_it=5000,
              graph_iters=8)
    for s in range(2):
        h_ref, u_ref, _, _ = O.newmark_skyline(
            om, E[s], rho=50.0, alphaR=30.0, betaR=2e-4, dt=dt, nsteps=nsteps, acc_dir=2,
            acc_mag=1.0, acc_period=period, acc_tlen=nsteps * dt, probe=probe)
        u, _, _, its = b.dynamic_state(0, s)
        assert rel(b.history(0, s), h_ref) < bar
        assert rel(u, u_ref) < bar
        assert its < 30 * nsteps
    b.close()
    mesh.close()


def test_solver_pcg_surface_with_multigrid(ctx):
    """analyzers.SolverPCG(preconditioner=MSA_PRECOND_MULTIGRID) on the
    12^3 cantilever with one modulus: the direct solution in < 25 products."""
    print("\n[case] SolverPCG multigrid 12^3", flush=True)
    model = msa.cantilever(12)
    E = np.full(model.n_elem, 2.1e5)
    solver = msa.SolverPCG(ctx, model, E, preconditioner=msa.MSA_PRECOND_MULTIGRID)
    u = solver.Solve(20000, 1e-12)
    assert solver.Status == msa.MSA_OK and solver.Iterations < 25
    levels, _ = oracle_levels(model, E)
    assert rel(u, spl.spsolve(levels[0].K.tocsc(), model.f)) < 1e-10
    solver.batch.close()
    solver.mesh.close()


@pytest.mark.parametrize("dynamics", [False, True])
def test_matrix_free_cycle_matches_stored_value_cycle(ctx, dynamics):
    """MIXED batches on uniform grids smooth with matrix-free products (cell
    moduli x the shared element matrix, plus the consistent mass in the
    dynamics mode); FP64 batches use the stored values.  One V-cycle of each
    on the same vector agree to FP32 accuracy, boundary cells included."""
    print(f"\n[case] matrix-free vs stored-value cycle dynamics={dynamics}", flush=True)
    model = msa.cantilever(12, 12, 8, lz=0.5)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    rng = np.random.default_rng(13)
    E = np.stack([lognormal(rng, model.n_elem) for _ in range(2)])
    r = rng.standard_normal(model.n_free)
    z = {}
    for prec in (L.MSA_FP64, L.MSA_MIXED):
        b = msa.Batch(ctx, [mesh], [2], prec, 0)
        b.set_modulus(0, E)
        b.set_rhs(0, model.f)
        if dynamics:
            b.set_dynamics(7.85e-9, 5.0, 1e-5, 0.25, 0.5, 1e-4)
        b.assemble(0.3)
        b.set_preconditioner(msa.MSA_PRECOND_MULTIGRID)
        b.set_smoothing_interval(0)   # the same (Gershgorin) interval in both precisions
        info = b.mg_info(0)
        assert info["levels"] == 2
        assert info["matrix_free"] == (prec == L.MSA_MIXED)
        z[prec] = b.mg_apply(0, 1, r)
        b.close()
    mesh.close()
    assert rel(z[L.MSA_MIXED], z[L.MSA_FP64]) < 2e-5


def _grid_variant(kind):
    """Uniform grids for the matrix-free forms: "long" (80 x 4 x 4: rows of
    80 free nodes, so wavefronts start and end inside rows and the lanes at
    both wave ends load their x - 1 / x + 1 columns themselves), "hole"
    (12 x 12 x 8 with one interior node fixed: the free nodes are not a
    box, so the general form with the grid-point table runs) and "roller"
    (z fixed on the z = 0 face as well: partially free blocks, box form,
    the ELL transfers)."""
    from types import SimpleNamespace
    from msolve_amd.core import enumerate_dofs
    if kind == "long":
        m = msa.cantilever(80, 4, 4, lx=4.0, ly=0.2, lz=0.2)
        return SimpleNamespace(xyz=m.xyz, conn=m.conn, dof=m.dof, n_free=m.n_free,
                               n_elem=m.n_elem, f=m.f)
    m = msa.cantilever(12, 12, 8, lz=0.5)
    nX, nY = 13, 13
    cons = np.asarray(m.constrained, np.uint8).reshape(-1, 3).copy()
    if kind == "hole":
        cons[6 + nX * (6 + nY * 4), :] = 1
    else:
        cons[np.arange(nX * nY), 2] = 1
    dof, n_free = enumerate_dofs(cons)
    xyz = np.asarray(m.xyz).reshape(-1, 3)
    f = np.zeros(n_free)
    tip = np.nonzero(np.isclose(xyz[:, 0], xyz[:, 0].max()))[0]
    for nd in tip:
        if dof[3 * nd + 2] >= 0:
            f[dof[3 * nd + 2]] = -100.0 / tip.size
    return SimpleNamespace(xyz=m.xyz, conn=m.conn, dof=dof, n_free=n_free, n_elem=m.n_elem, f=f)


@pytest.mark.parametrize("kind", ["long", "hole", "roller"])
def test_matrix_free_forms_match_oracle_cycle(ctx, kind):
    """One V-cycle of a MIXED batch (matrix-free products: box-grid form
    with the x columns shared between lanes, or the general form) and of an
    FP64 batch (stored values) against the oracle's cycle, on grids that
    exercise rows longer than a wavefront, a non-box set of free nodes and
    partially free blocks.  MIXED batches factor the coarsest level from an
    FP64 copy of its values: on the thin 80 x 4 x 4 beam the FP32 rounding
    of that operator, amplified by its condition number, would change the
    cycle by 1e-4 (the planted check below), so both precisions are held to
    the plain oracle cycle."""
    import scipy.sparse as sp
    print(f"\n[case] matrix-free forms vs oracle cycle grid={kind}", flush=True)
    model = _grid_variant(kind)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    rng = np.random.default_rng(17)
    E = lognormal(rng, model.n_elem)
    r = rng.standard_normal(model.n_free)
    def oracle(lam=None):
        levels, Ps = MG.hierarchy(np.asarray(model.xyz).reshape(-1, 3), model.conn, model.dof,
                                  model.n_free, E)
        if lam is not None:
            for L_, l in zip(levels[:-1], lam):
                L_.lam = float(l)
        return levels, Ps
    bar = {L.MSA_FP64: 1e-9, L.MSA_MIXED: 1e-6}
    if kind == "long":  # the bar is tight enough to see an FP32 coarsest factor
        levels, Ps = oracle()
        zg = MG.vcycle(levels, Ps, r)
        Kc = levels[-1].K
        levels[-1].K = sp.csr_matrix((Kc.data.astype(np.float32).astype(np.float64), Kc.indices,
                                      Kc.indptr), shape=Kc.shape)
        assert rel(MG.vcycle(levels, Ps, r), zg) > 10 * bar[L.MSA_MIXED]
    for prec in (L.MSA_FP64, L.MSA_MIXED):
        b = msa.Batch(ctx, [mesh], [1], prec, 0)
        b.set_modulus(0, E[None, :])
        b.set_rhs(0, model.f)
        b.assemble(0.3)
        b.set_preconditioner(msa.MSA_PRECOND_MULTIGRID)
        info = b.mg_info(0)
        z = b.mg_apply(0, 0, r)
        levels, Ps = oracle(b.mg_lambda(0, 0))   # the oracle's cycle on the GPU's interval
        zref = MG.vcycle(levels, Ps, r)
        b.close()
        assert info["levels"] == len(Ps)  # coarse levels
        assert info["matrix_free"] == (prec == L.MSA_MIXED)
        assert rel(z, zref) < bar[prec], (prec, rel(z, zref))
    mesh.close()


def _box_variant(kind):
    """Box grids of fully free blocks for the matrix-core smoother:
    "cube" 12 x 12 x 8 (two y blocks of 7 and 6 node rows), "wide" 64 x 6 x 4
    (64 cells in x, the kernel's limit; one y block of 7 rows), "clamped2"
    10 x 8 x 6 with the x = 0 and z = 0 faces clamped (the box starts at
    z = 1, so the first cell layer's bottom plane is not owned) and "rows"
    10 x 16 x 4 (17 node rows: 7 + 7 + 3).  Every count is even, so each
    grid coarsens at least once."""
    from types import SimpleNamespace
    from msolve_amd.core import enumerate_dofs
    if kind == "cube":
        m = msa.cantilever(12, 12, 8, lz=0.5)
    elif kind == "wide":
        m = msa.cantilever(64, 6, 4, lx=4.0, ly=0.4, lz=0.3)
    elif kind == "rows":
        m = msa.cantilever(10, 16, 4, ly=1.6, lz=0.4)
    else:
        m = msa.cantilever(10, 8, 6, lz=0.6)
        cons = np.asarray(m.constrained, np.uint8).reshape(-1, 3).copy()
        cons[np.arange(11 * 9), :] = 1            # the z = 0 face
        dof, n_free = enumerate_dofs(cons)
        xyz = np.asarray(m.xyz).reshape(-1, 3)
        f = np.zeros(n_free)
        tip = np.nonzero(np.isclose(xyz[:, 0], xyz[:, 0].max()))[0]
        for nd in tip:
            if dof[3 * nd + 2] >= 0:
                f[dof[3 * nd + 2]] = -100.0 / tip.size
        return SimpleNamespace(xyz=m.xyz, conn=m.conn, dof=dof, n_free=n_free,
                               n_elem=m.n_elem, f=f)
    return SimpleNamespace(xyz=m.xyz, conn=m.conn, dof=m.dof, n_free=m.n_free,
                           n_elem=m.n_elem, f=m.f)


@pytest.mark.parametrize("kind,dynamics", [("cube", False), ("wide", False), ("clamped2", False),
                                           ("rows", False), ("cube", True)])
def test_matrix_core_smoother_matches_valu_form_and_oracle(ctx, kind, dynamics):
    """One V-cycle of a MIXED batch with the finest (and coarse) box-grid
    products on the matrix cores against the same cycle with the vector-ALU
    products (both FP32: bar 1e-6) and, in statics, against the oracle's
    cycle (bar 1e-6, as the other matrix-free forms); then a mixed solve to
    |f-Ku|/|f| < 1e-10 on the matrix cores reaches the oracle's direct
    solution to 1e-8."""
    print(f"\n[case] matrix-core smoother grid={kind} dynamics={dynamics}", flush=True)
    model = _box_variant(kind)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    rng = np.random.default_rng(29)
    E = np.stack([lognormal(rng, model.n_elem) for _ in range(3)])
    r = rng.standard_normal(model.n_free)
    z = {}
    for mc in (True, False):
        b = msa.Batch(ctx, [mesh], [3], L.MSA_MIXED, 1)
        b.set_modulus(0, E)
        b.set_rhs(0, model.f)
        if dynamics:
            b.set_dynamics(7.85e-9, 5.0, 1e-5, 0.25, 0.5, 1e-4)
        b.set_matrix_cores(mc)
        b.assemble(0.3)
        b.set_preconditioner(msa.MSA_PRECOND_MULTIGRID)
        b.set_smoothing_interval(0)   # the Gershgorin interval: both forms and the oracle share it
        assert b.mg_info(0)["matrix_free"]
        z[mc] = [b.mg_apply(0, s, r) for s in range(3)]
        if mc and not dynamics:
            b.solve(tol=1e-10, max_it=2000, graph_iters=8)
            st = b.stats(0)
            assert (st["status"] == msa.MSA_OK).all(), st
            for s in range(3):
                ke, bad = O.ke_mesh(model.conn, model.xyz, E[s], 0.3)
                ri, data = O.skyline_assemble(model.conn, model.dof, model.n_free, ke)
                fac, rc = O.skyline_factorize(ri, data, 1e-15)
                assert bad == 0 and rc == 0
                assert rel(b.solution(0, s), O.skyline_solve(ri, fac, model.f)) < 1e-8
        b.close()
    mesh.close()
    for s in range(3):
        d = rel(z[True][s], z[False][s])
        print(f"[matrix cores] sample {s}: cycle vs vector-ALU form {d:.2e}", flush=True)
        assert d < 1e-6
    if not dynamics:
        levels, Ps = MG.hierarchy(np.asarray(model.xyz).reshape(-1, 3), model.conn, model.dof,
                                  model.n_free, E[1])
        assert rel(z[True][1], MG.vcycle(levels, Ps, r)) < 1e-6


MARGIN = {4: 1.3, 6: 1.2, 8: 1.2}


def test_lanczos_interval_brackets_the_spectrum(ctx):
    """The smoothing interval's upper end from 4 Lanczos steps (margin 1.3,
    the defaults), from 6 and from 8 (margin 1.2) against the spectrum of D^-1 K of each level, from the oracle's
    level operators (scipy eigsh): it lies above the top eigenvalue and
    within the margin of it, and below the Gershgorin bound (which is what
    0 Lanczos steps gives, to 1e-12 in FP64).  MIXED batches (FP32 vectors)
    find the FP64 value to 1e-3.  The tighter interval saves PCG
    iterations (FP64, 12^3)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    print("\n[case] Lanczos smoothing interval", flush=True)
    out = {}
    for prec in (L.MSA_FP64, L.MSA_MIXED):
        for steps in (4, 6, 8, 0):
            b, models, meshes, Es = make(ctx, [8, 12], 2, prec, seed=21)
            b.set_smoothing_interval(steps, MARGIN.get(steps, 1.2))
            b.solve(1e-11, 5000)
            for c in range(2):
                assert (b.stats(c)["status"] == msa.MSA_OK).all()
                for s in range(2):
                    out[(prec, steps, c, s)] = (b.mg_lambda(c, s), int(b.stats(c)["iters"][s]))
            close(b, meshes)
    for c, m in enumerate(models):
        for s in range(2):
            levels, _ = oracle_levels(m, Es[c][s])
            lam0, it0 = out[(L.MSA_FP64, 0, c, s)]
            for l, L_ in enumerate(levels[:-1]):
                dh = sp.diags(np.sqrt(L_.dinv))
                top = sla.eigsh(dh @ L_.K @ dh, k=1, which="LA", return_eigenvectors=False,
                                tol=1e-10)[0]
                assert abs(lam0[l] - L_.lam) <= 1e-12 * L_.lam
                for k in (4, 6, 8):
                    lam = out[(L.MSA_FP64, k, c, s)][0]
                    assert len(lam) == len(levels) - 1
                    print(f"[lanczos] class {c} sample {s} level {l} steps {k}: top {top:.5f} "
                          f"lam {lam[l]:.5f} gershgorin {L_.lam:.5f}", flush=True)
                    assert top * (1 - 1e-9) <= lam[l] <= min(MARGIN[k] * top, L_.lam) * (1 + 1e-12)
                    assert abs(out[(L.MSA_MIXED, k, c, s)][0][l] - lam[l]) <= 1e-3 * lam[l]
            if c == 1:
                for k in (4, 6, 8):
                    it = out[(L.MSA_FP64, k, c, s)][1]
                    assert it < it0, (k, it, it0)

