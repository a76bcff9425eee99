"""The Monte Carlo analyzer's outputs and stiffness-storage modes on the GPU
(MonteCarloAnalyzerWithStochasticMaterial.cs:18-24, :182-296, :298-402).

* Normal: the value, iteration and time files of SolveNormal and the
  generated randoms file; every simulation's displacement equals the
  oracle's skyline direct solve of the oracle's KL modulus (1e-9).
* StoreToDisk / StoreToDiskAndCalculate: each simulation's K written with
  SkylineMatrix2D.WriteToFile equals the oracle's skyline assembly of the
  same modulus (1e-13); StoreToDisk solves nothing.
* LoadFromDiskAndCalculate: the stored matrices read back in the order
  file's order, solved with PCG preconditioned per block by the band factor
  of the block's middle matrix (MakePreconditioner), give the Normal run's
  displacements of those matrices (1e-9), and the block's own preconditioner
  matrix converges in at most two products.
* msa_batch_set_matrix_csr: a loaded copy of an assembled K solves
  bit-identically to the assembled batch; it rejects mixed batches, entries
  outside the sparsity pattern, and the multigrid preconditioner."""
import os

import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd import _lib as L
from msolve_amd.analyzers import StiffnessMatrixProductionMode as Mode
from msolve_amd.matrix_io import read_skyline, skyline_to_dense
from msolve_amd.stochastic import kl_table, read_randoms_file
from oracle import oracle as O

pytestmark = pytest.mark.gpu

M = 8


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def oracle_direct(m, E):
    ke, bad = O.ke_mesh(m.conn, m.xyz, E, 0.3)
    assert bad == 0
    ri, data = O.skyline_assemble(m.conn, m.dof, m.n_free, ke)
    fac, rc = O.skyline_factorize(ri, data.copy(), 1e-15)
    assert rc == 0
    return ri, data, O.skyline_solve(ri, fac, m.f)


def analyzer(ctx, model, ns, table, **kw):
    return msa.MonteCarloAnalyzerWithStochasticMaterial(
        ctx, [model], [ns], table, M, meanE=2.1e5, cov=0.3, seed=4242, **kw)


def test_normal_mode_files_match_oracle(ctx, tmp_path):
    print("\n[case] MC analyzer normal mode files", flush=True)
    model = msa.cantilever(4)
    table = kl_table([1.0, 1.0, 1.0], 0.3, M)
    ns = 6
    log = str(tmp_path / "mc")
    an = analyzer(ctx, model, ns, table, fileNameForLogging=log)
    an.Initialize()
    an.Solve(tolerance=1e-12, maxIterations=5000, graph_iters=8)
    vals = [float(l) for l in open(f"{log}-{M}-0.txt").read().split()]
    iters = [int(l) for l in open(f"{log}-{M}-0-Iters.txt").read().split()]
    times = open(f"{log}-{M}-0-Times.txt").read().splitlines()
    assert len(vals) == ns and len(iters) == ns
    assert [t.split(":")[0] for t in times] == ["Elements", "Factorize", "Solution", "All",
                                                 "Total iterations"]
    assert int(times[-1].split(":")[1]) == sum(iters)
    st = an.Stats(0)
    assert iters == st["iters"].tolist() and (st["status"] == msa.MSA_OK).all()
    xi = read_randoms_file(str(tmp_path / f"randoms{M}.txt"), M, ns)
    assert np.array_equal(xi, an.randomNumbers)
    dof = model.n_free - 1
    for s in range(ns):
        Eo, bad = O.kl_field(model.conn, model.xyz, table, xi[s], model.domain, 2.1e5, 0.3)
        assert bad == 0
        _, _, ud = oracle_direct(model, Eo)
        u = an.SimulationSolution(s)
        assert rel(u, ud) < 1e-9
        assert abs(vals[s] - u[dof]) <= 1e-14 * abs(u[dof])   # 15 significant digits


@pytest.mark.parametrize("mode", [Mode.StoreToDisk, Mode.StoreToDiskAndCalculate])
def test_store_modes_write_reference_skyline(ctx, tmp_path, mode):
    print(f"\n[case] MC analyzer {mode.name}", flush=True)
    model = msa.cantilever(3)
    table = kl_table([1.0, 1.0, 1.0], 0.3, M)
    ns, start = 3, 5
    kpath = str(tmp_path / "K.txt")
    an = analyzer(ctx, model, ns, table, stiffnessMatrixProductionMode=mode,
                  stiffnessMatrixPath=kpath, simulationStartFrom=start,
                  randomsReadFileName=None)
    an.Initialize()
    an.Solve(tolerance=1e-12, maxIterations=5000)
    assert (an.values is None) == (mode == Mode.StoreToDisk)
    for s in range(ns):
        ri, data = read_skyline(str(tmp_path / f"KSub1Sim{start + s}.txt"), None)
        Eo, _ = O.kl_field(model.conn, model.xyz, table, an.randomNumbers[s], model.domain,
                           2.1e5, 0.3)
        ro, do, ud = oracle_direct(model, Eo)
        assert np.array_equal(ri, ro)
        Kg, Ko = skyline_to_dense(ri, data), skyline_to_dense(ro, do)
        assert np.abs(Kg - Ko).max() <= 1e-13 * np.abs(Ko).max()
        if mode == Mode.StoreToDiskAndCalculate:
            assert rel(an.SimulationSolution(start + s), ud) < 1e-9


def test_load_from_disk_with_block_preconditioner(ctx, tmp_path):
    print("\n[case] MC analyzer LoadFromDiskAndCalculate", flush=True)
    model = msa.cantilever(4)
    table = kl_table([1.0, 1.0, 1.0], 0.3, M)
    ns = 6
    kpath = str(tmp_path / "K.txt")
    log = str(tmp_path / "mc")
    store = analyzer(ctx, model, ns, table, stiffnessMatrixProductionMode=Mode.StoreToDiskAndCalculate,
                     stiffnessMatrixPath=kpath)
    store.Initialize()
    store.Solve(tolerance=1e-12, maxIterations=5000)
    u_of = {s: store.SimulationSolution(s) for s in range(ns)}
    # 6 simulations in blocks of 4: the second block's preconditioner is
    # order[6], a matrix beyond the simulations
    order = [3, 1, 4, 0, 5, 2, 1, 0]
    with open(f"{log}-Order.txt", "w") as fh:
        fh.write("\n".join(str(i) for i in order) + "\n")
    an = analyzer(ctx, model, ns, table, stiffnessMatrixProductionMode=Mode.LoadFromDiskAndCalculate,
                  stiffnessMatrixPath=kpath, fileNameForLogging=log, blockSize=4)
    an.Solve(tolerance=1e-12, maxIterations=5000, graph_iters=8)
    assert [mats for mats, _ in an._blocks] == [[4, 3, 1, 0], [1, 5, 2]]
    for i in range(ns):
        assert rel(an.SimulationSolution(i), u_of[order[i]]) < 1e-9
    assert an.iterations[2] <= 2       # order[2] is the first block's preconditioner matrix
    assert an.iterations.max() < store.iterations.max()
    vals = [float(l) for l in open(f"{log}-{M}-0.txt").read().split()]
    assert np.allclose(vals, [u_of[order[i]][model.n_free - 1] for i in range(ns)], rtol=1e-9)
    assert len(open(f"{log}-Times-4-0.txt").read().splitlines()) == 4
    # the reference's error when a block's preconditioner index is past the file
    with open(f"{log}-Order.txt", "w") as fh:
        fh.write("\n".join(str(i) for i in order[:6]) + "\n")
    bad = analyzer(ctx, model, ns, table, stiffnessMatrixProductionMode=Mode.LoadFromDiskAndCalculate,
                   stiffnessMatrixPath=kpath, fileNameForLogging=log, blockSize=4)
    with pytest.raises(msa.MsaError):
        bad.Solve(tolerance=1e-12)


def test_set_matrix_csr_equals_assembly_and_rejects(ctx):
    print("\n[case] set_matrix_csr", flush=True)
    model = msa.cantilever(4)
    rng = np.random.default_rng(6)
    s = np.sqrt(np.log(1.09))
    E = 2.1e5 * np.exp(-0.5 * s * s + s * rng.standard_normal((3, model.n_elem)))
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    sols = []
    for load in (False, True):
        b = msa.Batch(ctx, [mesh], [3], L.MSA_FP64, 0)
        if load:
            for smp in range(3):
                b.set_matrix_csr(0, smp, *csr[smp])
        else:
            b.set_modulus(0, E)
            b.assemble(0.3)
            csr = [b.matrix_csr(0, smp) for smp in range(3)]
        b.set_rhs(0, model.f)
        b.solve(tol=1e-12, max_it=5000, graph_iters=8)
        assert (b.stats(0)["status"] == msa.MSA_OK).all()
        sols.append([b.solution(0, smp) for smp in range(3)])
        if load:
            rp, col, val = csr[0]
            assert all(np.array_equal(b.matrix_csr(0, k)[2], csr[k][2]) for k in range(3))
            # an entry outside the pattern: row 0 against the last DOF
            far = model.n_free - 1
            assert far not in col[rp[0]:rp[1]]
            rp2 = rp.copy()
            rp2[1:] += 1
            col2 = np.insert(col, rp[1], far)
            val2 = np.insert(val, rp[1], 1.0)
            with pytest.raises(msa.MsaError):
                b.set_matrix_csr(0, 0, rp2, col2, val2)
            val2[rp[1]] = 0.0          # a zero there is skipped
            b.set_matrix_csr(0, 0, rp2, col2, val2)
            with pytest.raises(msa.MsaError, match="moduli"):
                b.set_preconditioner(L.MSA_PRECOND_MULTIGRID)   # its values need the moduli
                b.solve(tol=1e-10, max_it=100)
        b.close()
    for smp in range(3):
        assert np.array_equal(sols[0][smp], sols[1][smp])
    bm = msa.Batch(ctx, [mesh], [1], L.MSA_MIXED, 0)
    with pytest.raises(msa.MsaError):
        bm.set_matrix_csr(0, 0, *csr[0])
    bm.close()
    mesh.close()
