"""Host-side mirrors of the reference's plug-in interfaces for this path,
each backed by the HIP library through the C ABI.

  ElementStructuralStiffnessProvider  IElementMatrixProvider.Matrix
      (ISAAR.MSolve.PreProcessor/Providers/ElementStructuralStiffnessProvider
      .cs:11-38 -> Hexa8WithStochasticMaterial.StiffnessMatrix, H8WSM.cs:61-110)
  SolverPCG                           IIterativeSolver / ISolver over one model
      (ISAAR.MSolve.Solvers/PCGSkyline/SolverPCGSkyline.cs:28-157 with the
      Jacobi calculator of the commented-out :85-91,164)
  SolverSkyline                       ISolver, direct (ISAAR.MSolve.Solvers/
      Skyline/SolverSkyline.cs:55-205) on the GPU band Cholesky
  MonteCarloAnalyzerWithStochasticMaterial
      (ISAAR.MSolve.Analyzers/MonteCarloAnalyzerWithStochasticMaterial.cs:
      182-296): the per-sample loop as one batched device job.
  NewmarkDynamicAnalyzer
      (ISAAR.MSolve.Analyzers/NewmarkDynamicAnalyzer.cs:38-315 over
      ProblemStructural, ISAAR.MSolve.Problems/ProblemStructural.cs:104-211):
      linear implicit Newmark of many stochastic-modulus samples at once.
Argument meaning and defaults follow the reference (SolverPCG<T>.Solve() uses
20000 iterations and 1e-7, SolverPCGSkyline.cs:102-108); errors raise
MsaError instead of the reference's InvalidOperationException.
"""
from __future__ import annotations

import enum
import os
import time

import numpy as np

from . import _lib as L
from .core import Batch, Context, Mesh, ke_packed
from .matrix_io import read_skyline, write_skyline
from .stochastic import read_randoms_file, standard_normal, write_randoms_file


class ElementStructuralStiffnessProvider:
    """Batched element stiffness: Matrix(xyz8[n,8,3], E[n], nu) -> packed
    SymmetricMatrix2D data [n, 300] (upper triangle, row-major)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def Matrix(self, xyz8, E, nu=0.3):
        ke, status = ke_packed(self.ctx, xyz8, E, nu)
        if (status != L.MSA_OK).any():
            bad = int(np.nonzero(status != L.MSA_OK)[0][0])
            raise L.MsaError(L.MSA_NEG_DETJ, f"element {bad}: Jacobian determinant below 1e-8")
        return ke


class SolverPCG:
    """FP64 PCG on one model (x0 = 0, stop at |r|/|f| < tolerance); Jacobi
    preconditioner as the reference's SolverPCG, or preconditioner=
    MSA_PRECOND_MULTIGRID for the geometric V-cycle."""

    def __init__(self, ctx: Context, model, E, nu=0.3, tile=1, preconditioner=L.MSA_PRECOND_JACOBI):
        self.ctx, self.model, self.nu = ctx, model, nu
        self.E = np.ascontiguousarray(E, np.float64).ravel()
        self.tile = tile
        self.preconditioner = preconditioner
        self.mesh = None
        self.batch = None
        self.VectorX = None
        self.CurrentIteration = 0
        self.Status = None

    def Initialize(self):
        m = self.model
        self.mesh = Mesh(self.ctx, m.xyz, m.conn, m.dof, m.n_free)
        self.batch = Batch(self.ctx, [self.mesh], [1], L.MSA_FP64, self.tile)
        if self.preconditioner != L.MSA_PRECOND_JACOBI:
            self.batch.set_preconditioner(self.preconditioner)
        self.batch.set_modulus(0, self.E)
        self.batch.set_rhs(0, m.f)
        self.batch.assemble(self.nu)

    def Solve(self, maxIterations=20000, tolerance=1e-7, graph_iters=0):
        if self.batch is None:
            self.Initialize()
        self.batch.solve(tolerance, maxIterations, graph_iters=graph_iters)
        st = self.batch.stats(0)
        self.VectorX = self.batch.solution(0, 0)
        # the reference keeps the loop index at the break (one less than the
        # number of products on convergence, SolverPCG.cs:117-138)
        n = int(st["iters"][0])
        self.Status = int(st["status"][0])
        self.CurrentIteration = n - 1 if self.Status == L.MSA_OK and n > 0 else n
        self.Iterations = n
        self.RelativeResidual = float(st["relres"][0])
        return self.VectorX


class SolverSkyline:
    """Direct FP64 solve of one model (SolverSkyline, ISAAR.MSolve.Solvers/
    Skyline/SolverSkyline.cs:55-205): Initialize() assembles and factorizes
    (a banded Cholesky on the GPU in place of the skyline LDL^T), Solve()
    forward/back-substitutes; the solution is in VectorX."""

    def __init__(self, ctx: Context, model, E, nu=0.3):
        self.ctx, self.model, self.nu = ctx, model, nu
        self.E = np.ascontiguousarray(E, np.float64).ravel()
        self.mesh = None
        self.batch = None
        self.VectorX = None

    def Initialize(self):
        m = self.model
        self.mesh = Mesh(self.ctx, m.xyz, m.conn, m.dof, m.n_free)
        self.batch = Batch(self.ctx, [self.mesh], [1], L.MSA_FP64, 1)
        self.batch.set_modulus(0, self.E)
        self.batch.set_rhs(0, m.f)
        self.batch.set_solver(L.MSA_SOLVER_DIRECT)
        self.batch.assemble(self.nu)
        self.batch.factorize()

    def Solve(self):
        if self.batch is None:
            self.Initialize()
        self.batch.direct_solve()
        st = self.batch.stats(0)
        self.Status = int(st["status"][0])
        self.RelativeResidual = float(st["relres"][0])
        self.VectorX = self.batch.solution(0, 0)
        return self.VectorX


class StiffnessMatrixProductionMode(enum.IntEnum):
    """MonteCarloAnalyzerWithStochasticMaterial.cs:18-24."""
    Normal = 0
    StoreToDisk = 1
    LoadFromDiskAndCalculate = 2
    StoreToDiskAndCalculate = 3


def _timespan(seconds: float) -> str:
    """.NET TimeSpan.ToString(): [d.]hh:mm:ss[.fffffff]."""
    ticks = int(round(seconds * 1e7))
    sec, frac = divmod(ticks, 10 ** 7)
    d, sec = divmod(sec, 86400)
    h, sec = divmod(sec, 3600)
    m, sec = divmod(sec, 60)
    out = (f"{d}." if d else "") + f"{h:02d}:{m:02d}:{sec:02d}"
    return out + (f".{frac:07d}" if frac else "")


def _dotnet_double(x: float) -> str:
    """.NET Framework Double.ToString(): 15 significant digits, "E+XX" exponent."""
    t = format(float(x), ".15G")
    if "E" in t:
        mant, e = t.split("E")
        t = f"{mant}E{e[0]}{abs(int(e)):02d}"
    return t


class MonteCarloAnalyzerWithStochasticMaterial:
    """Monte Carlo over lognormal KL modulus samples, all samples solved
    together (MonteCarloAnalyzerWithStochasticMaterial.cs:26-405).

    `models` is one model per mesh class; `samples` the number of samples of
    each class; `table` the KL table (stochastic.kl_table or read_kl_file);
    `randomsReadFileName` optionally supplies xi in the reference's file
    layout, read from `simulationStartFrom` on.

    Outputs (:232-296), written when `fileNameForLogging` is given (None: no
    files, the results stay in memory):
      "<f>-<M>-<start>.txt"        the solution at the output DOF, one
                                   simulation per line (.NET Double.ToString);
      "<f>-<M>-<start>-Iters.txt"  PCG iterations per simulation (the
                                   reference writes an array it never fills:
                                   zeros; these are the counts);
      "<f>-<M>-<start>-Times.txt"  Elements / Factorize / Solution / All /
                                   Total iterations (TimeSpan strings);
      "randoms<M>.txt"             beside <f>, the generated xi, variable-major
                                   (Initialize, :186-205), when not read.
    `outputDOF` is one free-DOF index for every class or one per class; the
    reference hard-codes node 1, X (:236); None takes each class's last free
    DOF.  The phase buckets are batch totals: "element" = KL modulus + Ke +
    assembly, "factorize" = the preconditioner's values (first-sample band
    factor or multigrid levels; ~0 for Jacobi), "solution" = the PCG.

    `stiffnessMatrixProductionMode` (:18-24, :368-402):
      StoreToDisk                writes every simulation's K with
                                 SkylineMatrix2D.WriteToFile as
                                 "<path>/<name>Sub<c+1>Sim<i><ext>" and solves
                                 nothing; StoreToDiskAndCalculate also solves;
      LoadFromDiskAndCalculate   reads "<f>-Order.txt" (one matrix number a
                                 line), and for simulation i solves the stored
                                 matrix Order[i] with PCG preconditioned by the
                                 band factor of matrix Order[b + blockSize/2],
                                 b = the start of i's block of `blockSize`
                                 (SolveWithOrder / MakePreconditioner,
                                 :298-366); one model, an MSA_FP64 batch (the
                                 stored matrices replace the assembly), times
                                 in "<f>-Times-<blockSize>-<start>.txt" (all,
                                 element, factorize, solution).  Stored files
                                 are read in either layout (the reference
                                 writes ASCII and reads binary)."""

    def __init__(self, ctx: Context, models, samples, table, expansionOrder: int,
                 meanE=2.1e5, cov=0.5, nu=0.3, precision=L.MSA_FP64, tile=0, seed=20261202,
                 randomsReadFileName: str | None = None, simulationStartFrom: int = 0,
                 preconditioner=L.MSA_PRECOND_JACOBI, fileNameForLogging: str | None = None,
                 stiffnessMatrixProductionMode=StiffnessMatrixProductionMode.Normal,
                 stiffnessMatrixPath: str = "", blockSize: int = 5, outputDOF=None):
        self.ctx = ctx
        self.preconditioner = preconditioner
        self.models = list(models)
        self.samples = [int(s) for s in samples]
        self.table = table
        self.expansionOrder = expansionOrder
        self.meanE, self.cov, self.nu = meanE, cov, nu
        self.precision, self.tile, self.seed = precision, tile, seed
        self.randomsReadFileName = randomsReadFileName
        self.simulationStartFrom = simulationStartFrom
        self.fileNameForLogging = fileNameForLogging
        self.stiffnessMatrixProductionMode = StiffnessMatrixProductionMode(
            stiffnessMatrixProductionMode)
        self.stiffnessMatrixPath = stiffnessMatrixPath
        self.blockSize = int(blockSize)
        if len(self.samples) != len(self.models):
            raise L.MsaError(L.MSA_ERR_ARG, "one sample count per model")
        if outputDOF is None:
            self.outputDOF = [m.n_free - 1 for m in self.models]
        elif np.ndim(outputDOF) == 0:
            self.outputDOF = [int(outputDOF)] * len(self.models)
        else:
            self.outputDOF = [int(d) for d in outputDOF]
        for d, m in zip(self.outputDOF, self.models):
            if not 0 <= d < m.n_free:
                raise L.MsaError(L.MSA_ERR_ARG, f"output DOF {d} outside [0, {m.n_free})")
        if self.stiffnessMatrixProductionMode == StiffnessMatrixProductionMode.LoadFromDiskAndCalculate:
            if not fileNameForLogging:
                raise L.MsaError(L.MSA_ERR_ARG, "LoadFromDiskAndCalculate reads "
                                                "<fileNameForLogging>-Order.txt")
            if len(self.models) != 1 or precision != L.MSA_FP64:
                raise L.MsaError(L.MSA_ERR_ARG, "LoadFromDiskAndCalculate takes one model and "
                                                "an MSA_FP64 batch")
            if self.blockSize < 1:
                raise L.MsaError(L.MSA_ERR_ARG, f"blockSize {blockSize}")
        self.batch = None
        self.randomNumbers = None
        self.matrixOrder: list[int] = []
        self.times = {}
        self.values = None
        self.iterations = None

    # -------------------------------------------------------------- files
    def _log_name(self, suffix: str) -> str:
        return (f"{self.fileNameForLogging}-{self.expansionOrder}-{self.simulationStartFrom}"
                f"{suffix}.txt")

    def _matrix_name(self, cls: int, simulation: int) -> str:
        """"<path>/<name>Sub<id>Sim<i><ext>" of WriteMatricesToFile (:368-379)."""
        name = self.stiffnessMatrixPath or "K"
        path = os.path.dirname(name)
        base, ext = os.path.splitext(os.path.basename(name))
        return os.path.join(path, f"{base}Sub{cls + 1}Sim{simulation}{ext}")

    def _write_lines(self, path: str, lines):
        with open(path, "w", newline="\r\n") as fh:
            fh.write("".join(f"{l}\n" for l in lines))

    def _write_times(self, path: str, lines):
        self._write_lines(path, lines)

    # ----------------------------------------------------------- analyzer
    def Initialize(self):
        total = sum(self.samples)
        M = self.expansionOrder
        if self.randomsReadFileName:
            xi = read_randoms_file(self.randomsReadFileName, M)
            self.randomNumbers = xi[self.simulationStartFrom:self.simulationStartFrom + total]
        else:
            self.randomNumbers = standard_normal(self.seed, total, M)
            if self.fileNameForLogging:
                write_randoms_file(os.path.join(os.path.dirname(self.fileNameForLogging),
                                                f"randoms{M}.txt"), self.randomNumbers)
        if self.stiffnessMatrixProductionMode == StiffnessMatrixProductionMode.LoadFromDiskAndCalculate:
            self.BuildMatrices()
            return
        self.meshes = [Mesh(self.ctx, m.xyz, m.conn, m.dof, m.n_free) for m in self.models]
        self.batch = Batch(self.ctx, self.meshes, self.samples, self.precision, self.tile)
        if self.preconditioner != L.MSA_PRECOND_JACOBI:
            self.batch.set_preconditioner(self.preconditioner)
        off = 0
        for c, m in enumerate(self.models):
            self.batch.set_kl(c, self.table, m.domain, self.meanE, self.cov)
            self.batch.set_xi(c, self.randomNumbers[off:off + self.samples[c]])
            self.batch.set_rhs(c, m.f)
            off += self.samples[c]
        self.ctx.synchronize()

    def BuildMatrices(self):
        """LoadFromDiskAndCalculate: reads the matrix order (:128-139)."""
        if self.stiffnessMatrixProductionMode == StiffnessMatrixProductionMode.LoadFromDiskAndCalculate:
            with open(f"{self.fileNameForLogging}-Order.txt") as fh:
                self.matrixOrder = [int(l) for l in fh.read().split()]

    def Solve(self, tolerance=1e-10, maxIterations=20000, inner_tol=1e-4, max_outer=20,
              graph_iters=0):
        mode = self.stiffnessMatrixProductionMode
        if mode == StiffnessMatrixProductionMode.LoadFromDiskAndCalculate:
            if not self.matrixOrder:
                self.Initialize()
            return self._solve_with_order(tolerance, maxIterations, graph_iters)
        if self.batch is None:
            self.Initialize()
        return self._solve_normal(tolerance, maxIterations, inner_tol, max_outer, graph_iters)

    def _solve_normal(self, tolerance, maxIterations, inner_tol, max_outer, graph_iters):
        """SolveNormal (:232-296) over the whole batch."""
        mode = self.stiffnessMatrixProductionMode
        b, ctx = self.batch, self.ctx
        t_all = time.perf_counter()
        t = time.perf_counter()
        b.eval_kl()
        b.assemble(self.nu)
        ctx.synchronize()
        self.times = {"element": time.perf_counter() - t, "factorize": 0.0, "solution": 0.0}
        if mode in (StiffnessMatrixProductionMode.StoreToDisk,
                    StiffnessMatrixProductionMode.StoreToDiskAndCalculate):
            off = self.simulationStartFrom
            for c in range(len(self.models)):
                for s in range(self.samples[c]):
                    write_skyline(self._matrix_name(c, off + s), *b.matrix_skyline(c, s))
                off += self.samples[c]
        if mode == StiffnessMatrixProductionMode.StoreToDisk:
            self.times["all"] = time.perf_counter() - t_all
            return self
        t = time.perf_counter()
        b.prepare_preconditioner()
        ctx.synchronize()
        self.times["factorize"] = time.perf_counter() - t
        t = time.perf_counter()
        b.solve(tolerance, maxIterations, inner_tol, max_outer, graph_iters)
        ctx.synchronize()
        self.times["solution"] = time.perf_counter() - t
        self.times["all"] = time.perf_counter() - t_all
        values, iters = [], []
        for c in range(len(self.models)):
            st = b.stats(c)
            iters.extend(int(i) for i in st["iters"])
            values.extend(float(b.solution(c, s)[self.outputDOF[c]])
                          for s in range(self.samples[c]))
        self.values, self.iterations = np.array(values), np.array(iters)
        if self.fileNameForLogging:
            self._write_lines(self._log_name(""), (_dotnet_double(v) for v in values))
            self._write_lines(self._log_name("-Iters"), iters)
            self._write_times(self._log_name("-Times"), [
                f"Elements: {_timespan(self.times['element'])}",
                f"Factorize: {_timespan(self.times['factorize'])}",
                f"Solution: {_timespan(self.times['solution'])}",
                f"All: {_timespan(self.times['all'])}",
                f"Total iterations: {int(np.sum(iters))}"])
        return self

    def _solve_with_order(self, tolerance, maxIterations, graph_iters):
        """SolveWithOrder + MakePreconditioner (:298-366): one class per block
        of `blockSize` simulations; sample 0 of a class is the block's
        preconditioner matrix, whose band factor (MSA_PRECOND_FIRST_SAMPLE)
        preconditions every sample of the class."""
        bs, start, nsim = self.blockSize, self.simulationStartFrom, self.samples[0]
        order = self.matrixOrder
        model = self.models[0]
        if start + nsim > len(order):
            raise L.MsaError(L.MSA_ERR_ARG, f"{len(order)} matrices in the order file for "
                                            f"simulations {start}..{start + nsim - 1}")
        blocks = []   # (matrix numbers with the preconditioner's first, simulation per sample)
        for kb in range(start // bs, (start + nsim - 1) // bs + 1):
            pidx = kb * bs + bs // 2
            if pidx >= len(order):
                raise L.MsaError(L.MSA_ERR_ARG, f"block at {kb * bs}: preconditioner matrix "
                                                f"index {pidx} beyond the order file")
            sims = [i for i in range(max(kb * bs, start), min(kb * bs + bs, start + nsim))]
            if pidx in sims:
                sims.remove(pidx)
                blocks.append(([order[pidx]] + [order[i] for i in sims], [pidx] + sims))
            else:   # a partial block: the preconditioner matrix is solved too, not reported
                blocks.append(([order[pidx]] + [order[i] for i in sims], [None] + sims))
        t_all = time.perf_counter()
        self.meshes = [Mesh(self.ctx, model.xyz, model.conn, model.dof, model.n_free)]
        b = Batch(self.ctx, self.meshes * len(blocks), [len(mb) for mb, _ in blocks],
                  L.MSA_FP64, self.tile)
        self.batch = b
        b.set_preconditioner(L.MSA_PRECOND_FIRST_SAMPLE)
        t = time.perf_counter()
        for c, (mats, _) in enumerate(blocks):
            b.set_rhs(c, model.f)
            for s, mat in enumerate(mats):
                b.set_matrix_skyline(c, s, *read_skyline(self._matrix_name(0, mat), None))
        self.ctx.synchronize()
        self.times = {"element": time.perf_counter() - t}
        t = time.perf_counter()
        b.prepare_preconditioner()
        self.times["factorize"] = time.perf_counter() - t
        t = time.perf_counter()
        b.solve(tolerance, maxIterations, graph_iters=graph_iters)
        self.ctx.synchronize()
        self.times["solution"] = time.perf_counter() - t
        self.times["all"] = time.perf_counter() - t_all
        values = np.zeros(nsim)
        iters = np.zeros(nsim, np.int64)
        for c, (_, sims) in enumerate(blocks):
            st = b.stats(c)
            for s, i in enumerate(sims):
                if i is None:
                    continue
                values[i - start] = b.solution(c, s)[self.outputDOF[0]]
                iters[i - start] = st["iters"][s]
        self.values, self.iterations = values, iters
        self._blocks = blocks
        if self.fileNameForLogging:
            self._write_lines(self._log_name(""), (_dotnet_double(v) for v in values))
            self._write_times(f"{self.fileNameForLogging}-Times-{bs}-{start}.txt",
                              [_timespan(self.times[k])
                               for k in ("all", "element", "factorize", "solution")])
        return self

    def SimulationSolution(self, simulation: int):
        """The full solution of one simulation (index from simulationStartFrom)."""
        i = simulation
        if self.stiffnessMatrixProductionMode == StiffnessMatrixProductionMode.LoadFromDiskAndCalculate:
            for c, (_, sims) in enumerate(self._blocks):
                if i in sims:
                    return self.batch.solution(c, sims.index(i))
            raise L.MsaError(L.MSA_ERR_ARG, f"simulation {i} not solved")
        off = self.simulationStartFrom
        for c, n in enumerate(self.samples):
            if off <= i < off + n:
                return self.batch.solution(c, i - off)
            off += n
        raise L.MsaError(L.MSA_ERR_ARG, f"simulation {i} not solved")

    def Solution(self, cls: int, sample: int):
        return self.batch.solution(cls, sample)

    def Stats(self, cls: int):
        return self.batch.stats(cls)


def sinusoidal_history(period: float, timeStep: float, timeLength: float, magnitude: float,
                       nsteps: int):
    """MassAccelerationSinusoidalLoad[i] for i < nsteps
    (ISAAR.MSolve.PreProcessor/MassAccelerationSinusoidalLoad.cs:17-29):
    magnitude sin(2 pi / period * timeStep * i) for i < (int)(timeLength /
    timeStep), else 0."""
    steps = int(timeLength / timeStep)
    i = np.arange(nsteps, dtype=np.float64)
    out = magnitude * np.sin(2 * np.pi / period * timeStep * i)
    out[np.arange(nsteps) >= steps] = 0.0
    return out


class NewmarkDynamicAnalyzer:
    """Linear implicit Newmark over a batch of stochastic-modulus samples.

    `alpha`, `delta`, `timeStep`, `totalTime` as the reference constructor
    (NewmarkDynamicAnalyzer.cs:38-49; InitializeCoefficients throws for
    delta < 0.5 or alpha < (0.5 + delta)^2 / 4, here MsaError); `density`,
    `rayleighAlpha`, `rayleighBeta` are the Hexa8 element properties
    (Hexa8.cs:87-89); the base acceleration is a MassAccelerationSinusoidalLoad
    (period, magnitude) on DOF direction `accelerationDOF` (0/1/2 = X/Y/Z).
    The loop runs (int)(totalTime / timeStep) steps (:284) from rest; every
    effective-stiffness solve is a Jacobi-PCG started from the previous
    step's displacement (the reference's SolverPCG would iterate on a stale
    residual here; SolverSkyline is its correct solver and the oracle)."""

    def __init__(self, ctx: Context, models, E_per_class, alpha=0.25, delta=0.5,
                 timeStep=1e-4, totalTime=0.05, density=7.85e-9, rayleighAlpha=0.0,
                 rayleighBeta=0.0, period=0.05, magnitude=1.0, accelerationDOF=2, nu=0.3,
                 precision=L.MSA_MIXED, tile=0, static_loads=True,
                 preconditioner=L.MSA_PRECOND_JACOBI):
        self.ctx = ctx
        self.preconditioner = preconditioner
        self.models = list(models)
        self.E = [np.atleast_2d(np.asarray(E, np.float64)) for E in E_per_class]
        self.alpha, self.delta, self.timeStep, self.totalTime = alpha, delta, timeStep, totalTime
        self.density, self.rayleighAlpha, self.rayleighBeta = density, rayleighAlpha, rayleighBeta
        self.period, self.magnitude, self.accelerationDOF = period, magnitude, accelerationDOF
        self.nu, self.precision, self.tile, self.static_loads = nu, precision, tile, static_loads
        self.batch = None
        self.times = {}

    @property
    def steps(self) -> int:
        return int(self.totalTime / self.timeStep)

    def BuildMatrices(self, probes=None):
        self.meshes = [Mesh(self.ctx, m.xyz, m.conn, m.dof, m.n_free) for m in self.models]
        self.batch = Batch(self.ctx, self.meshes, [E.shape[0] for E in self.E], self.precision,
                           self.tile)
        if self.preconditioner != L.MSA_PRECOND_JACOBI:
            self.batch.set_preconditioner(self.preconditioner)
        for c, m in enumerate(self.models):
            self.batch.set_modulus(c, self.E[c])
            if self.static_loads:
                self.batch.set_rhs(c, m.f)
        self.batch.set_dynamics(self.density, self.rayleighAlpha, self.rayleighBeta, self.alpha,
                                self.delta, self.timeStep)
        for c, m in enumerate(self.models):
            dof = m.n_free - 1 if probes is None else probes[c]
            self.batch.set_probe(c, dof)
        self.batch.assemble(self.nu)

    def Solve(self, tolerance=1e-8, maxIterations=20000, graph_iters=16):
        if self.batch is None:
            self.BuildMatrices()
        acc = sinusoidal_history(self.period, self.timeStep, self.totalTime, self.magnitude,
                                 self.steps)
        t0 = time.perf_counter()
        self.batch.newmark(acc, self.accelerationDOF, tolerance, maxIterations, graph_iters)
        self.ctx.synchronize()
        self.times["all"] = time.perf_counter() - t0
        return self

    def History(self, cls: int, sample: int):
        return self.batch.history(cls, sample)

    def State(self, cls: int, sample: int):
        return self.batch.dynamic_state(cls, sample)
