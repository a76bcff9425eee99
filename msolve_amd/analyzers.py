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

import time

import numpy as np

from . import _lib as L
from .core import Batch, Context, Mesh, ke_packed
from .stochastic import standard_normal, read_randoms_file


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


class MonteCarloAnalyzerWithStochasticMaterial:
    """Monte Carlo over lognormal KL modulus samples, all samples solved
    together.  `models` is one model per mesh class; `samples` the number of
    samples of each class; `table` the KL table (stochastic.kl_table or
    read_kl_file); `randomsReadFileName` optionally supplies xi in the
    reference's file layout."""

    def __init__(self, ctx: Context, models, samples, table, expansionOrder: int,
                 meanE=2.1e5, cov=0.5, nu=0.3, precision=L.MSA_FP64, tile=0, seed=20261202,
                 randomsReadFileName: str | None = None, simulationStartFrom: int = 0,
                 preconditioner=L.MSA_PRECOND_JACOBI):
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
        self.batch = None
        self.randomNumbers = None
        self.times = {}

    def Initialize(self):
        total = sum(self.samples)
        M = self.expansionOrder
        if self.randomsReadFileName:
            xi = read_randoms_file(self.randomsReadFileName, M)
            self.randomNumbers = xi[self.simulationStartFrom:self.simulationStartFrom + total]
        else:
            self.randomNumbers = standard_normal(self.seed, total, M)
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

    def Solve(self, tolerance=1e-10, maxIterations=20000, inner_tol=1e-4, max_outer=20,
              graph_iters=0):
        if self.batch is None:
            self.Initialize()
        t0 = time.perf_counter()
        self.batch.run(self.nu, tolerance, maxIterations, inner_tol, max_outer, graph_iters)
        self.ctx.synchronize()
        self.times["all"] = time.perf_counter() - t0
        return self

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
