"""msolve_amd: MI355X (gfx950) implementation of MSolve's linear-elastic Hexa8
stiffness -> assembly -> PCG path, driven per Monte Carlo sample.

The compute runs in msolve_amd/_lib/libmsolve_amd.so (HIP, C ABI declared in
include/msolve_amd.h); this package is the host-side binding.
"""
from ._lib import (MSA_BREAKDOWN, MSA_FP64, MSA_MIXED, MSA_NAN_INPUT, MSA_NEG_DETJ,
                   MSA_NOT_CONVERGED, MSA_OK, MSA_PRECOND_FIRST_SAMPLE, MSA_PRECOND_JACOBI,
                   MSA_PRECOND_MULTIGRID,
                   MSA_REFINE_RELIABLE, MSA_REFINE_RESTART, MSA_SOLVER_DIRECT, MSA_SOLVER_PCG,
                   MsaError, STATUS_NAMES, lib)
from .analyzers import (ElementStructuralStiffnessProvider,
                        MonteCarloAnalyzerWithStochasticMaterial, NewmarkDynamicAnalyzer,
                        SolverPCG, SolverSkyline, sinusoidal_history)
from .core import (Batch, Context, Mesh, autotune_schedule, autotune_tile, enumerate_dofs,
                   ke_packed)
from .model import box_mesh, cantilever

__all__ = [
    "MSA_OK", "MSA_NOT_CONVERGED", "MSA_BREAKDOWN", "MSA_NEG_DETJ", "MSA_NAN_INPUT",
    "MSA_FP64", "MSA_MIXED", "MSA_REFINE_RELIABLE", "MSA_REFINE_RESTART", "MSA_SOLVER_PCG",
    "MSA_SOLVER_DIRECT", "MSA_PRECOND_JACOBI", "MSA_PRECOND_MULTIGRID",
    "MSA_PRECOND_FIRST_SAMPLE", "MsaError",
    "STATUS_NAMES", "lib",
    "Context", "Mesh", "Batch", "ke_packed", "enumerate_dofs", "autotune_tile",
    "autotune_schedule",
    "box_mesh", "cantilever", "ElementStructuralStiffnessProvider", "SolverPCG", "SolverSkyline",
    "MonteCarloAnalyzerWithStochasticMaterial", "NewmarkDynamicAnalyzer", "sinusoidal_history",
]
