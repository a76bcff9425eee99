/*
 * msolve_amd.h -- C ABI of the MI355X (gfx950) implementation of MSolve's
 * linear-elastic Hexa8 stiffness -> assembly -> PCG path, driven per Monte
 * Carlo sample.  Plain C types only; every entry point returns an int status
 * (MSA_OK = 0) and never throws.  Host arrays are owned by the caller and only
 * read/written during the call; device memory is owned by the msa_* objects.
 *
 * One msa_ctx per host thread (or per stream); objects created from a ctx are
 * used with that ctx only.  Calls on different contexts are independent.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the isaar/MSolve tree).
 */
#ifndef MSOLVE_AMD_H
#define MSOLVE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
    MSA_OK = 0,
    MSA_NOT_CONVERGED = 1, /* PCG hit max iterations (the reference is silent
                              here, SolverPCG.cs:117-138) */
    MSA_BREAKDOWN = 2,     /* p.q <= 0 or non-finite scalar in PCG */
    MSA_NEG_DETJ = 3,      /* Jacobian determinant < 1e-8 (rElHexa8.f90:203) */
    MSA_NAN_INPUT = 4,     /* non-finite modulus / load */
    MSA_ERR_ARG = -1,
    MSA_ERR_HIP = -2,
    MSA_ERR_NO_DEVICE = -3,
    MSA_ERR_OOM = -4,
    MSA_ERR_STATE = -5
};

enum { MSA_FP64 = 0, MSA_MIXED = 1 };
enum { MSA_REFINE_RESTART = 0, MSA_REFINE_RELIABLE = 1 };
enum { MSA_SOLVER_PCG = 0, MSA_SOLVER_DIRECT = 1 };
enum { MSA_PRECOND_JACOBI = 0, MSA_PRECOND_MULTIGRID = 1, MSA_PRECOND_FIRST_SAMPLE = 2 };

typedef struct msa_ctx msa_ctx;
typedef struct msa_mesh msa_mesh;
typedef struct msa_batch msa_batch;

/* Last error message of the calling thread (static storage). */
const char *msa_last_error(void);
int msa_version(void);

/* ------------------------------------------------- Fortran drop-in (ABI)
 * Same names, argument lists and layouts as the femelements.dll imports of
 * ISAAR.MSolve.PreProcessor/Elements/Hexa8.cs:27-55 (cdecl, everything by
 * reference), replacing FEMElements/rElHexa8.f90:422-464 (CalcH8GaussMatrices)
 * and :489-543 (CalcH8K).  Computed on the GPU of a lazily created default
 * context (device 0); one small launch per call.  Only iInt == 2 is
 * supported (the reference's Hexa8 uses 2, Hexa8.cs:17).  On error the
 * outputs are filled with NaN (the Fortran has no error channel).
 *   xyz[8][3]; w[8]; S[8][8]; DS[8][24]; J[8] Fortran (3,3); detJ[8];
 *   B[8][24][6] (= Fortran faB(6,24,8)); E[8][6][6]; K[300] packed upper.  */
void CALCH8GAUSSMATRICES(const int *iInt, const double *xyz, double *w,
                         double *S, double *DS, double *J, double *detJ,
                         double *B);
void CALCH8K(const int *iInt, const double *E, const double *B,
             const double *w, double *K);
/* The remaining femelements.dll imports of Hexa8.cs:35-61: CalcH8Strains
 * (rElHexa8.f90:635-648) strains[8][6] = B_M u for u[24]; CalcH8Forces
 * (:724-747) forces[24] = sum_M w[M] B_M^T stresses[M] for stresses[8][6];
 * CalcH8MLumped (:673-699) M[300] packed, sum(w) density / 8 on the 24
 * diagonal entries. */
void CALCH8STRAINS(const int *iInt, const double *B, const double *u,
                   double *strains);
void CALCH8FORCES(const int *iInt, const double *B, const double *w,
                  const double *stresses, double *forces);
void CALCH8MLUMPED(const int *iInt, const double *density, const double *w,
                   double *M);

/* --------------------------------------------------------------- context */
int msa_ctx_create(int device, msa_ctx **out);
/* Destroy every mesh and batch of the context first. */
int msa_ctx_destroy(msa_ctx *ctx);
/* The HIP stream every kernel of this context runs on (hipStream_t). */
void *msa_ctx_stream(msa_ctx *ctx);
int msa_ctx_synchronize(msa_ctx *ctx);

/* ---------------------------------------------------- model / numbering
 * Free-DOF numbering of Model.EnumerateGlobalDOFs / Subdomain.EnumerateDOFs
 * (ISAAR.MSolve.PreProcessor/Model.cs:165-215, Subdomain.cs:77-168): nodes in
 * ascending index, DOFs X, Y, Z, constrained -> -1 (eliminated).
 * constrained[3*node + d] != 0 marks a constraint. */
int msa_enumerate_dofs(int n_nodes, const unsigned char *constrained, int *dof,
                       int *n_free);

/* A Hexa8 mesh: node coordinates xyz[n_nodes][3], 0-based connectivity
 * conn[n_elem][8] in the Hexa8 node order (rElHexa8.f90:25-117), the free-DOF
 * map dof[n_nodes][3] from msa_enumerate_dofs (or the reference's own
 * NodalDOFsDictionary).  Builds the node-block sparsity of the global K, the
 * element -> block-slot map and an element colouring on the host, and
 * uploads mesh data to the device.  Replaces the per-sample skyline profile
 * + scatter bookkeeping of GlobalMatrixAssemblerSkyline.cs:14-103. */
int msa_mesh_create(msa_ctx *ctx, int n_nodes, const double *xyz, int n_elem,
                    const int *conn, const int *dof, int n_free,
                    msa_mesh **out);
/* info[0] free node blocks, [1] block nonzeros, [2] scalar nonzeros of the
 * free-DOF K (full symmetric), [3] colours, [4] max blocks per row,
 * [5] n_free, [6] n_elem, [7] n_nodes. */
int msa_mesh_info(const msa_mesh *mesh, long long *info);
int msa_mesh_destroy(msa_mesh *mesh);

/* ------------------------------------------------- element stiffness
 * Batched Hexa8WithStochasticMaterial.StiffnessMatrix / Hexa8.StiffnessMatrix
 * behind IElementMatrixProvider.Matrix (ISAAR.MSolve.PreProcessor/Elements/
 * Hexa8WithStochasticMaterial.cs:61-110, Hexa8.cs:368-401,
 * Providers/ElementStructuralStiffnessProvider.cs:11-38): 2x2x2 Gauss
 * B^T D B with isotropic D(E[e], nu).  xyz8[n_elem][8][3] element node
 * coordinates; ke300[n_elem][300] packed upper triangle row-major
 * (SymmetricMatrix2D.cs:22-49).  elem_status[n_elem] (may be NULL) gets
 * MSA_OK or MSA_NEG_DETJ.  FP64 on the GPU. */
int msa_ke_packed(msa_ctx *ctx, int n_elem, const double *xyz8,
                  const double *E, double nu, double *ke300, int *elem_status);
/* No counterpart in the reference.  msa_ke_packed without any device
 * allocation, for callers that cannot allocate (code under HIP graph
 * capture elsewhere in the process, tight serving loops, harnesses that
 * forbid hipMalloc): same arguments, same kernel and bit-identical output.
 * The device staging lives in one fixed array per GPU (about 10.7 MB, held
 * until the process ends), so n_elem is limited to 4096 per call; larger
 * n_elem returns MSA_ERR_ARG.  Calls that overlap on one GPU (several
 * threads or contexts on the same device) share that array and race: the
 * caller serialises them.  The context must come from msa_ctx_create, which
 * looks up the array's device address. */
int msa_ke_packed_noalloc(msa_ctx *ctx, int n_elem, const double *xyz8,
                          const double *E, double nu, double *ke300,
                          int *elem_status);

/* ----------------------------------------------------------- MC batch
 * A batch solves many stochastic-modulus samples at once: n_classes meshes
 * (uneven sizes allowed), n_samples[c] samples of mesh c.  Samples of a class
 * share the sparsity and the load vector; each has its own modulus field,
 * stiffness values and solution.  Replaces the per-sample loop of
 * MonteCarloAnalyzerWithStochasticMaterial.SolveNormal
 * (ISAAR.MSolve.Analyzers/MonteCarloAnalyzerWithStochasticMaterial.cs:
 * 232-296) with its StaticAnalyzer -> LinearAnalyzer -> SolverPCG chain.
 *   precision   MSA_FP64: FP64 element integration, assembly and PCG.
 *               MSA_MIXED: FP32 integration/assembly/inner PCG, FP64 outer
 *               residual f - K u (matrix-free), update and stopping test.
 *   (mixed refinement scheme: msa_batch_set_refinement)
 *   tile        samples per lane tile: 1, 2, 4, 8, 16, 32 or 64 (0 = 1, the
 *               fastest for whole solves on MI355X: samples converge at
 *               different iterations and a wider tile keeps iterating
 *               until its slowest sample is done). */
int msa_batch_create(msa_ctx *ctx, int n_classes, msa_mesh *const *meshes,
                     const int *n_samples, int precision, int tile,
                     msa_batch **out);
int msa_batch_destroy(msa_batch *b);
/* Device bytes held by the batch. */
long long msa_batch_bytes(const msa_batch *b);

/* Modulus per element per sample, E[n_samples][n_elem] (host). */
int msa_batch_set_modulus(msa_batch *b, int cls, const double *E);

/* Lognormal Karhunen-Loeve modulus field of
 * LognormalFileStochasticCoefficientsProvider.GetCoefficient
 * (ISAAR.MSolve.PreProcessor/Stochastic/
 * LognormalFileStochasticCoefficientsProvider.cs:134-193), evaluated at the
 * element centroid rule of Hexa8WithStochasticMaterial.cs:112-139.  Table in
 * the reference's KL-file fields (:90-109): w1d[3][kmax], a1d[3][kmax] per
 * 1-based 1-D mode, evpPos[M][3] (1-based), eigs[M]; domain[3] the
 * stochastic domain lengths (coordinates are centred by domain/2).
 * cov replaces the reference's hard-coded 0.5 (:115-117).  The per-element
 * basis is built once per class on the device. */
int msa_batch_set_kl(msa_batch *b, int cls, int M, int kmax,
                     const double *w1d, const double *a1d, const int *evpPos,
                     const double *eigs, const double *domain, double meanE,
                     double cov);
/* Standard-normal KL coefficients xi[n_samples][M] of class cls (the
 * RandomVariables of MC...Material.cs:261). */
int msa_batch_set_xi(msa_batch *b, int cls, const double *xi);
/* E = exp(ln meanE - sigma^2/2 + sum_i xi_i phi_i(centroid)) on the device,
 * all classes with a KL basis. */
int msa_batch_eval_kl(msa_batch *b);

/* Load vector f[n_free] of class cls, shared by its samples
 * (Subdomain.Forces, Model.AssignNodalLoads, Model.cs:227-238). */
int msa_batch_set_rhs(msa_batch *b, int cls, const double *f);

/* Element stiffness + coloured scatter into the block image of K for every
 * sample, plus the Jacobi diagonal (ProblemStructural.BuildKs ->
 * GlobalMatrixAssemblerSkyline.CalculateGlobalMatrix,
 * ISAAR.MSolve.Problems/ProblemStructural.cs:63-89, GMAS.cs:51-103). */
int msa_batch_assemble(msa_batch *b, double nu);

/* Solve K u = f for every sample (SolverPCG.Solve, ISAAR.MSolve.Solvers/PCG/
 * SolverPCG.cs:105-139, with the simple search vector calculator
 * SolverPCGSimpleSearchVectorCalculator.cs:18-41; x0 = 0, stop when
 * |r|/|f| < tol).  Preconditioner: Jacobi.  FP64: tol on the PCG residual,
 * max_it PCG iterations.  MIXED: tol on the FP64 residual |f - K u|/|f|,
 * inner FP32 PCG to inner_tol (relative) with at most max_it iterations,
 * at most max_outer refinements.  graph_iters > 0 captures that many PCG
 * iterations per hipGraph launch (even number), 0 launches eagerly.  Once
 * a sample of the batch has converged, the remaining iterations go in
 * blocks of 2 (MIXED with residual replacement: one replacement period),
 * polled after each, so the solve ends at most one short block after its
 * last sample. */
int msa_batch_solve(msa_batch *b, double tol, int max_it, double inner_tol,
                    int max_outer, int graph_iters);

/* Refinement scheme of MIXED solves (no counterpart in the reference, which
 * is FP64 only).  MSA_REFINE_RELIABLE (default): the FP32 PCG never restarts;
 * a sample whose FP32 residual has fallen by `drop` (default 0.1; 1e-4 with
 * the multigrid or first-sample preconditioner) since its last replacement
 * -- or whose estimated FP64 residual passes tol -- gets, at the next check
 * (every `period` iterations, even, default 8; 2 with those), its FP32
 * correction folded into the FP64 solution, its residual replaced by the
 * FP64 f - K u, the stopping test |f - K u|/|f| < tol, and the FP32
 * recurrence continued from that residual (search direction rescaled);
 * inner_tol and max_outer are unused.  MSA_REFINE_RESTART: classic
 * refinement, each outer step solves K d = r/|r| in FP32 to inner_tol from
 * zero (period and drop unused). */
int msa_batch_set_refinement(msa_batch *b, int method, int period, double drop);

/* Captured PCG iteration blocks (no counterpart in the reference): info[0]
 * captures so far (a solve's capture records the caller's block and the
 * short tail block together; a Newmark run adds its three step graphs),
 * [1] kernel nodes of the current block graph (0: none), [2] PCG
 * iterations per replay of it, [3] graph launches so far.  Every captured
 * graph is checked on the host before it is instantiated: kernel nodes only,
 * non-empty grids, workgroups within each kernel's launch bounds, LDS within
 * a CU; a failed check fails the solve with MSA_ERR_STATE.  Changing the
 * tolerance, the iteration limit, the refinement scheme or the
 * preconditioner re-captures at the next solve. */
int msa_batch_graph_info(msa_batch *b, long long *info);

/* Stored layout of class cls (no counterpart in the reference; the
 * reference's SkylineMatrix2D keeps the upper profile only): info[0] 1 if
 * the class is a box grid whose block rows hold their 27 neighbours in
 * stencil slots (offset (dx, dy, dz) in slot (dz+1)*9 + (dy+1)*3 + dx+1,
 * missing neighbours as zero blocks), [1] 1 if the PCG product reads only
 * the upper triangle of that image (one-sample tiles: the lower blocks are
 * the transposes of the upper blocks of the rows above, as in the skyline
 * product), [2] stored block slots per sample, [3] block rows. */
int msa_batch_layout_info(msa_batch *b, int cls, long long *info);

/* The upper-triangle storage and product of box-grid classes (enable = 1,
 * the default) or the full image: both triangles stored and streamed, the
 * products as on other meshes (enable = 0), for instance for a loaded
 * matrix that is not symmetric.  No counterpart in the reference (its
 * SkylineMatrix2D always keeps the upper profile).  A change of form
 * needs the values again: the batch counts as not assembled (assemble,
 * run, or load the matrices again). */
int msa_batch_set_symmetric_product(msa_batch *b, int enable);

/* The PCG product q = K p of box-grid classes computed matrix-free (enable
 * = 1, the default): MIXED batches of one-sample tiles with the multigrid
 * preconditioner, on the classes it coarsens (a class without coarser
 * levels keeps the stored product), whose finest smoothing products compute K v from
 * the cell moduli and the unit element matrix in FP32.  The product then
 * reads z, p_old and the moduli instead of the stored FP32 values; the
 * FP64 residual replacements (also matrix-free) fix the accuracy, and the
 * stored values are still assembled (export, coarse levels, warm starts).
 * enable = 0: the stored-value product.  Takes effect at the next solve.
 * Replaces, for those classes, the role of SkylineMatrix2D.Multiply
 * (SkylineMatrix2D.cs:253-269) in SolverPCG.cs:120. */
int msa_batch_set_matrix_free_product(msa_batch *b, int enable);

/* The PCG product form of class cls at the last solve: 0 the stored
 * values (full image), 1 the stored upper triangle, 2 matrix-free. */
int msa_batch_product_form(msa_batch *b, int cls, int *form);

/* Element integration precision of a MIXED batch, whose stiffness values are
 * stored in FP32: MSA_FP64 (default) integrates in FP64 and rounds each
 * assembled value; MSA_MIXED integrates in FP32.  FP32 integration leaves
 * ~4e-7 relative error in K (vs 6e-8 for rounding), which costs about a
 * quarter more PCG iterations under the FP64 residual replacement. */
int msa_batch_set_integration(msa_batch *b, int precision);

/* Modulus field -> assembly -> solve in one call (the whole MC job with all
 * inputs resident on the device). */
int msa_batch_run(msa_batch *b, double nu, double tol, int max_it,
                  double inner_tol, int max_outer, int graph_iters);

/* Solution of one sample in the reference free-DOF numbering, u[n_free]. */
int msa_batch_get_solution(msa_batch *b, int cls, int sample, double *u);
/* Per-sample results of class cls: iters[n] PCG iterations (matrix-vector
 * products; mixed: summed over inner solves), outer[n] refinements (FP64:
 * 0), relres[n] final relative residual (mixed: the FP64 |f-Ku|/|f|),
 * status[n].  Any pointer may be NULL. */
int msa_batch_get_stats(msa_batch *b, int cls, int *iters, int *outer,
                        double *relres, int *status);
/* Element results of one solved sample (Hexa8.CalculateStresses and
 * CalculateForces, ISAAR.MSolve.PreProcessor/Elements/Hexa8.cs:453-510, with
 * ElasticMaterial3D stresses, Materials/ElasticMaterial3D.cs:20-43): per
 * element and Gauss point (xi outer, zeta inner) strains[n_elem][8][6] =
 * B u_e (Voigt xx, yy, zz, xy, yz, zx, engineering shear), stresses
 * [n_elem][8][6] = D(E_e, nu) strain, internal forces forces[n_elem][24] =
 * sum w B^T sigma (= Ke u_e).  Any pointer may be NULL.  FP64 from the FP64
 * solution of the last solve. */
int msa_batch_element_results(msa_batch *b, int cls, int sample, double *strains,
                              double *stresses, double *forces);
/* The modulus field of one sample, E[n_elem]. */
int msa_batch_get_modulus(msa_batch *b, int cls, int sample, double *E);

/* Assembled K of one sample in the reference free-DOF numbering as full
 * symmetric CSR: rowptr[n_free+1], col[nnz], val[nnz] (nnz from
 * msa_mesh_info info[2]); MIXED batches return the FP32 values widened. */
int msa_batch_get_matrix_csr(msa_batch *b, int cls, int sample,
                             long long *rowptr, int *col, double *val);
/* Sets the stiffness of one sample from a full symmetric CSR in the
 * reference free-DOF numbering, in place of assembling it: the GPU side of
 * MonteCarloAnalyzerWithStochasticMaterial.ReadMatricesFromFile
 * (MonteCarloAnalyzerWithStochasticMaterial.cs:381-402, the
 * LoadFromDiskAndCalculate mode), whose SkylineMatrix2D.ReadFromFile image
 * a host converts to CSR.  MSA_FP64 batches without dynamics only (the
 * mixed path's FP64 residual and the multigrid levels are formed from the
 * moduli: a batch of loaded matrices solves with the Jacobi or the
 * first-sample preconditioner, or the direct solver).  Every nonzero entry
 * must lie in the mesh's sparsity pattern (MSA_ERR_ARG otherwise; zeros
 * outside it, such as a skyline profile's, are skipped); duplicates add.
 * The sample's other stored values become zero; the first call on a batch
 * never assembled zeroes every sample, so every sample must be set before
 * a solve.  D^-1 is refreshed; the preconditioner follows at the solve.
 * On box-grid classes with one-sample tiles (msa_batch_layout_info) the
 * products use the upper triangle of the loaded matrix, as the skyline
 * image it came from holds only the upper profile. */
int msa_batch_set_matrix_csr(msa_batch *b, int cls, int sample,
                             const long long *rowptr, const int *col,
                             const double *val);
/* Forms the preconditioner's values from the current stiffness now (the
 * first-sample band factor, or the multigrid levels and their coarsest
 * factors) instead of at the next solve: the counterpart of the
 * factorization in SolverPCG's Initialize
 * (MonteCarloSolverPCGDirectMatrixCalculator.cs:38-80), so a caller can time
 * it apart.  No-op when they are current or with the Jacobi preconditioner. */
int msa_batch_prepare_preconditioner(msa_batch *b);
/* y = K x for one sample with the batch's SpMV kernel (x, y in the
 * reference numbering, FP64 on the host). */
int msa_batch_spmv(msa_batch *b, int cls, int sample, const double *x,
                   double *y);

/* Kernel timing with HIP events on the context stream (eager launches while
 * enabled).  kind: 0 SpMV+dot kernel, 1 update kernel, 2 assembly,
 * 3 FP64 residual, 4 other, 5 multigrid smoothing products of the finest
 * level.  ms_total[6], launches[6]. */
int msa_batch_set_kernel_timing(msa_batch *b, int enable);
int msa_batch_get_kernel_timing(msa_batch *b, double *ms_total,
                                long long *launches);
/* Algorithmic bytes of one launch of the SpMV+dot kernel over every sample
 * of the batch: bytes_per_launch[0] the kernel's (values + shared indices +
 * vectors; on box grids with one-sample tiles the upper triangle of the
 * values, diagonal included, and no indices; matrix-free classes: z and
 * p_old read and p written by the p pass, p and q by the product, and the
 * FP64 cell moduli), bytes_per_launch[1] a plain SpMV's
 * (all values + indices + one vector read and one written). */
int msa_batch_spmv_bytes(msa_batch *b, double *bytes_per_launch);
/* Measurement (no counterpart in the reference): times `reps` back-to-back
 * launches of the batch's SpMV + dot kernel with every sample active (the
 * state at the start of a PCG solve, which it sets up: the PCG iterates, and
 * so an FP64 batch's solution, are overwritten) between one pair of events
 * on the context stream; *ms_per_launch = elapsed / reps.  Free of the
 * per-launch event overhead of msa_batch_set_kernel_timing, which matters for
 * launches of about 100 microseconds. */
int msa_batch_time_spmv(msa_batch *b, int reps, double *ms_per_launch);

/* ----------------------------------------------------------- direct solver
 * Banded Cholesky K = L L^T of every sample's assembled FP64 stiffness on the
 * GPU: the SkylineMatrix2D.Factorize / Solve pair behind SolverSkyline
 * (ISAAR.MSolve.Matrices/SkylineMatrix2D.cs:63-212,
 * ISAAR.MSolve.Solvers/Skyline/SolverSkyline.cs:90-205) restated as a
 * blocked band Cholesky (64-column blocks: diagonal factor, panel solve,
 * trailing update of the band window) in the free-DOF order with the
 * smallest half bandwidth among the block order, reverse Cuthill-McKee and
 * the lexicographic coordinate orders.  FP64 batches only.  Storage per
 * sample: n * W doubles (msa_batch_direct_info).  A non-positive pivot gives
 * the sample MSA_BREAKDOWN. */
int msa_batch_factorize(msa_batch *b);
/* Zero-energy modes of the factorization (SkylineMatrix2D.Factorize with
 * its zems / zemCols lists, SkylineMatrix2D.cs:63-158): with tol > 0, a
 * pivot of magnitude under tol is a zero-energy mode: the pivot becomes 1,
 * its row and column of the factor are cleared, and the mode (the cleared
 * row, back-substituted through the factor) is kept, up to max_modes (1..64)
 * per sample; a sample with more gets MSA_BREAKDOWN.  The solve then holds
 * the mode's DOF at its load value as the reference's does.  tol = 0
 * (default) turns it off: a non-positive pivot is MSA_BREAKDOWN.  Takes
 * effect at the next factorization. */
int msa_batch_set_zem(msa_batch *b, double tol, int max_modes);
/* The zero-energy modes of one sample's last factorization: *n_modes found,
 * cols[min(n, max_modes)] their DOFs (the reference's zemCols) and
 * modes[min(n, max_modes)][n_free] the vectors (its zems), both in the
 * reference free-DOF numbering.  The factorization runs in the band order:
 * the modes span the same null space as the reference's, and equal its
 * vectors where the band order is the DOF order (the natural order wins
 * ties, as on box grids). */
int msa_batch_get_zem(msa_batch *b, int cls, int sample, int max_modes, int *n_modes,
                      int *cols, double *modes);
/* u = K^-1 f for every sample from the factors (factorizing first when the
 * assembled values changed); msa_batch_get_stats then reports 0 iterations,
 * |f - K u|/|f| and the status. */
int msa_batch_direct_solve(msa_batch *b);
/* MSA_SOLVER_PCG (default) or MSA_SOLVER_DIRECT for msa_batch_solve,
 * msa_batch_run and the steps of msa_batch_newmark (the effective stiffness
 * is factorized once and every step is two band sweeps). */
int msa_batch_set_solver(msa_batch *b, int solver);
/* info[0] band dimension n (free DOFs), [1] half bandwidth, [2] storage
 * width W, [3] factor bytes per sample, for class cls. */
int msa_batch_direct_info(msa_batch *b, int cls, long long *info);

/* ----------------------------------------------------------- preconditioner
 * Preconditioner of the PCG solves (msa_batch_solve, msa_batch_newmark).
 * Replaces the choice made by SolverPCG's search-vector calculator
 * (SolverPCG.cs:105-139 with SolverPCGSimpleSearchVectorCalculator.cs:18-41,
 * Jacobi) and the first-sample LDL^T preconditioner of the Monte Carlo
 * calculator (MonteCarloSolverPCGDirectMatrixCalculator.cs:38-80, applied in
 * MCPCGDirect.cs:48-51): both reduce PCG iterations with a matrix reused
 * across solves; here every sample gets its own multigrid hierarchy.
 *
 * MSA_PRECOND_JACOBI (default): z = D^-1 r.
 * MSA_PRECOND_MULTIGRID: z = one geometric multigrid V-cycle on r, for every
 * class whose mesh is a structured tensor-product grid of hexahedra with
 * element counts that halve (each level keeps >= 2 elements per direction,
 * at most max_levels coarse levels, <= 0: as many as possible) and whose
 * coarsest level has at most 12000 free DOFs; other classes keep Jacobi.
 * Coarse levels are rediscretized on the coarsened grid (nodes inherit the
 * constraints, element modulus = mean of its 8 children) and assembled by
 * the batch's kernels.  Per level: 2 Chebyshev-Jacobi steps on
 * [lam/ratio, lam] (lam = Gershgorin bound of D^-1 K per sample; ratio 0
 * selects 12), restriction P^T, coarse correction, prolongation P
 * (trilinear), 2 Chebyshev-Jacobi steps; the coarsest level is solved by
 * band Cholesky per sample.  The cycle is symmetric positive definite, so
 * PCG keeps its convergence guarantee; the solution is the same, within the
 * tolerance, as with Jacobi.  The hierarchy is built by the first call with
 * MSA_PRECOND_MULTIGRID (later calls must pass the same max_levels) and
 * re-assembled by every msa_batch_assemble.
 * MSA_PRECOND_FIRST_SAMPLE: the reference's Monte Carlo preconditioner
 * itself (MonteCarloSolverPCGDirectMatrixCalculator.cs:38-80): the band
 * Cholesky factor (= LDL^T) of the first sample's assembled K of each class,
 * computed at assembly and shared by every sample of the class; z = L^-T
 * L^-1 r by forward and backward band sweeps.  One factor of n x (band + 64)
 * doubles per class (msa_batch_direct_info); max_levels and ratio unused.
 * Unless msa_batch_set_refinement was called, MIXED batches check for
 * residual replacements every 2 iterations with a 1e-3 trigger under
 * multigrid and the first-sample factor (8 and 0.1 under Jacobi). */
int msa_batch_set_preconditioner(msa_batch *b, int kind, int max_levels, double ratio);
/* Implementation of the matrix-free smoothing products on box grids of
 * fully free blocks (no counterpart in the reference): enable != 0 (the
 * default) computes each cell's element product on the matrix cores
 * (v_mfma_f32_16x16x4f32 over 16 cells at a time, for grids of at most 64
 * cells in x), 0 keeps the FP32 vector-ALU form.  Both are FP32 and agree to
 * FP32 rounding; other classes are unaffected. */
int msa_batch_set_matrix_cores(msa_batch *b, int enable);
/* info[0] coarse levels of class cls (0: Jacobi), [1] free DOFs of the
 * coarsest level, [2] free DOFs summed over the coarse levels, [3] half
 * bandwidth of the coarsest level, [4] 1 if the finest level's smoothing
 * products are matrix-free, [5] the number of coarse levels that are.
 * Matrix-free products are used by MIXED batches with one-sample tiles on
 * grids of equal, equally oriented cells (all elements share one element
 * matrix): K v from the cell moduli and that matrix in FP32, without
 * reading the stored values. */
int msa_batch_mg_info(msa_batch *b, int cls, long long *info);
/* Upper end lam of each level's Chebyshev smoothing interval [lam/ratio,
 * lam] (no counterpart in the reference, whose PCG preconditioner is Jacobi
 * or the first sample's factor): with lanczos_steps > 0 (default 4,
 * margin 1.3) lam = min(Gershgorin bound, margin x the top Ritz value of
 * lanczos_steps Lanczos steps on K x = lambda D x), per sample and level,
 * formed with the preconditioner's values; 0 keeps the Gershgorin bound.
 * The Ritz value converges to the top of the spectrum from below, so the
 * margin keeps lam above it; the Gershgorin bound caps it. */
int msa_batch_set_smoothing_interval(msa_batch *b, int lanczos_steps, double margin);
int msa_batch_get_smoothing_interval(msa_batch *b, int *lanczos_steps, double *margin);
/* lam[max_levels] of one sample: level 0 (the class) first, 0 past the
 * coarsest smoothed level. */
int msa_batch_mg_lambda(msa_batch *b, int cls, int sample, int max_levels, double *lam);
/* z = B r (one V-cycle) for one sample of class cls, r and z in the
 * reference numbering.  Uses the batch's PCG work vectors: call it between
 * solves. */
int msa_batch_mg_apply(msa_batch *b, int cls, int sample, const double *r, double *z);

/* ----------------------------------------------------------- dynamics
 * Linear implicit Newmark time integration of every sample of the batch
 * (NewmarkDynamicAnalyzer, ISAAR.MSolve.Analyzers/NewmarkDynamicAnalyzer.cs
 * with ProblemStructural, ISAAR.MSolve.Problems/ProblemStructural.cs).
 *
 * msa_batch_set_dynamics switches the batch to the effective stiffness
 * K + a0 M + a1 C (ProblemStructural.CalculateEffectiveMatrix, :138-157):
 * consistent mass of density rho (Hexa8.CalculateConsistentMass,
 * ISAAR.MSolve.PreProcessor/Elements/Hexa8.cs:403-436), Rayleigh damping
 * C = rayleigh_alpha M + rayleigh_beta K (Hexa8.DampingMatrix, :444-451),
 * Newmark parameters alpha (>= (0.5 + delta)^2 / 4) and delta (>= 0.5) and
 * time step dt (InitializeCoefficients, NewmarkDynamicAnalyzer.cs:53-77;
 * invalid parameters give MSA_ERR_ARG as the reference throws).  Call it
 * before msa_batch_assemble.  Loads: the class load vector (msa_batch_set_rhs,
 * optional, constant in time) plus the mass-acceleration history load. */
int msa_batch_set_dynamics(msa_batch *b, double rho, double rayleigh_alpha,
                           double rayleigh_beta, double newmark_alpha,
                           double newmark_delta, double dt);
/* Reference free DOF of class cls whose displacement is recorded after every
 * step (msa_batch_get_history); -1 records nothing. */
int msa_batch_set_probe(msa_batch *b, int cls, int dof);
/* Runs nsteps steps from rest (NewmarkDynamicAnalyzer.Solve, :280-299): per
 * step the forces F = f + acc[step] * M_full 1_dir on the free rows
 * (Model.AssignMassAccelerationHistoryLoads, PreProcessor/Model.cs:266-280;
 * acc_dir 0/1/2 = X/Y/Z, -1: no history load), the right-hand side F +
 * M (a0 u + a2 v + a3 a) + C (a1 u + a4 v + a5 a) (CalculateRHSImplicit,
 * :189-206), a PCG solve to |F_eff - K_eff u|/|F_eff| < tol started from the
 * previous displacement (the reference's PCG reuses a stale residual here,
 * so its SolverSkyline is the oracle), and the velocity/acceleration update
 * (:301-315).  acc[nsteps] is on the host (MassAccelerationSinusoidalLoad
 * values, PreProcessor/MassAccelerationSinusoidalLoad.cs:17-20).
 * warm_start selects the first iterate of every step's PCG: 0 zero, 1 the
 * previous displacement (recommended: ~8-17% fewer iterations than zero on
 * the 20 Hz base-acceleration workload), 2 the predictor u + dt v + dt^2/2 a
 * (worse there: the undamped average-acceleration scheme leaves step-to-step
 * oscillation of the stiff modes in v and a). */
int msa_batch_newmark(msa_batch *b, int nsteps, const double *acc, int acc_dir,
                      double tol, int max_it, int graph_iters, int warm_start);
/* Probe displacement after each step of the last msa_batch_newmark call,
 * hist[nsteps]. */
int msa_batch_get_history(msa_batch *b, int cls, int sample, double *hist);
/* Displacement, velocity and acceleration of one sample after the last step
 * (reference numbering, [n_free] each; any pointer may be NULL) and the PCG
 * iterations summed over the steps. */
int msa_batch_get_dynamic_state(msa_batch *b, int cls, int sample, double *u,
                                double *v, double *a, long long *iters);

/* ----------------------------------------------------------- autotuner
 * Times the SpMV+dot kernel of mesh `mesh` for every lane tile
 * (1..64 samples per wave) with n_samples synthetic samples and writes the
 * fastest tile to *best_tile; ms_per_tile[7] receives the timings (tiles
 * 1, 2, 4, 8, 16, 32, 64; negative = skipped). */
int msa_autotune_tile(msa_ctx *ctx, msa_mesh *mesh, int n_samples,
                      int precision, int *best_tile, double *ms_per_tile);

/* Persistent launch (the batched-PCG launch of the solve path; no
 * counterpart in the reference, whose SolverPCG.Solve loops on the host,
 * SolverPCG.cs:105-149): with enable, a Jacobi-preconditioned PCG solve
 * (MSA_FP64 batches, and the inner solves of MSA_REFINE_RESTART) runs as
 * ONE cooperative launch of co-resident workgroups that loops over the
 * iterations itself: the four phases of an iteration (search direction +
 * product, alpha, update, test + beta) are separated by grid-wide barriers,
 * every workgroup leaves when no sample is active.  The per-item work and
 * the summation order are those of the graph-captured iteration, so the
 * results are the same bits.  One-sample tiles only (MSA_ERR_STATE
 * otherwise); other preconditioners and the kernel-timing mode keep the
 * captured iteration blocks.  info[0] persistent launches so far, info[1]
 * workgroups of the last one. */
int msa_batch_set_persistent(msa_batch *b, int enable);
int msa_batch_persistent_info(msa_batch *b, long long *info);

/* Whole-solve schedule autotuner (no counterpart in the reference; the batch
 * schedule of the product replacing SkylineMatrix2D.Multiply inside
 * SolverPCG.Solve, SolverPCG.cs:105-149).  For each candidate (lane tile T,
 * slot unroll U) of msa_autotune_candidates it builds a trial batch of the
 * given meshes with samples_per_class samples each (fixed lognormal moduli,
 * COV 0.3; unit load on every free DOF), the given precision and
 * preconditioner, and times one assembly + PCG solve (tolerance 1e-9 MIXED,
 * 1e-10 FP64) after a warm-up; ms_per_sample[MSA_TUNE_CANDIDATES] (may be
 * NULL) receives the time per sample (negative: the candidate could not be
 * built or did not converge).  The fastest is written to *best_tile and
 * *best_unroll: create the batch with that tile and call
 * msa_batch_set_schedule with that unroll. */
#define MSA_TUNE_CANDIDATES 8
int msa_autotune_schedule(msa_ctx *ctx, int n_classes, msa_mesh *const *meshes,
                          int samples_per_class, int precision, int precond,
                          int *best_tile, int *best_unroll, double *ms_per_sample);
/* The candidates, tiles[MSA_TUNE_CANDIDATES] and unrolls[...]. */
int msa_autotune_candidates(int *tiles, int *unrolls);
/* Slots per group of the SpMV's rows (1..4; lane tiles T > 1 only 3, the
 * default): the loads of a group's slots issue together.  Results do not
 * depend on it (the slots are summed in the same order). */
int msa_batch_set_schedule(msa_batch *b, int unroll);

#ifdef __cplusplus
}
#endif
#endif
