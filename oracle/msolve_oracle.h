/*
 * msolve_oracle.h -- CPU restatement of the MSolve linear-elastic Hexa8 /
 * assembly / PCG path, used ONLY as test infrastructure (the checker in
 * tests/, the smoke check and the CPU-baseline leg of bench.py).  It is
 * never linked into, loaded by, or called from the msolve_amd product.
 *
 * Every function cites the reference file:line it restates.  Array layouts
 * follow the reference's C#/Fortran conventions unless stated otherwise.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   orc_h8_gauss_matrices / orc_h8_k / orc_ke_mesh  -- sympy-exact brick Ke
 *       fixtures (tests/golden/), mpmath 50-digit distorted-hex Ke with the
 *       13-digit Gauss constant, symmetry, 6 rigid-body zero modes, PSD,
 *       linearity in E, Ke(sX) = s Ke(X).
 *   orc_skyline_*   -- dense numpy Cholesky / scipy spsolve on small meshes,
 *       exact profile heights vs a brute-force rule, K == scipy COO assembly.
 *   orc_pcg_*       -- converges to the direct solution; residual history
 *       recomputed independently; Jacobi vs LDL^T preconditioners agree.
 *   orc_kl_field    -- closed-form lognormal moments, direct summation in
 *       numpy of the restated formula on a hand-built table.
 *   orc_mesh_box / orc_enumerate_dofs / orc_cantilever_loads -- brute-force
 *       rules on tiny meshes (node coordinates, DOF counts, load sums).
 */
#ifndef MSOLVE_ORACLE_H
#define MSOLVE_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Mesh + numbering + loads ------------------------------------------------ */
int orc_mesh_box(int nex, int ney, int nez, double dx, double dy, double dz,
                 double *xyz, int *conn);
int orc_enumerate_dofs(int n_nodes, const unsigned char *constrained, int *dof);
void orc_cantilever_bc(int nex, int ney, int nez, unsigned char *constrained);
int orc_cantilever_loads(int nex, int ney, int nez, const int *dof,
                         double total, double *f);

/* Element (rElHexa8.f90) --------------------------------------------------- */
void orc_h8_gauss_matrices(int iInt, const double *xyz, double *w, double *S,
                           double *DS, double *J, double *detJ, double *B);
void orc_h8_k(int iInt, const double *E, const double *B, const double *w,
              double *K300);
void orc_elastic_D(double E, double nu, double *D36);
int orc_ke_mesh(int n_elem, const int *conn, const double *xyz,
                const double *E_elem, double nu, double *ke300);
void orc_ke_element_csharp(const double *xyz, double E, double nu, double *K300);
void orc_h8_mass_csharp(const double *xyz, double rho, double *M300);

/* Skyline (GlobalMatrixAssemblerSkyline.cs, SkylineMatrix2D.cs) ------------ */
int64_t orc_skyline_rowindex(int n_elem, const int *conn, const int *dof,
                             int n_free, int64_t *rowIndex);
void orc_skyline_assemble(int n_elem, const int *conn, const int *dof,
                          const double *ke300, const int64_t *rowIndex,
                          double *data);
void orc_skyline_multiply(int n, const int64_t *rowIndex, const double *data,
                          const double *x, double *y);
int orc_skyline_factorize(int n, const int64_t *rowIndex, double *data,
                          double tol);
int orc_skyline_factorize_zem(int n, const int64_t *rowIndex, double *data,
                              double tol, int max_modes, int *cols, double *zems);
void orc_skyline_solve(int n, const int64_t *rowIndex, const double *data,
                       const double *f, double *x);

/* CSR (deviation used where the skyline profile is infeasible) ------------- */
int64_t orc_csr_count(int n_elem, const int *conn, const int *dof, int n_free,
                      int64_t *rowptr);
void orc_csr_fill(int n_elem, const int *conn, const int *dof, int n_free,
                  const int64_t *rowptr, int *col);
void orc_csr_assemble(int n_elem, const int *conn, const int *dof, int n_free,
                      const double *ke300, const int64_t *rowptr,
                      const int *col, double *val);
void orc_csr_multiply(int n, const int64_t *rowptr, const int *col,
                      const double *val, const double *x, double *y);

/* PCG (SolverPCG.cs, SolverPCGSimpleSearchVectorCalculator.cs) ------------- */
enum { ORC_PRECOND_NONE = 0, ORC_PRECOND_JACOBI = 1, ORC_PRECOND_SKYLINE = 2 };
int orc_pcg_skyline(int n, const int64_t *rowIndex, const double *K,
                    int precond, const int64_t *P_rowIndex, const double *P,
                    const double *f, double *x, int max_it, double tol,
                    int *iters, double *err_hist);
int orc_pcg_csr(int n, const int64_t *rowptr, const int *col, const double *val,
                int precond, const double *f, double *x, int max_it,
                double tol, int *iters, double *err_hist);

/* Lognormal KL field (LognormalFileStochasticCoefficientsProvider.cs) ------ */
void orc_element_centroid(const double *xyz8, double *c);
int orc_kl_field(int n_elem, const int *conn, const double *xyz, int M,
                 int kmax, const double *w1d, const double *a1d,
                 const int *evpPos, const double *eigs, const double *domain,
                 const double *xi, double meanE, double cov, double *E_out);

/* Reference Monte Carlo loop (MonteCarloAnalyzerWithStochasticMaterial.cs) - */
int orc_mc_reference(int nex, int ney, int nez, double nu, int M, int kmax,
                     const double *w1d, const double *a1d, const int *evpPos,
                     const double *eigs, double meanE, double cov,
                     int n_samples, const double *xi, double tol, int max_it,
                     double *u_out, int *iters, double *times);

int orc_mc_csr_jacobi(int nex, int ney, int nez, double nu, int M, int kmax,
                      const double *w1d, const double *a1d, const int *evpPos,
                      const double *eigs, double meanE, double cov,
                      int n_samples, const double *xi, double tol, int max_it,
                      double *u_out, int *iters, double *times);

/* Element results (rElHexa8.f90 CalcH8Strains/Forces/MLumped) -------------- */
void orc_h8_strains(int iInt, const double *B, const double *u, double *strains);
void orc_h8_forces(int iInt, const double *B, const double *w, const double *stresses,
                   double *forces);
void orc_h8_mlumped(int iInt, double density, const double *w, double *M300);
void orc_elastic_stress(double E, double nu, const double *strain, double *stress);
void orc_element_results(int n_elem, const int *conn, const double *xyz, const int *dof,
                         const double *u, const double *E, double nu, double *strains,
                         double *stresses, double *forces);

/* Newmark dynamics (NewmarkDynamicAnalyzer.cs, ProblemStructural.cs) ------- */
double orc_sinusoidal_acc(int i, double mag, double period, double dt, double tlen);
int orc_newmark_skyline(int n_elem, const int *conn, const double *xyz, const int *dof,
                        int n, const double *E, double nu, double rho, double alphaR,
                        double betaR, double nb_alpha, double nb_delta, double dt, int nsteps,
                        const double *f_static, int acc_dir, double acc_mag,
                        double acc_period, double acc_tlen, int probe, double *probe_hist,
                        double *u_out, double *v_out, double *a_out);

#ifdef __cplusplus
}
#endif
#endif
