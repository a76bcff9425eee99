/*
 * msolve_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, serial restatement of the reference MSolve algorithm for the
 * linear-elastic Hexa8 -> skyline assembly -> PCG Monte Carlo path.  It is the
 * checker for the HIP product (msolve_amd) and the CPU baseline of bench.py.
 * Nothing in msolve_amd links, loads or calls this file.
 *
 * Citations are to the reference tree (isaar/MSolve):
 *   rElHexa8.f90                = FEMElements/FEMElements/rElHexa8.f90
 *   Hexa8.cs, H8WSM.cs          = ISAAR.MSolve.PreProcessor/Elements/...
 *   GMAS.cs                     = ISAAR.MSolve.PreProcessor/GlobalMatrixAssemblerSkyline.cs
 *   SkylineMatrix2D.cs          = ISAAR.MSolve.Matrices/SkylineMatrix2D.cs
 *   SolverPCG.cs                = ISAAR.MSolve.Solvers/PCG/SolverPCG.cs
 *   Lognormal...cs              = ISAAR.MSolve.PreProcessor/Stochastic/LognormalFileStochasticCoefficientsProvider.cs
 *   MC...Material.cs            = ISAAR.MSolve.Analyzers/MonteCarloAnalyzerWithStochasticMaterial.cs
 *   MCPCGDirect.cs              = ISAAR.MSolve.Analyzers/MonteCarloSolverPCGDirectMatrixCalculator.cs
 *
 * Deviations from the reference, all deliberate:
 *   - skyline rowIndex is int64 (the reference's int32 overflows at 64^3);
 *   - dot products are summed serially (the reference sums AffinityCount
 *     chunks in parallel, Vector.cs:107-131, so its last bits depend on the
 *     core count);
 *   - orc_pcg_* report the number of matrix-vector products performed
 *     (the reference stores currentIteration, one less on convergence,
 *     SolverPCG.cs:117-138).
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math).
 */
#define _POSIX_C_SOURCE 199309L
#include "msolve_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* Mesh, DOF numbering, loads                                                */
/* ------------------------------------------------------------------------ */

/* Structured brick mesh.  Node order x-fastest, then y, then z, coordinates
 * accumulated by repeated addition exactly as Hexa8MeshGenerator.BuildNodes
 * (Hexa8MeshGenerator.cs:530-546).  Elements z-outer, y, x-inner with the
 * node order of Hexa8MeshGenerator.cs:1080-1089 (bottom face counter-
 * clockwise, then top face).  conn is 0-based. */
int orc_mesh_box(int nex, int ney, int nez, double dx, double dy, double dz,
                 double *xyz, int *conn)
{
    const int nX = nex + 1, nY = ney + 1, nZ = nez + 1;
    double p[3] = {0.0, 0.0, 0.0};
    int node = 0;
    for (int i = 0; i < nZ; i++) {
        for (int j = 0; j < nY; j++) {
            for (int k = 0; k < nX; k++) {
                xyz[3 * node + 0] = p[0];
                xyz[3 * node + 1] = p[1];
                xyz[3 * node + 2] = p[2];
                node++;
                p[0] += dx;
            }
            p[0] = 0;
            p[1] += dy;
        }
        p[1] = 0;
        p[2] += dz;
    }
    int e = 0;
    for (int i = 0; i < nez; i++)
        for (int j = 0; j < ney; j++)
            for (int k = 0; k < nex; k++) {
                int *c = conn + 8 * e++;
                c[0] = nX * nY * i + nX * j + k;
                c[1] = nX * nY * i + nX * j + k + 1;
                c[2] = nX * nY * i + nX * (j + 1) + k + 1;
                c[3] = nX * nY * i + nX * (j + 1) + k;
                c[4] = nX * nY * (i + 1) + nX * j + k;
                c[5] = nX * nY * (i + 1) + nX * j + k + 1;
                c[6] = nX * nY * (i + 1) + nX * (j + 1) + k + 1;
                c[7] = nX * nY * (i + 1) + nX * (j + 1) + k;
            }
    return e;
}

/* Free-DOF numbering: nodes in ascending ID (Subdomain.cs:149-168 sorts the
 * IDs), per node the DOF types X, Y, Z in enumeration order; a constrained DOF
 * gets -1 and is eliminated (Model.cs:165-215, Subdomain.cs:77-135).
 * constrained[3*node+d] != 0 marks a constraint.  Returns the free count. */
int orc_enumerate_dofs(int n_nodes, const unsigned char *constrained, int *dof)
{
    int total = 0;
    for (int n = 0; n < n_nodes; n++)
        for (int d = 0; d < 3; d++)
            dof[3 * n + d] = constrained[3 * n + d] ? -1 : total++;
    return total;
}

/* Cantilever boundary condition of the workload: every DOF of the x = 0
 * face is constrained. */
void orc_cantilever_bc(int nex, int ney, int nez, unsigned char *constrained)
{
    const int nX = nex + 1, nY = ney + 1, nZ = nez + 1;
    memset(constrained, 0, (size_t)3 * nX * nY * nZ);
    for (int i = 0; i < nZ; i++)
        for (int j = 0; j < nY; j++) {
            int n = nX * nY * i + nX * j;
            constrained[3 * n] = constrained[3 * n + 1] = constrained[3 * n + 2] = 1;
        }
}

/* Total load `total` in Z, split equally over the x = max face nodes; each
 * node's entry is ASSIGNED (Model.cs:227-238, `Forces[dof] = Amount / ...`,
 * one subdomain).  f has n_free entries and is cleared first.  Returns the
 * number of loaded nodes. */
int orc_cantilever_loads(int nex, int ney, int nez, const int *dof,
                         double total, double *f)
{
    const int nX = nex + 1, nY = ney + 1, nZ = nez + 1;
    int n_free = 0;
    for (int i = 0; i < 3 * nX * nY * nZ; i++)
        if (dof[i] >= 0) n_free++;
    memset(f, 0, sizeof(double) * (size_t)n_free);
    const int count = nY * nZ;
    const double amount = total / count;
    for (int i = 0; i < nZ; i++)
        for (int j = 0; j < nY; j++) {
            int n = nX * nY * i + nX * j + nex;
            int d = dof[3 * n + 2];
            if (d >= 0) f[d] = amount;
        }
    return count;
}

/* ------------------------------------------------------------------------ */
/* Hexa8 element: rElHexa8.f90                                               */
/* ------------------------------------------------------------------------ */

/* Gauss-Legendre table, Fortran DATA GP/GW (rElHexa8.f90:435-445), column
 * iInt-1 holds the iInt-point rule with the 13-digit truncated constants. */
static const double GP[4][4] = {
    {0.0, 0.0, 0.0, 0.0},
    {-.5773502691896, .5773502691896, 0.0, 0.0},
    {-.7745966692415, 0.0, .7745966692415, 0.0},
    {-.8611363115941, -.3399810435849, .3399810435849, .8611363115941}};
static const double GW[4][4] = {
    {2.0, 0.0, 0.0, 0.0},
    {1.0, 1.0, 0.0, 0.0},
    {.5555555555556, .8888888888889, .5555555555556, 0.0},
    {.3478548451375, .6521451548625, .6521451548625, .3478548451375}};

/* CalcH8NablaShape (rElHexa8.f90:25-69).  DS[0..7] = dN/dxi, [8..15] dN/deta,
 * [16..23] dN/dzeta, with the sqrt(1/8) factor applied to both terms. */
static void h8_nabla_shape(double xi, double eta, double zeta, double *DS)
{
    const double s = 0.35355339059327376220042218105242;
    double xiP = (1.0 + xi) * s, etaP = (1.0 + eta) * s, zetaP = (1.0 + zeta) * s;
    double xiM = (1.0 - xi) * s, etaM = (1.0 - eta) * s, zetaM = (1.0 - zeta) * s;
    DS[0] = -etaM * zetaM;  DS[1] = -DS[0];
    DS[2] = etaP * zetaM;   DS[3] = -DS[2];
    DS[4] = -etaM * zetaP;  DS[5] = -DS[4];
    DS[6] = etaP * zetaP;   DS[7] = -DS[6];
    DS[8] = -xiM * zetaM;   DS[9] = -xiP * zetaM;
    DS[10] = -DS[9];        DS[11] = -DS[8];
    DS[12] = -xiM * zetaP;  DS[13] = -xiP * zetaP;
    DS[14] = -DS[13];       DS[15] = -DS[12];
    DS[16] = -xiM * etaM;   DS[17] = -xiP * etaM;
    DS[18] = -xiP * etaP;   DS[19] = -xiM * etaP;
    DS[20] = -DS[16];       DS[21] = -DS[17];
    DS[22] = -DS[18];       DS[23] = -DS[19];
}

/* CalcH8Shape (rElHexa8.f90:93-117). */
static void h8_shape(double xi, double eta, double zeta, double *S)
{
    const double h = 0.5;
    double xiP = (1.0 + xi) * h, etaP = (1.0 + eta) * h, zetaP = (1.0 + zeta) * h;
    double xiM = (1.0 - xi) * h, etaM = (1.0 - eta) * h, zetaM = (1.0 - zeta) * h;
    S[0] = xiM * etaM * zetaM; S[1] = xiP * etaM * zetaM;
    S[2] = xiP * etaP * zetaM; S[3] = xiM * etaP * zetaM;
    S[4] = xiM * etaM * zetaP; S[5] = xiP * etaM * zetaP;
    S[6] = xiP * etaP * zetaP; S[7] = xiM * etaP * zetaP;
}

/* CalcH8JDetJ (rElHexa8.f90:141-208).  xyz is the C# double[8,3] (node-major)
 * = Fortran faXYZ(3,8).  Jf is Fortran J(3,3) column-major: Jf[i + 3j] =
 * J(i+1, j+1) = sum_a DS_i(a) x_j(a), i natural, j physical.  Returns 0, or
 * -1001 when detJ < 1e-8 (the Fortran prints and continues). */
static int h8_jdetj(const double *xyz, const double *DS, double *Jf, double *det)
{
#define X(j, a) xyz[3 * (a) + (j)]
#define JJ(i, j) Jf[(i) + 3 * (j)]
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const double *d = DS + 8 * i;
            JJ(i, j) = d[0] * X(j, 0) + d[1] * X(j, 1) + d[2] * X(j, 2) +
                       d[3] * X(j, 3) + d[4] * X(j, 4) + d[5] * X(j, 5) +
                       d[6] * X(j, 6) + d[7] * X(j, 7);
        }
    double d1 = JJ(0, 0) * (JJ(1, 1) * JJ(2, 2) - JJ(2, 1) * JJ(1, 2));
    double d2 = -JJ(0, 1) * (JJ(1, 0) * JJ(2, 2) - JJ(2, 0) * JJ(1, 2));
    double d3 = JJ(0, 2) * (JJ(1, 0) * JJ(2, 1) - JJ(2, 0) * JJ(1, 1));
    *det = d1 + d2 + d3;
    return (*det < 0.00000001) ? -1001 : 0;
#undef X
#undef JJ
}

/* CalcH8B (rElHexa8.f90:232-397): Voigt rows [xx, yy, zz, xy, yz, zx] with
 * engineering shear; columns node-major (u, v, w).  Bf is Fortran B(6,24):
 * Bf[r + 6c]. */
static void h8_b(double det, const double *Jf, const double *DS, double *Bf)
{
#define JJ(i, j) Jf[(i) + 3 * (j)]
#define BB(r, c) Bf[(r) + 6 * (c)]
    double inv = 1.0 / det;
    double Ji[3][3];
    Ji[0][0] = (JJ(1, 1) * JJ(2, 2) - JJ(2, 1) * JJ(1, 2)) * inv;
    Ji[1][0] = (JJ(2, 0) * JJ(1, 2) - JJ(1, 0) * JJ(2, 2)) * inv;
    Ji[2][0] = (JJ(1, 0) * JJ(2, 1) - JJ(2, 0) * JJ(1, 1)) * inv;
    Ji[0][1] = (JJ(2, 1) * JJ(0, 2) - JJ(0, 1) * JJ(2, 2)) * inv;
    Ji[1][1] = (JJ(0, 0) * JJ(2, 2) - JJ(2, 0) * JJ(0, 2)) * inv;
    Ji[2][1] = (JJ(2, 0) * JJ(0, 1) - JJ(2, 1) * JJ(0, 0)) * inv;
    Ji[0][2] = (JJ(0, 1) * JJ(1, 2) - JJ(1, 1) * JJ(0, 2)) * inv;
    Ji[1][2] = (JJ(1, 0) * JJ(0, 2) - JJ(0, 0) * JJ(1, 2)) * inv;
    Ji[2][2] = (JJ(0, 0) * JJ(1, 1) - JJ(1, 0) * JJ(0, 1)) * inv;
    for (int c = 0; c < 144; c++) Bf[c] = 0.0;
    for (int a = 0; a < 8; a++) {
        double gx = Ji[0][0] * DS[a] + Ji[0][1] * DS[8 + a] + Ji[0][2] * DS[16 + a];
        double gy = Ji[1][0] * DS[a] + Ji[1][1] * DS[8 + a] + Ji[1][2] * DS[16 + a];
        double gz = Ji[2][0] * DS[a] + Ji[2][1] * DS[8 + a] + Ji[2][2] * DS[16 + a];
        int u = 3 * a, v = 3 * a + 1, w = 3 * a + 2;
        BB(0, u) = gx;
        BB(1, v) = gy;
        BB(2, w) = gz;
        BB(3, u) = gy; BB(3, v) = gx;
        BB(4, v) = gz; BB(4, w) = gy;
        BB(5, u) = gz; BB(5, w) = gx;
    }
#undef JJ
#undef BB
}

/* CalcH8GaussMatrices (rElHexa8.f90:422-464).  Gauss points ordered xi outer,
 * zeta inner; weight = GWx*GWy*GWz*detJ.  Layouts (C# arrays handed to the
 * Fortran, Hexa8WithStochasticMaterial.cs:84-91): w[M], S[M][8], DS[M][24],
 * J[M] as Fortran (3,3) column-major, detJ[M], B[M][24][6]. */
void orc_h8_gauss_matrices(int iInt, const double *xyz, double *w, double *S,
                           double *DS, double *J, double *detJ, double *B)
{
    int M = 0;
    for (int ix = 0; ix < iInt; ix++)
        for (int iy = 0; iy < iInt; iy++)
            for (int iz = 0; iz < iInt; iz++) {
                double xi = GP[iInt - 1][ix], eta = GP[iInt - 1][iy],
                       zeta = GP[iInt - 1][iz];
                h8_nabla_shape(xi, eta, zeta, DS + 24 * M);
                h8_shape(xi, eta, zeta, S + 8 * M);
                (void)h8_jdetj(xyz, DS + 24 * M, J + 9 * M, detJ + M);
                h8_b(detJ[M], J + 9 * M, DS + 24 * M, B + 144 * M);
                w[M] = GW[iInt - 1][ix] * GW[iInt - 1][iy] * GW[iInt - 1][iz] * detJ[M];
                M++;
            }
}

/* CalcH8K (rElHexa8.f90:489-543): per Gauss point, EB = E*B(:,I), then for
 * J >= I, K(KS) += (B(:,J).EB) * w.  E is Fortran faE(6,6,M): E[36M + r + 6c]
 * = E(r+1, c+1); K300 is the upper triangle packed row-major. */
void orc_h8_k(int iInt, const double *E, const double *B, const double *w,
              double *K300)
{
    const int n3 = iInt * iInt * iInt;
    for (int i = 0; i < 300; i++) K300[i] = 0.0;
    for (int M = 0; M < n3; M++) {
        const double *Em = E + 36 * M;
        const double *Bm = B + 144 * M;
        int KS = 0;
        for (int I = 0; I < 24; I++) {
            double EB[6];
            for (int r = 0; r < 6; r++)
                EB[r] = Em[r + 0] * Bm[0 + 6 * I] + Em[r + 6] * Bm[1 + 6 * I] +
                        Em[r + 12] * Bm[2 + 6 * I] + Em[r + 18] * Bm[3 + 6 * I] +
                        Em[r + 24] * Bm[4 + 6 * I] + Em[r + 30] * Bm[5 + 6 * I];
            for (int Jc = I; Jc < 24; Jc++) {
                double st = Bm[0 + 6 * Jc] * EB[0] + Bm[1 + 6 * Jc] * EB[1] +
                            Bm[2 + 6 * Jc] * EB[2] + Bm[3 + 6 * Jc] * EB[3] +
                            Bm[4 + 6 * Jc] * EB[4] + Bm[5 + 6 * Jc] * EB[5];
                K300[KS] = K300[KS] + st * w[M];
                KS++;
            }
        }
    }
}

/* Isotropic D (ElasticMaterial3D.cs:20-44; StochasticElasticMaterial3D.cs:
 * 26-53 with the sampled modulus in place of E).  Row-major 6x6. */
void orc_elastic_D(double E, double nu, double *D)
{
    double e1 = E / (1.0 + nu);
    double e2 = e1 * nu / (1.0 - 2.0 * nu);
    double e3 = e1 + e2;
    double e4 = e1 * 0.5;
    for (int i = 0; i < 36; i++) D[i] = 0.0;
    D[0] = e3; D[1] = e2; D[2] = e2;
    D[6] = e2; D[7] = e3; D[8] = e2;
    D[12] = e2; D[13] = e2; D[14] = e3;
    D[21] = e4; D[28] = e4; D[35] = e4;
}

/* Hexa8WithStochasticMaterial.StiffnessMatrix (H8WSM.cs:61-110): the same D
 * at all 8 Gauss points (the modulus is evaluated at the element centroid,
 * H8WSM.cs:65-70,112-139), then CalcH8GaussMatrices + CalcH8K.  Returns the
 * number of elements with a Jacobian determinant below 1e-8. */
int orc_ke_mesh(int n_elem, const int *conn, const double *xyz,
                const double *E_elem, double nu, double *ke300)
{
    int bad = 0;
    for (int e = 0; e < n_elem; e++) {
        double x8[24], w[8], S[64], DS[192], J[72], detJ[8], B[1152], E[288], D[36];
        for (int a = 0; a < 8; a++)
            for (int j = 0; j < 3; j++) x8[3 * a + j] = xyz[3 * conn[8 * e + a] + j];
        orc_h8_gauss_matrices(2, x8, w, S, DS, J, detJ, B);
        for (int m = 0; m < 8; m++)
            if (detJ[m] < 0.00000001) { bad++; break; }
        orc_elastic_D(E_elem[e], nu, D);
        for (int m = 0; m < 8; m++) memcpy(E + 36 * m, D, sizeof(D));
        orc_h8_k(2, E, B, w, ke300 + 300 * (size_t)e);
    }
    return bad;
}

/* The deterministic C# element (Hexa8.cs:284-401): derivatives with the
 * plain 1/8 factor, Jacobian3D (Jacobian3D.cs:53-117), B from the inverse
 * (Hexa8.cs:256-282), K += (b_j . D b_i) * w accumulated per Gauss point. */
void orc_ke_element_csharp(const double *xyz, double E, double nu, double *K300)
{
    static const double sgn[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                     {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    double D[36];
    orc_elastic_D(E, nu, D);
    for (int i = 0; i < 300; i++) K300[i] = 0.0;
    for (int ix = 0; ix < 2; ix++)
        for (int iy = 0; iy < 2; iy++)
            for (int iz = 0; iz < 2; iz++) {
                double xi = GP[1][ix], eta = GP[1][iy], zeta = GP[1][iz];
                double dxi[8], deta[8], dzeta[8];
                for (int a = 0; a < 8; a++) {
                    dxi[a] = sgn[a][0] * 0.125 * (1.0 + sgn[a][1] * eta) * (1.0 + sgn[a][2] * zeta);
                    deta[a] = sgn[a][1] * 0.125 * (1.0 + sgn[a][0] * xi) * (1.0 + sgn[a][2] * zeta);
                    dzeta[a] = sgn[a][2] * 0.125 * (1.0 + sgn[a][0] * xi) * (1.0 + sgn[a][1] * eta);
                }
                double Jm[3][3] = {{0}};
                for (int a = 0; a < 8; a++)
                    for (int j = 0; j < 3; j++) {
                        Jm[0][j] += xyz[3 * a + j] * dxi[a];
                        Jm[1][j] += xyz[3 * a + j] * deta[a];
                        Jm[2][j] += xyz[3 * a + j] * dzeta[a];
                    }
                double det = Jm[0][0] * (Jm[1][1] * Jm[2][2] - Jm[2][1] * Jm[1][2]) -
                             Jm[0][1] * (Jm[1][0] * Jm[2][2] - Jm[2][0] * Jm[1][2]) +
                             Jm[0][2] * (Jm[1][0] * Jm[2][1] - Jm[2][0] * Jm[1][1]);
                double inv = 1.0 / det, Ji[3][3];
                Ji[0][0] = (Jm[1][1] * Jm[2][2] - Jm[2][1] * Jm[1][2]) * inv;
                Ji[0][1] = (Jm[2][1] * Jm[0][2] - Jm[0][1] * Jm[2][2]) * inv;
                Ji[0][2] = (Jm[0][1] * Jm[1][2] - Jm[1][1] * Jm[0][2]) * inv;
                Ji[1][0] = (Jm[2][0] * Jm[1][2] - Jm[1][0] * Jm[2][2]) * inv;
                Ji[1][1] = (Jm[0][0] * Jm[2][2] - Jm[2][0] * Jm[0][2]) * inv;
                Ji[1][2] = (Jm[1][0] * Jm[0][2] - Jm[0][0] * Jm[1][2]) * inv;
                Ji[2][0] = (Jm[1][0] * Jm[2][1] - Jm[2][0] * Jm[1][1]) * inv;
                Ji[2][1] = (Jm[2][0] * Jm[0][1] - Jm[2][1] * Jm[0][0]) * inv;
                Ji[2][2] = (Jm[0][0] * Jm[1][1] - Jm[1][0] * Jm[0][1]) * inv;
                double b[6][24];
                memset(b, 0, sizeof(b));
                for (int a = 0; a < 8; a++) {
                    b[0][3 * a] = Ji[0][0] * dxi[a] + Ji[0][1] * deta[a] + Ji[0][2] * dzeta[a];
                    b[1][3 * a + 1] = Ji[1][0] * dxi[a] + Ji[1][1] * deta[a] + Ji[1][2] * dzeta[a];
                    b[2][3 * a + 2] = Ji[2][0] * dxi[a] + Ji[2][1] * deta[a] + Ji[2][2] * dzeta[a];
                    b[3][3 * a] = b[1][3 * a + 1];
                    b[3][3 * a + 1] = b[0][3 * a];
                    b[4][3 * a + 1] = b[2][3 * a + 2];
                    b[4][3 * a + 2] = b[1][3 * a + 1];
                    b[5][3 * a] = b[2][3 * a + 2];
                    b[5][3 * a + 2] = b[0][3 * a];
                }
                double wf = GW[1][ix] * GW[1][iy] * GW[1][iz] * det;
                int KS = 0;
                for (int i = 0; i < 24; i++) {
                    double eb[6];
                    for (int r = 0; r < 6; r++)
                        eb[r] = D[6 * r] * b[0][i] + D[6 * r + 1] * b[1][i] + D[6 * r + 2] * b[2][i] +
                                D[6 * r + 3] * b[3][i] + D[6 * r + 4] * b[4][i] + D[6 * r + 5] * b[5][i];
                    for (int j = i; j < 24; j++) {
                        double st = b[0][j] * eb[0] + b[1][j] * eb[1] + b[2][j] * eb[2] +
                                    b[3][j] * eb[3] + b[4][j] * eb[4] + b[5][j] * eb[5];
                        K300[KS++] += st * wf;
                    }
                }
            }
}

/* Hexa8.CalculateConsistentMass (Hexa8.cs:403-436): per Gauss point of the
 * C# rule (xi outer, zeta inner; weight = w_xi w_eta w_zeta detJ, detJ of
 * Jacobian3D as in orc_ke_element_csharp) and shape values of CalcH8Shape
 * (Hexa8.cs:137-158), M[3a+d][3b+d] += N_a N_b weight rho for b >= a; the
 * three direction copies are equal.  Output: packed upper 300. */
void orc_h8_mass_csharp(const double *xyz, double rho, double *M300)
{
    static const double sgn[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                     {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    double m[8][8];
    memset(m, 0, sizeof(m));
    for (int ix = 0; ix < 2; ix++)
        for (int iy = 0; iy < 2; iy++)
            for (int iz = 0; iz < 2; iz++) {
                double xi = GP[1][ix], eta = GP[1][iy], zeta = GP[1][iz];
                double dxi[8], deta[8], dzeta[8];
                for (int a = 0; a < 8; a++) {
                    dxi[a] = sgn[a][0] * 0.125 * (1.0 + sgn[a][1] * eta) * (1.0 + sgn[a][2] * zeta);
                    deta[a] = sgn[a][1] * 0.125 * (1.0 + sgn[a][0] * xi) * (1.0 + sgn[a][2] * zeta);
                    dzeta[a] = sgn[a][2] * 0.125 * (1.0 + sgn[a][0] * xi) * (1.0 + sgn[a][1] * eta);
                }
                double Jm[3][3] = {{0}};
                for (int a = 0; a < 8; a++)
                    for (int j = 0; j < 3; j++) {
                        Jm[0][j] += xyz[3 * a + j] * dxi[a];
                        Jm[1][j] += xyz[3 * a + j] * deta[a];
                        Jm[2][j] += xyz[3 * a + j] * dzeta[a];
                    }
                double det = Jm[0][0] * (Jm[1][1] * Jm[2][2] - Jm[2][1] * Jm[1][2]) -
                             Jm[0][1] * (Jm[1][0] * Jm[2][2] - Jm[2][0] * Jm[1][2]) +
                             Jm[0][2] * (Jm[1][0] * Jm[2][1] - Jm[2][0] * Jm[1][1]);
                double wd = GW[1][ix] * GW[1][iy] * GW[1][iz] * det * rho;
                double N[8];
                for (int a = 0; a < 8; a++)
                    N[a] = ((1.0 + sgn[a][0] * xi) * 0.5) * ((1.0 + sgn[a][1] * eta) * 0.5) *
                           ((1.0 + sgn[a][2] * zeta) * 0.5);
                for (int a = 0; a < 8; a++)
                    for (int b = a; b < 8; b++) m[a][b] += N[a] * N[b] * wd;
            }
    for (int i = 0; i < 300; i++) M300[i] = 0.0;
    for (int a = 0; a < 8; a++)
        for (int b = a; b < 8; b++)
            for (int d = 0; d < 3; d++) {
                int r = 3 * a + d, c = 3 * b + d;
                M300[300 - ((24 - r + 1) * (24 - r) / 2) + c - r] = m[a][b];
            }
}

/* SymmetricMatrix2D.IndexOf (SymmetricMatrix2D.cs:44-49) for the 24x24
 * packed upper triangle. */
static inline int sym_index(int row, int col)
{
    int r = row < col ? row : col, c = row < col ? col : row;
    return 300 - ((24 - r + 1) * (24 - r) / 2) + c - r;
}

/* ------------------------------------------------------------------------ */
/* Skyline storage                                                           */
/* ------------------------------------------------------------------------ */

/* CalculateRowIndex (GMAS.cs:14-44): height[d] = max over elements holding d
 * of d - minDOF(e); rowIndex[0] = 0, rowIndex[1] = 1, rowIndex[i+1] =
 * rowIndex[i] + height[i] + 1.  Returns the number of stored entries. */
int64_t orc_skyline_rowindex(int n_elem, const int *conn, const int *dof,
                             int n_free, int64_t *rowIndex)
{
    int *h = (int *)calloc((size_t)n_free + 1, sizeof(int));
    for (int e = 0; e < n_elem; e++) {
        int mn = 0x7fffffff;
        for (int a = 0; a < 8; a++)
            for (int d = 0; d < 3; d++) {
                int g = dof[3 * conn[8 * e + a] + d];
                if (g != -1 && g < mn) mn = g;
            }
        for (int a = 0; a < 8; a++)
            for (int d = 0; d < 3; d++) {
                int g = dof[3 * conn[8 * e + a] + d];
                if (g != -1 && g - mn > h[g]) h[g] = g - mn;
            }
    }
    rowIndex[0] = 0;
    if (n_free > 0) rowIndex[1] = 1;
    for (int i = 1; i < n_free; i++) rowIndex[i + 1] = rowIndex[i] + h[i] + 1;
    free(h);
    return rowIndex[n_free];
}

/* CalculateGlobalMatrix scatter (GMAS.cs:60-98): element order, rows then
 * columns of Ke, entries with dofRow >= dofCol >= 0 added at
 * data[rowIndex[dofRow] + dofRow - dofCol].  data must be zeroed. */
void orc_skyline_assemble(int n_elem, const int *conn, const int *dof,
                          const double *ke300, const int64_t *rowIndex,
                          double *data)
{
    for (int e = 0; e < n_elem; e++) {
        const double *K = ke300 + 300 * (size_t)e;
        for (int i = 0; i < 24; i++) {
            int gr = dof[3 * conn[8 * e + i / 3] + i % 3];
            if (gr == -1) continue;
            for (int j = 0; j < 24; j++) {
                int gc = dof[3 * conn[8 * e + j / 3] + j % 3];
                if (gc == -1) continue;
                int height = gr - gc;
                if (height >= 0) data[rowIndex[gr] + height] += K[sym_index(i, j)];
            }
        }
    }
}

/* SkylineMatrix2D.Multiply (SkylineMatrix2D.cs:224-286), scale 1, cleared
 * output: symmetric product over the full profile. */
void orc_skyline_multiply(int n, const int64_t *rowIndex, const double *d,
                          const double *x, double *y)
{
    memset(y, 0, sizeof(double) * (size_t)n);
    int64_t pos = 0;
    for (int i = 0; i < n; i++) {
        int64_t height = rowIndex[i + 1] - rowIndex[i];
        if (height <= 0) continue;
        y[i] += 1.0 * d[pos] * x[i];
        pos++;
        for (int64_t j = 0; j < height - 1; j++) {
            int row = (int)(i - j - 1);
            y[row] += 1.0 * d[pos] * x[i];
            y[i] += 1.0 * d[pos] * x[row];
            pos++;
        }
    }
}

/* SkylineMatrix2D.Factorize (SkylineMatrix2D.cs:63-158): in-place active-
 * column LDL^T.  A pivot under `tol` used as a divisor raises (returns -1);
 * a pivot under `tol` on the diagonal being formed is a zero-energy mode:
 * it is set to 1, its column/row cleared, and (for the first `max_modes`)
 * its column index kept in cols[] and its vector in zems[k*n ..]: 1 at the
 * column, minus the cleared row of L before it (:114-124), then
 * back-substituted through L^T after the factorization (:138-157).
 * Returns the number of zero-energy modes found. */
int orc_skyline_factorize_zem(int n, const int64_t *rowIndex, double *d, double tol,
                              int max_modes, int *cols, double *zems)
{
    int kfix = 0;
    const int64_t len = rowIndex[n];
    for (int nn = 0; nn < n; nn++) {
        int64_t KN = rowIndex[nn];
        int64_t KL = KN + 1;
        int64_t KU = rowIndex[nn + 1] - 1;
        int64_t KH = KU - KL;
        if (KH < 0) continue;
        int64_t K;
        if (KH > 0) {
            K = nn - KH;
            int64_t KLT = KU;
            for (int64_t j = 0; j < KH; j++) {
                KLT--;
                int64_t KI = rowIndex[K];
                int64_t ND = rowIndex[K + 1] - KI - 1;
                if (ND > 0) {
                    int64_t KK = (j + 1 < ND) ? j + 1 : ND;
                    double C = 0;
                    for (int64_t l = 1; l <= KK; l++) C += d[KI + l] * d[KLT + l];
                    d[KLT] -= C;
                }
                K++;
            }
        }
        K = nn;
        double B = 0;
        for (int64_t KK = KL; KK <= KU; KK++) {
            K--;
            int64_t KI = rowIndex[K];
            if (fabs(d[KI]) < tol) return -1;
            double C = d[KK] / d[KI];
            B += C * d[KK];
            d[KK] = C;
        }
        d[KN] -= B;
        if (fabs(d[KN]) < tol) {
            d[KN] = 1;
            double *z = kfix < max_modes ? zems + (size_t)kfix * n : NULL;
            if (z) {
                memset(z, 0, sizeof(double) * (size_t)n);
                z[nn] = 1;
                cols[kfix] = nn;
            }
            int j1 = nn;
            for (int64_t i1 = KN + 1; i1 <= KU; i1++) {
                j1--;
                if (z) z[j1] = -d[i1];
                d[i1] = 0;
            }
            for (int irest = nn + 1; irest < n; irest++) {
                int64_t m1 = rowIndex[irest] + irest - nn;
                if (m1 <= rowIndex[irest + 1] && m1 < len) d[m1] = 0;
            }
            kfix++;
        }
    }
    const int kept = kfix < max_modes ? kfix : max_modes;
    for (int f = 0; f < kept && n > 1; f++) {
        double *z = zems + (size_t)f * n;
        for (int nn = n - 1; nn >= 1; nn--) {
            int64_t KL = rowIndex[nn] + 1, KU = rowIndex[nn + 1] - 1;
            int k = nn;
            for (int64_t KK = KL; KK <= KU; KK++) {
                k--;
                z[k] -= d[KK] * z[nn];
            }
        }
    }
    return kfix;
}

int orc_skyline_factorize(int n, const int64_t *rowIndex, double *d, double tol)
{
    return orc_skyline_factorize_zem(n, rowIndex, d, tol, 0, NULL, NULL);
}

/* SkylineMatrix2D.Solve (SkylineMatrix2D.cs:161-212): forward reduction,
 * diagonal scaling, back substitution on the factorized profile. */
void orc_skyline_solve(int n, const int64_t *rowIndex, const double *d,
                       const double *f, double *x)
{
    memcpy(x, f, sizeof(double) * (size_t)n);
    for (int nn = 0; nn < n; nn++) {
        int64_t KL = rowIndex[nn] + 1, KU = rowIndex[nn + 1] - 1;
        if (KU >= KL) {
            int k = nn;
            double C = 0;
            for (int64_t KK = KL; KK <= KU; KK++) {
                k--;
                C += d[KK] * x[k];
            }
            x[nn] -= C;
        }
    }
    for (int nn = 0; nn < n; nn++) x[nn] /= d[rowIndex[nn]];
    int nn = n - 1;
    for (int l = 1; l < n; l++) {
        int64_t KL = rowIndex[nn] + 1, KU = rowIndex[nn + 1] - 1;
        if (KU >= KL) {
            int k = nn;
            for (int64_t KK = KL; KK <= KU; KK++) {
                k--;
                x[k] -= d[KK] * x[nn];
            }
        }
        nn--;
    }
}

/* ------------------------------------------------------------------------ */
/* CSR over free DOFs (the same element scatter into full symmetric storage) */
/* ------------------------------------------------------------------------ */

typedef struct { int *ptr, *idx; } node_adj;

/* Node adjacency through shared elements, sorted ascending per node. */
static node_adj build_node_adj(int n_elem, const int *conn, int n_nodes)
{
    int *cnt = (int *)calloc((size_t)n_nodes + 1, sizeof(int));
    for (int e = 0; e < n_elem; e++)
        for (int a = 0; a < 8; a++) cnt[conn[8 * e + a] + 1]++;
    for (int i = 0; i < n_nodes; i++) cnt[i + 1] += cnt[i];
    int *inc = (int *)malloc(sizeof(int) * (size_t)cnt[n_nodes]);
    int *fillp = (int *)malloc(sizeof(int) * (size_t)n_nodes);
    memcpy(fillp, cnt, sizeof(int) * (size_t)n_nodes);
    for (int e = 0; e < n_elem; e++)
        for (int a = 0; a < 8; a++) inc[fillp[conn[8 * e + a]]++] = e;
    node_adj A;
    A.ptr = (int *)calloc((size_t)n_nodes + 1, sizeof(int));
    int cap = 64 * n_nodes + 64;
    A.idx = (int *)malloc(sizeof(int) * (size_t)cap);
    int *tmp = (int *)malloc(sizeof(int) * 8 * 1024);
    for (int n = 0; n < n_nodes; n++) {
        int m = 0;
        for (int p = cnt[n]; p < cnt[n + 1]; p++)
            for (int a = 0; a < 8; a++) tmp[m++] = conn[8 * inc[p] + a];
        /* insertion sort + unique (m <= 8 * incident elements) */
        for (int i = 1; i < m; i++) {
            int v = tmp[i], j = i - 1;
            while (j >= 0 && tmp[j] > v) { tmp[j + 1] = tmp[j]; j--; }
            tmp[j + 1] = v;
        }
        int u = 0;
        for (int i = 0; i < m; i++)
            if (u == 0 || tmp[i] != tmp[u - 1]) tmp[u++] = tmp[i];
        if (A.ptr[n] + u > cap) {
            cap = 2 * cap + u;
            A.idx = (int *)realloc(A.idx, sizeof(int) * (size_t)cap);
        }
        memcpy(A.idx + A.ptr[n], tmp, sizeof(int) * (size_t)u);
        A.ptr[n + 1] = A.ptr[n] + u;
    }
    free(tmp); free(fillp); free(inc); free(cnt);
    return A;
}

static int max_node(int n_elem, const int *conn)
{
    int m = -1;
    for (int i = 0; i < 8 * n_elem; i++) if (conn[i] > m) m = conn[i];
    return m + 1;
}

/* Row pointers of the full symmetric CSR pattern; returns nnz. */
int64_t orc_csr_count(int n_elem, const int *conn, const int *dof, int n_free,
                      int64_t *rowptr)
{
    int n_nodes = max_node(n_elem, conn);
    node_adj A = build_node_adj(n_elem, conn, n_nodes);
    memset(rowptr, 0, sizeof(int64_t) * ((size_t)n_free + 1));
    for (int n = 0; n < n_nodes; n++) {
        int cols = 0;
        for (int p = A.ptr[n]; p < A.ptr[n + 1]; p++)
            for (int d = 0; d < 3; d++) cols += dof[3 * A.idx[p] + d] >= 0;
        for (int d = 0; d < 3; d++) {
            int g = dof[3 * n + d];
            if (g >= 0) rowptr[g + 1] = cols;
        }
    }
    for (int i = 0; i < n_free; i++) rowptr[i + 1] += rowptr[i];
    free(A.ptr); free(A.idx);
    return rowptr[n_free];
}

void orc_csr_fill(int n_elem, const int *conn, const int *dof, int n_free,
                  const int64_t *rowptr, int *col)
{
    (void)n_free;
    int n_nodes = max_node(n_elem, conn);
    node_adj A = build_node_adj(n_elem, conn, n_nodes);
    for (int n = 0; n < n_nodes; n++)
        for (int d = 0; d < 3; d++) {
            int g = dof[3 * n + d];
            if (g < 0) continue;
            int64_t k = rowptr[g];
            for (int p = A.ptr[n]; p < A.ptr[n + 1]; p++)
                for (int q = 0; q < 3; q++) {
                    int c = dof[3 * A.idx[p] + q];
                    if (c >= 0) col[k++] = c;
                }
        }
    free(A.ptr); free(A.idx);
}

static int64_t csr_find(const int64_t *rowptr, const int *col, int r, int c)
{
    int64_t lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) / 2;
        if (col[mid] == c) return mid;
        if (col[mid] < c) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* The GMAS.cs:60-98 scatter into full symmetric CSR (both triangles). */
void orc_csr_assemble(int n_elem, const int *conn, const int *dof, int n_free,
                      const double *ke300, const int64_t *rowptr,
                      const int *col, double *val)
{
    memset(val, 0, sizeof(double) * (size_t)rowptr[n_free]);
    for (int e = 0; e < n_elem; e++) {
        const double *K = ke300 + 300 * (size_t)e;
        for (int i = 0; i < 24; i++) {
            int gr = dof[3 * conn[8 * e + i / 3] + i % 3];
            if (gr == -1) continue;
            for (int j = 0; j < 24; j++) {
                int gc = dof[3 * conn[8 * e + j / 3] + j % 3];
                if (gc == -1) continue;
                val[csr_find(rowptr, col, gr, gc)] += K[sym_index(i, j)];
            }
        }
    }
}

void orc_csr_multiply(int n, const int64_t *rowptr, const int *col,
                      const double *val, const double *x, double *y)
{
    for (int i = 0; i < n; i++) {
        double s = 0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) s += val[k] * x[col[k]];
        y[i] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* PCG: SolverPCG.Solve (SolverPCG.cs:105-139) with the simple search vector  */
/* calculator (SolverPCGSimpleSearchVectorCalculator.cs:18-41)                */
/* ------------------------------------------------------------------------ */

typedef void (*op_fn)(const void *ctx, const double *in, double *out);

static double dot(int n, const double *a, const double *b)
{
    double s = 0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* x0 = 0, r = f, detf = |f| (SolverPCGSkyline.cs:75-100 with
 * MCPCGDirect.cs:53-80).  Returns 0 on convergence, 1 otherwise. */
static int pcg_core(int n, op_fn matvec, const void *mctx, op_fn prec,
                    const void *pctx, const double *f, double *x, int max_it,
                    double tol, int *iters, double *err_hist)
{
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    double *z = (double *)calloc((size_t)n, sizeof(double));
    double *p = (double *)calloc((size_t)n, sizeof(double));
    double *q = (double *)calloc((size_t)n, sizeof(double));
    memcpy(r, f, sizeof(double) * (size_t)n);
    memset(x, 0, sizeof(double) * (size_t)n);
    double detf = sqrt(dot(n, f, f));
    int it = 0, converged = 0;
    double zr_old = 0;
    if (sqrt(dot(n, r, r)) > 1e-25) {
        prec(pctx, r, z);
        for (it = 0; it < max_it; it++) {
            double zr = dot(n, z, r);
            if (it > 0) {
                double b = zr / zr_old;
                for (int i = 0; i < n; i++) p[i] = z[i] + b * p[i];
            } else {
                memcpy(p, z, sizeof(double) * (size_t)n);
            }
            zr_old = zr;
            matvec(mctx, p, q);
            double a = zr / dot(n, p, q);
            double detz = 0;
            for (int i = 0; i < n; i++) {
                r[i] -= a * q[i];
                x[i] += a * p[i];
                detz += r[i] * r[i];
            }
            double err = sqrt(detz) / detf;
            if (err_hist) err_hist[it] = err;
            if (err < tol) { converged = 1; it++; break; }
            prec(pctx, r, z);
        }
    } else {
        converged = 1;
    }
    *iters = it;
    free(r); free(z); free(p); free(q);
    return converged ? 0 : 1;
}

typedef struct { int n; const int64_t *ri; const double *d; } sky_ctx;
typedef struct { int n; const int64_t *rp; const int *col; const double *val; } csr_ctx;
typedef struct { int n; const double *diag; } jac_ctx;

static void sky_mv(const void *c, const double *in, double *out)
{
    const sky_ctx *s = (const sky_ctx *)c;
    orc_skyline_multiply(s->n, s->ri, s->d, in, out);
}
static void sky_prec(const void *c, const double *in, double *out)
{
    const sky_ctx *s = (const sky_ctx *)c;
    orc_skyline_solve(s->n, s->ri, s->d, in, out);
}
static void csr_mv(const void *c, const double *in, double *out)
{
    const csr_ctx *s = (const csr_ctx *)c;
    orc_csr_multiply(s->n, s->rp, s->col, s->val, in, out);
}
static void jac_prec(const void *c, const double *in, double *out)
{
    const jac_ctx *s = (const jac_ctx *)c;
    for (int i = 0; i < s->n; i++) out[i] = in[i] / s->diag[i];
}
static void id_prec(const void *c, const double *in, double *out)
{
    const jac_ctx *s = (const jac_ctx *)c;
    memcpy(out, in, sizeof(double) * (size_t)s->n);
}

/* PCG with the skyline product (MCPCGDirect.cs:48-51).  precond:
 * ORC_PRECOND_SKYLINE = solve with the factorized profile P (MCPCGDirect.cs:
 * 38-46, the first sample's LDL^T), ORC_PRECOND_JACOBI = 1/diag(K),
 * ORC_PRECOND_NONE = identity. */
int orc_pcg_skyline(int n, const int64_t *rowIndex, const double *K,
                    int precond, const int64_t *P_rowIndex, const double *P,
                    const double *f, double *x, int max_it, double tol,
                    int *iters, double *err_hist)
{
    sky_ctx mk = {n, rowIndex, K};
    sky_ctx pk = {n, P_rowIndex, P};
    double *diag = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) diag[i] = K[rowIndex[i]];
    jac_ctx jk = {n, diag};
    int rc;
    if (precond == ORC_PRECOND_SKYLINE)
        rc = pcg_core(n, sky_mv, &mk, sky_prec, &pk, f, x, max_it, tol, iters, err_hist);
    else if (precond == ORC_PRECOND_JACOBI)
        rc = pcg_core(n, sky_mv, &mk, jac_prec, &jk, f, x, max_it, tol, iters, err_hist);
    else
        rc = pcg_core(n, sky_mv, &mk, id_prec, &jk, f, x, max_it, tol, iters, err_hist);
    free(diag);
    return rc;
}

/* Jacobi (or unpreconditioned) PCG on the CSR image. */
int orc_pcg_csr(int n, const int64_t *rowptr, const int *col, const double *val,
                int precond, const double *f, double *x, int max_it,
                double tol, int *iters, double *err_hist)
{
    csr_ctx mk = {n, rowptr, col, val};
    double *diag = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) {
        int64_t k = csr_find(rowptr, col, i, i);
        diag[i] = k >= 0 ? val[k] : 1.0;
    }
    jac_ctx jk = {n, diag};
    int rc = pcg_core(n, csr_mv, &mk, precond == ORC_PRECOND_JACOBI ? jac_prec : id_prec,
                      &jk, f, x, max_it, tol, iters, err_hist);
    free(diag);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Lognormal Karhunen-Loeve field                                            */
/* ------------------------------------------------------------------------ */

/* Hexa8WithStochasticMaterial.GetStochasticPoints (H8WSM.cs:112-139):
 * bounding-box minimum plus half the largest pairwise node distance per
 * axis. */
void orc_element_centroid(const double *x8, double *c)
{
    double X = 0, Y = 0, Z = 0;
    double mnx = x8[0], mny = x8[1], mnz = x8[2];
    for (int i = 0; i < 8; i++) {
        mnx = mnx > x8[3 * i] ? x8[3 * i] : mnx;
        mny = mny > x8[3 * i + 1] ? x8[3 * i + 1] : mny;
        mnz = mnz > x8[3 * i + 2] ? x8[3 * i + 2] : mnz;
        for (int j = i + 1; j < 8; j++) {
            double ax = fabs(x8[3 * j] - x8[3 * i]);
            double ay = fabs(x8[3 * j + 1] - x8[3 * i + 1]);
            double az = fabs(x8[3 * j + 2] - x8[3 * i + 2]);
            X = X < ax ? ax : X;
            Y = Y < ay ? ay : Y;
            Z = Z < az ? az : Z;
        }
    }
    c[0] = mnx + X / 2;
    c[1] = mny + Y / 2;
    c[2] = mnz + Z / 2;
}

/* Lognormal modulus at each element centroid (Lognormal...cs:134-160,
 * 173-193): E = exp(ln(meanE) - sigma^2/2 + sum_i xi_i * sigma * sqrt(eigs_i)
 * * prod_d a * (index even ? sin : cos)(w * (x_d - L_d/2))), with
 * sigma = sqrt(ln(cov^2 + 1)) (Lognormal...cs:115-117, mean 1; the reference
 * hard-codes cov = 0.5, here it is an argument).  The 1-D table is
 * w1d[d*kmax + k], a1d[d*kmax + k] for 1-based mode index k+1; evpPos[3i+d]
 * holds the 1-based mode index of axis d in 3-D term i.  Returns the number of
 * non-finite moduli. */
int orc_kl_field(int n_elem, const int *conn, const double *xyz, int M,
                 int kmax, const double *w1d, const double *a1d,
                 const int *evpPos, const double *eigs, const double *domain,
                 const double *xi, double meanE, double cov, double *E_out)
{
    double sig = sqrt(log(pow(cov / 1.0, 2) + 1));
    int bad = 0;
    for (int e = 0; e < n_elem; e++) {
        double x8[24], c[3];
        for (int a = 0; a < 8; a++)
            for (int j = 0; j < 3; j++) x8[3 * a + j] = xyz[3 * conn[8 * e + a] + j];
        orc_element_centroid(x8, c);
        double res = log(meanE) - 0.5 * pow(sig, 2);
        for (int i = 0; i < M; i++) {
            double fcoef = sig * sqrt(eigs[i]);
            for (int d = 0; d < 3; d++) {
                int index = evpPos[3 * i + d];
                double w = w1d[d * kmax + index - 1];
                double a = a1d[d * kmax + index - 1];
                double x = c[d] - domain[d] / 2;
                fcoef *= index % 2 == 0 ? a * sin(w * x) : a * cos(w * x);
            }
            fcoef += 0.0; /* meanGaussian (Lognormal...cs:118,157) */
            res += xi[i] * fcoef;
        }
        E_out[e] = exp(res);
        if (!isfinite(E_out[e])) bad++;
    }
    return bad;
}

/* ------------------------------------------------------------------------ */
/* Reference Monte Carlo loop                                                */
/* ------------------------------------------------------------------------ */

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* MonteCarloAnalyzerWithStochasticMaterial.SolveNormal (MC...Material.cs:
 * 232-296) on the cantilever of the workload (unit cube, x = 0 clamped,
 * total load -100 in Z on the x = 1 face): per sample set xi, rebuild K
 * (KL modulus at centroids -> Ke with memoized (w, B) per element,
 * H8WSM.cs:80-106 -> skyline profile + scatter, recomputed every sample as
 * GMAS.cs:51-103), then Initialize (first sample only: clone K and factorize
 * with tolerance 1e-8, MCPCGDirect.cs:53-80) and Solve (PCG preconditioned by
 * that factor, SolverPCG.cs:105-139, tolerance `tol` instead of the
 * hard-wired 1e-7).  u_out[s*n_free ...] receives each solution (may be
 * NULL), iters[s] the product count; times = {element, factorize, solution,
 * all} seconds (the buckets of MC...Material.cs:250-291).  Returns n_free, or
 * a negative error. */
int orc_mc_reference(int nex, int ney, int nez, double nu, int M, int kmax,
                     const double *w1d, const double *a1d, const int *evpPos,
                     const double *eigs, double meanE, double cov,
                     int n_samples, const double *xi, double tol, int max_it,
                     double *u_out, int *iters, double *times)
{
    const int nX = nex + 1, nY = ney + 1, nZ = nez + 1;
    const int n_nodes = nX * nY * nZ, n_elem = nex * ney * nez;
    double *xyz = (double *)malloc(sizeof(double) * 3 * (size_t)n_nodes);
    int *conn = (int *)malloc(sizeof(int) * 8 * (size_t)n_elem);
    unsigned char *cons = (unsigned char *)malloc(3 * (size_t)n_nodes);
    int *dof = (int *)malloc(sizeof(int) * 3 * (size_t)n_nodes);
    orc_mesh_box(nex, ney, nez, 1.0 / nex, 1.0 / ney, 1.0 / nez, xyz, conn);
    orc_cantilever_bc(nex, ney, nez, cons);
    int n = orc_enumerate_dofs(n_nodes, cons, dof);
    double *f = (double *)malloc(sizeof(double) * (size_t)n);
    orc_cantilever_loads(nex, ney, nez, dof, -100.0, f);
    const double domain[3] = {1.0, 1.0, 1.0};

    /* memoized integration data per element (Hexa8Memoizer, H8WSM.cs:11-27) */
    double *Wm = (double *)malloc(sizeof(double) * 8 * (size_t)n_elem);
    double *Bm = (double *)malloc(sizeof(double) * 1152 * (size_t)n_elem);
    double *E = (double *)malloc(sizeof(double) * (size_t)n_elem);
    double *ke = (double *)malloc(sizeof(double) * 300 * (size_t)n_elem);
    int64_t *ri = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    double *P = NULL, *K = NULL, *x = (double *)malloc(sizeof(double) * (size_t)n);
    int64_t *Pri = NULL;
    int memo = 0;
    times[0] = times[1] = times[2] = times[3] = 0;
    double t_all = now_s();
    int rc_all = n;
    for (int s = 0; s < n_samples; s++) {
        double t0 = now_s();
        orc_kl_field(n_elem, conn, xyz, M, kmax, w1d, a1d, evpPos, eigs, domain,
                     xi + (size_t)s * M, meanE, cov, E);
        for (int e = 0; e < n_elem; e++) {
            double *w = Wm + 8 * (size_t)e, *B = Bm + 1152 * (size_t)e;
            if (!memo) {
                double x8[24], S[64], DS[192], J[72], dj[8];
                for (int a = 0; a < 8; a++)
                    for (int j = 0; j < 3; j++) x8[3 * a + j] = xyz[3 * conn[8 * e + a] + j];
                orc_h8_gauss_matrices(2, x8, w, S, DS, J, dj, B);
            }
            double D[36], Eg[288];
            orc_elastic_D(E[e], nu, D);
            for (int m = 0; m < 8; m++) memcpy(Eg + 36 * m, D, sizeof(D));
            orc_h8_k(2, Eg, B, w, ke + 300 * (size_t)e);
        }
        memo = 1;
        int64_t len = orc_skyline_rowindex(n_elem, conn, dof, n, ri);
        free(K);
        K = (double *)calloc((size_t)len, sizeof(double));
        if (!K) { rc_all = -2; break; }
        orc_skyline_assemble(n_elem, conn, dof, ke, ri, K);
        double t1 = now_s();
        times[0] += t1 - t0;
        if (!P) {
            P = (double *)malloc(sizeof(double) * (size_t)len);
            Pri = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
            memcpy(P, K, sizeof(double) * (size_t)len);
            memcpy(Pri, ri, sizeof(int64_t) * ((size_t)n + 1));
            if (orc_skyline_factorize(n, Pri, P, 1e-8) != 0) { rc_all = -3; break; }
        }
        double t2 = now_s();
        times[1] += t2 - t1;
        int it = 0;
        orc_pcg_skyline(n, ri, K, ORC_PRECOND_SKYLINE, Pri, P, f, x, max_it, tol, &it, NULL);
        times[2] += now_s() - t2;
        if (iters) iters[s] = it;
        if (u_out) memcpy(u_out + (size_t)s * n, x, sizeof(double) * (size_t)n);
    }
    times[3] = now_s() - t_all;
    free(xyz); free(conn); free(cons); free(dof); free(f); free(Wm); free(Bm);
    free(E); free(ke); free(ri); free(P); free(Pri); free(K); free(x);
    return rc_all;
}

/* The Monte Carlo loop of orc_mc_reference with the storage and
 * preconditioner swapped for the CSR image and Jacobi (the deviation used
 * where the skyline profile is infeasible, 32^3 and up): per sample KL
 * modulus -> Ke (memoized w, B) -> CSR scatter -> Jacobi PCG to `tol`.
 * The CSR pattern is built once.  Same outputs as orc_mc_reference. */
int orc_mc_csr_jacobi(int nex, int ney, int nez, double nu, int M, int kmax,
                      const double *w1d, const double *a1d, const int *evpPos,
                      const double *eigs, double meanE, double cov,
                      int n_samples, const double *xi, double tol, int max_it,
                      double *u_out, int *iters, double *times)
{
    const int nX = nex + 1, nY = ney + 1, nZ = nez + 1;
    const int n_nodes = nX * nY * nZ, n_elem = nex * ney * nez;
    double *xyz = (double *)malloc(sizeof(double) * 3 * (size_t)n_nodes);
    int *conn = (int *)malloc(sizeof(int) * 8 * (size_t)n_elem);
    unsigned char *cons = (unsigned char *)malloc(3 * (size_t)n_nodes);
    int *dof = (int *)malloc(sizeof(int) * 3 * (size_t)n_nodes);
    orc_mesh_box(nex, ney, nez, 1.0 / nex, 1.0 / ney, 1.0 / nez, xyz, conn);
    orc_cantilever_bc(nex, ney, nez, cons);
    int n = orc_enumerate_dofs(n_nodes, cons, dof);
    double *f = (double *)malloc(sizeof(double) * (size_t)n);
    orc_cantilever_loads(nex, ney, nez, dof, -100.0, f);
    const double domain[3] = {1.0, 1.0, 1.0};
    int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    int64_t nnz = orc_csr_count(n_elem, conn, dof, n, rp);
    int *col = (int *)malloc(sizeof(int) * (size_t)nnz);
    orc_csr_fill(n_elem, conn, dof, n, rp, col);
    double *val = (double *)malloc(sizeof(double) * (size_t)nnz);
    double *Wm = (double *)malloc(sizeof(double) * 8 * (size_t)n_elem);
    double *Bm = (double *)malloc(sizeof(double) * 1152 * (size_t)n_elem);
    double *E = (double *)malloc(sizeof(double) * (size_t)n_elem);
    double *ke = (double *)malloc(sizeof(double) * 300 * (size_t)n_elem);
    double *x = (double *)malloc(sizeof(double) * (size_t)n);
    times[0] = times[1] = times[2] = times[3] = 0;
    double t_all = now_s();
    for (int e = 0; e < n_elem; e++) {
        double x8[24], S[64], DS[192], J[72], dj[8];
        for (int a = 0; a < 8; a++)
            for (int j = 0; j < 3; j++) x8[3 * a + j] = xyz[3 * conn[8 * e + a] + j];
        orc_h8_gauss_matrices(2, x8, Wm + 8 * (size_t)e, S, DS, J, dj, Bm + 1152 * (size_t)e);
    }
    for (int s = 0; s < n_samples; s++) {
        double t0 = now_s();
        orc_kl_field(n_elem, conn, xyz, M, kmax, w1d, a1d, evpPos, eigs, domain,
                     xi + (size_t)s * M, meanE, cov, E);
        for (int e = 0; e < n_elem; e++) {
            double D[36], Eg[288];
            orc_elastic_D(E[e], nu, D);
            for (int m = 0; m < 8; m++) memcpy(Eg + 36 * m, D, sizeof(D));
            orc_h8_k(2, Eg, Bm + 1152 * (size_t)e, Wm + 8 * (size_t)e, ke + 300 * (size_t)e);
        }
        orc_csr_assemble(n_elem, conn, dof, n, ke, rp, col, val);
        double t1 = now_s();
        times[0] += t1 - t0;
        int it = 0;
        orc_pcg_csr(n, rp, col, val, ORC_PRECOND_JACOBI, f, x, max_it, tol, &it, NULL);
        times[2] += now_s() - t1;
        if (iters) iters[s] = it;
        if (u_out) memcpy(u_out + (size_t)s * n, x, sizeof(double) * (size_t)n);
    }
    times[3] = now_s() - t_all;
    free(xyz); free(conn); free(cons); free(dof); free(f); free(rp); free(col); free(val);
    free(Wm); free(Bm); free(E); free(ke); free(x);
    return n;
}


/* ------------------------------------------------------------------------ */
/* Element results                                                           */
/* ------------------------------------------------------------------------ */

/* CalcH8Strains (rElHexa8.f90:635-648): strains(:, M) = B(:, :, M) u per
 * Gauss point; B is Fortran (6, 24, M) (column-major: B[(M*24 + c)*6 + r]),
 * strains (6, M). */
void orc_h8_strains(int iInt, const double *B, const double *u, double *strains)
{
    const int M3 = iInt * iInt * iInt;
    for (int m = 0; m < M3; m++)
        for (int r = 0; r < 6; r++) {
            double acc = 0.0;
            for (int c = 0; c < 24; c++) acc += B[((size_t)m * 24 + c) * 6 + r] * u[c];
            strains[m * 6 + r] = acc;
        }
}

/* CalcH8Forces (rElHexa8.f90:724-747): forces = sum_M w(M) B(:,:,M)^T
 * stresses(:, M), Gauss points in the order of the loops iX, iY, iZ. */
void orc_h8_forces(int iInt, const double *B, const double *w, const double *stresses,
                   double *forces)
{
    const int M3 = iInt * iInt * iInt;
    for (int c = 0; c < 24; c++) forces[c] = 0.0;
    for (int m = 0; m < M3; m++)
        for (int c = 0; c < 24; c++) {
            double acc = 0.0;
            for (int r = 0; r < 6; r++) acc += B[((size_t)m * 24 + c) * 6 + r] * stresses[m * 6 + r];
            forces[c] = forces[c] + w[m] * acc;
        }
}

/* CalcH8MLumped (rElHexa8.f90:673-699): diagonal mass sum(w) rho / 8 on the
 * 24 diagonal entries of the packed upper 300 (positions 1, 25, 48, ...). */
void orc_h8_mlumped(int iInt, double density, const double *w, double *M300)
{
    const int M3 = iInt * iInt * iInt;
    for (int i = 0; i < 300; i++) M300[i] = 0.0;
    double term = 0.0;
    for (int m = 0; m < M3; m++) term = term + w[m];
    term = term * density * 0.125;
    int pos = 0;
    for (int s1 = 24; s1 >= 1; s1--) {
        M300[pos] = term;
        pos += s1;
    }
}

/* Stresses of ElasticMaterial3D (ElasticMaterial3D.cs:20-43): D(E, nu) times
 * the strains, D as built there (fE1 = E/(1+nu), fE2 = fE1 nu/(1-2nu),
 * fE3 = fE1 + fE2, fE4 = fE1/2), the product summed over the 6 columns. */
void orc_elastic_stress(double E, double nu, const double *strain, double *stress)
{
    const double e1 = E / (1 + nu), e2 = e1 * nu / (1 - 2 * nu), e3 = e1 + e2, e4 = e1 * 0.5;
    const double D[6][6] = {{e3, e2, e2, 0, 0, 0}, {e2, e3, e2, 0, 0, 0}, {e2, e2, e3, 0, 0, 0},
                            {0, 0, 0, e4, 0, 0},   {0, 0, 0, 0, e4, 0},   {0, 0, 0, 0, 0, e4}};
    for (int i = 0; i < 6; i++) {
        double acc = 0.0;
        for (int j = 0; j < 6; j++) acc += D[i][j] * strain[j];
        stress[i] = acc;
    }
}

/* Per-element results of a solved model, as Hexa8.CalculateStresses +
 * CalculateForces (Hexa8.cs:453-510) evaluate them element by element:
 * local displacements (constrained DOFs 0), Gauss matrices, strains per
 * Gauss point, stresses D(E_e) strain, internal forces sum w B^T sigma.
 * strains/stresses [n_elem][8][6], forces [n_elem][24] (any may be NULL). */
void orc_element_results(int n_elem, const int *conn, const double *xyz, const int *dof,
                         const double *u, const double *E, double nu, double *strains,
                         double *stresses, double *forces)
{
    double w[8], S[64], DS[192], J[72], dJ[8], B[1152], x8[24], ue[24], st[48], sg[48], fe[24];
    for (int e = 0; e < n_elem; e++) {
        for (int a = 0; a < 8; a++)
            for (int d = 0; d < 3; d++) {
                const int n = conn[8 * e + a];
                x8[3 * a + d] = xyz[3 * n + d];
                const int g = dof[3 * n + d];
                ue[3 * a + d] = g >= 0 ? u[g] : 0.0;
            }
        orc_h8_gauss_matrices(2, x8, w, S, DS, J, dJ, B);
        orc_h8_strains(2, B, ue, st);
        for (int m = 0; m < 8; m++) orc_elastic_stress(E[e], nu, st + 6 * m, sg + 6 * m);
        orc_h8_forces(2, B, w, sg, fe);
        if (strains) memcpy(strains + (size_t)e * 48, st, sizeof(st));
        if (stresses) memcpy(stresses + (size_t)e * 48, sg, sizeof(sg));
        if (forces) memcpy(forces + (size_t)e * 24, fe, sizeof(fe));
    }
}

/* ------------------------------------------------------------------------ */
/* Newmark dynamics                                                          */
/* ------------------------------------------------------------------------ */

/* Time step i of MassAccelerationSinusoidalLoad (MassAccelerationSinusoidalLoad.cs:17-29):
 * mag sin(2 pi / period * dt * i) for i < (int)(tlen / dt), else 0. */
double orc_sinusoidal_acc(int i, double mag, double period, double dt, double tlen)
{
    const int steps = (int)(tlen / dt);
    const double pi = 3.14159265358979323846; /* Math.PI, the double nearest pi */
    return i < steps ? mag * sin(2 * pi / period * dt * i) : 0.0;
}

/* Linear implicit Newmark of NewmarkDynamicAnalyzer: coefficients
 * (NewmarkDynamicAnalyzer.cs:53-77; alpha = beta of the textbook, delta =
 * gamma), effective matrix K + a0 M + a1 C (ProblemStructural.cs:138-157 with
 * SkylineMatrix2D.LinearCombination, SkylineMatrix2D.cs:310-340), element
 * damping C_e = alphaR M_e + betaR K_e (Hexa8.DampingMatrix, Hexa8.cs:444-451),
 * consistent mass (orc_h8_mass_csharp), per step: forces cleared, static
 * nodal loads assigned, the mass-acceleration history load M_e a_e added on
 * free rows (Model.cs:266-280; a_e holds acc(i) on every DOF of direction
 * acc_dir, constrained nodes included, Hexa8.cs:512-530), RHS = F + M (a0 u +
 * a2 v + a3 a) + C (a1 u + a4 v + a5 a) (NewmarkDynamicAnalyzer.cs:189-206),
 * skyline LDL^T solve (the only reference solver that is correct here,
 * SolverSkyline), then the update of NewmarkDynamicAnalyzer.cs:301-315.
 * Steps run i = 0 .. nsteps-1 (NewmarkDynamicAnalyzer.cs:284) from rest.
 * probe_hist[i] = u[probe] after step i.  Returns 0, or < 0 on failure. */
int orc_newmark_skyline(int n_elem, const int *conn, const double *xyz, const int *dof,
                        int n, const double *E, double nu, double rho, double alphaR,
                        double betaR, double nb_alpha, double nb_delta, double dt, int nsteps,
                        const double *f_static, int acc_dir, double acc_mag,
                        double acc_period, double acc_tlen, int probe, double *probe_hist,
                        double *u_out, double *v_out, double *a_out)
{
    if (nb_delta < 0.5 || nb_alpha < 0.25 * (0.5 + nb_delta) * (0.5 + nb_delta)) return -1;
    const double a0 = 1 / (nb_alpha * dt * dt), a1 = nb_delta / (nb_alpha * dt);
    const double a2 = 1 / (nb_alpha * dt), a3 = 1 / (2 * nb_alpha) - 1;
    const double a4 = nb_delta / nb_alpha - 1, a5 = dt * 0.5 * (nb_delta / nb_alpha - 2);
    const double a6 = dt * (1 - nb_delta), a7 = nb_delta * dt;
    double *ke = (double *)malloc(sizeof(double) * 300 * (size_t)n_elem);
    double *me = (double *)malloc(sizeof(double) * 300 * (size_t)n_elem);
    double *ce = (double *)malloc(sizeof(double) * 300 * (size_t)n_elem);
    int64_t *ri = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    if (orc_ke_mesh(n_elem, conn, xyz, E, nu, ke) != 0) {
        free(ke); free(me); free(ce); free(ri);
        return -2;
    }
    for (int e = 0; e < n_elem; e++) {
        double x8[24];
        for (int a = 0; a < 8; a++)
            for (int j = 0; j < 3; j++) x8[3 * a + j] = xyz[3 * conn[8 * e + a] + j];
        orc_h8_mass_csharp(x8, rho, me + 300 * (size_t)e);
        for (int i = 0; i < 300; i++)
            ce[300 * (size_t)e + i] = alphaR * me[300 * (size_t)e + i] + betaR * ke[300 * (size_t)e + i];
    }
    const int64_t len = orc_skyline_rowindex(n_elem, conn, dof, n, ri);
    double *K = (double *)calloc((size_t)len, sizeof(double));
    double *Mg = (double *)calloc((size_t)len, sizeof(double));
    double *Cg = (double *)calloc((size_t)len, sizeof(double));
    orc_skyline_assemble(n_elem, conn, dof, ke, ri, K);
    orc_skyline_assemble(n_elem, conn, dof, me, ri, Mg);
    orc_skyline_assemble(n_elem, conn, dof, ce, ri, Cg);
    for (int64_t j = 0; j < len; j++) K[j] = 1.0 * K[j] + a0 * Mg[j] + a1 * Cg[j];
    int rc = orc_skyline_factorize(n, ri, K, 1e-15);
    /* mass-acceleration load for acc = 1: element M_e times the unit vector
     * of direction acc_dir, free rows only */
    double *bacc = (double *)calloc((size_t)n, sizeof(double));
    for (int e = 0; e < n_elem && acc_dir >= 0; e++) {
        const double *m = me + 300 * (size_t)e;
        for (int i = 0; i < 24; i++) {
            const int g = dof[3 * conn[8 * e + i / 3] + i % 3];
            if (g < 0) continue;
            double fi = 0;
            for (int j = 0; j < 24; j++)
                if (j % 3 == acc_dir) fi += m[sym_index(i, j)] * 1.0;
            bacc[g] += fi;
        }
    }
    double *u = (double *)calloc((size_t)n, sizeof(double));  /* previous displacement */
    double *v = (double *)calloc((size_t)n, sizeof(double));  /* current displacement */
    double *v1 = (double *)calloc((size_t)n, sizeof(double)); /* velocity */
    double *v2 = (double *)calloc((size_t)n, sizeof(double)); /* acceleration */
    double *uu = (double *)malloc(sizeof(double) * (size_t)n);
    double *uc = (double *)malloc(sizeof(double) * (size_t)n);
    double *uum = (double *)malloc(sizeof(double) * (size_t)n);
    double *ucc = (double *)malloc(sizeof(double) * (size_t)n);
    double *rhs = (double *)malloc(sizeof(double) * (size_t)n);
    double *x = (double *)malloc(sizeof(double) * (size_t)n);
    for (int step = 0; rc == 0 && step < nsteps; step++) {
        const double ag = orc_sinusoidal_acc(step, acc_mag, acc_period, dt, acc_tlen);
        for (int i = 0; i < n; i++) {
            uu[i] = a0 * v[i] + a2 * v1[i] + a3 * v2[i];
            uc[i] = a1 * v[i] + a4 * v1[i] + a5 * v2[i];
        }
        orc_skyline_multiply(n, ri, Mg, uu, uum);
        orc_skyline_multiply(n, ri, Cg, uc, ucc);
        for (int i = 0; i < n; i++) {
            const double F = (f_static ? f_static[i] : 0.0) + bacc[i] * ag;
            rhs[i] = F + uum[i] + ucc[i];
        }
        orc_skyline_solve(n, ri, K, rhs, x);
        for (int i = 0; i < n; i++) {
            u[i] = v[i];
            v[i] = x[i];
            const double vv = v2[i];
            v2[i] = a0 * (v[i] - u[i]) - a2 * v1[i] - a3 * vv;
            v1[i] += a6 * vv + a7 * v2[i];
        }
        if (probe_hist) probe_hist[step] = v[probe];
    }
    if (u_out) memcpy(u_out, v, sizeof(double) * (size_t)n);
    if (v_out) memcpy(v_out, v1, sizeof(double) * (size_t)n);
    if (a_out) memcpy(a_out, v2, sizeof(double) * (size_t)n);
    free(ke); free(me); free(ce); free(ri); free(K); free(Mg); free(Cg); free(bacc);
    free(u); free(v); free(v1); free(v2); free(uu); free(uc); free(uum); free(ucc);
    free(rhs); free(x);
    return rc == 0 ? 0 : -3;
}
