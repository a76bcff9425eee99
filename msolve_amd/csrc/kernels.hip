// HIP kernels for gfx950 (CDNA4): Hexa8 element stiffness, coloured scatter
// into the sliced block-ELL image of K, fused Jacobi-PCG steps over many
// samples at once, mixed-precision iterative refinement and the lognormal
// Karhunen-Loeve modulus field.  Layouts are described in msa_common.hpp.
//
// Wavefront mapping of the PCG kernels: lane = rr*T + s covers R = 64/T block
// rows (rr) times T samples (s).  Every value load reads 64 consecutive words;
// every gathered vector load reads T consecutive words of one block column.
// Per-sample dot products are reduced wave -> workgroup -> a partial per
// work item; k_reduce, launched next, sums the partials of a (class, tile) in
// a fixed order, so results do not depend on scheduling.
#include <algorithm>
#include <cmath>

#include <hip/hip_cooperative_groups.h>

#include "msa_common.hpp"

namespace msa {

// ------------------------------------------------------------ helpers

// Pointers kept in ClassDev are generic; the hot kernels cast them to the
// global address space so loads and stores compile to global_* (counted by
// vmcnt only) instead of flat_* instructions.
#define MSA_GLOBAL __attribute__((address_space(1)))
template <typename X>
__device__ __forceinline__ MSA_GLOBAL X *gbl(X *p)
{
    return (MSA_GLOBAL X *)p;
}

// Workgroups are dispatched round-robin over the 8 XCDs; give every XCD a
// contiguous range of work items so neighbouring slices share its L2.
__device__ __forceinline__ int xcd_item(int nitems)
{
    const int per = gridDim.x >> 3;
    const int item = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    return item < nitems ? item : -1;
}

static inline int xcd_grid(int nitems) { return ((nitems + 7) / 8) * 8; }

template <int T, typename V>
__device__ __forceinline__ V reduce_rows(V v)
{
#pragma unroll
    for (int off = T; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Wave + workgroup reduction of NV per-sample sums; writes NV*T partials.
template <int T, int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double *out)
{
    __shared__ double red[kWaves][NV][T];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = reduce_rows<T>(v[k]);
    if (lane < T) {
#pragma unroll
        for (int k = 0; k < NV; k++) red[wave][k][lane] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < T) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double acc = red[0][k][threadIdx.x];
#pragma unroll
            for (int w = 1; w < kWaves; w++) acc += red[w][k][threadIdx.x];
            out[k * T + threadIdx.x] = acc;
        }
    }
}


// Three components of one node in one-sample tiles (T = 1 layout: v[3*block
// + d]) move as a single 12- or 24-byte access.
template <typename real>
struct vec3 {
    real x, y, z;
};

template <typename real>
__device__ __forceinline__ vec3<real> ld3(const MSA_GLOBAL real *p, size_t i)
{
    const MSA_GLOBAL real *q = p + 3 * i;
    return vec3<real>{q[0], q[1], q[2]};
}

// The 9 entries of one stored block (slot, row lane rr, sample lane s);
// NT: non-temporal loads (values streamed once per product).  One-sample
// tiles read 16-byte groups (val_at).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <typename real, int T, bool NT = false>
__device__ __forceinline__ void load_block(const MSA_GLOBAL real *val, size_t slot, int rr, int s,
                                           real (&m)[9])
{
    if constexpr (T == 1) {
        const MSA_GLOBAL real *b = val + slot * 576;
        if constexpr (sizeof(real) == 4) {
            const MSA_GLOBAL f32x4 *p = (const MSA_GLOBAL f32x4 *)(b + rr * 4);
            f32x4 g0, g1;
            if constexpr (NT) {
                g0 = __builtin_nontemporal_load(p);
                g1 = __builtin_nontemporal_load(p + 64);
            } else {
                g0 = p[0];
                g1 = p[64];
            }
            m[0] = g0.x; m[1] = g0.y; m[2] = g0.z; m[3] = g0.w;
            m[4] = g1.x; m[5] = g1.y; m[6] = g1.z; m[7] = g1.w;
        } else {
            const MSA_GLOBAL f64x2 *p = (const MSA_GLOBAL f64x2 *)(b + rr * 2);
#pragma unroll
            for (int g = 0; g < 4; g++) {
                f64x2 d;
                if constexpr (NT)
                    d = __builtin_nontemporal_load(p + 64 * g);
                else
                    d = p[64 * g];
                m[2 * g] = d.x;
                m[2 * g + 1] = d.y;
            }
        }
        if constexpr (NT)
            m[8] = __builtin_nontemporal_load(b + 512 + rr);
        else
            m[8] = b[512 + rr];
    } else {
#pragma unroll
        for (int e = 0; e < 9; e++) {
            const MSA_GLOBAL real *p = val + val_at<real, T>(slot, e, rr, s);
            if constexpr (NT)
                m[e] = __builtin_nontemporal_load(p);
            else
                m[e] = *p;
        }
    }
}

// Store of one block in the one-sample-tile layout (16-byte groups).
template <typename real>
__device__ __forceinline__ void store_block(MSA_GLOBAL real *val, size_t slot, int rr,
                                            const real (&m)[9])
{
    MSA_GLOBAL real *b = val + slot * 576;
    if constexpr (sizeof(real) == 4) {
        MSA_GLOBAL f32x4 *p = (MSA_GLOBAL f32x4 *)(b + rr * 4);
        p[0] = f32x4{m[0], m[1], m[2], m[3]};
        p[64] = f32x4{m[4], m[5], m[6], m[7]};
    } else {
        MSA_GLOBAL f64x2 *p = (MSA_GLOBAL f64x2 *)(b + rr * 2);
#pragma unroll
        for (int g = 0; g < 4; g++) p[64 * g] = f64x2{m[2 * g], m[2 * g + 1]};
    }
    b[512 + rr] = m[8];
}

// Stored block j of a block row (slot base + j) whose block column is col.
// Classes that keep only the upper triangle (ClassDev::sym: stencil order,
// one-sample tiles) do not store their lower slots: lower slot j < 13 is
// read as the transpose of upper slot 26 - j of row col (which holds
// K(col, br)).  col < 0 (padding) reads the slot's own zeros.
template <typename real, int T, bool NT = false>
__device__ __forceinline__ void load_row_block(const ClassDev &C, const MSA_GLOBAL real *val,
                                               size_t base, int j, int col, int rr, int s,
                                               real (&m)[9])
{
    if constexpr (T == 1) {
        if (C.sym && j < 13 && col >= 0) {
            real t[9];
            load_block<real, 1, NT>(val, (size_t)(col >> 6) * 27 + (26 - j), col & 63, 0, t);
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++) m[3 * p + q] = t[3 * q + p];
            return;
        }
    }
    load_block<real, T, NT>(val, base + j, rr, s, m);
}

// Position of entry (p, q) of block j of row (rr, s) in a tile's values,
// with the lower-slot rule of load_row_block.
template <typename real, int T>
__device__ __forceinline__ size_t row_entry(const ClassDev &C, size_t base, int j, int col, int rr,
                                            int s, int p, int q)
{
    if constexpr (T == 1) {
        if (C.sym && j < 13 && col >= 0)
            return val_at<real, 1>((size_t)(col >> 6) * 27 + (26 - j), 3 * q + p, col & 63, 0);
    }
    return val_at<real, T>(base + j, 3 * p + q, rr, s);
}

template <typename real>
__device__ __forceinline__ void st3(MSA_GLOBAL real *p, size_t i, real x, real y, real z)
{
    MSA_GLOBAL real *q = p + 3 * i;
    q[0] = x;
    q[1] = y;
    q[2] = z;
}

// Right-hand side of row (br, d) for lane s of a tile at vb: the class's
// shared load vector, or the per-sample one of the dynamics mode.
__device__ __forceinline__ double fval(const ClassDev &C, size_t vb, int br, int d, int T, int s)
{
    return C.fs ? C.fs[vb + ((size_t)br * 3 + d) * T + s] : C.f[br * 3 + d];
}

// Samples a multigrid cycle serves: 0 every sample, 1 active, 2 due for a
// residual replacement, 3 refining.
__device__ __forceinline__ bool mg_act(const SampleState &S, int mode)
{
    switch (mode) {
    case 0: return true;
    case 1: return S.active;
    case 2: return S.ir_active && S.ru_due;
    default: return S.ir_active;
    }
}

__device__ __forceinline__ double fnorm_of(const ClassDev &C, const SampleState &S)
{
    return C.fs ? S.fnorm : C.fnorm;
}

// ------------------------------------------------------------ Hexa8 stiffness

// 13-digit Gauss abscissa of the reference (GaussQuadrature.cs:73-81,
// rElHexa8.f90:435-439) and sqrt(1/8) of CalcH8NablaShape (rElHexa8.f90:30).
constexpr double kGauss = 0.5773502691896;
constexpr double kSq125 = 0.35355339059327376220042218105242;

// One wavefront computes the unit-modulus element stiffness
// Khat = sum_gp w_gp B^T Dhat B (2x2x2 Gauss, Dhat = D(E = 1, nu)) and stages
// it in LDS as 64 node-pair blocks Kh[(a*8 + b)*9 + 3p + q].
// Phase 1, lane = gp*8 + a: shape derivatives at Gauss point gp for node a,
//   Jacobian by an 8-lane reduction, determinant/inverse as CalcH8JDetJ and
//   CalcH8B (rElHexa8.f90:141-250), gradient g_a of N_a -> LDS.
// Phase 2, lane = a*8 + b: the 3x3 block (a, b) of B^T D B for isotropic D:
//   lam g_a,p g_b,q + mu g_a,q g_b,p + mu delta_pq (g_a . g_b), summed over
//   the Gauss points with weight detJ (all Gauss weights are 1).
// xa is the coordinate of node a = lane & 7 (every lane loads its own node).
// Coordinates are taken relative to node 0 (the Jacobian is translation-
// invariant; this keeps the FP32 variant accurate on small elements).
// Returns true in lanes whose Gauss point has detJ < 1e-8.
// With Nl/Mh (64 entries each) it also stages the unit-density consistent
// mass Mh[a*8 + b] = sum_gp N_a N_b detJ (Hexa8.cs:403-436, shape values of
// CalcH8Shape, Hexa8.cs:137-158); the element mass is rho Mh on each of the
// three displacement directions.
template <typename kreal>
__device__ bool wave_khat(const double xa[3], double nu, kreal *Kh, kreal *g, kreal *w,
                          kreal *Nl = nullptr, kreal *Mh = nullptr)
{
    const int lane = threadIdx.x & 63;
    const int gp = lane >> 3, a = lane & 7;
    const double x0 = __shfl(xa[0], lane & 56, 64);
    const double y0 = __shfl(xa[1], lane & 56, 64);
    const double z0 = __shfl(xa[2], lane & 56, 64);
    const kreal X[3] = {(kreal)(xa[0] - x0), (kreal)(xa[1] - y0), (kreal)(xa[2] - z0)};
    const kreal sx = ((a + 1) & 2) ? 1 : -1;
    const kreal sy = (a & 2) ? 1 : -1;
    const kreal sz = (a & 4) ? 1 : -1;
    const kreal xi = ((gp >> 2) & 1) ? (kreal)kGauss : (kreal)-kGauss;
    const kreal eta = ((gp >> 1) & 1) ? (kreal)kGauss : (kreal)-kGauss;
    const kreal zeta = (gp & 1) ? (kreal)kGauss : (kreal)-kGauss;
    const kreal s = (kreal)kSq125;
    const kreal fx = (1 + sx * xi) * s, fy = (1 + sy * eta) * s, fz = (1 + sz * zeta) * s;
    const kreal dn[3] = {sx * (fy * fz), sy * (fx * fz), sz * (fx * fy)};
    kreal J[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            kreal v = dn[i] * X[j];
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            v += __shfl_xor(v, 4, 64);
            J[i][j] = v;
        }
    const kreal det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                      J[0][1] * (J[1][0] * J[2][2] - J[2][0] * J[1][2]) +
                      J[0][2] * (J[1][0] * J[2][1] - J[2][0] * J[1][1]);
    const kreal inv = (kreal)1 / det;
    kreal Ji[3][3];
    Ji[0][0] = (J[1][1] * J[2][2] - J[2][1] * J[1][2]) * inv;
    Ji[1][0] = (J[2][0] * J[1][2] - J[1][0] * J[2][2]) * inv;
    Ji[2][0] = (J[1][0] * J[2][1] - J[2][0] * J[1][1]) * inv;
    Ji[0][1] = (J[2][1] * J[0][2] - J[0][1] * J[2][2]) * inv;
    Ji[1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) * inv;
    Ji[2][1] = (J[2][0] * J[0][1] - J[2][1] * J[0][0]) * inv;
    Ji[0][2] = (J[0][1] * J[1][2] - J[1][1] * J[0][2]) * inv;
    Ji[1][2] = (J[1][0] * J[0][2] - J[0][0] * J[1][2]) * inv;
    Ji[2][2] = (J[0][0] * J[1][1] - J[1][0] * J[0][1]) * inv;
#pragma unroll
    for (int j = 0; j < 3; j++)
        g[lane * 3 + j] = Ji[j][0] * dn[0] + Ji[j][1] * dn[1] + Ji[j][2] * dn[2];
    if (a == 0) w[gp] = det;
    if (Mh)
        Nl[lane] = ((1 + sx * xi) * (kreal)0.5) * ((1 + sy * eta) * (kreal)0.5) *
                   ((1 + sz * zeta) * (kreal)0.5);
    __syncthreads();

    const int pa = lane >> 3, pb = lane & 7;
    const kreal e1 = (kreal)(1.0 / (1.0 + nu));
    const kreal lam = e1 * (kreal)nu / (kreal)(1.0 - 2.0 * nu);
    const kreal mu = e1 * (kreal)0.5;
    kreal K[9];
#pragma unroll
    for (int k = 0; k < 9; k++) K[k] = 0;
#pragma unroll
    for (int q8 = 0; q8 < 8; q8++) {
        const kreal wq = w[q8];
        const kreal *ga = g + (q8 * 8 + pa) * 3;
        const kreal *gb = g + (q8 * 8 + pb) * 3;
        const kreal A0 = ga[0], A1 = ga[1], A2 = ga[2];
        const kreal B0 = gb[0], B1 = gb[1], B2 = gb[2];
        const kreal mdot = mu * (A0 * B0 + A1 * B1 + A2 * B2);
        const kreal A[3] = {A0, A1, A2}, B[3] = {B0, B1, B2};
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
                kreal t = lam * A[p] * B[q] + mu * A[q] * B[p];
                if (p == q) t += mdot;
                K[p * 3 + q] += wq * t;
            }
    }
#pragma unroll
    for (int k = 0; k < 9; k++) Kh[lane * 9 + k] = K[k];
    if (Mh) {
        kreal m = 0;
#pragma unroll
        for (int q8 = 0; q8 < 8; q8++) m += (Nl[q8 * 8 + pa] * Nl[q8 * 8 + pb]) * w[q8];
        Mh[lane] = m;
    }
    __syncthreads();
    return det < (kreal)0.00000001;
}

__device__ __forceinline__ int sym_index(int r, int c)
{
    if (r > c) {
        int t = r;
        r = c;
        c = t;
    }
    return 300 - ((24 - r + 1) * (24 - r)) / 2 + c - r;
}

// Packed 300-entry Ke of independent elements (IElementMatrixProvider.Matrix).
__global__ __launch_bounds__(64) void k_ke_packed(int n_elem, const double *__restrict__ xyz8,
                                                  const double *__restrict__ E, double nu,
                                                  double *__restrict__ ke300,
                                                  int *__restrict__ status)
{
    __shared__ double Kh[576];
    __shared__ double g[192];
    __shared__ double w[8];
    const int e = blockIdx.x;
    if (e >= n_elem) return;
    const int lane = threadIdx.x & 63, a = lane & 7;
    double xa[3] = {xyz8[(size_t)e * 24 + a * 3], xyz8[(size_t)e * 24 + a * 3 + 1],
                    xyz8[(size_t)e * 24 + a * 3 + 2]};
    const bool neg = wave_khat<double>(xa, nu, Kh, g, w);
    const unsigned long long anyneg = __ballot(neg);
    const double Ee = E[e];
    const int pa = lane >> 3, pb = lane & 7;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int I = 3 * pa + p, Jc = 3 * pb + q;
            if (Jc >= I) ke300[(size_t)e * 300 + sym_index(I, Jc)] = Ee * Kh[lane * 9 + p * 3 + q];
        }
    if (status && lane == 0) status[e] = anyneg ? MSA_NEG_DETJ : MSA_OK;
}

// CALCH8GAUSSMATRICES drop-in: one lane per Gauss point, the Fortran formulas
// with absolute coordinates (rElHexa8.f90:25-464).  out layout per Gauss
// point M: w | S[8] | DS[24] | J[9] (Fortran column-major) | detJ | B[144].
__global__ void k_gauss_matrices(const double *__restrict__ xyz, double *__restrict__ out)
{
    const int M = threadIdx.x;
    if (M >= 8) return;
    const double xi = ((M >> 2) & 1) ? kGauss : -kGauss;
    const double eta = ((M >> 1) & 1) ? kGauss : -kGauss;
    const double zeta = (M & 1) ? kGauss : -kGauss;
    double *w = out + 0, *S = out + 8, *DS = out + 72, *J = out + 264, *dJ = out + 336,
           *B = out + 344;
    double xiP = (1.0 + xi) * kSq125, etaP = (1.0 + eta) * kSq125, zetaP = (1.0 + zeta) * kSq125;
    double xiM = (1.0 - xi) * kSq125, etaM = (1.0 - eta) * kSq125, zetaM = (1.0 - zeta) * kSq125;
    double d[24];
    d[0] = -etaM * zetaM; d[1] = -d[0]; d[2] = etaP * zetaM; d[3] = -d[2];
    d[4] = -etaM * zetaP; d[5] = -d[4]; d[6] = etaP * zetaP; d[7] = -d[6];
    d[8] = -xiM * zetaM; d[9] = -xiP * zetaM; d[10] = -d[9]; d[11] = -d[8];
    d[12] = -xiM * zetaP; d[13] = -xiP * zetaP; d[14] = -d[13]; d[15] = -d[12];
    d[16] = -xiM * etaM; d[17] = -xiP * etaM; d[18] = -xiP * etaP; d[19] = -xiM * etaP;
    d[20] = -d[16]; d[21] = -d[17]; d[22] = -d[18]; d[23] = -d[19];
    for (int k = 0; k < 24; k++) DS[M * 24 + k] = d[k];
    const double h = 0.5;
    double xp = (1.0 + xi) * h, ep = (1.0 + eta) * h, zp = (1.0 + zeta) * h;
    double xm = (1.0 - xi) * h, em = (1.0 - eta) * h, zm = (1.0 - zeta) * h;
    double sh[8] = {xm * em * zm, xp * em * zm, xp * ep * zm, xm * ep * zm,
                    xm * em * zp, xp * em * zp, xp * ep * zp, xm * ep * zp};
    for (int k = 0; k < 8; k++) S[M * 8 + k] = sh[k];
    double Jm[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double v = 0;
            for (int a = 0; a < 8; a++) v += d[8 * i + a] * xyz[3 * a + j];
            Jm[i][j] = v;
            J[M * 9 + i + 3 * j] = v;
        }
    double det = Jm[0][0] * (Jm[1][1] * Jm[2][2] - Jm[2][1] * Jm[1][2]) -
                 Jm[0][1] * (Jm[1][0] * Jm[2][2] - Jm[2][0] * Jm[1][2]) +
                 Jm[0][2] * (Jm[1][0] * Jm[2][1] - Jm[2][0] * Jm[1][1]);
    dJ[M] = det;
    double inv = 1.0 / det, Ji[3][3];
    Ji[0][0] = (Jm[1][1] * Jm[2][2] - Jm[2][1] * Jm[1][2]) * inv;
    Ji[1][0] = (Jm[2][0] * Jm[1][2] - Jm[1][0] * Jm[2][2]) * inv;
    Ji[2][0] = (Jm[1][0] * Jm[2][1] - Jm[2][0] * Jm[1][1]) * inv;
    Ji[0][1] = (Jm[2][1] * Jm[0][2] - Jm[0][1] * Jm[2][2]) * inv;
    Ji[1][1] = (Jm[0][0] * Jm[2][2] - Jm[2][0] * Jm[0][2]) * inv;
    Ji[2][1] = (Jm[2][0] * Jm[0][1] - Jm[2][1] * Jm[0][0]) * inv;
    Ji[0][2] = (Jm[0][1] * Jm[1][2] - Jm[1][1] * Jm[0][2]) * inv;
    Ji[1][2] = (Jm[1][0] * Jm[0][2] - Jm[0][0] * Jm[1][2]) * inv;
    Ji[2][2] = (Jm[0][0] * Jm[1][1] - Jm[1][0] * Jm[0][1]) * inv;
    double *Bm = B + M * 144;
    for (int k = 0; k < 144; k++) Bm[k] = 0.0;
    for (int a = 0; a < 8; a++) {
        double gx = Ji[0][0] * d[a] + Ji[0][1] * d[8 + a] + Ji[0][2] * d[16 + a];
        double gy = Ji[1][0] * d[a] + Ji[1][1] * d[8 + a] + Ji[1][2] * d[16 + a];
        double gz = Ji[2][0] * d[a] + Ji[2][1] * d[8 + a] + Ji[2][2] * d[16 + a];
        int u = 3 * a;
        Bm[0 + 6 * u] = gx;
        Bm[1 + 6 * (u + 1)] = gy;
        Bm[2 + 6 * (u + 2)] = gz;
        Bm[3 + 6 * u] = gy; Bm[3 + 6 * (u + 1)] = gx;
        Bm[4 + 6 * (u + 1)] = gz; Bm[4 + 6 * (u + 2)] = gy;
        Bm[5 + 6 * u] = gz; Bm[5 + 6 * (u + 2)] = gx;
    }
    w[M] = 1.0 * 1.0 * 1.0 * det;
}

// CALCH8K drop-in (rElHexa8.f90:489-543): one lane per packed entry, the
// Fortran summation order over Gauss points and strain components.
__global__ void k_h8k(const double *__restrict__ E, const double *__restrict__ B,
                      const double *__restrict__ w, double *__restrict__ K)
{
    const int ks = blockIdx.x * blockDim.x + threadIdx.x;
    if (ks >= 300) return;
    int I = 0, rem = ks;
    while (rem >= 24 - I) {
        rem -= 24 - I;
        I++;
    }
    const int Jc = I + rem;
    double acc = 0;
    for (int M = 0; M < 8; M++) {
        const double *Em = E + 36 * M, *Bm = B + 144 * M;
        double st = 0;
        double EB[6];
        for (int r = 0; r < 6; r++)
            EB[r] = Em[r] * Bm[6 * I] + Em[r + 6] * Bm[1 + 6 * I] + Em[r + 12] * Bm[2 + 6 * I] +
                    Em[r + 18] * Bm[3 + 6 * I] + Em[r + 24] * Bm[4 + 6 * I] +
                    Em[r + 30] * Bm[5 + 6 * I];
        st = Bm[6 * Jc] * EB[0] + Bm[1 + 6 * Jc] * EB[1] + Bm[2 + 6 * Jc] * EB[2] +
             Bm[3 + 6 * Jc] * EB[3] + Bm[4 + 6 * Jc] * EB[4] + Bm[5 + 6 * Jc] * EB[5];
        acc = acc + st * w[M];
    }
    K[ks] = acc;
}

// CalcH8Strains (rElHexa8.f90:635-648): strains(r, M) = sum_c B(r, c, M) u(c),
// lane = M*8 + r (r < 6); B Fortran (6, 24, 8).
__global__ void k_h8strains(const double *__restrict__ B, const double *__restrict__ u,
                            double *__restrict__ strains)
{
    const int M = threadIdx.x >> 3, r = threadIdx.x & 7;
    if (M >= 8 || r >= 6) return;
    double acc = 0.0;
    for (int c = 0; c < 24; c++) acc += B[(M * 24 + c) * 6 + r] * u[c];
    strains[M * 6 + r] = acc;
}

// CalcH8Forces (rElHexa8.f90:724-747): forces(c) = sum_M w(M) sum_r B(r, c, M)
// stresses(r, M), lane = c.
__global__ void k_h8forces(const double *__restrict__ B, const double *__restrict__ w,
                           const double *__restrict__ stresses, double *__restrict__ forces)
{
    const int c = threadIdx.x;
    if (c >= 24) return;
    double f = 0.0;
    for (int M = 0; M < 8; M++) {
        double acc = 0.0;
        for (int r = 0; r < 6; r++) acc += B[(M * 24 + c) * 6 + r] * stresses[M * 6 + r];
        f = f + w[M] * acc;
    }
    forces[c] = f;
}

// CalcH8MLumped (rElHexa8.f90:673-699): sum(w) rho / 8 on the 24 diagonal
// entries of the packed upper 300, zero elsewhere.
__global__ void k_h8mlumped(double density, const double *__restrict__ w, double *__restrict__ M)
{
    double term = 0.0;
    for (int m = 0; m < 8; m++) term = term + w[m];
    term = term * density * 0.125;
    for (int i = threadIdx.x; i < 300; i += blockDim.x) {
        // diagonal positions: 0, 24, 47, ... (row I starts at sum_{k<I} (24 - k))
        int I = 0, start = 0;
        while (start + (24 - I) <= i) {
            start += 24 - I;
            I++;
        }
        M[i] = (i == start) ? term : 0.0;
    }
}

// Element results of one sample (Hexa8.CalculateStresses + CalculateForces,
// Hexa8.cs:453-510 with CalcH8Strains/CalcH8Forces and ElasticMaterial3D):
// per element (one wavefront) and Gauss point M the strains B_M u_e, the
// stresses D(E_e, nu) strain and the internal forces sum_M w_M B_M^T
// sigma_M (w_M = detJ, the 2-point Gauss weights are 1).  u is the sample's
// FP64 solution (internal rows, lane s of its tile); constrained DOFs are 0.
template <int T>
__global__ __launch_bounds__(64) void k_element_results(const ClassDev *__restrict__ cp,
                                                        const double *__restrict__ u, int s,
                                                        int tile, double nu,
                                                        double *__restrict__ strains,
                                                        double *__restrict__ stresses,
                                                        double *__restrict__ forces)
{
    __shared__ double Kh[576];
    __shared__ double g[192];
    __shared__ double w[8];
    __shared__ double ue[24];
    __shared__ double eps[48], sig[48];
    const ClassDev &C = *cp;
    const int e = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int *en = C.conn + (size_t)e * 8;
    const int na = en[lane & 7];
    const double xa[3] = {C.xyz[3 * na], C.xyz[3 * na + 1], C.xyz[3 * na + 2]};
    (void)wave_khat<double>(xa, nu, Kh, g, w);
    if (lane < 24) {
        const int nd = en[lane / 3], d = lane % 3;
        const int b = C.node_block[nd];
        ue[lane] = (b >= 0 && ((C.node_free[nd] >> d) & 1)) ? u[((size_t)b * 3 + d) * T + s] : 0.0;
    }
    __syncthreads();
    const int M = lane >> 3, r = lane & 7;
    if (r < 6) {
        // B(r, 3a+j): rows xx, yy, zz, xy, yz, zx of the Voigt order
        double acc = 0.0;
        for (int a = 0; a < 8; a++) {
            const double *ga = g + (M * 8 + a) * 3;
            const double ux = ue[3 * a], uy = ue[3 * a + 1], uz = ue[3 * a + 2];
            switch (r) {
            case 0: acc += ga[0] * ux; break;
            case 1: acc += ga[1] * uy; break;
            case 2: acc += ga[2] * uz; break;
            case 3: acc += ga[1] * ux; acc += ga[0] * uy; break;
            case 4: acc += ga[2] * uy; acc += ga[1] * uz; break;
            default: acc += ga[2] * ux; acc += ga[0] * uz; break;
            }
        }
        eps[M * 6 + r] = acc;
    }
    __syncthreads();
    const double Ee = C.E[((size_t)tile * C.n_elem + e) * T + s];
    if (r < 6) {
        // ElasticMaterial3D.GetConstitutiveMatrix (ElasticMaterial3D.cs:20-43)
        const double e1 = Ee / (1 + nu), e2 = e1 * nu / (1 - 2 * nu), e3 = e1 + e2, e4 = e1 * 0.5;
        const double *ep = eps + M * 6;
        double v;
        if (r < 3)
            v = (r == 0 ? e3 : e2) * ep[0] + (r == 1 ? e3 : e2) * ep[1] + (r == 2 ? e3 : e2) * ep[2];
        else
            v = e4 * ep[r];
        sig[M * 6 + r] = v;
        if (strains) strains[(size_t)e * 48 + M * 6 + r] = ep[r];
        if (stresses) stresses[(size_t)e * 48 + M * 6 + r] = v;
    }
    __syncthreads();
    if (forces && lane < 24) {
        const int a = lane / 3, j = lane % 3;
        double f = 0.0;
        for (int m = 0; m < 8; m++) {
            const double *ga = g + (m * 8 + a) * 3;
            const double *sg = sig + m * 6;
            double acc;
            if (j == 0)
                acc = ga[0] * sg[0] + ga[1] * sg[3] + ga[2] * sg[5];
            else if (j == 1)
                acc = ga[1] * sg[1] + ga[0] * sg[3] + ga[2] * sg[4];
            else
                acc = ga[2] * sg[2] + ga[1] * sg[4] + ga[0] * sg[5];
            f = f + w[m] * acc;
        }
        forces[(size_t)e * 24 + lane] = f;
    }
}

// Khat (and, with Mh, the unit-density mass) of element e in LDS: from the
// class's FP64 element cache when it is built, else computed in place.
__device__ __forceinline__ void stage_khat(const ClassDev &C, int e, const double xa[3], double nu,
                                           double *Kh, double *g, double *w, double *Nl,
                                           double *Mh)
{
    const int lane = threadIdx.x & 63;
    if (C.kc && (!Mh || C.mc)) {
        const double *k = C.kc + (size_t)e * 576;
        for (int i = lane; i < 576; i += 64) Kh[i] = k[i];
        if (Mh) Mh[lane] = C.mc[(size_t)e * 64 + lane];
        __syncthreads();
    } else {
        stage_khat(C, e, xa, nu, Kh, g, w, Nl, Mh);
    }
}

__device__ __forceinline__ int colour_class(const ColourLaunch &cl, int item)
{
    int c = 0;
    while (c + 1 < cl.ncls && item >= cl.prefix[c + 1]) c++;
    return c;
}

// FP32-integration assembly (msa_batch_set_integration(MSA_MIXED)): one
// wavefront per element of the current colour (all classes) and tile group
// (blockIdx.y strides over the tiles), Khat integrated in FP32 and staged in
// LDS, then val += E_s * Khat for every sample of its tiles.  Elements of
// one colour share no node, so the read-modify-writes are race-free.
template <int T>
__global__ __launch_bounds__(64) void k_assemble(const ClassDev *__restrict__ cls, ColourLaunch cl,
                                                 double nu)
{
    constexpr int R = 64 / T;
    __shared__ float Kh[576];
    __shared__ float g[192];
    __shared__ float w[8];
    __shared__ float Nl[64], Mh[64];
    const int item = blockIdx.x;
    const int c = colour_class(cl, item);
    if (item >= cl.prefix[cl.ncls]) return;
    const ClassDev &C = cls[c];
    const int e = C.colour_elems[C.colour_ptr[cl.colour] + (item - cl.prefix[c])];
    const int lane = threadIdx.x & 63;
    const int *en = C.conn + (size_t)e * 8;
    const int na = en[lane & 7];
    const double xa[3] = {C.xyz[3 * na], C.xyz[3 * na + 1], C.xyz[3 * na + 2]};
    const bool neg = wave_khat<float>(xa, nu, Kh, g, w, Nl, C.dyn ? Mh : nullptr);
    if (__ballot(neg) && lane == 0 && blockIdx.y == 0) atomicAdd(C.negdet, 1);
    // effective stiffness of the dynamics mode: cK E Khat + cM Mhat on the
    // block diagonals (ProblemStructural.cs:138-157)
    const float cK = C.dyn ? (float)C.cK : 1.f, cM = (float)C.cM;
    const uint8_t *em = C.emap + (size_t)e * 64;
    const int s = lane % T;
    for (int tile = (int)blockIdx.y; tile < C.ntiles; tile += gridDim.y) {
        if (tile * T + s >= C.nsamp) continue;
        float *val = (float *)C.val + (size_t)tile * C.tile_vals;
        const float Es = (float)C.E[((size_t)tile * C.n_elem + e) * T + s] * cK;
        for (int pe = lane / T; pe < 576; pe += R) {
            const int pair = pe / 9, ent = pe - 9 * pair;
            const int j = em[pair];
            if (j == 255) continue;
            const int a = pair >> 3, b = pair & 7;
            const int p = ent / 3, q = ent - 3 * (ent / 3);
            const int nA = en[a], nB = en[b];
            if (!((C.node_free[nA] >> p) & 1) || !((C.node_free[nB] >> q) & 1)) continue;
            const int br = C.node_block[nA];
            const int slot = C.slice_ptr[br / R] + j;
            const size_t idx = val_at<float, T>(slot, ent, br % R, s);
            if (C.dyn && p == q)
                val[idx] += Es * Kh[pe] + cM * Mh[pair];
            else
                val[idx] += Es * Kh[pe];
        }
    }
}

// FP64 element cache of a class: Khat of every element (unit modulus) and,
// with_mass, the unit-density consistent mass; one wavefront per element.
// Counts elements with a Gauss-point detJ < 1e-8 in negdet.
__global__ __launch_bounds__(64) void k_khat_cache(const ClassDev *__restrict__ cp, double nu,
                                                   int with_mass)
{
    __shared__ double Kh[576];
    __shared__ double g[192];
    __shared__ double w[8];
    __shared__ double Nl[64], Mh[64];
    const ClassDev &C = *cp;
    const int e = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int *en = C.conn + (size_t)e * 8;
    const int na = en[lane & 7];
    const double xa[3] = {C.xyz[3 * na], C.xyz[3 * na + 1], C.xyz[3 * na + 2]};
    const bool neg = wave_khat<double>(xa, nu, Kh, g, w, Nl, with_mass ? Mh : nullptr);
    if (__ballot(neg) && lane == 0) atomicAdd(C.negdet, 1);
    MSA_GLOBAL double *k = gbl(C.kc + (size_t)e * 576);
    for (int i = lane; i < 576; i += 64) k[i] = Kh[i];
    if (with_mass) gbl(C.mc)[(size_t)e * 64 + lane] = Mh[lane];
}

// Gather assembly (GlobalMatrixAssemblerSkyline.cs:51-103 restated per
// stored value): every value of block (row, slot) is the FP64 sum over the
// elements holding both nodes of E_e cK Khat_e (+ cM Mhat_e on the block
// diagonal in the dynamics mode), in ascending element order, rounded once
// to `real` and written once.  Entries of constrained components are zero.
template <typename real, int T>
__global__ __launch_bounds__(256) void k_assemble_gather(const ClassDev *__restrict__ cls,
                                                         const int4 *__restrict__ items,
                                                         int nitems)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    if (tile * T + s >= C.nsamp) return;
    const MSA_GLOBAL double *__restrict__ E = gbl((const double *)C.E + (size_t)tile * C.n_elem * T);
    MSA_GLOBAL real *__restrict__ val = gbl((real *)C.val + (size_t)tile * C.tile_vals);
    const MSA_GLOBAL int *__restrict__ gmap = gbl(C.gmap);
    const MSA_GLOBAL int *__restrict__ cidx = gbl(C.cidx);
    const MSA_GLOBAL double *__restrict__ kc = gbl((const double *)C.kc);
    const MSA_GLOBAL double *__restrict__ mc = gbl((const double *)C.mc);
    const double cK = C.dyn ? C.cK : 1.0, cM = C.dyn ? C.cM : 0.0;
    // uniform structured grids (one element matrix for every element, up
    // to rounding): element 0's Khat (and Mhat) from LDS instead of the
    // per-element cache
    __shared__ double kl[640];
    const bool uni = C.mf;
    if (uni) {
        for (int q = threadIdx.x; q < 576; q += blockDim.x) kl[q] = kc[q];
        if (cM != 0.0)
            for (int q = threadIdx.x; q < 64; q += blockDim.x) kl[576 + q] = mc[q];
        __syncthreads();
    }
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * R + rr;
        if (br >= C.nb) continue;
        const int base = C.slice_ptr[slice], W = C.slice_ptr[slice + 1] - base;
        const uint8_t fr = C.block_free[br];
        for (int j = 0; j < W; j++) {
            const int slot = base + j;
            const int col = cidx[(size_t)slot * R + rr];
            double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (col >= 0) {
                // the slot's contribution codes load together (a code < 0
                // ends the list: the rest are skipped, not waited on)
                int codes[8];
#pragma unroll
                for (int gi = 0; gi < 8; gi++) codes[gi] = gmap[((size_t)slot * 8 + gi) * R + rr];
#pragma unroll 2
                for (int gi = 0; gi < 8; gi++) {
                    const int code = codes[gi];
                    if (code < 0) continue;
                    const int e = code >> 6, ab = code & 63;
                    const double Es = E[(size_t)e * T + s] * cK;
                    const MSA_GLOBAL double *k = kc + (size_t)e * 576 + ab * 9;
#pragma unroll
                    for (int q = 0; q < 9; q++) acc[q] += Es * k[q];
                    if (cM != 0.0) {
                        const double m = cM * mc[(size_t)e * 64 + ab];
                        acc[0] += m;
                        acc[4] += m;
                        acc[8] += m;
                    }
                }
                const uint8_t fc = C.block_free[col];
#pragma unroll
                for (int q = 0; q < 9; q++)
                    if (!((fr >> (q / 3)) & 1) || !((fc >> (q % 3)) & 1)) acc[q] = 0.0;
            }
#pragma unroll
            for (int q = 0; q < 9; q++) val[val_at<real, T>(slot, q, rr, s)] = (real)acc[q];
        }
    }
}

// ------------------------------------------------------------ PCG kernels

// One-sample tiles: the gather assembly of G consecutive samples in one
// pass, so the gather lists and the cached Khat entries of a stored block
// are read once per G samples instead of once per sample.  Each sample's
// sum runs in the same (ascending element) order as k_assemble_gather, so
// the values are bitwise the same.  Items: (class, tile group, chunk).
template <typename real, int G>
__global__ __launch_bounds__(256) void k_assemble_gather_g(const ClassDev *__restrict__ cls,
                                                           const int4 *__restrict__ items,
                                                           int nitems)
{
    constexpr int R = 64;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile0 = it.y * G, chunk = it.z;
    const int nt = min(G, C.ntiles - tile0);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane;
    const MSA_GLOBAL double *__restrict__ E = gbl((const double *)C.E + (size_t)tile0 * C.n_elem);
    MSA_GLOBAL real *__restrict__ val = gbl((real *)C.val + (size_t)tile0 * C.tile_vals);
    const MSA_GLOBAL int *__restrict__ gmap = gbl(C.gmap);
    const MSA_GLOBAL int *__restrict__ cidx = gbl(C.cidx);
    const MSA_GLOBAL double *__restrict__ kc = gbl((const double *)C.kc);
    const MSA_GLOBAL double *__restrict__ mc = gbl((const double *)C.mc);
    const double cK = C.dyn ? C.cK : 1.0, cM = C.dyn ? C.cM : 0.0;
    // uniform structured grids (one element matrix for every element, up
    // to rounding): element 0's Khat (and Mhat) from LDS instead of the
    // per-element cache
    __shared__ double kl[640];
    const bool uni = C.mf;
    if (uni) {
        for (int q = threadIdx.x; q < 576; q += blockDim.x) kl[q] = kc[q];
        if (cM != 0.0)
            for (int q = threadIdx.x; q < 64; q += blockDim.x) kl[576 + q] = mc[q];
        __syncthreads();
    }
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    // upper-triangle classes store slots 13..26 only (load_row_block reads
    // a lower slot from the upper one of its column's row)
    const int j0 = C.sym ? 13 : 0;
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * R + rr;
        if (br >= C.nb) continue;
        const int base = C.slice_ptr[slice], W = C.slice_ptr[slice + 1] - base;
        const uint8_t fr = C.block_free[br];
        for (int j = j0; j < W; j++) {
            const int slot = base + j;
            const int col = cidx[(size_t)slot * R + rr];
            double acc[G][9];
#pragma unroll
            for (int g = 0; g < G; g++)
#pragma unroll
                for (int q = 0; q < 9; q++) acc[g][q] = 0.0;
            if (col >= 0) {
                // the slot's contribution codes load together (a code < 0
                // ends the list: the rest are skipped, not waited on)
                int codes[8];
#pragma unroll
                for (int gi = 0; gi < 8; gi++) codes[gi] = gmap[((size_t)slot * 8 + gi) * R + rr];
#pragma unroll 2
                for (int gi = 0; gi < 8; gi++) {
                    const int code = codes[gi];
                    if (code < 0) continue;
                    const int e = code >> 6, ab = code & 63;
                    double k[9], m = 0.0;
                    if (uni) {
#pragma unroll
                        for (int q = 0; q < 9; q++) k[q] = kl[ab * 9 + q];
                        if (cM != 0.0) m = cM * kl[576 + ab];
                    } else {
                        const MSA_GLOBAL double *kp = kc + (size_t)e * 576 + ab * 9;
#pragma unroll
                        for (int q = 0; q < 9; q++) k[q] = kp[q];
                        if (cM != 0.0) m = cM * mc[(size_t)e * 64 + ab];
                    }
#pragma unroll
                    for (int g = 0; g < G; g++) {
                        if (g >= nt) break;
                        const double Es = E[(size_t)g * C.n_elem + e] * cK;
#pragma unroll
                        for (int q = 0; q < 9; q++) acc[g][q] += Es * k[q];
                        if (cM != 0.0) {
                            acc[g][0] += m;
                            acc[g][4] += m;
                            acc[g][8] += m;
                        }
                    }
                }
                const uint8_t fc = C.block_free[col];
#pragma unroll
                for (int g = 0; g < G; g++)
#pragma unroll
                    for (int q = 0; q < 9; q++)
                        if (!((fr >> (q / 3)) & 1) || !((fc >> (q % 3)) & 1)) acc[g][q] = 0.0;
            }
#pragma unroll
            for (int g = 0; g < G; g++) {
                if (g >= nt) break;
                real m[9];
#pragma unroll
                for (int q = 0; q < 9; q++) m[q] = (real)acc[g][q];
                store_block<real>(val + (size_t)g * C.tile_vals, slot, rr, m);
                if constexpr (sizeof(real) == 4) {
                    if (C.val64)  // coarsest level of a MIXED batch: the FP64 sums too
                        store_block<double>(gbl(C.val64) + (size_t)(tile0 + g) * C.tile_vals, slot, rr,
                                            acc[g]);
                }
            }
        }
    }
}

template <typename real, int T>
__global__ __launch_bounds__(256) void k_dinv(const ClassDev *__restrict__ cls,
                                              const int4 *__restrict__ items, int nitems)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    if (tile * T + s >= C.nsamp) return;
    const real *val = (const real *)C.val + (size_t)tile * C.tile_vals;
    real *dinv = (real *)C.dinv + (size_t)tile * C.tile_vec;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * R + rr;
        if (br >= C.nb) continue;
        const int slot = C.slice_ptr[slice] + C.diag_pos[br];
        const uint8_t fr = C.block_free[br];
        real w[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const real v = val[val_at<real, T>(slot, 4 * d, rr, s)];
            w[d] = (((fr >> d) & 1) && v != (real)0) ? (real)1 / v : (real)0;
        }
        if constexpr (T == 1) {
            st3(gbl(dinv), br, w[0], w[1], w[2]);  // one 12- or 24-byte store
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++) dinv[((size_t)br * 3 + d) * T + s] = w[d];
        }
    }
}

// Start of a PCG solve with x0 = 0 (SolverPCGSkyline.Initialize +
// SolverPCG.Solve prologue, SolverPCGSkyline.cs:75-100, SolverPCG.cs:112-114):
// r = b, z = D^-1 r, p = 0, |b|.  mode 0: b = f (FP64 solve);
// mode 1: b = r64 / |r64| (inner solve of the refinement).
template <typename real, int T>
__global__ __launch_bounds__(256) void k_pcg_init(const ClassDev *__restrict__ cls,
                                                  const int4 *__restrict__ items, int nitems,
                                                  int mode, int *active_count)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const bool act = sample < C.nsamp && (mode == 0 || C.st[sample].ir_active);
    const double sc = (act && mode) ? 1.0 / C.st[sample].scale : 1.0;
    const size_t vb = (size_t)tile * C.tile_vec;
    real *x = (real *)C.x + vb, *r = (real *)C.r + vb, *z = (real *)C.z + vb;
    real *p0 = (real *)C.p0 + vb, *p1 = (real *)C.p1 + vb;
    const real *dinv = (const real *)C.dinv + vb;
    const double *r64 = C.r64 ? C.r64 + vb : nullptr;
    double acc[2] = {0, 0};
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    if (act) {
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            if (br >= C.nb) continue;
            if constexpr (T == 1) {
                // each vector's three components move as one access; the
                // same arithmetic and summation order as below
                real bq[3];
                if (mode) {
                    const vec3<double> t = ld3(gbl(r64), br);
                    bq[0] = (real)(t.x * sc);
                    bq[1] = (real)(t.y * sc);
                    bq[2] = (real)(t.z * sc);
                } else {
#pragma unroll
                    for (int d = 0; d < 3; d++) bq[d] = (real)fval(C, vb, br, d, T, s);
                }
                const vec3<real> dv = ld3(gbl(dinv), br);
                const real z0 = dv.x * bq[0], z1 = dv.y * bq[1], z2 = dv.z * bq[2];
                st3(gbl(r), br, bq[0], bq[1], bq[2]);
                st3(gbl(x), br, (real)0, (real)0, (real)0);
                st3(gbl(z), br, z0, z1, z2);
                st3(gbl(p0), br, (real)0, (real)0, (real)0);
                st3(gbl(p1), br, (real)0, (real)0, (real)0);
                acc[0] += (double)(bq[0] * z0);
                acc[1] += (double)(bq[0] * bq[0]);
                acc[0] += (double)(bq[1] * z1);
                acc[1] += (double)(bq[1] * bq[1]);
                acc[0] += (double)(bq[2] * z2);
                acc[1] += (double)(bq[2] * bq[2]);
                continue;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const size_t o = ((size_t)br * 3 + d) * T + s;
                const real b = mode ? (real)(r64[o] * sc) : (real)fval(C, vb, br, d, T, s);
                const real zv = dinv[o] * b;
                r[o] = b;
                x[o] = 0;
                z[o] = zv;
                p0[o] = 0;
                p1[o] = 0;
                acc[0] += (double)(b * zv);
                acc[1] += (double)(b * b);
            }
        }
    }
    block_partials<T, 2>(acc, C.partial + ((size_t)tile * C.nchunks + chunk) * 3 * T);
}

// Whole-wave DPP shifts of one value: lane l takes lane l - 1's (shr) or
// l + 1's (shl) src; lane 0 (shr) / 63 (shl) keeps old.  Doubles move as
// two 32-bit halves.
template <int CTRL, typename real>
__device__ __forceinline__ real dpp_wave(real old, real src)
{
    if constexpr (sizeof(real) == 4) {
        return __builtin_bit_cast(real, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                    __builtin_bit_cast(int, src),
                                                                    CTRL, 0xf, 0xf, false));
    } else {
        const unsigned long long o = __builtin_bit_cast(unsigned long long, old);
        const unsigned long long v = __builtin_bit_cast(unsigned long long, src);
        const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)v, CTRL, 0xf,
                                                   0xf, false);
        const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(v >> 32),
                                                   CTRL, 0xf, 0xf, false);
        return __builtin_bit_cast(real, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    }
}
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;

// Rows [s0, s1) slices of a stencil class (one-sample tiles) in the
// symmetric product q = K p, p = z + beta p_old: every row's 14 upper
// blocks (slots 13..26) times p at br + off, and its 13 lower blocks as the
// transposes of the upper blocks stored in the rows br - off (slot s of row
// j holds K(j, j + off_s)).  The values cross HBM about once: a row's
// upper blocks are loaded through the caches and re-read transposed by the
// rows above them, which the same and the neighbouring workgroups run at
// about the same time, so most re-reads hit L2.  (SkylineMatrix2D.Multiply, SkylineMatrix2D.cs:253-269,
// likewise applies each stored entry of the upper profile to both
// triangles, by scattering into the column's rows; here every row gathers,
// with no write conflicts.)
//
// The 27 positions a row needs, forward and backward, are its 3x3x3
// stencil: stencil position n = (dz+1)*9 + (dy+1)*3 + dx+1 at row
// br + off(n); n > 13 is upper slot n of this row, n < 13 the transpose of
// upper slot 26 - n of row br + off(n), n = 13 the diagonal.  Lane l holds
// row br = 64 * slice + l, so each lane loads z and p_old only at its own x
// for the 9 (dy, dz) columns and takes x - 1 / x + 1 from lanes l - 1 / l + 1
// by DPP shifts (lanes 0 / 63 load theirs): 18 gathers per row instead of
// 54.  Column m and its mirror 8 - m go together, m from 4 down to 0.
// Summation: each (+off, -off) pair first, the pairs in ascending distance
// k (n = 13 +- k), the diagonal last.  A row's terms nearly cancel
// (rigid-body modes); this order keeps the FP32 rounding of the sum at or
// below that of the column-ordered sum, where a diagonal-first order raised
// it by about half and cost multigrid-PCG iterations.  The whole wave runs
// every slice (DPP needs every lane): rows past nb compute on the last row
// and store nothing.  Missing neighbours hold zero blocks; positions
// outside [0, nb) are clamped (their blocks are zero) and a backward row
// above the first is masked.  Returns the lane's p.q partial.
template <typename real>
__device__ __forceinline__ real pcg_a_sym_rows(const ClassDev &C, int s0, int s1, int wave,
                                               int rr, bool act, real beta,
                                               const MSA_GLOBAL real *__restrict__ z,
                                               const MSA_GLOBAL real *__restrict__ po,
                                               MSA_GLOBAL real *__restrict__ pn,
                                               MSA_GLOBAL real *__restrict__ q,
                                               const MSA_GLOBAL real *__restrict__ val)
{
    using V3 = vec3<real>;
    const int sy = C.sym_sy, sz = C.sym_sz, nb1 = C.nb - 1;
    real pq = 0;
    if (!act) return pq;  // one sample per tile: uniform over the wave
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * 64 + rr;
        const int brc = min(br, nb1);
        // p at stencil column m = (dy+1) + 3 (dz+1): c[0] at x - 1, c[1] at x,
        // c[2] at x + 1
        auto column = [&](int m, real (&c)[3][3]) {
            const int mo = sy * (m % 3 - 1) + sz * (m / 3 - 1);
            const V3 zc = ld3(z, min(max(brc + mo, 0), nb1));
            const V3 pc = ld3(po, min(max(brc + mo, 0), nb1));
            const real own[3] = {zc.x + beta * pc.x, zc.y + beta * pc.y, zc.z + beta * pc.z};
            real edge[3] = {0, 0, 0};
            if (rr == 0 || rr == 63) {
                const int r1 = min(max(brc + mo + (rr == 0 ? -1 : 1), 0), nb1);
                const V3 ze = ld3(z, r1), pe = ld3(po, r1);
                edge[0] = ze.x + beta * pe.x;
                edge[1] = ze.y + beta * pe.y;
                edge[2] = ze.z + beta * pe.z;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                c[1][d] = own[d];
                c[0][d] = dpp_wave<kDppWaveShr1>(edge[d], own[d]);
                c[2][d] = dpp_wave<kDppWaveShl1>(edge[d], own[d]);
            }
        };
        // the block of stencil position n times v (transposed for n < 13)
        auto term = [&](int n, const real (&v)[3], real (&r)[3]) {
            const int su = n >= 13 ? n : 26 - n;
            const int off = (n % 3 - 1) + sy * ((n / 3) % 3 - 1) + sz * (n / 9 - 1);
            const int j = n >= 13 ? brc : brc + off;
            const int jc = max(j, 0);
            real m[9];
            load_block<real, 1, false>(val, (size_t)(jc >> 6) * 27 + su, jc & 63, 0, m);
            const real w0 = j >= 0 ? v[0] : (real)0, w1 = j >= 0 ? v[1] : (real)0,
                       w2 = j >= 0 ? v[2] : (real)0;
            if (n < 13) {
                r[0] = m[0] * w0 + m[3] * w1 + m[6] * w2;
                r[1] = m[1] * w0 + m[4] * w1 + m[7] * w2;
                r[2] = m[2] * w0 + m[5] * w1 + m[8] * w2;
            } else {
                r[0] = m[0] * w0 + m[1] * w1 + m[2] * w2;
                r[1] = m[3] * w0 + m[4] * w1 + m[5] * w2;
                r[2] = m[6] * w0 + m[7] * w1 + m[8] * w2;
            }
        };
        real a[3], d[3], own[3];
        {
            real c[3][3], t12[3], t14[3];
            column(4, c);
            term(14, c[2], t14);
            term(12, c[0], t12);
            term(13, c[1], d);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                a[k] = 0;
                a[k] += t14[k] + t12[k];
                own[k] = c[1][k];
            }
        }
#pragma unroll 1
        for (int m = 3; m >= 0; m--) {
            real cl[3][3], ch[3][3], rl[3][3], rh[3][3];
            column(m, cl);
            column(8 - m, ch);
#pragma unroll
            for (int dx = 2; dx >= 0; dx--) {
                term(3 * m + dx, cl[dx], rl[dx]);
                term(26 - 3 * m - dx, ch[2 - dx], rh[dx]);
            }
#pragma unroll
            for (int dx = 2; dx >= 0; dx--)
#pragma unroll
                for (int k = 0; k < 3; k++) a[k] += rh[dx][k] + rl[dx][k];
        }
#pragma unroll
        for (int k = 0; k < 3; k++) a[k] += d[k];
        if (br < C.nb) {
            st3(pn, br, own[0], own[1], own[2]);
            st3(q, br, a[0], a[1], a[2]);
            pq += own[0] * a[0] + own[1] * a[1] + own[2] * a[2];
        }
    }
    return pq;
}

// The matrix-free product of a uniform-grid class (defined with the
// smoothing kernels below).
template <bool BOX, bool PG = false>
__device__ __forceinline__ void mf_product(const ClassDev &C, const MSA_GLOBAL float *v,
                                           const double *Es, int br, float &y0, float &y1,
                                           float &y2, float *pown = nullptr);

// Search direction + product + curvature (SolverPCGSimpleSearchVectorCalculator
// .cs:18-41, MCPCGDirect.cs:48-51): p = z + beta p_old on the fly for every
// gathered column, q = K p, own rows of p written to the other buffer,
// pq = p.q per sample; the last workgroup of each tile sets alpha = rz/pq.
// Stencil classes of one-sample tiles take the symmetric product above.
// U: slots per group of a row (the loads of a group issue together; the
// schedule the autotuner picks); every U sums the slots in the same order.
template <typename real, int T, int U = 3>
__device__ __forceinline__ void pcg_a_item(const ClassDev *__restrict__ cls, const int4 it,
                                           int pbuf)
{
    constexpr int R = 64 / T;
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const bool act = sample < C.nsamp && C.st[sample].active;
    const real beta = act ? (real)C.st[sample].beta : (real)0;
    const size_t vb = (size_t)tile * C.tile_vec;
    using V3 = vec3<real>;
    const MSA_GLOBAL real *__restrict__ z = gbl((const real *)C.z + vb);
    const MSA_GLOBAL real *__restrict__ po = gbl((const real *)(pbuf ? C.p1 : C.p0) + vb);
    MSA_GLOBAL real *__restrict__ pn = gbl((real *)(pbuf ? C.p0 : C.p1) + vb);
    MSA_GLOBAL real *__restrict__ q = gbl((real *)C.q + vb);
    const MSA_GLOBAL real *__restrict__ val = gbl((const real *)C.val + (size_t)tile * C.tile_vals);
    const MSA_GLOBAL int *__restrict__ cidx = gbl(C.cidx);
    const MSA_GLOBAL int *__restrict__ sptr = gbl(C.slice_ptr);
    real pq = 0;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    // One slot: gathered p = z + beta p_old of block column col, then the 3x3
    // block times it.  Padding slots (cidx < 0) hold zero values; they read
    // block 0 instead of branching, so the loads of several slots overlap.
    auto gather = [&](int col, real &v0, real &v1, real &v2) {
        if constexpr (T == 1) {
            const V3 zc = ld3(z, col);
            const V3 pc = ld3(po, col);
            v0 = zc.x + beta * pc.x;
            v1 = zc.y + beta * pc.y;
            v2 = zc.z + beta * pc.z;
        } else {
            const size_t xo = (size_t)col * 3 * T + s;
            v0 = z[xo] + beta * po[xo];
            v1 = z[xo + T] + beta * po[xo + T];
            v2 = z[xo + 2 * T] + beta * po[xo + 2 * T];
        }
    };
    // stiffness values are streamed once per product: load them
    // non-temporally so L2 keeps the reused vector gathers
    auto block = [&](int slot, real (&m)[9]) { load_block<real, T, true>(val, slot, rr, s, m); };
    bool sym = false;
    if constexpr (T == 1) sym = C.sym != 0;
    if (sym) {
        if (__ballot(act)) pq = pcg_a_sym_rows<real>(C, s0, s1, wave, rr, act, beta, z, po, pn, q, val);
    } else if (__ballot(act)) {
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            const int base = sptr[slice];
            const int W = sptr[slice + 1] - base;
            if (act && br < C.nb) {
                real a0 = 0, a1 = 0, a2 = 0;
                int j = 0;
                // the next group's column indices load with this group's
                // values and gathers (one memory round trip per group)
                // (indices past the row's last slot re-read that slot: no
                // branch, so nothing waits on them early)
                const int last = max(W - 1, 0);
                int nxt[U];
#pragma unroll
                for (int u = 0; u < U; u++) nxt[u] = max(cidx[(size_t)(base + min(u, last)) * R + rr], 0);
                for (; j + U <= W; j += U) {
                    int col[U];
#pragma unroll
                    for (int u = 0; u < U; u++) col[u] = nxt[u];
                    real v[U][3], m[U][9];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        gather(col[u], v[u][0], v[u][1], v[u][2]);
                        block(base + j + u, m[u]);
                    }
#pragma unroll
                    for (int u = 0; u < U; u++)
                        nxt[u] = max(cidx[(size_t)(base + min(j + U + u, last)) * R + rr], 0);
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        a0 += m[u][0] * v[u][0] + m[u][1] * v[u][1] + m[u][2] * v[u][2];
                        a1 += m[u][3] * v[u][0] + m[u][4] * v[u][1] + m[u][5] * v[u][2];
                        a2 += m[u][6] * v[u][0] + m[u][7] * v[u][1] + m[u][8] * v[u][2];
                    }
                }
                for (; j < W; j++) {
                    const int col = max(cidx[(size_t)(base + j) * R + rr], 0);
                    real v0, v1, v2, m[9];
                    gather(col, v0, v1, v2);
                    block(base + j, m);
                    a0 += m[0] * v0 + m[1] * v1 + m[2] * v2;
                    a1 += m[3] * v0 + m[4] * v1 + m[5] * v2;
                    a2 += m[6] * v0 + m[7] * v1 + m[8] * v2;
                }
                real n0, n1, n2;
                gather(br, n0, n1, n2);
                if constexpr (T == 1) {
                    st3(pn, br, n0, n1, n2);
                    st3(q, br, a0, a1, a2);
                } else {
                    const size_t o = (size_t)br * 3 * T + s;
                    pn[o] = n0;
                    pn[o + T] = n1;
                    pn[o + 2 * T] = n2;
                    q[o] = a0;
                    q[o + T] = a1;
                    q[o + 2 * T] = a2;
                }
                pq += n0 * a0 + n1 * a1 + n2 * a2;
            }
        }
    }
    double acc[1] = {(double)pq};
    block_partials<T, 1>(acc, C.partial + ((size_t)tile * C.nchunks + chunk) * 3 * T);
}

template <typename real, int T, int U = 3>
__global__ __launch_bounds__(256) void k_pcg_a(const ClassDev *__restrict__ cls,
                                               const int4 *__restrict__ items, int nitems,
                                               int pbuf, int *active_count)
{
    const int item = xcd_item(nitems);
    if (item < 0) return;
    pcg_a_item<real, T, U>(cls, items[item], pbuf);
}

// The same step for box-grid classes whose product is matrix-free
// (mf_pcg: MIXED, one-sample tiles, multigrid set up), in two passes over
// an item's rows.  k_pcg_p writes p = z + beta p_old; k_pcg_a_mf then
// computes q = K p from the cell moduli and the unit element matrix (the
// operator of the finest smoothing products) with one gather per node,
// and the p.q partials.  No stored value is read.  An item covers
// max(1, it.w) consecutive chunks.
__global__ __launch_bounds__(256) void k_pcg_p(const ClassDev *__restrict__ cls,
                                               const int4 *__restrict__ items, int nitems,
                                               int pbuf)
{
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, nch = max(1, it.w);
    if (tile >= C.nsamp || !C.st[tile].active) return;  // one sample per tile
    const float beta = (float)C.st[tile].beta;
    const size_t vb = (size_t)tile * C.tile_vec;
    const MSA_GLOBAL float *__restrict__ z = gbl((const float *)C.z + vb);
    const MSA_GLOBAL float *__restrict__ po = gbl((const float *)(pbuf ? C.p1 : C.p0) + vb);
    MSA_GLOBAL float *__restrict__ pn = gbl((float *)(pbuf ? C.p0 : C.p1) + vb);
    const int r0 = it.z * C.chunk * 64, r1 = min((it.z + nch) * C.chunk * 64, C.nb);
    for (int br = r0 + (int)threadIdx.x; br < r1; br += blockDim.x) {
        const vec3<float> zc = ld3(z, br), pc = ld3(po, br);
        st3(pn, br, zc.x + beta * pc.x, zc.y + beta * pc.y, zc.z + beta * pc.z);
    }
}

__global__ __launch_bounds__(256) void k_pcg_a_mf(const ClassDev *__restrict__ cls,
                                                  const int4 *__restrict__ items, int nitems,
                                                  int pbuf)
{
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z, nch = max(1, it.w);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool act = tile < C.nsamp && C.st[tile].active;  // one sample per tile
    const size_t vb = (size_t)tile * C.tile_vec;
    const MSA_GLOBAL float *__restrict__ p = gbl((const float *)(pbuf ? C.p0 : C.p1) + vb);
    MSA_GLOBAL float *__restrict__ q = gbl((float *)C.q + vb);
    const double *Es = C.E + (size_t)tile * C.n_elem;
    for (int ck = chunk; ck < chunk + nch; ck++) {
        float pq = 0.f;
        const int s0 = ck * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
        if (act) {
            for (int slice = s0 + wave; slice < s1; slice += kWaves) {
                const int br = slice * 64 + lane;
                if (br >= C.nb) continue;
                float y[3], own[3];
                mf_product<true, true>(C, p, Es, br, y[0], y[1], y[2], own);
                // a constrained component's row is zero in K (as in the
                // stored product); classes of fully free blocks have none
                if (!__builtin_amdgcn_readfirstlane(C.mf_allfree)) {
                    const uint8_t fr = C.block_free[br];
                    y[0] = (fr & 1) ? y[0] : 0.f;
                    y[1] = (fr & 2) ? y[1] : 0.f;
                    y[2] = (fr & 4) ? y[2] : 0.f;
                }
                st3(q, br, y[0], y[1], y[2]);
                pq += own[0] * y[0] + own[1] * y[1] + own[2] * y[2];
            }
        }
        double acc[1] = {(double)pq};
        block_partials<1, 1>(acc, C.partial + ((size_t)tile * C.nchunks + ck) * 3);
        // the next chunk's partials reuse the same LDS rows
        if (ck + 1 < chunk + nch) __syncthreads();
    }
}

// Update + preconditioner + stopping test (SolverPCG.cs:122-136):
// x += alpha p, r -= alpha q, z = D^-1 r, rz = r.z, rr = r.r; the last
// workgroup of each tile checks |r|/|b| < tol and forms beta.
template <typename real, int T>
__device__ __forceinline__ void pcg_b_item(const ClassDev *__restrict__ cls, const int4 it,
                                           int pbuf)
{
    constexpr int R = 64 / T;
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const bool act = sample < C.nsamp && C.st[sample].active;
    const real alpha = act ? (real)C.st[sample].alpha : (real)0;
    const size_t vb = (size_t)tile * C.tile_vec;
    MSA_GLOBAL real *__restrict__ x = gbl((real *)C.x + vb);
    MSA_GLOBAL real *__restrict__ r = gbl((real *)C.r + vb);
    MSA_GLOBAL real *__restrict__ z = gbl((real *)C.z + vb);
    const MSA_GLOBAL real *__restrict__ pn = gbl((const real *)(pbuf ? C.p0 : C.p1) + vb);
    const MSA_GLOBAL real *__restrict__ q = gbl((const real *)C.q + vb);
    const MSA_GLOBAL real *__restrict__ dinv = gbl((const real *)C.dinv + vb);
    real rz = 0, rr2 = 0;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    if (act) {
#pragma unroll 2
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            if (br >= C.nb) continue;
            if constexpr (T == 1) {
                using V = vec3<real>;
                const V xv = ld3(x, br), pv = ld3(pn, br), rv = ld3(r, br);
                const V qv = ld3(q, br), dv = ld3(dinv, br);
                const V xn{xv.x + alpha * pv.x, xv.y + alpha * pv.y, xv.z + alpha * pv.z};
                const V rn{rv.x - alpha * qv.x, rv.y - alpha * qv.y, rv.z - alpha * qv.z};
                const V zn{dv.x * rn.x, dv.y * rn.y, dv.z * rn.z};
                st3(x, br, xn.x, xn.y, xn.z);
                st3(r, br, rn.x, rn.y, rn.z);
                st3(z, br, zn.x, zn.y, zn.z);
                rz += rn.x * zn.x + rn.y * zn.y + rn.z * zn.z;
                rr2 += rn.x * rn.x + rn.y * rn.y + rn.z * rn.z;
            } else {
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const size_t o = ((size_t)br * 3 + d) * T + s;
                    const real xv = x[o] + alpha * pn[o];
                    const real rv = r[o] - alpha * q[o];
                    const real zv = dinv[o] * rv;
                    x[o] = xv;
                    r[o] = rv;
                    z[o] = zv;
                    rz += rv * zv;
                    rr2 += rv * rv;
                }
            }
        }
    }
    double acc[2] = {(double)rz, (double)rr2};
    block_partials<T, 2>(acc, C.partial + ((size_t)tile * C.nchunks + chunk) * 3 * T);
}

template <typename real, int T>
__global__ __launch_bounds__(256) void k_pcg_b(const ClassDev *__restrict__ cls,
                                               const int4 *__restrict__ items, int nitems,
                                               int pbuf, double tol, int max_it,
                                               int *active_count)
{
    const int item = xcd_item(nitems);
    if (item < 0) return;
    pcg_b_item<real, T>(cls, items[item], pbuf);
}

// y = K x for every lane of one tile (tests and the single-product entry).
template <typename real, int T>
__global__ __launch_bounds__(256) void k_spmv(const ClassDev *__restrict__ cls,
                                              const int4 *__restrict__ items, int nitems,
                                              const double *__restrict__ xin,
                                              double *__restrict__ yout)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const real *val = (const real *)C.val + (size_t)tile * C.tile_vals;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * R + rr;
        if (br >= C.nb) continue;
        const int base = C.slice_ptr[slice], W = C.slice_ptr[slice + 1] - base;
        double a0 = 0, a1 = 0, a2 = 0;
        for (int j = 0; j < W; j++) {
            const int slot = base + j;
            const int col = C.cidx[(size_t)slot * R + rr];
            if (col < 0) continue;
            const size_t xo = (size_t)col * 3 * T + s;
            const double v0 = xin[xo], v1 = xin[xo + T], v2 = xin[xo + 2 * T];
            real m[9];
            load_row_block<real, T>(C, gbl(val), base, j, col, rr, s, m);
            a0 += (double)m[0] * v0 + (double)m[1] * v1 + (double)m[2] * v2;
            a1 += (double)m[3] * v0 + (double)m[4] * v1 + (double)m[5] * v2;
            a2 += (double)m[6] * v0 + (double)m[7] * v1 + (double)m[8] * v2;
        }
        const size_t o = (size_t)br * 3 * T + s;
        yout[o] = a0;
        yout[o + T] = a1;
        yout[o + 2 * T] = a2;
    }
}

// ------------------------------------------------------------ refinement

// u64 = 0 (warm = 0; warm = 1 keeps the previous solution as the start),
// r64 = f, outer state (mixed precision only).
template <int T>
__global__ __launch_bounds__(256) void k_ir_begin(const ClassDev *__restrict__ cls,
                                                  const int4 *__restrict__ items, int nitems,
                                                  int *ir_count, int warm)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const size_t vb = (size_t)tile * C.tile_vec;
    if (sample < C.nsamp) {
        const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            if (br >= C.nb) continue;
            if constexpr (T == 1) {
                // the loads first, then one 24-byte store per vector
                const double f0 = fval(C, vb, br, 0, T, s), f1 = fval(C, vb, br, 1, T, s),
                             f2 = fval(C, vb, br, 2, T, s);
                if (!warm) st3(gbl(C.u64 + vb), br, 0.0, 0.0, 0.0);
                st3(gbl(C.r64 + vb), br, f0, f1, f2);
                continue;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const size_t o = vb + ((size_t)br * 3 + d) * T + s;
                if (!warm) C.u64[o] = 0;
                C.r64[o] = fval(C, vb, br, d, T, s);
            }
        }
    }
    if (chunk == 0 && threadIdx.x < T) {
        const int smp = tile * T + threadIdx.x;
        if (smp < C.nsamp) {
            SampleState &S = C.st[smp];
            const double fn = fnorm_of(C, S);
            S.r64norm = fn;
            S.scale = fn;
            S.outer = 0;
            S.iters = 0;
            S.total_iters = 0;
            S.relres = 1.0;
            S.active = 0;
            if (!isfinite(fn)) {
                S.status = MSA_NAN_INPUT;
                S.ir_active = 0;
            } else if (fn > 1e-25) {
                S.status = MSA_NOT_CONVERGED;
                S.ir_active = 1;
                atomicAdd(ir_count, 1);
            } else {
                S.status = MSA_OK;
                S.relres = 0;
                S.ir_active = 0;
            }
        }
    }
}

// u64 += |r64| * x (FP64 update of the refinement), r64 = f.
template <int T>
__global__ __launch_bounds__(256) void k_ir_update(const ClassDev *__restrict__ cls,
                                                   const int4 *__restrict__ items, int nitems)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    if (sample >= C.nsamp || !C.st[sample].ir_active) return;
    const double sc = C.st[sample].scale;
    const size_t vb = (size_t)tile * C.tile_vec;
    const float *x = (const float *)C.x + vb;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * R + rr;
        if (br >= C.nb) continue;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const size_t o = ((size_t)br * 3 + d) * T + s;
            C.u64[vb + o] += sc * (double)x[o];
            C.r64[vb + o] = fval(C, vb, br, d, T, s);
        }
    }
}

// Element operator on one sample, row i = 3a + p of the element:
// y[row] += cKe * (Khat xk)_i + cMe * sum_b Mhat[a][b] xm[b, p]; rows and
// columns of constrained DOFs are skipped (vectors hold free rows only).
__device__ __forceinline__ void apply_rows(const ClassDev &C, const int *en, const double *Kh,
                                           const double *Mh, double cKe, double cMe,
                                           const double *xk, const double *xm, double *y,
                                           size_t vb, int T, int s, int i)
{
    const int ia = i / 3, p = i - 3 * (i / 3);
    const int nd = en[ia];
    const int b = C.node_block[nd];
    if (b < 0 || !((C.node_free[nd] >> p) & 1)) return;
    double yk = 0, ym = 0;
#pragma unroll
    for (int jn = 0; jn < 8; jn++) {
        const int nj = en[jn];
        const int bj = C.node_block[nj];
        const uint8_t fr = C.node_free[nj];
        if (cKe != 0) {
            const double *blk = Kh + (ia * 8 + jn) * 9 + p * 3;
#pragma unroll
            for (int d = 0; d < 3; d++)
                if (bj >= 0 && ((fr >> d) & 1))
                    yk += blk[d] * xk[vb + ((size_t)bj * 3 + d) * T + s];
        }
        if (cMe != 0 && bj >= 0 && ((fr >> p) & 1))
            ym += Mh[ia * 8 + jn] * xm[vb + ((size_t)bj * 3 + p) * T + s];
    }
    y[vb + ((size_t)b * 3 + p) * T + s] += cKe * yk + cMe * ym;
}

// FP64 residual r64 -= K u64, matrix-free: one wavefront per element of the
// current colour recomputes Khat in FP64 (LDS) and applies E_s Khat to the
// element's displacements of every active sample (dynamics mode: the
// effective matrix cK E_s Khat + cM Mhat).
// need_due = 1: only samples with a pending replacement (reliable updates);
// one-sample tiles take them from the class's due list, two per pass (lanes
// 0-23 and 32-55).
template <int T>
__global__ __launch_bounds__(64) void k_residual64(const ClassDev *__restrict__ cls,
                                                   ColourLaunch cl, double nu, int need_due)
{
    constexpr int R = 64 / T;
    __shared__ double Kh[576];
    __shared__ double g[192];
    __shared__ double w[8];
    __shared__ double Nl[64], Mh[64];
    const int item = blockIdx.x;
    const int c = colour_class(cl, item);
    if (item >= cl.prefix[cl.ncls]) return;
    const ClassDev &C = cls[c];
    const int e = C.colour_elems[C.colour_ptr[cl.colour] + (item - cl.prefix[c])];
    const int lane = threadIdx.x & 63;
    const int ndue = need_due ? *C.due : 0;
    if (need_due && ndue == 0) return;
    // block y handles due-list entries 2y, 2y+1, 2y+2Y, ... (one-sample tiles)
    // or tiles y, y+Y, ... (Y = gridDim.y)
    const int Y = gridDim.y, y = blockIdx.y;
    bool any = false;
    if (T == 1 && need_due) {
        for (int j = lane;; j += 64) {
            const int k = 2 * y + (j & 1) + 2 * Y * (j >> 1);
            if (k >= ndue) break;
            const SampleState &S = C.st[C.due[1 + k]];
            any |= S.ir_active && S.ru_due;
        }
    } else {
        for (int j = lane;; j += 64) {
            const int tile = y + (j / T) * Y;
            if (tile >= C.ntiles) break;
            const int smp = tile * T + j % T;
            if (smp < C.nsamp) any |= C.st[smp].ir_active && (!need_due || C.st[smp].ru_due);
        }
    }
    if (!__ballot(any)) return;
    const int *en = C.conn + (size_t)e * 8;
    const int na = en[lane & 7];
    const double xa[3] = {C.xyz[3 * na], C.xyz[3 * na + 1], C.xyz[3 * na + 2]};
    stage_khat(C, e, xa, nu, Kh, g, w, Nl, C.dyn ? Mh : nullptr);
    const double cK = C.dyn ? C.cK : 1.0, cMe = C.dyn ? -C.cM : 0.0;
    if constexpr (T == 1) {
        if (need_due) {
            // Two samples per pass (lanes 0-23 and 32-55, lane = element row
            // i = 3a + p).  Each lane keeps its row of Khat (and Mhat) in
            // registers for all samples; the element's 24 displacements of a
            // sample are loaded once and shared by wave shuffles.
            __shared__ int colb[24];
            if (lane < 24) {
                const int nd = en[lane / 3], d = lane % 3;
                const int bj = C.node_block[nd];
                colb[lane] = (bj >= 0 && ((C.node_free[nd] >> d) & 1)) ? 3 * bj + d : -1;
            }
            __syncthreads();
            const int half = lane >> 5, i = lane & 31;
            const int ii = i < 24 ? i : 0;
            const int row = i < 24 ? colb[i] : -1;
            const int ia = ii / 3, p = ii - 3 * (ii / 3);
            double krow[24], mrow[8];
#pragma unroll
            for (int jn = 0; jn < 8; jn++) {
#pragma unroll
                for (int d = 0; d < 3; d++) krow[3 * jn + d] = Kh[(ia * 8 + jn) * 9 + p * 3 + d];
                mrow[jn] = cMe != 0 ? Mh[ia * 8 + jn] : 0.0;
            }
            const int base = half << 5;
            for (int k0 = 2 * y; k0 < ndue; k0 += 2 * Y) {
                const int k = k0 + half;
                int smp = 0;
                bool ok = k < ndue;
                if (ok) {
                    smp = C.due[1 + k];
                    ok = C.st[smp].ir_active && C.st[smp].ru_due;
                }
                double uv = 0;
                if (ok && row >= 0) uv = C.u64[(size_t)smp * C.tile_vec + row];
                double yk = 0, ym = 0;
#pragma unroll
                for (int j = 0; j < 24; j++) yk += krow[j] * __shfl(uv, base + j, 64);
                if (cMe != 0) {
#pragma unroll
                    for (int jn = 0; jn < 8; jn++) ym += mrow[jn] * __shfl(uv, base + 3 * jn + p, 64);
                }
                if (ok && row >= 0) {
                    const double Es = C.E[(size_t)smp * C.n_elem + e];
                    C.r64[(size_t)smp * C.tile_vec + row] += -(Es * cK) * yk + cMe * ym;
                }
            }
            return;
        }
    }
    const int s = lane % T, sub = lane / T;
    for (int tile = y; tile < C.ntiles; tile += Y) {
        const int smp = tile * T + s;
        if (smp >= C.nsamp || !C.st[smp].ir_active || (need_due && !C.st[smp].ru_due)) continue;
        const size_t vb = (size_t)tile * C.tile_vec;
        const double Es = C.E[((size_t)tile * C.n_elem + e) * T + s];
        for (int i = sub; i < 24; i += R)
            apply_rows(C, en, Kh, Mh, -(Es * cK), cMe, C.u64, C.u64, C.r64, vb, T, s, i);
    }
}

// |r64| per sample and the refinement stopping test |f - K u|/|f| < tol.
template <int T>
__global__ __launch_bounds__(256) void k_ir_norm(const ClassDev *__restrict__ cls,
                                                 const int4 *__restrict__ items, int nitems,
                                                 double tol, int max_outer, int *ir_count)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const bool act = sample < C.nsamp && C.st[sample].ir_active;
    const size_t vb = (size_t)tile * C.tile_vec;
    double acc[1] = {0};
    if (act) {
        const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            if (br >= C.nb) continue;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double v = C.r64[vb + ((size_t)br * 3 + d) * T + s];
                acc[0] += v * v;
            }
        }
    }
    block_partials<T, 1>(acc, C.partial + ((size_t)tile * C.nchunks + chunk) * 3 * T);
}

// ------------------------------------------------------------ reliable updates
// Mixed precision without restarts: the FP32 PCG keeps its search direction
// while every `period` iterations the accumulated FP32 correction is folded
// into the FP64 solution, the residual is replaced by the FP64 f - K u, the
// stopping test runs on that FP64 residual and the FP32 recurrence restarts
// from it (search direction rescaled to the new residual scale).

// u64 += scale * x32, x32 = 0, r64 = f for every sample still refining.
template <int T>
__global__ __launch_bounds__(256) void k_ru_fold(const ClassDev *__restrict__ cls,
                                                 const int4 *__restrict__ items, int nitems)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    if (*C.due == 0 || sample >= C.nsamp || !C.st[sample].ir_active || !C.st[sample].ru_due)
        return;
    const double sc = C.st[sample].scale;
    const size_t vb = (size_t)tile * C.tile_vec;
    float *x = (float *)C.x + vb;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * R + rr;
        if (br >= C.nb) continue;
        if constexpr (T == 1) {
            // each vector's three components move as one access
            MSA_GLOBAL double *u = gbl(C.u64 + vb), *r = gbl(C.r64 + vb);
            MSA_GLOBAL float *xg = gbl(x);
            const vec3<double> uv = ld3(u, br);
            const vec3<float> xv = ld3(xg, br);
            st3(u, br, uv.x + sc * (double)xv.x, uv.y + sc * (double)xv.y, uv.z + sc * (double)xv.z);
            st3(xg, br, 0.f, 0.f, 0.f);
            st3(r, br, fval(C, vb, br, 0, T, s), fval(C, vb, br, 1, T, s), fval(C, vb, br, 2, T, s));
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const size_t o = ((size_t)br * 3 + d) * T + s;
                C.u64[vb + o] += sc * (double)x[o];
                x[o] = 0.f;
                C.r64[vb + o] = fval(C, vb, br, d, T, s);
            }
        }
    }
}

// |r64| and the FP64 stopping test; sets the rescale ratio for the restart.
template <int T>
__global__ __launch_bounds__(256) void k_ru_norm(const ClassDev *__restrict__ cls,
                                                 const int4 *__restrict__ items, int nitems,
                                                 double tol, int max_it, int *ir_count,
                                                 int *active_count)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const bool act = sample < C.nsamp && C.st[sample].ir_active && C.st[sample].ru_due;
    const size_t vb = (size_t)tile * C.tile_vec;
    double acc[1] = {0};
    if (act) {
        const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            if (br >= C.nb) continue;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double v = C.r64[vb + ((size_t)br * 3 + d) * T + s];
                acc[0] += v * v;
            }
        }
    }
    block_partials<T, 1>(acc, C.partial + ((size_t)tile * C.nchunks + chunk) * 3 * T);
}

// Restart of the FP32 recurrence from the FP64 residual: r = r64/|r64|,
// z = D^-1 r, p_old *= old_scale/new_scale, rz = r.z.
template <int T>
__global__ __launch_bounds__(256) void k_ru_restart(const ClassDev *__restrict__ cls,
                                                    const int4 *__restrict__ items, int nitems)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const bool act = sample < C.nsamp && C.st[sample].ir_active && C.st[sample].ru_due;
    const size_t vb = (size_t)tile * C.tile_vec;
    double acc[1] = {0};
    if (act) {
        const double inv = 1.0 / C.st[sample].scale;
        const float ratio = (float)C.st[sample].pq;
        float *r = (float *)C.r + vb, *z = (float *)C.z + vb, *p0 = (float *)C.p0 + vb;
        const float *dinv = (const float *)C.dinv + vb;
        const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            if (br >= C.nb) continue;
            if constexpr (T == 1) {
                // each vector's three components move as one access; the
                // same arithmetic and summation order as the general form
                const vec3<double> rd = ld3(gbl((const double *)C.r64 + vb), br);
                const vec3<float> dv = ld3(gbl(dinv), br), pv = ld3(gbl((const float *)p0), br);
                const float r0 = (float)(rd.x * inv), r1 = (float)(rd.y * inv), r2 = (float)(rd.z * inv);
                const float z0 = dv.x * r0, z1 = dv.y * r1, z2 = dv.z * r2;
                st3(gbl(r), br, r0, r1, r2);
                st3(gbl(z), br, z0, z1, z2);
                st3(gbl(p0), br, pv.x * ratio, pv.y * ratio, pv.z * ratio);
                acc[0] += (double)(r0 * z0);
                acc[0] += (double)(r1 * z1);
                acc[0] += (double)(r2 * z2);
            } else {
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const size_t o = ((size_t)br * 3 + d) * T + s;
                    const float rv = (float)(C.r64[vb + o] * inv);
                    const float zv = dinv[o] * rv;
                    r[o] = rv;
                    z[o] = zv;
                    p0[o] *= ratio;
                    acc[0] += (double)(rv * zv);
                }
            }
        }
    }
    block_partials<T, 1>(acc, C.partial + ((size_t)tile * C.nchunks + chunk) * 3 * T);
}

// ------------------------------------------------------------ Newmark dynamics
// Linear implicit Newmark (NewmarkDynamicAnalyzer.cs) for every sample of a
// batch.  The stored matrix is the effective stiffness cK K + cM M; each time
// step forms its right-hand side on the device (k_dyn_prep, k_dyn_rhs),
// solves starting from the previous displacement and updates velocity and
// acceleration (k_dyn_update).

__global__ void k_set_int(int *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_step_advance(int *step)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *step += 1;
}

// Mass-acceleration load of a unit base acceleration in direction dir
// (Hexa8.CalculateAccelerationForces, Hexa8.cs:512-530: M_e times a vector
// holding 1 on every DOF of that direction, constrained nodes included;
// added into the free rows, Model.cs:266-280): bacc[row(a, dir)] +=
// rho sum_b Mhat[a][b].  One wavefront per element of the colour.
__global__ __launch_bounds__(64) void k_dyn_bacc(const ClassDev *__restrict__ cls,
                                                 ColourLaunch cl, double nu, int dir)
{
    __shared__ double Kh[576];
    __shared__ double g[192];
    __shared__ double w[8];
    __shared__ double Nl[64], Mh[64];
    const int item = blockIdx.x;
    const int c = colour_class(cl, item);
    if (item >= cl.prefix[cl.ncls]) return;
    const ClassDev &C = cls[c];
    const int e = C.colour_elems[C.colour_ptr[cl.colour] + (item - cl.prefix[c])];
    const int lane = threadIdx.x & 63;
    const int *en = C.conn + (size_t)e * 8;
    const int na = en[lane & 7];
    const double xa[3] = {C.xyz[3 * na], C.xyz[3 * na + 1], C.xyz[3 * na + 2]};
    stage_khat(C, e, xa, nu, Kh, g, w, Nl, Mh);
    if (lane < 8) {
        const int b = C.node_block[en[lane]];
        if (b >= 0 && ((C.node_free[en[lane]] >> dir) & 1)) {
            double f = 0;
            for (int jn = 0; jn < 8; jn++) f += C.rho * Mh[lane * 8 + jn];
            C.bacc[(size_t)b * 3 + dir] += f;
        }
    }
}

// Start of a time history: u = v = a = 0 (from rest), summed iterations 0.
template <int T>
__global__ __launch_bounds__(256) void k_dyn_reset(const ClassDev *__restrict__ cls,
                                                   const int4 *__restrict__ items, int nitems,
                                                   int mixed)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const size_t vb = (size_t)tile * C.tile_vec;
    double *u = mixed ? C.u64 : (double *)C.x;
    if (sample < C.nsamp) {
        const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            if (br >= C.nb) continue;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const size_t o = vb + ((size_t)br * 3 + d) * T + s;
                u[o] = 0;
                C.vel[o] = 0;
                C.acc[o] = 0;
            }
        }
    }
    if (chunk == 0 && threadIdx.x < T) {
        const int smp = tile * T + threadIdx.x;
        if (smp < C.nsamp) C.st[smp].cum_iters = 0;
    }
}

// Right-hand side operands of a step (NewmarkDynamicAnalyzer.cs:189-206):
// uold = u, wm = (a0 u + a2 v + a3 a) + alphaR (a1 u + a4 v + a5 a),
// wk = betaR (a1 u + a4 v + a5 a), fs = f_static + acc(step) bacc; k_dyn_rhs
// then adds rho M wm + K wk (the reference's M uu + C uc with C = alphaR M +
// betaR K).  start == 2 replaces u by the predictor u + dt v + dt^2/2 a as
// the first iterate of the step's solve.
template <int T>
__global__ __launch_bounds__(256) void k_dyn_prep(const ClassDev *__restrict__ cls,
                                                  const int4 *__restrict__ items, int nitems,
                                                  int mixed, DynCoef dc)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    if (sample >= C.nsamp) return;
    const size_t vb = (size_t)tile * C.tile_vec;
    double *u = mixed ? C.u64 : (double *)C.x;
    const double ag = dc.acc ? dc.acc[*dc.step] : 0.0;
    const double h2 = 0.5 * dc.dt * dc.dt;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * R + rr;
        if (br >= C.nb) continue;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const size_t o = vb + ((size_t)br * 3 + d) * T + s;
            const double uu = u[o], vv = C.vel[o], aa = C.acc[o];
            C.uold[o] = uu;
            const double w1 = dc.a0 * uu + dc.a2 * vv + dc.a3 * aa;
            const double w2 = dc.a1 * uu + dc.a4 * vv + dc.a5 * aa;
            C.wm[o] = w1 + dc.alphaR * w2;
            C.wk[o] = dc.betaR * w2;
            C.fs[o] = C.f[br * 3 + d] + C.bacc[br * 3 + d] * ag;
            // start of the step's solve (the solver keeps u as its first iterate)
            if (dc.start == 2) u[o] = uu + dc.dt * vv + h2 * aa;
        }
    }
}

// fs += rho Mhat wm + E_s Khat wk per element of the colour, every sample
// (wk is zero, and skipped, without stiffness-proportional damping).
template <int T>
__global__ __launch_bounds__(64) void k_dyn_rhs(const ClassDev *__restrict__ cls,
                                                ColourLaunch cl, double nu, int stiff)
{
    constexpr int R = 64 / T;
    __shared__ double Kh[576];
    __shared__ double g[192];
    __shared__ double w[8];
    __shared__ double Nl[64], Mh[64];
    const int item = blockIdx.x;
    const int c = colour_class(cl, item);
    if (item >= cl.prefix[cl.ncls]) return;
    const ClassDev &C = cls[c];
    const int e = C.colour_elems[C.colour_ptr[cl.colour] + (item - cl.prefix[c])];
    const int lane = threadIdx.x & 63;
    const int *en = C.conn + (size_t)e * 8;
    const int na = en[lane & 7];
    const double xa[3] = {C.xyz[3 * na], C.xyz[3 * na + 1], C.xyz[3 * na + 2]};
    stage_khat(C, e, xa, nu, Kh, g, w, Nl, Mh);
    const int Y = gridDim.y, y = blockIdx.y;
    if constexpr (T == 1) {
        const int half = lane >> 5, i = lane & 31;
        if (i < 24)
            for (int smp = 2 * y + half; smp < C.nsamp; smp += 2 * Y) {
                const double Es = stiff ? C.E[(size_t)smp * C.n_elem + e] : 0.0;
                apply_rows(C, en, Kh, Mh, Es, C.rho, C.wk, C.wm, C.fs,
                           (size_t)smp * C.tile_vec, 1, 0, i);
            }
    } else {
        const int s = lane % T, sub = lane / T;
        for (int tile = y; tile < C.ntiles; tile += Y) {
            const int smp = tile * T + s;
            if (smp >= C.nsamp) continue;
            const size_t vb = (size_t)tile * C.tile_vec;
            const double Es = stiff ? C.E[((size_t)tile * C.n_elem + e) * T + s] : 0.0;
            for (int i = sub; i < 24; i += R)
                apply_rows(C, en, Kh, Mh, Es, C.rho, C.wk, C.wm, C.fs, vb, T, s, i);
        }
    }
}

// fs = f for every sample (a static solve of a dynamics-mode batch).
template <int T>
__global__ __launch_bounds__(256) void k_fs_static(const ClassDev *__restrict__ cls,
                                                   const int4 *__restrict__ items, int nitems)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    if (tile * T + s >= C.nsamp) return;
    const size_t vb = (size_t)tile * C.tile_vec;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * R + rr;
        if (br >= C.nb) continue;
#pragma unroll
        for (int d = 0; d < 3; d++) C.fs[vb + ((size_t)br * 3 + d) * T + s] = C.f[br * 3 + d];
    }
}

// |fs|^2 partials per sample (reduced by OP_FNORM).
template <int T>
__global__ __launch_bounds__(256) void k_fs_norm(const ClassDev *__restrict__ cls,
                                                 const int4 *__restrict__ items, int nitems)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const size_t vb = (size_t)tile * C.tile_vec;
    double acc[1] = {0};
    if (sample < C.nsamp) {
        const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            if (br >= C.nb) continue;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double v = C.fs[vb + ((size_t)br * 3 + d) * T + s];
                acc[0] += v * v;
            }
        }
    }
    block_partials<T, 1>(acc, C.partial + ((size_t)tile * C.nchunks + chunk) * 3 * T);
}

// Warm-started FP64 PCG start: x kept (the previous step's displacement),
// r = fs - K x, z = D^-1 r, p = 0; partials r.z and r.r (OP_INIT2).
template <int T>
__global__ __launch_bounds__(256) void k_pcg_init_warm(const ClassDev *__restrict__ cls,
                                                       const int4 *__restrict__ items, int nitems)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const size_t vb = (size_t)tile * C.tile_vec;
    const double *val = (const double *)C.val + (size_t)tile * C.tile_vals;
    const double *x = (const double *)C.x + vb;
    double *r = (double *)C.r + vb, *z = (double *)C.z + vb;
    double *p0 = (double *)C.p0 + vb, *p1 = (double *)C.p1 + vb;
    const double *dinv = (const double *)C.dinv + vb;
    double acc[2] = {0, 0};
    if (sample < C.nsamp) {
        const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            if (br >= C.nb) continue;
            const int base = C.slice_ptr[slice], W = C.slice_ptr[slice + 1] - base;
            double kx[3] = {0, 0, 0};
            for (int j = 0; j < W; j++) {
                const int slot = base + j;
                const int col = C.cidx[(size_t)slot * R + rr];
                if (col < 0) continue;
                const size_t xo = (size_t)col * 3 * T + s;
                const double v0 = x[xo], v1 = x[xo + T], v2 = x[xo + 2 * T];
                double m[9];
                load_row_block<double, T>(C, gbl(val), base, j, col, rr, s, m);
                kx[0] += m[0] * v0 + m[1] * v1 + m[2] * v2;
                kx[1] += m[3] * v0 + m[4] * v1 + m[5] * v2;
                kx[2] += m[6] * v0 + m[7] * v1 + m[8] * v2;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const size_t o = ((size_t)br * 3 + d) * T + s;
                const double rv = fval(C, vb, br, d, T, s) - kx[d];
                const double zv = dinv[o] * rv;
                r[o] = rv;
                z[o] = zv;
                p0[o] = 0;
                p1[o] = 0;
                acc[0] += rv * zv;
                acc[1] += rv * rv;
            }
        }
    }
    block_partials<T, 2>(acc, C.partial + ((size_t)tile * C.nchunks + chunk) * 3 * T);
}

// Velocity and acceleration update (NewmarkDynamicAnalyzer.cs:301-315):
// a_new = a0 (u - u_old) - a2 v - a3 a, v += a6 a + a7 a_new; the probe row
// of every sample is recorded; iterations of the step are summed.
template <int T>
__global__ __launch_bounds__(256) void k_dyn_update(const ClassDev *__restrict__ cls,
                                                    const int4 *__restrict__ items, int nitems,
                                                    int mixed, DynCoef dc)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const size_t vb = (size_t)tile * C.tile_vec;
    const double *u = mixed ? C.u64 : (const double *)C.x;
    const int step = *dc.step;
    if (sample < C.nsamp) {
        const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            if (br >= C.nb) continue;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const size_t o = vb + ((size_t)br * 3 + d) * T + s;
                const double un = u[o], ao = C.acc[o];
                const double an = dc.a0 * (un - C.uold[o]) - dc.a2 * C.vel[o] - dc.a3 * ao;
                C.vel[o] = C.vel[o] + (dc.a6 * ao + dc.a7 * an);
                C.acc[o] = an;
                if (C.hist && C.probe == br * 3 + d && step < C.hist_len)
                    C.hist[(size_t)sample * C.hist_len + step] = un;
            }
        }
    }
    if (chunk == 0 && threadIdx.x < T) {
        const int smp = tile * T + threadIdx.x;
        if (smp < C.nsamp) {
            SampleState &S = C.st[smp];
            S.cum_iters += mixed ? S.total_iters : S.iters;
        }
    }
}

// ------------------------------------------------------------ direct solver
// Banded Cholesky K = L L^T of each sample's FP64 stiffness (the GPU
// counterpart of SkylineMatrix2D.Factorize / Solve, SkylineMatrix2D.cs:
// 63-212, used by SolverSkyline.cs:90-205), blocked by kBandNB = 64 columns:
// per block the diagonal factor (k_chol_diag), the panel solve below it
// within the band (k_chol_panel) and the trailing update of the band window
// (k_chol_update).  Lower band storage band[j*bW + (i - j)], bW >= bandwidth
// + 64, so every panel is a full rectangle (entries outside the true band
// stay zero).  grid.y = sample of the class.

// Lower band entries of every sample's assembled K (FP64 batch values).
template <typename real, int T>
__global__ __launch_bounds__(256) void k_band_scatter(const ClassDev *__restrict__ cls,
                                                      const int4 *__restrict__ items, int nitems)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    if (!C.band) return;
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    if (sample >= C.nsamp) return;
    if (C.band_ss == 0 && sample != 0) return;  // one factor shared by all samples: sample 0's
    const real *val = (const real *)C.val + (size_t)tile * C.tile_vals;
    double *AB = C.band + (size_t)sample * C.band_ss;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * R + rr;
        if (br >= C.nb) continue;
        const int base = C.slice_ptr[slice], W = C.slice_ptr[slice + 1] - base;
        for (int j = 0; j < W; j++) {
            const int slot = base + j;
            const int col = C.cidx[(size_t)slot * R + rr];
            if (col < 0) continue;
            for (int p = 0; p < 3; p++) {
                const int i = C.binv[3 * br + p];
                if (i < 0) continue;
                for (int q = 0; q < 3; q++) {
                    const int jj = C.binv[3 * col + q];
                    if (jj < 0 || jj > i) continue;
                    AB[(size_t)jj * C.bW + (i - jj)] =
                        C.val64 ? C.val64[(size_t)tile * C.tile_vals + row_entry<double, T>(C, base, j, col, rr, s, p, q)]
                                : (double)val[row_entry<real, T>(C, base, j, col, rr, s, p, q)];
                }
            }
        }
    }
}

// In-place Cholesky of the nk x nk diagonal block at column k0; one
// workgroup of 64 threads (thread = block row) per sample.
//
// Zero-energy modes (C.zv set; SkylineMatrix2D.Factorize's kFix branch,
// SkylineMatrix2D.cs:110-130): a pivot d = K_nn - sum_k L_nk^2 with |d| <
// zem_tol is the LDL^T pivot of a singular direction.  The factor then takes
// L_nn = 1 and clears row n (saved first: w_n = 1, w_k = -L_nk; in LDL^T terms
// -L'_nk sqrt(D_k), which is what the reference's back-substitution through
// D^1/2 L'^T amounts to) and column n below the diagonal (the reference
// zeroes K's column n before the later rows use it, so their L entries in
// column n come out zero).  Row n's entries left of this block were used
// only for updates of row and column n, both discarded, so clearing them
// afterwards gives the reference's factor.  The panel kernel clears column
// n of its rows from zmask.
__global__ __launch_bounds__(64) void k_chol_diag(const ClassDev *__restrict__ cp, int k0, int nk)
{
    __shared__ double A[kBandNB][kBandNB + 1];
    __shared__ int fail, zem, zk;
    __shared__ unsigned long long zbits;
    const ClassDev &C = *cp;
    const int smp = blockIdx.y;
    const int t = threadIdx.x;
    double *AB = C.band + (size_t)smp * C.band_ss;
    const bool zon = C.zv != nullptr;
    if (t == 0) {
        fail = 0;
        zbits = 0ull;
    }
    if (t < nk)
        for (int c = 0; c <= t; c++) A[t][c] = AB[(size_t)(k0 + c) * C.bW + (t - c)];
    __syncthreads();
    for (int c = 0; c < nk; c++) {
        if (t == c) {
            double d = A[c][c];
            zem = 0;
            if (zon && isfinite(d) && fabs(d) < C.zem_tol) {
                zem = 1;
                zk = C.zcount[smp]++;
                if (zk >= C.zem_max) fail = 1;  // more modes than kept
                zbits |= 1ull << c;
                d = 1.0;
            } else if (!(d > 0.0) || !isfinite(d)) {
                fail = 1;
                d = 1.0;
            }
            A[c][c] = sqrt(d);
        }
        __syncthreads();
        if (zem) {
            const int n = k0 + c;
            double *w = zk < C.zem_max ? C.zv + ((size_t)smp * C.zem_max + zk) * C.bn : nullptr;
            if (w && t == 0) {
                w[n] = 1.0;
                C.zcol[(size_t)smp * C.zem_max + zk] = n;
            }
            // row n left of this block: columns k in (n - bW, k0)
            for (int k = max(0, n - C.bW + 1) + t; k < k0; k += 64) {
                double *e = AB + (size_t)k * C.bW + (n - k);
                if (w) w[k] = -*e;
                *e = 0.0;
            }
            if (t < c) {
                if (w) w[k0 + t] = -A[c][t];
                A[c][t] = 0.0;
            }
            if (t > c && t < nk) A[t][c] = 0.0;
        } else if (t > c && t < nk) {
            A[t][c] /= A[c][c];
        }
        __syncthreads();
        if (t > c && t < nk)
            for (int c2 = c + 1; c2 <= t; c2++) A[t][c2] -= A[t][c] * A[c2][c];
        __syncthreads();
    }
    if (t < nk)
        for (int c = 0; c <= t; c++) AB[(size_t)(k0 + c) * C.bW + (t - c)] = A[t][c];
    if (t == 0 && fail) C.bfail[smp] = 1;
    if (t == 0 && zon) C.zmask[(size_t)smp * ((C.bn + kBandNB - 1) / kBandNB) + k0 / kBandNB] = zbits;
}

// Panel rows i in [r0, r0 + m): L21 = A21 L11^-T, 64 rows per workgroup of
// 256 threads: the 64 x 64 panel tile and the packed L11 in LDS, then a
// column sweep (scale column c, update the columns right of it).  Columns
// c >= nk of a last, partial block are left zero.
__global__ __launch_bounds__(256) void k_chol_panel(const ClassDev *__restrict__ cp, int k0, int nk,
                                                    int r0, int m)
{
    __shared__ double P[kBandNB][kBandNB + 1];
    __shared__ double Lp[kBandNB * (kBandNB + 1) / 2];  // row c, column c2 <= c at c(c+1)/2 + c2
    const ClassDev &C = *cp;
    const int smp = blockIdx.y;
    const int tid = threadIdx.x;
    double *AB = C.band + (size_t)smp * C.band_ss;
    for (int idx = tid; idx < kBandNB * kBandNB; idx += 256) {
        const int c = idx >> 6, c2 = idx & 63;
        if (c2 <= c)
            Lp[c * (c + 1) / 2 + c2] = c < nk ? AB[(size_t)(k0 + c2) * C.bW + (c - c2)]
                                              : (c == c2 ? 1.0 : 0.0);
    }
    const int row0 = r0 + blockIdx.x * kBandNB;
    for (int idx = tid; idx < kBandNB * kBandNB; idx += 256) {
        const int c = idx >> 6, r = idx & 63;
        const int i = row0 + r;
        P[r][c] = (c < nk && i < r0 + m) ? AB[(size_t)(k0 + c) * C.bW + (i - k0 - c)] : 0.0;
    }
    const unsigned long long zbits =
        C.zv ? C.zmask[(size_t)smp * ((C.bn + kBandNB - 1) / kBandNB) + k0 / kBandNB] : 0ull;
    __syncthreads();
    for (int c = 0; c < nk; c++) {
        // a zero-energy column of L is zero below the diagonal
        if (tid < kBandNB) P[tid][c] = ((zbits >> c) & 1ull) ? 0.0 : P[tid][c] / Lp[c * (c + 1) / 2 + c];
        __syncthreads();
        const int rest = kBandNB * (nk - c - 1);
        for (int idx = tid; idx < rest; idx += 256) {
            const int r = idx & 63, c2 = c + 1 + (idx >> 6);
            P[r][c2] -= P[r][c] * Lp[c2 * (c2 + 1) / 2 + c];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < kBandNB * kBandNB; idx += 256) {
        const int c = idx >> 6, r = idx & 63;
        const int i = row0 + r;
        if (c < nk && i < r0 + m) AB[(size_t)(k0 + c) * C.bW + (i - k0 - c)] = P[r][c];
    }
}

// Trailing update of the band window: A[i][j] -= sum_c L[i][k0+c] L[j][k0+c]
// for r0 <= j <= i < r0 + m; one 256-thread workgroup per 64 x 64 tile (I, J),
// J <= I, each thread a 4 x 4 sub-block; the panel columns are staged in LDS
// in two halves of 32.
__global__ __launch_bounds__(256) void k_chol_update(const ClassDev *__restrict__ cp, int k0, int nk,
                                                     int r0, int m)
{
    __shared__ double Li[kBandNB][33], Lj[kBandNB][33];
    const ClassDev &C = *cp;
    const int smp = blockIdx.y;
    double *AB = C.band + (size_t)smp * C.band_ss;
    // tile index -> (I, J) with J <= I
    const int tix = blockIdx.x;
    int I = (int)((sqrt(8.0 * tix + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= tix) I++;
    while (I * (I + 1) / 2 > tix) I--;
    const int J = tix - I * (I + 1) / 2;
    const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
    double acc[4][4] = {};
    for (int h = 0; h < 2; h++) {
        for (int e = threadIdx.x; e < kBandNB * 32; e += 256) {
            const int r = e >> 5, c = (e & 31) + 32 * h;
            const int ir = r0 + I * kBandNB + r, jr = r0 + J * kBandNB + r;
            Li[r][c & 31] = (c < nk && ir < r0 + m) ? AB[(size_t)(k0 + c) * C.bW + (ir - k0 - c)] : 0.0;
            Lj[r][c & 31] = (c < nk && jr < r0 + m) ? AB[(size_t)(k0 + c) * C.bW + (jr - k0 - c)] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int c = 0; c < 32; c++) {
            double a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                a[u] = Li[ti * 4 + u][c];
                b[u] = Lj[tj * 4 + u][c];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[u][v] += a[u] * b[v];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = r0 + I * kBandNB + ti * 4 + u, j = r0 + J * kBandNB + tj * 4 + v;
            if (i < r0 + m && j < r0 + m && i >= j) AB[(size_t)j * C.bW + (i - j)] -= acc[u][v];
        }
}

// by[s][i] = right-hand side of band row i (class load vector, or the
// per-sample one of the dynamics mode).
template <int T>
__global__ void k_band_rhs(const ClassDev *__restrict__ cp, int from_fs)
{
    const ClassDev &C = *cp;
    const int smp = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.bn) return;
    const int row = C.bperm[i];
    const int tile = smp / T, s = smp % T;
    C.by[(size_t)smp * C.bn + i] =
        from_fs ? C.fs[(size_t)tile * C.tile_vec + (size_t)row * T + s] : C.f[row];
}

// Forward step of L y = b for block k0: y_blk = L11^-1 y_blk (workgroup 0
// of the x grid, 64 threads), then rows [r0, r0 + m) minus L21 y_blk.  Two
// launches per block: `phase` 0 is the diagonal solve, 1 the update.
__global__ __launch_bounds__(64) void k_band_fwd(const ClassDev *__restrict__ cp, int k0, int nk,
                                                 int r0, int m, int phase, int mode)
{
    __shared__ double yb[kBandNB];
    const ClassDev &C = *cp;
    const int smp = blockIdx.y;
    if (mode >= 0 && !mg_act(C.mg_st[smp], mode)) return;  // coarsest multigrid level
    const int t = threadIdx.x;
    const double *AB = C.band + (size_t)smp * C.band_ss;
    double *y = C.by + (size_t)smp * C.bn;
    if (phase == 0) {
        if (t < nk) yb[t] = y[k0 + t];
        __syncthreads();
        for (int c = 0; c < nk; c++) {
            if (t == c) yb[c] /= AB[(size_t)(k0 + c) * C.bW];
            __syncthreads();
            if (t > c && t < nk) yb[t] -= AB[(size_t)(k0 + c) * C.bW + (t - c)] * yb[c];
            __syncthreads();
        }
        if (t < nk) y[k0 + t] = yb[t];
        return;
    }
    if (t < nk) yb[t] = y[k0 + t];
    __syncthreads();
    const int i = r0 + blockIdx.x * 64 + t;
    if (i >= r0 + m) return;
    double acc = 0.0;
    for (int c = 0; c < nk; c++) acc += AB[(size_t)(k0 + c) * C.bW + (i - k0 - c)] * yb[c];
    y[i] -= acc;
}

// Backward step of L^T x = y for block k0: phase 0 writes partial sums
// bpart[g][c] = sum over rows i of chunk g of L[i][k0+c] x[i]; phase 1
// (one workgroup) sums them in chunk order and back-substitutes L11^T.
__global__ __launch_bounds__(64) void k_band_bwd(const ClassDev *__restrict__ cp, int k0, int nk,
                                                 int r0, int m, int phase, int mode)
{
    __shared__ double xb[kBandNB];
    const ClassDev &C = *cp;
    const int smp = blockIdx.y;
    if (mode >= 0 && !mg_act(C.mg_st[smp], mode)) return;  // coarsest multigrid level
    const int t = threadIdx.x;
    const double *AB = C.band + (size_t)smp * C.band_ss;
    double *x = C.by + (size_t)smp * C.bn;
    const int nchunk = (C.bW + kBandChunkRows - 1) / kBandChunkRows;
    double *part = C.bpart + (size_t)smp * nchunk * kBandNB;
    if (phase == 0) {
        const int g = blockIdx.x;
        const int i0 = r0 + g * kBandChunkRows, i1 = min(r0 + m, i0 + kBandChunkRows);
        if (t < nk) {
            double acc = 0.0;
            for (int i = i0; i < i1; i++) acc += AB[(size_t)(k0 + t) * C.bW + (i - k0 - t)] * x[i];
            part[g * kBandNB + t] = acc;
        }
        return;
    }
    const int ng = m > 0 ? (m + kBandChunkRows - 1) / kBandChunkRows : 0;
    if (t < nk) {
        double v = x[k0 + t];
        for (int g = 0; g < ng; g++) v -= part[g * kBandNB + t];
        xb[t] = v;
    }
    __syncthreads();
    for (int c = nk - 1; c >= 0; c--) {
        if (t == c) xb[c] /= AB[(size_t)(k0 + c) * C.bW];
        __syncthreads();
        if (t < c) xb[t] -= AB[(size_t)(k0 + t) * C.bW + (c - t)] * xb[c];
        __syncthreads();
    }
    if (t < nk) x[k0 + t] = xb[t];
}

// Both band sweeps (L y = b, then L^T x = y) of one sample in one
// workgroup, for bands of at most kBandOneMax rows: the same 64-column
// blocks and the same order of operations as k_band_fwd / k_band_bwd (so the
// same results), without a launch per block.  y lives in LDS; each block's
// diagonal triangle is staged in LDS (rows padded to 65 against bank
// conflicts) and solved by wave 0 with the right-hand side in registers.
__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, l);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

__global__ __launch_bounds__(256) void k_band_solve_one(const ClassDev *__restrict__ cp, int mode)
{
    __shared__ double yl[kBandOneMax];
    __shared__ double Ld[kBandNB][kBandNB + 1];
    __shared__ double part[kBandNB];
    const ClassDev &C = *cp;
    const int smp = blockIdx.x;
    if (mode >= 0 && !mg_act(C.mg_st[smp], mode)) return;  // coarsest multigrid level
    const int n = C.bn, W = C.bW;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const double *AB = C.band + (size_t)smp * C.band_ss;
    double *y = C.by + (size_t)smp * n;
    for (int i = t; i < n; i += 256) yl[i] = y[i];
    // Ld[c][j] = AB[(k0 + c) W + j] = L[k0 + c + j][k0 + c], j < nk - c
    auto stage = [&](int k0, int nk) {
        for (int e = t; e < kBandNB * kBandNB; e += 256) {
            const int c = e >> 6, j = e & 63;
            Ld[c][j] = (c < nk && j < nk - c) ? AB[(size_t)(k0 + c) * W + j] : 0.0;
        }
    };
    __syncthreads();
    for (int k0 = 0; k0 < n; k0 += kBandNB) {
        const int nk = min(kBandNB, n - k0), r0 = k0 + nk, m = min(n - r0, W - kBandNB);
        stage(k0, nk);
        __syncthreads();
        if (wave == 0) {
            double v = lane < nk ? yl[k0 + lane] : 0.0;
            for (int c = 0; c < nk; c++) {
                const double yc = readlane_f64(v, c) / Ld[c][0];
                if (lane == c) v = yc;
                if (lane > c && lane < nk) v -= Ld[c][lane - c] * yc;
            }
            if (lane < nk) yl[k0 + lane] = v;
        }
        __syncthreads();
        // (unrolled so that the loads of a group issue together; the sum
        // keeps its order)
        for (int i = r0 + t; i < r0 + m; i += 256) {
            double acc = 0.0;
#pragma unroll 16
            for (int c = 0; c < nk; c++) acc += AB[(size_t)(k0 + c) * W + (i - k0 - c)] * yl[k0 + c];
            yl[i] -= acc;
        }
        __syncthreads();
    }
    const int last = ((n - 1) / kBandNB) * kBandNB;
    for (int k0 = last; k0 >= 0; k0 -= kBandNB) {
        const int nk = min(kBandNB, n - k0), r0 = k0 + nk, m = min(n - r0, W - kBandNB);
        stage(k0, nk);
        // column partials of L21^T x in chunks of kBandChunkRows rows,
        // subtracted chunk by chunk (the order of k_band_bwd)
        if (wave == 1 && lane < nk) {
            double v = yl[k0 + lane];
            for (int i0 = r0; i0 < r0 + m; i0 += kBandChunkRows) {
                const int i1 = min(r0 + m, i0 + kBandChunkRows);
                double acc = 0.0;
#pragma unroll 16
                for (int i = i0; i < i1; i++)
                    acc += AB[(size_t)(k0 + lane) * W + (i - k0 - lane)] * yl[i];
                v -= acc;
            }
            part[lane] = v;
        }
        __syncthreads();
        if (wave == 0) {
            double v = lane < nk ? part[lane] : 0.0;
            for (int c = nk - 1; c >= 0; c--) {
                const double xc = readlane_f64(v, c) / Ld[c][0];
                if (lane == c) v = xc;
                if (lane < c) v -= Ld[lane][c - lane] * xc;
            }
            if (lane < nk) yl[k0 + lane] = v;
        }
        __syncthreads();
    }
    for (int i = t; i < n; i += 256) y[i] = yl[i];
}

// Solution back to the internal rows: x[row] = by[band(row)], constrained
// components 0 (FP64 batches: the PCG solution vector).
template <int T>
__global__ void k_band_out(const ClassDev *__restrict__ cp)
{
    const ClassDev &C = *cp;
    const int smp = blockIdx.y;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= 3 * C.nb) return;
    const int tile = smp / T, s = smp % T;
    const int i = C.binv[row];
    ((double *)C.x)[(size_t)tile * C.tile_vec + (size_t)row * T + s] =
        i >= 0 ? C.by[(size_t)smp * C.bn + i] : 0.0;
}

// ------------------------------------------------------------ KL field

// phi[e][i] = sigma sqrt(eigs_i) prod_d a * (index even ? sin : cos)
// (w (c_d - L_d/2)) at the element centroid (Lognormal...cs:134-160,
// Hexa8WithStochasticMaterial.cs:112-139).
__global__ void k_kl_basis(const ClassDev *__restrict__ cp, int n_elem, int M, int kmax,
                           const double *__restrict__ w1d, const double *__restrict__ a1d,
                           const int *__restrict__ evpPos, const double *__restrict__ eigs,
                           const double *__restrict__ domain, double sigma)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_elem) return;
    const ClassDev &C = *cp;
    double x8[24];
    for (int a = 0; a < 8; a++)
        for (int j = 0; j < 3; j++) x8[3 * a + j] = C.xyz[3 * C.conn[(size_t)e * 8 + a] + j];
    double X = 0, Y = 0, Z = 0, mnx = x8[0], mny = x8[1], mnz = x8[2];
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
    const double cxyz[3] = {mnx + X / 2, mny + Y / 2, mnz + Z / 2};
    for (int i = 0; i < M; i++) {
        double f = sigma * sqrt(eigs[i]);
        for (int d = 0; d < 3; d++) {
            const int index = evpPos[3 * i + d];
            const double wv = w1d[d * kmax + index - 1];
            const double av = a1d[d * kmax + index - 1];
            const double x = cxyz[d] - domain[d] / 2;
            f *= (index % 2 == 0) ? av * sin(wv * x) : av * cos(wv * x);
        }
        C.phi[(size_t)e * M + i] = f;
    }
}

// E[tile][e][s] = exp(c0 + sum_i xi_s,i phi_e,i) (Lognormal...cs:173-193).
template <int T>
__global__ void k_kl_eval(const ClassDev *__restrict__ cp, int ntiles, int n_elem, int M)
{
    // one thread: element e, lane s of kKlTiles consecutive tiles; the
    // element's row of phi is read once for all of them
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ngroups = (ntiles + kKlTiles - 1) / kKlTiles;
    const size_t total = (size_t)ngroups * n_elem * T;
    if (id >= total) return;
    const ClassDev &C = *cp;
    const int s = (int)(id % T);
    const size_t te = id / T;
    const int e = (int)(te % n_elem);
    const int t0 = (int)(te / n_elem) * kKlTiles;
    const int nt = min(kKlTiles, ntiles - t0);
    double res[kKlTiles];
    const double *xs[kKlTiles];
#pragma unroll
    for (int g = 0; g < kKlTiles; g++) {
        res[g] = C.c0;
        const int smp = min((t0 + g) * T + s, C.nsamp - 1);
        xs[g] = C.xi + (size_t)smp * M;
    }
    const double *ph = C.phi + (size_t)e * M;
    for (int i = 0; i < M; i++) {
        const double p = ph[i];
#pragma unroll
        for (int g = 0; g < kKlTiles; g++) res[g] += xs[g][i] * p;
    }
#pragma unroll
    for (int g = 0; g < kKlTiles; g++) {
        if (g >= nt) break;
        const int smp = (t0 + g) * T + s;
        C.E[((size_t)(t0 + g) * n_elem + e) * T + s] = smp < C.nsamp ? exp(res[g]) : 0.0;
    }
}


// ------------------------------------------------------------ reductions
// Every PCG-type kernel leaves per-sample partial sums, one per work item
// (chunk).  k_reduce (one 1024-thread workgroup per class tile, launched
// right after) sums them in a fixed order -- strided per thread, then a
// binary tree over the thread groups -- and applies the per-sample state
// update of `op`.  The kernel boundary makes the partials visible; no
// floating-point atomics, so results are reproducible run to run.

// A pending FP64 replacement is flagged per sample and the sample appended
// to the class's due list once (between list resets no flag goes 1 -> 0 ->
// 1, so the list never exceeds nsamp entries).  The list is reset by the
// reductions that clear every flag (OP_INIT*, OP_RURESTART); no sample is
// appended by those.
__device__ __forceinline__ void set_due(const ClassDev &C, SampleState &S, int smp)
{
    if (S.ru_due) return;
    S.ru_due = 1;
    C.due[1 + atomicAdd(C.due, 1)] = smp;
}

__device__ void apply_op(const ClassDev &C, int smp, const double *v, const ReduceArgs &a)
{
    SampleState &S = C.st[smp];
    switch (a.op) {
    case OP_INIT0:
    case OP_INIT1: {  // start of a PCG solve: rz = r.z, |b|
        const bool on = (a.op == OP_INIT0) || S.ir_active;
        S.iters = 0;
        S.beta = 0;
        S.ru_due = 0;
        if (!on) {
            S.active = 0;
            break;
        }
        S.rz = v[0];
        S.bnorm = sqrt(v[1]);
        S.relres = 1.0;
        if (!isfinite(v[0]) || !isfinite(v[1])) {
            S.status = MSA_NAN_INPUT;
            S.active = 0;
        } else if (S.bnorm > 1e-25) {
            S.status = MSA_NOT_CONVERGED;
            S.active = 1;
            atomicAdd(a.active_count, 1);
        } else {
            S.status = MSA_OK;
            S.relres = 0;
            S.active = 0;
        }
        break;
    }
    case OP_INIT2: {  // warm-started FP64 solve: |r0|/|f| against the step's |f|
        S.iters = 0;
        S.beta = 0;
        S.ru_due = 0;
        S.rz = v[0];
        S.bnorm = S.fnorm;
        const double err = S.fnorm > 1e-25 ? sqrt(v[1]) / S.fnorm : 0.0;
        S.relres = err;
        if (!isfinite(v[0]) || !isfinite(v[1]) || !isfinite(S.fnorm)) {
            S.status = MSA_NAN_INPUT;
            S.active = 0;
        } else if (S.fnorm > 1e-25 && !(err < a.tol)) {
            S.status = MSA_NOT_CONVERGED;
            S.active = 1;
            atomicAdd(a.active_count, 1);
        } else {
            S.status = MSA_OK;
            S.active = 0;
        }
        break;
    }
    case OP_FNORM:
        S.fnorm = sqrt(v[0]);
        break;
    case OP_DIRECT: {  // partials of k_pcg_init_warm: v[1] = |f - K x|^2
        const double fn = fnorm_of(C, S);
        S.iters = 0;
        S.total_iters = 0;
        S.outer = 0;
        S.active = 0;
        S.relres = fn > 1e-25 ? sqrt(v[1]) / fn : 0.0;
        S.status = (C.bfail && C.bfail[smp]) || !isfinite(S.relres) ? MSA_BREAKDOWN : MSA_OK;
        break;
    }
    case OP_A:  // alpha = rz / p.q
        if (!S.active) break;
        S.pq = v[0];
        if (!(v[0] > 0) || !isfinite(v[0])) {
            S.status = MSA_BREAKDOWN;
            S.active = 0;
            atomicSub(a.active_count, 1);
            if (S.ir_active) set_due(C, S, smp);  // reliable updates: final fold + test
        } else {
            S.alpha = S.rz / v[0];
        }
        break;
    case OP_B: {  // |r|/|b| < tol, else beta = rz_new / rz (SolverPCG.cs:130-136)
        if (!S.active) break;
        S.iters += 1;
        S.rr = v[1];
        const double err = sqrt(v[1]) / S.bnorm;
        S.relres = err;
        if (!isfinite(err) || !isfinite(v[0])) {
            S.status = MSA_BREAKDOWN;
            S.active = 0;
            atomicSub(a.active_count, 1);
            if (a.ru_drop > 0) set_due(C, S, smp);  // final fold of the correction
        } else if (err < a.tol) {
            S.status = MSA_OK;
            S.active = 0;
            atomicSub(a.active_count, 1);
        } else if (S.iters >= a.max_it) {
            S.status = MSA_NOT_CONVERGED;
            S.active = 0;
            atomicSub(a.active_count, 1);
            if (a.ru_drop > 0) set_due(C, S, smp);
        } else {
            S.beta = v[0] / S.rz;
            S.rz = v[0];
            // reliable updates: the FP32 residual (normalised to 1 at the last
            // replacement) fell by ru_drop, or the FP64 stopping test may pass
            if (a.ru_drop > 0) {
                const double rn = sqrt(v[1]);
                if (rn < a.ru_drop || rn * S.scale < a.ru_tol * fnorm_of(C, S)) set_due(C, S, smp);
            }
        }
        break;
    }
    case OP_IRNORM: {  // restarted refinement: |f - K u|/|f| < tol
        if (!S.ir_active) break;
        const double nr = sqrt(v[0]);
        S.total_iters += S.iters;
        S.outer += 1;
        S.r64norm = nr;
        const double err = nr / fnorm_of(C, S);
        S.relres = err;
        if (!isfinite(err)) {
            S.status = MSA_BREAKDOWN;
            S.ir_active = 0;
            atomicSub(a.ir_count, 1);
        } else if (err < a.tol) {
            S.status = MSA_OK;
            S.ir_active = 0;
            atomicSub(a.ir_count, 1);
        } else if (S.outer >= a.max_outer) {
            S.status = MSA_NOT_CONVERGED;
            S.ir_active = 0;
            atomicSub(a.ir_count, 1);
        } else {
            S.scale = nr;
        }
        break;
    }
    case OP_RUNORM: {  // residual replacement: FP64 stopping test, rescale ratio
        if (!S.ir_active || !S.ru_due) break;
        const double nr = sqrt(v[0]);
        S.outer += 1;
        S.total_iters = S.iters;
        S.r64norm = nr;
        const double err = nr / fnorm_of(C, S);
        S.relres = err;
        int done = 0;
        if (!isfinite(err)) {
            S.status = MSA_BREAKDOWN;
            done = 1;
        } else if (err < a.tol) {
            S.status = MSA_OK;
            done = 1;
        } else if (S.iters >= a.max_it || !S.active) {
            if (S.status != MSA_BREAKDOWN) S.status = MSA_NOT_CONVERGED;
            done = 1;
        }
        if (done) {
            S.ir_active = 0;
            S.ru_due = 0;
            atomicSub(a.ir_count, 1);
            if (S.active) {
                S.active = 0;
                atomicSub(a.active_count, 1);
            }
        } else {
            S.pq = S.scale / nr;  // rescale ratio of the search direction
            S.scale = nr;
        }
        break;
    }
    case OP_RURESTART:
        if (S.ir_active && S.ru_due) {
            S.rz = v[0];
            S.ru_due = 0;
        }
        break;
    }
}

// The partial sums of one tile (chunk by chunk) reduced and applied.  The
// summation is that of 1024 threads (thread v sums chunks v/T, +2G, ... and
// +G, ..., then a tree over G = 1024/T); NTHR threads play 1024/NTHR of them
// each, in any order (no step reads what it writes), so every NTHR gives the
// same bits.  red: [2][1024] in LDS.
template <int T, int NTHR>
__device__ __forceinline__ void reduce_tile(const ClassDev *__restrict__ cls, const int2 ct,
                                            const ReduceArgs &a, double (*red)[1024])
{
    constexpr int G = 1024 / T, VPT = 1024 / NTHR;
    const ClassDev &C = cls[ct.x];
    const int tile = ct.y;
    const bool two = a.op == OP_INIT0 || a.op == OP_INIT1 || a.op == OP_INIT2 || a.op == OP_B ||
                     a.op == OP_DIRECT;
#pragma unroll
    for (int q = 0; q < VPT; q++) {
        const int v = threadIdx.x + NTHR * q;
        const int t = v % T, g = v / T;
        const double *base = C.partial + (size_t)tile * C.nchunks * 3 * T + t;
        double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
        int c = g;
        for (; c + G < C.nchunks; c += 2 * G) {
            a0 += base[(size_t)c * 3 * T];
            b0 += base[(size_t)(c + G) * 3 * T];
            if (two) {
                a1 += base[(size_t)c * 3 * T + T];
                b1 += base[(size_t)(c + G) * 3 * T + T];
            }
        }
        if (c < C.nchunks) {
            a0 += base[(size_t)c * 3 * T];
            if (two) a1 += base[(size_t)c * 3 * T + T];
        }
        red[0][v] = a0 + b0;
        red[1][v] = a1 + b1;
    }
    __syncthreads();
    for (int stride = G / 2; stride >= 1; stride >>= 1) {
#pragma unroll
        for (int q = 0; q < VPT; q++) {
            const int v = threadIdx.x + NTHR * q;
            if (v / T < stride) {
                red[0][v] += red[0][v + stride * T];
                red[1][v] += red[1][v + stride * T];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < T) {
        const int t = threadIdx.x;
        const int smp = tile * T + t;
        const double v[2] = {red[0][t], red[1][t]};
        if (smp < C.nsamp) apply_op(C, smp, v, a);
    }
    // these ops clear every replacement flag of the class and append none
    if ((a.op == OP_INIT0 || a.op == OP_INIT1 || a.op == OP_INIT2 || a.op == OP_RURESTART) &&
        tile == 0 && threadIdx.x == 0)
        *C.due = 0;
}

template <int T>
__global__ __launch_bounds__(1024) void k_reduce(const ClassDev *__restrict__ cls,
                                                 const int2 *__restrict__ tiles, ReduceArgs a)
{
    __shared__ double red[2][1024];
    reduce_tile<T, 1024>(cls, tiles[blockIdx.x], a, red);
}

// The whole Jacobi-PCG of a batch in one cooperative launch: every
// iteration's four phases (search direction + product, alpha, update, test
// and beta) separated by grid-wide barriers, the same per-item and per-tile
// work and summation order as the four kernels of the captured iteration
// (so the same bits).  Work items go to XCDs in contiguous ranges, as
// xcd_item does.  Every workgroup reads the active-sample count after the
// same barrier, so all leave the loop together; `iters` bounds it.
template <typename real, int T, int U>
__global__ __launch_bounds__(256) void k_pcg_persist(const ClassDev *__restrict__ cls,
                                                     const int4 *__restrict__ items, int nitems,
                                                     const int2 *__restrict__ tiles, int ntiles,
                                                     ReduceArgs ra, ReduceArgs rb, int pbuf0,
                                                     int iters)
{
    __shared__ double red[2][1024];
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const int G = gridDim.x, gx = G >> 3, xcd = blockIdx.x & 7, m = blockIdx.x >> 3;
    const int per = (nitems + 7) >> 3;
    const int i0 = xcd * per + m, i1 = min((xcd + 1) * per, nitems);
    auto barrier = [&] {
        __threadfence();
        grid.sync();
        __threadfence();
    };
    for (int k = 0; k < iters; k++) {
        // the same value in every workgroup (read between the same two barriers)
        if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(rb.active_count, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) <= 0)
            break;
        const int pbuf = (pbuf0 + k) & 1;
        for (int i = i0; i < i1; i += gx) {
            pcg_a_item<real, T, U>(cls, items[i], pbuf);
            __syncthreads();
        }
        barrier();
        for (int t = blockIdx.x; t < ntiles; t += G) {
            reduce_tile<T, 256>(cls, tiles[t], ra, red);
            __syncthreads();
        }
        barrier();
        for (int i = i0; i < i1; i += gx) {
            pcg_b_item<real, T>(cls, items[i], pbuf);
            __syncthreads();
        }
        barrier();
        for (int t = blockIdx.x; t < ntiles; t += G) {
            reduce_tile<T, 256>(cls, tiles[t], rb, red);
            __syncthreads();
        }
        barrier();
    }
}

// ------------------------------------------------------------ multigrid

// Geometric multigrid V-cycle used as the PCG preconditioner z = B r.
// Level 0 is the solved class (b = its PCG residual r), coarser levels are
// rediscretizations of the mesh coarsened by two (b = restricted residual,
// stored in the level's r).  Per level: two Chebyshev-Jacobi steps from
// zero, residual, restriction, coarse correction, prolongation, two
// Chebyshev-Jacobi steps.  The same polynomial before and after keeps B
// symmetric; its interval [lam/ratio, lam] uses the Gershgorin bound lam of
// D^-1 K, so the smoother contracts every mode and B stays positive
// definite.  Buffers of a level: z holds D^-1 b on entry and the level's
// result on exit; A = q and B = p_pbuf are free between PCG updates.

// Chebyshev coefficients of the two-step smoother on [lam/ratio, lam]:
// x1 = x0 + d1, d1 = D^-1 r0 / theta; x2 = x1 + c1 d1 + c2 D^-1 r1.
struct ChebCoef {
    double ith, c1, c2;
};

__device__ __forceinline__ ChebCoef cheb_coef(double lam, double ratio)
{
    const double hi = lam, lo = lam / ratio;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
    const double sigma = theta / delta, rho0 = 1.0 / sigma, rho1 = 1.0 / (2.0 * sigma - rho0);
    return ChebCoef{1.0 / theta, rho1 * rho0, 2.0 * rho1 / delta};
}

template <typename real, int T>
__device__ __forceinline__ void gather3(const MSA_GLOBAL real *v, int col, int s, real &v0,
                                        real &v1, real &v2)
{
    if constexpr (T == 1) {
        const vec3<real> c = ld3(v, col);
        v0 = c.x;
        v1 = c.y;
        v2 = c.z;
    } else {
        const size_t o = (size_t)col * 3 * T + s;
        v0 = v[o];
        v1 = v[o + T];
        v2 = v[o + 2 * T];
    }
}

// y = K v for block row br (slots [base, base + W) of its slice): groups of
// three slots, padding columns clamped to block 0 (their values are zero);
// lower slots of upper-triangle classes by load_row_block.
template <typename real, int T>
__device__ __forceinline__ void row_product(const ClassDev &C, const MSA_GLOBAL real *val,
                                            const MSA_GLOBAL int *cidx, const MSA_GLOBAL real *v,
                                            int base, int W, int rr, int s, real &a0, real &a1,
                                            real &a2)
{
    constexpr int R = 64 / T;
    a0 = a1 = a2 = 0;
    int j = 0;
    for (; j + 3 <= W; j += 3) {
        int col[3];
#pragma unroll
        for (int u = 0; u < 3; u++) col[u] = cidx[(size_t)(base + j + u) * R + rr];
        real x[3][3], m[3][9];
#pragma unroll
        for (int u = 0; u < 3; u++) {
            gather3<real, T>(v, max(col[u], 0), s, x[u][0], x[u][1], x[u][2]);
            load_row_block<real, T, true>(C, val, base, j + u, col[u], rr, s, m[u]);
        }
#pragma unroll
        for (int u = 0; u < 3; u++) {
            a0 += m[u][0] * x[u][0] + m[u][1] * x[u][1] + m[u][2] * x[u][2];
            a1 += m[u][3] * x[u][0] + m[u][4] * x[u][1] + m[u][5] * x[u][2];
            a2 += m[u][6] * x[u][0] + m[u][7] * x[u][1] + m[u][8] * x[u][2];
        }
    }
    for (; j < W; j++) {
        const int col = cidx[(size_t)(base + j) * R + rr];
        real x0, x1, x2, m[9];
        gather3<real, T>(v, max(col, 0), s, x0, x1, x2);
        load_row_block<real, T, true>(C, val, base, j, col, rr, s, m);
        a0 += m[0] * x0 + m[1] * x1 + m[2] * x2;
        a1 += m[3] * x0 + m[4] * x1 + m[5] * x2;
        a2 += m[6] * x0 + m[7] * x1 + m[8] * x2;
    }
}

// Range-checked loads: a raw buffer descriptor (stride 0) over `bytes`
// bytes at p; a load at a byte offset >= bytes returns zeros (the buffer
// range check of the CDNA memory pipeline), so a missing neighbour is read
// as zero without a select.  p and bytes must be wave-uniform.
// (the typed LLVM intrinsics are bound directly: the clang builtins for
// 64/96/128-bit buffer loads of this toolchain return a single dword)
typedef float f32x3 __attribute__((ext_vector_type(3)));
__device__ f32x3 buffer_load_f32x3(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                   int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.v3f32");
__device__ double buffer_load_f64(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                  int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.f64");
__device__ f64x2 buffer_load_f64x2(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                   int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f64");
__device__ __forceinline__ __amdgpu_buffer_rsrc_t wave_buffer(const void *p, unsigned bytes)
{
    const unsigned long long a = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    void *q = (void *)(((unsigned long long)hi << 32) | lo);
    // dword 3 of the descriptor: 32-bit data format (gfx9 family)
    return __builtin_amdgcn_make_buffer_rsrc(q, (short)0, (int)__builtin_amdgcn_readfirstlane(bytes),
                                             0x00020000);
}

// Moduli of the 8 cells around grid node (i, j, k) (c = cx + 2 cy + 4 cz;
// cell (cx, cy, cz) exists when cx = 0 and i > 0, or cx = 1 and i < nx - 1,
// likewise in y, z), times cK; missing cells read modulus zero (offset past
// the end of the modulus buffer).  cw: 1 where the cell exists, else 0 (a
// value per lane rather than a lane mask, which would hold scalar registers
// through the products).
__device__ __forceinline__ void mf_moduli(const ClassDev &C, const double *Es, int i, int j, int k,
                                          int nx, int ny, const bool (&lo)[3],
                                          const bool (&hi)[3], float cK, float (&Ec)[8],
                                          float (&cw)[8])
{
    bool cin[8];
    const int ne = __builtin_amdgcn_readfirstlane(C.n_elem);
    const __amdgpu_buffer_rsrc_t eb = wave_buffer((const void *)Es, (unsigned)ne * 8u);
    const unsigned eout = (unsigned)ne * 8u;
    const int cy1 = nx - 1, cz1 = (nx - 1) * (ny - 1);
    const int c0 = (i - 1) + cy1 * (j - 1) + cz1 * (k - 1);
    int ce[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
        cin[c] = (cx ? hi[0] : lo[0]) && (cy ? hi[1] : lo[1]) && (cz ? hi[2] : lo[2]);
        ce[c] = cin[c] ? c0 + cx + cy * cy1 + cz * cz1 : 0;
    }
    if (!__builtin_amdgcn_readfirstlane(C.mf_cell_id)) {  // element numbers from the table
        const MSA_GLOBAL int *cell = gbl(C.mf_cell);
#pragma unroll
        for (int c = 0; c < 8; c++) ce[c] = cell[ce[c]];
    }
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const double e = buffer_load_f64(eb, (int)(cin[c] ? (unsigned)ce[c] * 8u : eout), 0, 0);
        Ec[c] = (float)e * cK;
        cw[c] = cin[c] ? 1.f : 0.f;
    }
}

// The element matrix is the same for every lane: constant-address loads
// with wave-uniform indices (scalar loads, SGPR operands of the FMAs).  An
// empty asm redefines the pointer on every call, so the 576 loads are not
// hoisted out of the caller's row loop (they would not fit the scalar
// registers).
using cfloat = __attribute__((address_space(4))) const float;
__device__ __forceinline__ cfloat *mf_khat(const ClassDev &C)
{
    unsigned long long kp = (unsigned long long)C.mf_k;
    asm volatile("" : "+s"(kp));
    return (cfloat *)kp;
}

// Consistent mass of the cells inside the grid (dynamics): acc += cM sum
// over existing cells of Mhat[a][b] v_b (component by component).
__device__ __forceinline__ void mf_mass(cfloat *mh, const float (&xv)[27][3], const float (&cw)[8],
                                        float cM, float (&acc)[3])
{
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
        const int a = (1 - cx) + 2 * (1 - cy) + 4 * (1 - cz);
        float m[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int o = 0; o < 8; o++) {
            const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
            const float mab = mh[a * 8 + o];
            const float *xb = xv[((cz + oz) * 3 + (cy + oy)) * 3 + (cx + ox)];
#pragma unroll
            for (int p = 0; p < 3; p++) m[p] = __builtin_fmaf(mab, xb[p], m[p]);
        }
        const float w = cw[c] * cM;
#pragma unroll
        for (int p = 0; p < 3; p++) acc[p] = __builtin_fmaf(w, m[p], acc[p]);
    }
}

// y = K v for block row br of a uniform structured grid class, from the
// moduli of the (up to) 8 cells around the node and the shared unit element
// matrix (FP32, stored in corner order): y = sum_cells cK E_c Khat[a][b] v_b
// (+ cM Mhat[a][b] v_b on the component itself), a = the node's corner in
// the cell, b over the cell's corners.  Neighbours outside the grid or
// constrained (block -1) are read at an offset past the end of the vector
// and come back as zeros.
//
// Blocks forming a lexicographic sub-box of the grid (x fastest): lane l of
// the wave holds block br = (first block of the wave) + l, so its x - 1 and
// x + 1 neighbours are held by lanes l - 1 and l + 1.  Each lane loads the 9
// neighbours of its own x column; the columns at x -/+ 1 come from the
// neighbouring lanes by whole-wave DPP shifts (lanes 0 and 63 load their
// missing column themselves), and the x -/+ 1 part of each cell is switched
// off at the ends of a row (where the neighbouring lane holds another row).
//
// BOX selects the box-grid form at compile time: kernels holding both
// forms would share one scalar-register budget for their element-matrix
// loads.
//
// PG (box grids): the PCG product K p, summed over the corners'
// differences from the node's value (below); pown receives the node's own
// value of v.
template <bool BOX, bool PG>
__device__ __forceinline__ void mf_product(const ClassDev &C, const MSA_GLOBAL float *v,
                                           const double *Es, int br, float &y0, float &y1,
                                           float &y2, float *pown)
{
    static_assert(BOX || !PG, "the difference form is a box-grid form");
    const int g = gbl(C.mf_bpos)[br];
    const int i = g & 1023, j = (g >> 10) & 1023, k = g >> 20;
    const int nx = __builtin_amdgcn_readfirstlane(C.mf_nx);
    const int ny = __builtin_amdgcn_readfirstlane(C.mf_ny);
    const int nz = __builtin_amdgcn_readfirstlane(C.mf_nz);
    const int nb = __builtin_amdgcn_readfirstlane(C.nb);
    const __amdgpu_buffer_rsrc_t vb = wave_buffer((const void *)v, (unsigned)nb * 12u);
    const unsigned vout = (unsigned)nb * 12u;
    const bool lo[3] = {i > 0, j > 0, k > 0}, hi[3] = {i < nx - 1, j < ny - 1, k < nz - 1};
    const float cK = C.dyn ? (float)C.cK : 1.f, cM = C.dyn ? (float)C.cM : 0.f;
    cfloat *kh = mf_khat(C);
    float acc[3] = {0.f, 0.f, 0.f};
    // one node's three components at byte offset o (past the end: zeros)
    auto node = [&](unsigned o) -> f32x3 { return buffer_load_f32x3(vb, (int)o, 0, 0); };
    if constexpr (BOX) {
        const int b0x = __builtin_amdgcn_readfirstlane(C.mf_box_lo[0]);
        const int b0y = __builtin_amdgcn_readfirstlane(C.mf_box_lo[1]);
        const int b0z = __builtin_amdgcn_readfirstlane(C.mf_box_lo[2]);
        const int b1x = __builtin_amdgcn_readfirstlane(C.mf_box_hi[0]);
        const int b1y = __builtin_amdgcn_readfirstlane(C.mf_box_hi[1]);
        const int b1z = __builtin_amdgcn_readfirstlane(C.mf_box_hi[2]);
        const bool blo[3] = {i > b0x, j > b0y, k > b0z}, bhi[3] = {i < b1x, j < b1y, k < b1z};
        const int sy = b1x - b0x + 1, sz = sy * (b1y - b0y + 1);
        const unsigned own = (unsigned)br * 12u;
        const int lane = threadIdx.x & 63;
        // own column (dx = 0): m = dy + 3 dz
        unsigned off[9];
        float x0[9][3], xf[9][3];
#pragma unroll
        for (int m = 0; m < 9; m++) {
            const int dy = m % 3, dz = m / 3;
            const bool in = (dy != 0 || blo[1]) && (dy != 2 || bhi[1]) && (dz != 0 || blo[2]) &&
                            (dz != 2 || bhi[2]);
            const int rel = 12 * ((dy - 1) * sy + (dz - 1) * sz);  // wave-uniform
            off[m] = in ? own + (unsigned)rel : vout;
        }
#pragma unroll
        for (int m = 0; m < 9; m++) {
            const f32x3 w = node(off[m]);
            x0[m][0] = w.x;
            x0[m][1] = w.y;
            x0[m][2] = w.z;
        }
        if constexpr (PG) {
            if (pown) {
                pown[0] = x0[4][0];
                pown[1] = x0[4][1];
                pown[2] = x0[4][2];
            }
        }
        // lane 0: the column at x - 1, lane 63: at x + 1 (zeros past the
        // box ends; other lanes take them from their neighbours)
#pragma unroll
        for (int m = 0; m < 9; m++) xf[m][0] = xf[m][1] = xf[m][2] = 0.f;
        if (lane == 0 || lane == 63) {
            const bool okx = lane == 0 ? blo[0] : bhi[0];
            const unsigned sh = lane == 0 ? (unsigned)-12 : 12u;
#pragma unroll
            for (int m = 0; m < 9; m++) {
                const f32x3 w = node((okx && off[m] != vout) ? off[m] + sh : vout);
                xf[m][0] = w.x;
                xf[m][1] = w.y;
                xf[m][2] = w.z;
            }
        }
        float Ec[8], cw[8];
        mf_moduli(C, Es, i, j, k, nx, ny, lo, hi, cK, Ec, cw);
        // x - 1 and x + 1 parts count where the neighbour is in the box
        const float fm = blo[0] ? 1.f : 0.f, fp = bhi[0] ? 1.f : 0.f;
        // column values: dx = 0 own, dx = -1 / +1 from lane l - 1 / l + 1
        // (lanes 0 / 63: their own loads), times fm / fp
        auto col = [&](int m, int dx, int p) -> float {
            if (dx == 1) return x0[m][p];
            const int o = __builtin_bit_cast(int, xf[m][p]), s = __builtin_bit_cast(int, x0[m][p]);
            const int v = dx == 0 ? __builtin_amdgcn_update_dpp(o, s, 0x138, 0xf, 0xf, false)
                                  : __builtin_amdgcn_update_dpp(o, s, 0x130, 0xf, 0xf, false);
            return __builtin_bit_cast(float, v);
        };
        // PG: each cell's block row times the corners' differences from the
        // node's own value, Khat[a][.] (x_o - x_a), which equals Khat[a][.] x
        // since the blocks of a row sum to zero (a translation stores no
        // energy).  Smooth search directions make K p a small difference of
        // large terms; the differences are small, so the FP32 product keeps
        // the accuracy the PCG needs.  A corner outside the box counts as
        // x_o = 0 (its difference -x_a).
        float xn[3] = {0.f, 0.f, 0.f};
        if constexpr (PG) {
            xn[0] = x0[4][0];
            xn[1] = x0[4][1];
            xn[2] = x0[4][2];
        }
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
            const int a = (1 - cx) + 2 * (1 - cy) + 4 * (1 - cz);  // the node's corner
            const float f = cx ? fp : fm;
            float t0[3] = {0.f, 0.f, 0.f}, ts[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int o = 0; o < 8; o++) {
                if (PG && o == a) continue;  // the node itself: difference 0
                const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
                const int m = (cy + oy) + 3 * (cz + oz);
                cfloat *kab = kh + (a * 8 + o) * 9;
                const int dx = cx + ox;  // 0: x - 1, 1: own column, 2: x + 1
                float xb[3];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const float v = col(m, dx, q);
                    xb[q] = !PG ? v : (dx == 1 ? v - xn[q] : __builtin_fmaf(f, v, -xn[q]));
                }
                float(&tt)[3] = (PG || dx == 1) ? t0 : ts;
#pragma unroll
                for (int p = 0; p < 3; p++) {
                    tt[p] = __builtin_fmaf(kab[3 * p], xb[0], tt[p]);
                    tt[p] = __builtin_fmaf(kab[3 * p + 1], xb[1], tt[p]);
                    tt[p] = __builtin_fmaf(kab[3 * p + 2], xb[2], tt[p]);
                }
            }
#pragma unroll
            for (int p = 0; p < 3; p++)
                acc[p] = PG ? __builtin_fmaf(Ec[c], t0[p], acc[p])
                            : __builtin_fmaf(Ec[c], __builtin_fmaf(f, ts[p], t0[p]), acc[p]);
            // keep each cell's 72 scalar loads next to their use
            __builtin_amdgcn_sched_barrier(0);
        }
        if (cM != 0.f) {
            float xs[27][3];  // (dx, dy, dz), x fastest
#pragma unroll
            for (int m = 0; m < 9; m++)
#pragma unroll
                for (int dx = 0; dx < 3; dx++)
#pragma unroll
                    for (int p = 0; p < 3; p++)
                        xs[m * 3 + dx][p] = col(m, dx, p) * (dx == 0 ? fm : (dx == 2 ? fp : 1.f));
            mf_mass(kh + 576, xs, cw, cM, acc);
        }
        y0 = acc[0];
        y1 = acc[1];
        y2 = acc[2];
    } else {
        // general grids: the 27 neighbours from the grid-point table
        const MSA_GLOBAL int *gblock = gbl(C.mf_gblock);
        unsigned off[27];
        {
            const int gp0 = i + nx * (j + ny * k), sy = nx, sz = nx * ny;
#pragma unroll
            for (int n = 0; n < 27; n++) {
                const int dx = n % 3, dy = (n / 3) % 3, dz = n / 9;
                const bool in = (dx != 0 || lo[0]) && (dx != 2 || hi[0]) && (dy != 0 || lo[1]) &&
                                (dy != 2 || hi[1]) && (dz != 0 || lo[2]) && (dz != 2 || hi[2]);
                const int gp = gp0 + (dx - 1) + (dy - 1) * sy + (dz - 1) * sz;
                const int b = gblock[in ? gp : gp0];
                off[n] = (in && b >= 0) ? (unsigned)b * 12u : vout;
            }
        }
        float xv[27][3];
#pragma unroll
        for (int n = 0; n < 27; n++) {
            const f32x3 w = buffer_load_f32x3(vb, (int)off[n], 0, 0);
            xv[n][0] = w.x;
            xv[n][1] = w.y;
            xv[n][2] = w.z;
        }
        float Ec[8], cw[8];
        mf_moduli(C, Es, i, j, k, nx, ny, lo, hi, cK, Ec, cw);
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
            const int a = (1 - cx) + 2 * (1 - cy) + 4 * (1 - cz);  // the node's corner
            float t[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int o = 0; o < 8; o++) {
                const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
                cfloat *kab = kh + (a * 8 + o) * 9;
                const float *xb = xv[((cz + oz) * 3 + (cy + oy)) * 3 + (cx + ox)];
                // one FMA per term (no product-sum trees)
#pragma unroll
                for (int p = 0; p < 3; p++) {
                    t[p] = __builtin_fmaf(kab[3 * p], xb[0], t[p]);
                    t[p] = __builtin_fmaf(kab[3 * p + 1], xb[1], t[p]);
                    t[p] = __builtin_fmaf(kab[3 * p + 2], xb[2], t[p]);
                }
            }
#pragma unroll
            for (int p = 0; p < 3; p++) acc[p] = __builtin_fmaf(Ec[c], t[p], acc[p]);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (cM != 0.f) mf_mass(kh + 576, xv, cw, cM, acc);
        y0 = acc[0];
        y1 = acc[1];
        y2 = acc[2];
    }
}

// FP32 and FP64 copies of element 0's unit Khat (and Mhat) for the
// matrix-free products, in corner order: block (oa, ob) of corners oa, ob
// is Khat's block (corner_node[oa], corner_node[ob]).
__global__ void k_mf_khat(const ClassDev *__restrict__ cp)
{
    const ClassDev &C = *cp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < 576) {
        const int blk = i / 9, r = i - 9 * (i / 9);
        const int a = C.mf_corner[blk >> 3], b = C.mf_corner[blk & 7];
        v = C.kc[(a * 8 + b) * 9 + r];
        const int p = r / 3, q = r - 3 * (r / 3);
        C.mf_k64[kMfK64T + (((blk >> 3) * 3 + q) * 8 + (blk & 7)) * 3 + p] = v;
    } else if (i < 640) {
        const int blk = i - 576;
        const int a = C.mf_corner[blk >> 3], b = C.mf_corner[blk & 7];
        v = C.mc ? C.mc[a * 8 + b] : 0.0;
    }
    if (i < 640) {
        C.mf_k[i] = (float)v;
        C.mf_k64[i] = v;
    }
}

// FP64 residual r64 -= (cK E K + cM M) u64 of the samples due for a
// replacement, matrix-free on a uniform grid class (the stored-value-free
// counterpart of k_residual64's due path): every free row of a due sample
// sums its 8 cells, each cell's 8 corners gathered from u64 (FP64), the
// element matrix in LDS.  One-sample tiles: a work item is one sample, so
// items of samples that are not due exit before staging anything.
__global__ __launch_bounds__(256) void k_residual64_mf(const ClassDev *__restrict__ cls,
                                                       const int4 *__restrict__ items, int nitems)
{
    constexpr int R = 64;
    __shared__ double kh[640];
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int sample = tile;
    if (*C.due == 0 || sample >= C.nsamp || !C.st[sample].ir_active || !C.st[sample].ru_due) return;
    for (int t = threadIdx.x; t < 640; t += blockDim.x) kh[t] = C.mf_k64[t];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nx = C.mf_nx, ny = C.mf_ny, nz = C.mf_nz;
    const MSA_GLOBAL int *gblock = gbl(C.mf_gblock);
    const MSA_GLOBAL int *cell = gbl(C.mf_cell);
    const MSA_GLOBAL double *E = gbl((const double *)C.E + (size_t)tile * C.n_elem);
    const MSA_GLOBAL double *u = gbl((const double *)C.u64 + (size_t)tile * C.tile_vec);
    MSA_GLOBAL double *r = gbl(C.r64 + (size_t)tile * C.tile_vec);
    const double cK = C.dyn ? C.cK : 1.0, cM = C.dyn ? C.cM : 0.0;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * R + lane;
        if (br >= C.nb) continue;
        const int g = gbl(C.mf_bpos)[br];
        const int i = g & 1023, j = (g >> 10) & 1023, k = g >> 20;
        double y[3] = {0.0, 0.0, 0.0};
        // box grids: corner blocks by arithmetic and range-checked loads (a
        // corner outside the box reads zeros), so no load waits on a table
        // lookup; two cells per step keep 16 corner loads in flight
        const bool box = C.mf_box;
        const int b0x = C.mf_box_lo[0], b0y = C.mf_box_lo[1], b0z = C.mf_box_lo[2];
        const int sxb = C.mf_box_hi[0] - b0x + 1, syb = C.mf_box_hi[1] - b0y + 1;
        const int bzn = C.mf_box_hi[2] - b0z + 1;
        const __amdgpu_buffer_rsrc_t ub = wave_buffer((const void *)u, (unsigned)C.nb * 24u);
        const unsigned uout = (unsigned)C.nb * 24u;
#pragma unroll 2
        for (int c = 0; c < 8; c++) {
            const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
            const int ci = i - 1 + cx, cj = j - 1 + cy, ck = k - 1 + cz;
            if (ci < 0 || ci >= nx - 1 || cj < 0 || cj >= ny - 1 || ck < 0 || ck >= nz - 1) continue;
            const double Ec = E[cell[ci + (nx - 1) * (cj + (ny - 1) * ck)]] * cK;
            const int a = (1 - cx) + 2 * (1 - cy) + 4 * (1 - cz);
            double uc[8][3];
#pragma unroll
            for (int o = 0; o < 8; o++) {
                const int gx = ci + (o & 1), gy = cj + ((o >> 1) & 1), gz = ck + (o >> 2);
                if (box) {
                    const int x = gx - b0x, yy = gy - b0y, z = gz - b0z;
                    const bool in = (unsigned)x < (unsigned)sxb && (unsigned)yy < (unsigned)syb &&
                                    (unsigned)z < (unsigned)bzn;
                    const unsigned off = in ? (unsigned)(x + sxb * (yy + syb * z)) * 24u : uout;
                    const f64x2 v = buffer_load_f64x2(ub, (int)off, 0, 0);
                    uc[o][0] = v.x;
                    uc[o][1] = v.y;
                    uc[o][2] = buffer_load_f64(ub, (int)(in ? off + 16u : uout), 0, 0);
                } else {
                    const int nb = gblock[gx + nx * (gy + ny * gz)];
                    const vec3<double> v = ld3(u, nb >= 0 ? nb : 0);
                    uc[o][0] = nb >= 0 ? v.x : 0.0;
                    uc[o][1] = nb >= 0 ? v.y : 0.0;
                    uc[o][2] = nb >= 0 ? v.z : 0.0;
                }
            }
            double t3[3] = {0.0, 0.0, 0.0}, m3[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int o = 0; o < 8; o++) {
                const double *kab = kh + (a * 8 + o) * 9;
#pragma unroll
                for (int p = 0; p < 3; p++)
                    t3[p] += kab[3 * p] * uc[o][0] + kab[3 * p + 1] * uc[o][1] + kab[3 * p + 2] * uc[o][2];
                const double mab = kh[576 + a * 8 + o];
#pragma unroll
                for (int p = 0; p < 3; p++) m3[p] += mab * uc[o][p];
            }
#pragma unroll
            for (int p = 0; p < 3; p++) y[p] += Ec * t3[p] + cM * m3[p];
        }
        const uint8_t fr = C.block_free[br];
        // the row's three components move as one 24-byte access; fixed
        // components are written back unchanged
        const vec3<double> rv = ld3((const MSA_GLOBAL double *)r, br);
        st3(r, br, (fr & 1) ? rv.x - y[0] : rv.x, (fr & 2) ? rv.y - y[1] : rv.y,
            (fr & 4) ? rv.z - y[2] : rv.z);
    }
}

// The same residual on box grids (free blocks = a lexicographic sub-box of
// the grid, one-sample tiles), one input component q at a time: each lane
// loads component q of the 9 (dy, dz) columns at its own x once and takes
// the x - 1 / x + 1 columns from lanes l - 1 / l + 1 by DPP shifts (lanes
// 0 and 63 load theirs), zero past a row end; the element matrix comes by
// scalar loads of its transposed copy, 24 + 8 values per (cell, q).  The
// general form above gathers the 8 corners of every cell (64 loads per
// row).  Items cover max(1, it.w) chunks.
using cdouble = __attribute__((address_space(4))) const double;
__device__ __forceinline__ cdouble *mf_khat64(const ClassDev &C)
{
    unsigned long long kp = (unsigned long long)C.mf_k64;
    asm volatile("" : "+s"(kp));
    return (cdouble *)kp;
}

__global__ __launch_bounds__(256) void k_residual64_mf_box(const ClassDev *__restrict__ cls,
                                                           const int4 *__restrict__ items,
                                                           int nitems)
{
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z, nch = max(1, it.w);
    const int sample = tile;
    if (*C.due == 0 || sample >= C.nsamp || !C.st[sample].ir_active || !C.st[sample].ru_due) return;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nx = __builtin_amdgcn_readfirstlane(C.mf_nx);
    const int ny = __builtin_amdgcn_readfirstlane(C.mf_ny);
    const int nz = __builtin_amdgcn_readfirstlane(C.mf_nz);
    const int nb = __builtin_amdgcn_readfirstlane(C.nb);
    const int b0x = __builtin_amdgcn_readfirstlane(C.mf_box_lo[0]);
    const int b0y = __builtin_amdgcn_readfirstlane(C.mf_box_lo[1]);
    const int b0z = __builtin_amdgcn_readfirstlane(C.mf_box_lo[2]);
    const int b1x = __builtin_amdgcn_readfirstlane(C.mf_box_hi[0]);
    const int b1y = __builtin_amdgcn_readfirstlane(C.mf_box_hi[1]);
    const int b1z = __builtin_amdgcn_readfirstlane(C.mf_box_hi[2]);
    const int sy = b1x - b0x + 1, sz = sy * (b1y - b0y + 1);
    const MSA_GLOBAL double *u = gbl((const double *)C.u64 + (size_t)tile * C.tile_vec);
    MSA_GLOBAL double *r = gbl(C.r64 + (size_t)tile * C.tile_vec);
    const double *Es = C.E + (size_t)tile * C.n_elem;
    const __amdgpu_buffer_rsrc_t ub = wave_buffer((const void *)u, (unsigned)nb * 24u);
    const unsigned uout = (unsigned)nb * 24u;
    const int ne = __builtin_amdgcn_readfirstlane(C.n_elem);
    const __amdgpu_buffer_rsrc_t eb = wave_buffer((const void *)Es, (unsigned)ne * 8u);
    const unsigned eout = (unsigned)ne * 8u;
    const double cK = C.dyn ? C.cK : 1.0, cM = C.dyn ? C.cM : 0.0;
    const int s0 = chunk * C.chunk, s1 = min((chunk + nch) * C.chunk, C.nslices);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * 64 + lane;
        if (br >= nb) continue;
        const int g = gbl(C.mf_bpos)[br];
        const int i = g & 1023, j = (g >> 10) & 1023, k = g >> 20;
        const bool lo[3] = {i > 0, j > 0, k > 0}, hi[3] = {i < nx - 1, j < ny - 1, k < nz - 1};
        const bool blo[3] = {i > b0x, j > b0y, k > b0z}, bhi[3] = {i < b1x, j < b1y, k < b1z};
        const unsigned own = (unsigned)br * 24u;
        unsigned off[9];
#pragma unroll
        for (int m = 0; m < 9; m++) {
            const int dy = m % 3, dz = m / 3;
            const bool in = (dy != 0 || blo[1]) && (dy != 2 || bhi[1]) && (dz != 0 || blo[2]) &&
                            (dz != 2 || bhi[2]);
            off[m] = in ? own + (unsigned)(24 * ((dy - 1) * sy + (dz - 1) * sz)) : uout;
        }
        // the 8 cells' moduli (cells outside the grid: 0) and mass weights
        double Ec[8], Mc[8];
        {
            const int cy1 = nx - 1, cz1 = (nx - 1) * (ny - 1);
            const int c0 = (i - 1) + cy1 * (j - 1) + cz1 * (k - 1);
            const bool ident = __builtin_amdgcn_readfirstlane(C.mf_cell_id) != 0;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
                const bool cin = (cx ? hi[0] : lo[0]) && (cy ? hi[1] : lo[1]) && (cz ? hi[2] : lo[2]);
                int ce = cin ? c0 + cx + cy * cy1 + cz * cz1 : 0;
                if (!ident) ce = gbl(C.mf_cell)[ce];
                const double e = buffer_load_f64(eb, (int)(cin ? (unsigned)ce * 8u : eout), 0, 0);
                Ec[c] = e * cK;
                Mc[c] = cin ? cM : 0.0;
            }
        }
        // x - 1 / x + 1 values count where the neighbour is in the box
        const double fm = blo[0] ? 1.0 : 0.0, fp = bhi[0] ? 1.0 : 0.0;
        double y[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
        for (int q = 0; q < 3; q++) {
            double xq[9], xm[9], xp[9], ym = 0.0;
#pragma unroll
            for (int m = 0; m < 9; m++)
                xq[m] = buffer_load_f64(ub, (int)(off[m] != uout ? off[m] + 8u * q : uout), 0, 0);
            // lane 0: the column at x - 1, lane 63: at x + 1 (the other
            // lanes' values are replaced by the shifts)
#pragma unroll
            for (int m = 0; m < 9; m++) xm[m] = 0.0;
            if (lane == 0 || lane == 63) {
                const bool okx = lane == 0 ? blo[0] : bhi[0];
                const unsigned sh = lane == 0 ? (unsigned)-24 : 24u;
#pragma unroll
                for (int m = 0; m < 9; m++)
                    xm[m] = buffer_load_f64(
                        ub, (int)((okx && off[m] != uout) ? off[m] + sh + 8u * q : uout), 0, 0);
            }
#pragma unroll
            for (int m = 0; m < 9; m++) {
                xp[m] = fp * dpp_wave<kDppWaveShl1>(xm[m], xq[m]);
                xm[m] = fm * dpp_wave<kDppWaveShr1>(xm[m], xq[m]);
            }
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
                const int a = (1 - cx) + 2 * (1 - cy) + 4 * (1 - cz);  // the node's corner
                double t[3] = {0.0, 0.0, 0.0}, mq = 0.0;
#pragma unroll
                for (int h = 0; h < 2; h++) {  // 4 corners (16 scalar values) at a time
                    cdouble *kt = mf_khat64(C) + kMfK64T + (a * 3 + q) * 24;
                    cdouble *mh = mf_khat64(C) + 576 + a * 8;
#pragma unroll
                    for (int o = 4 * h; o < 4 * h + 4; o++) {
                        const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
                        const int m = (cy + oy) + 3 * (cz + oz), dx = cx + ox;
                        const double v = dx == 0 ? xm[m] : (dx == 1 ? xq[m] : xp[m]);
#pragma unroll
                        for (int p = 0; p < 3; p++) t[p] = __builtin_fma(kt[o * 3 + p], v, t[p]);
                        mq = __builtin_fma(mh[o], v, mq);
                    }
                    if (h == 0) __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int p = 0; p < 3; p++) y[p] = __builtin_fma(Ec[c], t[p], y[p]);
                ym = __builtin_fma(Mc[c], mq, ym);
                // keep each (cell, component)'s 32 scalar loads next to their use
                __builtin_amdgcn_sched_barrier(0);
            }
            // the mass couples each component to itself only
            if (q == 0)
                y[0] += ym;
            else if (q == 1)
                y[1] += ym;
            else
                y[2] += ym;
        }
        const uint8_t fr = C.block_free[br];
        const vec3<double> rv = ld3((const MSA_GLOBAL double *)r, br);
        st3(r, br, (fr & 1) ? rv.x - y[0] : rv.x, (fr & 2) ? rv.y - y[1] : rv.y,
            (fr & 4) ? rv.z - y[2] : rv.z);
    }
}

// Row update of one smoothing step at one-sample tiles, given y = K src of
// block row br: the row's three components of each vector move as one
// 12- or 24-byte access.  PRE A = (1 + c1) z/theta + c2 D^-1 (b - K z/theta),
// RESID B = b - K A, POST1 B = A + D^-1 (b - K A)/theta, POST2
// z = B + c1 (B - A) + c2 D^-1 (b - K B) (and r.z).  mg_load1 reads the
// row's operands, which do not depend on y: the matrix-free product issues
// it before its own loads, so these reads overlap its arithmetic.
template <typename real>
struct MgRow {
    real bv[3], dv[3], xa[3], xb[3];
};

template <typename real, int STEP>
__device__ __forceinline__ MgRow<real> mg_load1(const MSA_GLOBAL real *b,
                                                const MSA_GLOBAL real *dinv,
                                                const MSA_GLOBAL real *z,
                                                const MSA_GLOBAL real *A,
                                                const MSA_GLOBAL real *B, int br)
{
    MgRow<real> w;
    auto get = [&](const MSA_GLOBAL real *p, real (&v)[3]) {
        const vec3<real> t = ld3(p, br);
        v[0] = t.x;
        v[1] = t.y;
        v[2] = t.z;
    };
    get(b, w.bv);
    if constexpr (STEP != MG_RESID) get(dinv, w.dv);
    if constexpr (STEP == MG_PRE) get(z, w.xa);
    if constexpr (STEP == MG_POST1 || STEP == MG_POST2) get(A, w.xa);
    if constexpr (STEP == MG_POST2) get(B, w.xb);
    return w;
}

template <typename real, int STEP>
__device__ __forceinline__ void mg_finish1(const MgRow<real> &w, MSA_GLOBAL real *dst, int br,
                                           const real (&y)[3], real ith, real c1, real c2,
                                           real &rz)
{
    real out[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if constexpr (STEP == MG_PRE) {
            const real x1 = w.xa[d] * ith;
            out[d] = x1 + c1 * x1 + c2 * w.dv[d] * (w.bv[d] - y[d] * ith);
        } else if constexpr (STEP == MG_RESID) {
            out[d] = w.bv[d] - y[d];
        } else if constexpr (STEP == MG_POST1) {
            out[d] = w.xa[d] + ith * w.dv[d] * (w.bv[d] - y[d]);
        } else {
            out[d] = w.xb[d] + c1 * (w.xb[d] - w.xa[d]) + c2 * w.dv[d] * (w.bv[d] - y[d]);
            rz += w.bv[d] * out[d];
        }
    }
    st3(dst, br, out[0], out[1], out[2]);
}

// One smoothing-level product with its row update (STEP, see MgStep):
//   MG_PRE   A = (1 + c1) x1 + c2 D^-1 (b - K x1),  x1 = z / theta
//   MG_RESID B = b - K A
//   MG_POST1 B = A + D^-1 (b - K A) / theta
//   MG_POST2 z = B + c1 (B - A) + c2 D^-1 (b - K B); level 0 also writes
//            the partial sums of r.z (slot 0 of the PCG partials)
template <typename real, int T, int STEP, int MF = 0>
__global__ __launch_bounds__(256) void k_mg_smooth(const ClassDev *__restrict__ cls,
                                                   const int4 *__restrict__ items, int nitems,
                                                   int pbuf, int mode, double ratio)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const bool act = sample < C.nsamp && mg_act(cls[C.mg_root].st[sample], mode);
    const ChebCoef cc = cheb_coef(act ? C.mg_lam[sample] : 1.0, ratio);
    const real ith = (real)cc.ith, c1 = (real)cc.c1, c2 = (real)cc.c2;
    const size_t vb = (size_t)tile * C.tile_vec;
    const MSA_GLOBAL real *__restrict__ b = gbl((const real *)C.r + vb);
    const MSA_GLOBAL real *__restrict__ dinv = gbl((const real *)C.dinv + vb);
    MSA_GLOBAL real *__restrict__ z = gbl((real *)C.z + vb);
    MSA_GLOBAL real *__restrict__ A = gbl((real *)C.q + vb);
    MSA_GLOBAL real *__restrict__ B = gbl((real *)(pbuf ? C.p1 : C.p0) + vb);
    const MSA_GLOBAL real *src = STEP == MG_PRE ? z : (STEP == MG_POST2 ? B : A);
    MSA_GLOBAL real *dst = STEP == MG_PRE ? A : (STEP == MG_POST2 ? z : B);
    const MSA_GLOBAL real *__restrict__ val = gbl((const real *)C.val + (size_t)tile * C.tile_vals);
    const MSA_GLOBAL int *__restrict__ cidx = gbl(C.cidx);
    const MSA_GLOBAL int *__restrict__ sptr = gbl(C.slice_ptr);
    // an item covers max(1, it.w) consecutive chunks (box-grid lists group
    // them, so one workgroup's set-up serves several slices per wave)
    const int nch = max(1, it.w);
    for (int ck = chunk; ck < chunk + nch; ck++) {
        real rz = 0;
        const int s0 = ck * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
        if (__ballot(act)) {
            for (int slice = s0 + wave; slice < s1; slice += kWaves) {
                const int br = slice * R + rr;
                const int base = sptr[slice];
                const int W = sptr[slice + 1] - base;
                if (!act || br >= C.nb) continue;
                real y[3];
                if constexpr (T == 1) {
                    const MgRow<real> w = mg_load1<real, STEP>(b, dinv, z, A, B, br);
                    if constexpr (MF != 0)
                        mf_product<MF == 2, MF == 2>(C, (const MSA_GLOBAL float *)src,
                                                     C.E + (size_t)tile * C.n_elem, br, y[0],
                                                     y[1], y[2]);
                    else
                        row_product<real, T>(C, val, cidx, src, base, W, rr, s, y[0], y[1], y[2]);
                    mg_finish1<real, STEP>(w, dst, br, y, ith, c1, c2, rz);
                    continue;
                }
                row_product<real, T>(C, val, cidx, src, base, W, rr, s, y[0], y[1], y[2]);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const size_t o = ((size_t)br * 3 + d) * T + s;
                    const real bv = b[o], dv = dinv[o];
                    real out;
                    if constexpr (STEP == MG_PRE) {
                        const real x1 = z[o] * ith;
                        out = x1 + c1 * x1 + c2 * dv * (bv - y[d] * ith);
                    } else if constexpr (STEP == MG_RESID) {
                        out = bv - y[d];
                    } else if constexpr (STEP == MG_POST1) {
                        out = A[o] + ith * dv * (bv - y[d]);
                    } else {
                        const real xb = B[o];
                        out = xb + c1 * (xb - A[o]) + c2 * dv * (bv - y[d]);
                        rz += bv * out;
                    }
                    dst[o] = out;
                }
            }
        }
        if constexpr (STEP == MG_POST2) {
            if (C.mg_depth == 0) {
                double acc[1] = {(double)rz};
                block_partials<T, 1>(acc, C.partial + ((size_t)tile * C.nchunks + ck) * 3 * T);
                // the next chunk's partials reuse the same LDS rows
                if (ck + 1 < chunk + nch) __syncthreads();
            }
        }
    }
}

// ------------------------------------------------ smoothing-interval estimate
// The Chebyshev smoother works on [lam/ratio, lam], lam the top of the
// spectrum of D^-1 K.  The Gershgorin bound is rigorous but loose on these
// grids (5.36 against a spectrum top near 3.1), which puts the interval
// above the spectrum and weakens the cycle.  A few Lanczos steps on
// K x = lambda D x per sample and level give the top Ritz value, a lower
// bound that converges from below; the interval then takes
// min(Gershgorin, margin x Ritz).  Level buffers (free before a solve):
// v = q, v_prev = z, y = p0 = b - K v with b = r = 0 (the RESID smoothing
// step), so K v = -p0.  All vectors are D-normalised: v.D.v = 1, D = 1/dinv
// on the free components (dinv = 0 marks a fixed one).
__device__ __forceinline__ float lz_hash(unsigned x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return (float)(x & 0xffffffu) * (1.0f / 8388608.0f) - 1.0f;  // [-1, 1)
}

// Unnormalised vectors u_k (|u_k|_D = n_k, v_k = u_k / n_k), so that a step
// takes two passes besides its product y = b - K u_k = -K u_k:
// op 0: u_0 = hashed values on the free components, u_prev = 0, b = 0;
//       partials of u.D.u
// op 2: partials of u_k.K u_k = -(u.y)   (alpha_k = that / n_k^2)
// op 3: w = (D^-1 K u_k - alpha_k u_k) / n_k - (n_k / n_k-1) u_k-1,
//       u_prev = u_k, u = w (= u_k+1); partials of w.D.w (n_k+1^2)
// lz[k] = alpha_k, lz[kLzMax + k] = n_k: the tridiagonal's off-diagonal
// between k - 1 and k is n_k.
template <typename real, int T>
__global__ __launch_bounds__(256) void k_lz_step(const ClassDev *__restrict__ cls,
                                                 const int4 *__restrict__ items, int nitems,
                                                 int op, int k)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const bool act = sample < C.nsamp;
    const size_t vb = (size_t)tile * C.tile_vec;
    real *__restrict__ u = (real *)C.q + vb;
    real *__restrict__ y = (real *)C.p0 + vb;
    real *__restrict__ up = (real *)C.z + vb;
    real *__restrict__ b = (real *)C.r + vb;
    const real *__restrict__ dinv = (const real *)C.dinv + vb;
    real alpha = 0, inv_n = 0, ratio = 0;
    if (act && op == 3) {
        const double *lz = C.mg_lz + (size_t)sample * (2 * kLzMax + 1);
        const double n = lz[kLzMax + k];
        alpha = (real)lz[k];
        inv_n = n > 0.0 ? (real)(1.0 / n) : (real)0;
        ratio = (k > 0 && lz[kLzMax + k - 1] > 0.0) ? (real)(n / lz[kLzMax + k - 1]) : (real)0;
    }
    double acc = 0.0;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    if (act) {
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            if (br >= C.nb) continue;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const size_t o = ((size_t)br * 3 + d) * T + s;
                const real dv = dinv[o];
                if (op == 0) {
                    const real h = dv != (real)0 ? (real)lz_hash((unsigned)(br * 3 + d) * 2654435761u +
                                                                 (unsigned)sample * 40503u)
                                                 : (real)0;
                    u[o] = h;
                    up[o] = 0;
                    b[o] = 0;
                    if (dv != (real)0) acc += (double)h * (double)h / (double)dv;
                } else if (op == 2) {
                    acc -= (double)u[o] * (double)y[o];
                } else {
                    const real uk = u[o];
                    const real w = dv != (real)0 ? (-dv * y[o] - alpha * uk) * inv_n - ratio * up[o]
                                                 : (real)0;
                    up[o] = uk;
                    u[o] = w;
                    if (dv != (real)0) acc += (double)w * (double)w / (double)dv;
                }
            }
        }
    }
    double a1[1] = {acc};
    block_partials<T, 1>(a1, C.partial + ((size_t)tile * C.nchunks + chunk) * 3 * T);
}

// Per sample of the classes in `list` (grid.y), one wave per sample: the
// chunk partials summed lane-strided, then across the wave in a fixed
// order.  which 0: n_0 = sqrt(u.D.u); 1: alpha_k = sum / n_k^2;
// 2: n_k+1 = sqrt(w.D.w).
template <int T>
__global__ __launch_bounds__(256) void k_lz_sum(const ClassDev *__restrict__ cls,
                                                const int *__restrict__ list, int which, int k)
{
    const ClassDev &C = cls[list[blockIdx.y]];
    const int lane = threadIdx.x & 63;
    const int smp = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (smp >= C.nsamp) return;
    const int tile = smp / T, t = smp % T;
    const double *p = C.partial + (size_t)tile * C.nchunks * 3 * T + t;
    double sum = 0.0;
    for (int c = lane; c < C.nchunks; c += 64) sum += p[(size_t)c * 3 * T];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) sum += __shfl_down(sum, off, 64);
    if (lane != 0) return;
    double *lz = C.mg_lz + (size_t)smp * (2 * kLzMax + 1);
    if (which == 0) {
        lz[kLzMax] = sqrt(fmax(sum, 0.0));
    } else if (which == 1) {
        const double n = lz[kLzMax + k];
        lz[k] = n > 0.0 ? sum / (n * n) : 0.0;
    } else {
        lz[kLzMax + k + 1] = sqrt(fmax(sum, 0.0));
    }
}

// Top eigenvalue of the Lanczos tridiagonal (alpha_0..alpha_m-1, off-
// diagonals n_1..n_m-1; m = steps, or fewer where an n vanishes: an
// invariant subspace) by bisection on the Sturm count; the interval's top
// becomes min(Gershgorin bound, margin x Ritz value).  A Ritz value that is
// not finite and positive leaves the Gershgorin bound.
__global__ void k_lz_eig(const ClassDev *__restrict__ cls, const int *__restrict__ list, int steps,
                         double margin)
{
    const ClassDev &C = cls[list[blockIdx.y]];
    const int smp = blockIdx.x * blockDim.x + threadIdx.x;
    if (smp >= C.nsamp) return;
    const double *lz = C.mg_lz + (size_t)smp * (2 * kLzMax + 1);
    const double *al = lz, *nb = lz + kLzMax;
    if (!(nb[0] > 0.0)) return;
    int m = steps;
    for (int j = 1; j < steps; j++)
        if (!(nb[j] > 1e-12 * fabs(al[0]))) {
            m = j;
            break;
        }
    double lo = 1e300, hi = -1e300;
    for (int j = 0; j < m; j++) {
        const double bl = j > 0 ? fabs(nb[j]) : 0.0;
        const double bu = j + 1 < m ? fabs(nb[j + 1]) : 0.0;
        lo = fmin(lo, al[j] - bl - bu);
        hi = fmax(hi, al[j] + bl + bu);
    }
    if (!(hi > lo) || !isfinite(hi) || !isfinite(lo)) return;
    for (int it = 0; it < 200 && hi - lo > 1e-13 * fabs(hi); it++) {
        const double x = 0.5 * (lo + hi);
        int below = 0;  // eigenvalues below x: negative pivots of T - x I
        double d = 1.0;
        for (int j = 0; j < m; j++) {
            d = (al[j] - x) - (j > 0 ? nb[j] * nb[j] / d : 0.0);
            if (d == 0.0) d = -1e-300;
            if (d < 0.0) below++;
        }
        if (below == m)
            hi = x;
        else
            lo = x;
    }
    if (isfinite(hi) && hi > 0.0) C.mg_lam[smp] = fmin(C.mg_lam[smp], margin * hi);
}

// The matrix-free smoothing products of a box grid of fully free blocks
// with the element contraction on the matrix cores (MIXED, one-sample
// tiles).  A workgroup owns kMfBY rows of nodes (all x) of one sample and
// marches through the grid's cell layers in z.  Per layer:
//   1. the layer's top node plane (of the product's input vector) and the
//      cells' moduli go to LDS (the bottom plane is already there);
//   2. every cell's 24 outputs Y_c = Khat (cK E_c U_c) (+ Mhat (cM U_c) in
//      dynamics) by v_mfma_f32_16x16x4f32: 16 cells per product, Khat's 24
//      rows as two 16-row blocks (the second half-used), K in 6 (12) steps
//      of 4; U_c, the cell's 8 corners x 3 components, gathered from LDS;
//   3. each owned node of the layer's bottom plane sums the 4 cells below it
//      (kept from the previous layer) and the 4 above, in a fixed order, and
//      takes the step's row update (mg_finish1, as k_mg_smooth); the sums
//      of the 4 cells below the next plane are kept for the next layer.
// Cells in the rows next to the owned ones are computed by both
// neighbouring workgroups, so every node's sum is formed in one place
// (no atomics; results do not depend on scheduling).  Level 0's POST2 writes
// its r.z partial in slot (sample, y block); y block 0 clears the slots
// past the last block.
constexpr int kMfBY = kMfmaRows;          // owned node rows per workgroup
constexpr int kMfMaxCX = kMfmaMaxCells;   // cells in x a workgroup handles
constexpr int kMfMaxCY = kMfBY + 1;
constexpr int kMfNNX = kMfMaxCX + 1, kMfNNY = kMfMaxCY + 1;
constexpr int kMfYS = 25;                 // LDS floats per cell (odd: conflict-free node reads)

template <int STEP>
__global__ __launch_bounds__(256) void k_mg_smooth_mfma(const ClassDev *__restrict__ cls,
                                                        const int4 *__restrict__ items, int nitems,
                                                        int pbuf, int mode, double ratio)
{
    __shared__ float Ul[2][kMfNNX * kMfNNY * 3];
    __shared__ float El[kMfMaxCX * kMfMaxCY];
    __shared__ float Yl[kMfMaxCX * kMfMaxCY * kMfYS];
    __shared__ float Al[(kMfMaxCX + 1) * kMfBY * 3];
    __shared__ float red[4];
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int sample = it.y, yb = it.z;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // box of free blocks, cell ranges touching this workgroup's nodes
    const int nx = C.mf_nx, ny = C.mf_ny, nz = C.mf_nz;
    const int b0x = C.mf_box_lo[0], b0y = C.mf_box_lo[1], b0z = C.mf_box_lo[2];
    const int b1x = C.mf_box_hi[0], b1y = C.mf_box_hi[1], b1z = C.mf_box_hi[2];
    const int nyb = (b1y - b0y + kMfBY) / kMfBY;
    const int y0 = b0y + yb * kMfBY, y1 = min(y0 + kMfBY - 1, b1y);
    const int cxl = max(b0x - 1, 0), cxh = min(b1x, nx - 2);
    const int cyl = max(y0 - 1, 0), cyh = min(y1, ny - 2);
    const int czl = max(b0z - 1, 0), czh = min(b1z, nz - 2);
    const int NCX = cxh - cxl + 1, NCY = cyh - cyl + 1, NNX = NCX + 1, NNY = NCY + 1;
    const int NOX = b1x - b0x + 1, NOY = y1 - y0 + 1, ncell = NCX * NCY;
    const int nnode = NNX * NNY, nown = NOX * NOY;
    const int sy = b1x - b0x + 1, sz = sy * (b1y - b0y + 1);
    const bool part = STEP == MG_POST2 && C.mg_depth == 0;
    double *pout = C.partial + (size_t)sample * C.nchunks * 3;
    if (part && yb == 0)
        for (int k = nyb + tid; k < C.nchunks; k += 256) pout[(size_t)k * 3] = 0.0;
    const bool act = sample < C.nsamp && mg_act(cls[C.mg_root].st[sample], mode);
    if (!act) {
        if (part && tid == 0) pout[(size_t)yb * 3] = 0.0;
        return;
    }
    const ChebCoef cc = cheb_coef(C.mg_lam[sample], ratio);
    const float ith = (float)cc.ith, c1 = (float)cc.c1, c2 = (float)cc.c2;
    const size_t vb = (size_t)sample * C.tile_vec;
    const MSA_GLOBAL float *__restrict__ b = gbl((const float *)C.r + vb);
    const MSA_GLOBAL float *__restrict__ dinv = gbl((const float *)C.dinv + vb);
    MSA_GLOBAL float *__restrict__ z = gbl((float *)C.z + vb);
    MSA_GLOBAL float *__restrict__ A = gbl((float *)C.q + vb);
    MSA_GLOBAL float *__restrict__ B = gbl((float *)(pbuf ? C.p1 : C.p0) + vb);
    const MSA_GLOBAL float *src = STEP == MG_PRE ? z : (STEP == MG_POST2 ? B : A);
    MSA_GLOBAL float *dst = STEP == MG_PRE ? A : (STEP == MG_POST2 ? z : B);
    const MSA_GLOBAL float *xsa = STEP == MG_PRE ? z : A;  // the update's first iterate
    const double *Es = C.E + (size_t)sample * C.n_elem;
    const float cK = C.dyn ? (float)C.cK : 1.f, cM = C.dyn ? (float)C.cM : 0.f;
    const bool mass = cM != 0.f;
    const bool cellid = C.mf_cell_id != 0;
    const int *celltab = C.mf_cell;
    // matrix-core operands of Khat (and Mhat x I3): block m, step s holds
    // row 16 m + (lane & 15), column 4 s + (lane >> 4) in corner order
    const int j16 = lane & 15, g = lane >> 4;
    float ak[2][6], am[2][6];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int st = 0; st < 6; st++) {
            const int i = 16 * m + j16, k = 4 * st + g;
            const bool in = i < 24;
            ak[m][st] = in ? C.mf_k[((i / 3) * 8 + k / 3) * 9 + 3 * (i % 3) + k % 3] : 0.f;
            am[m][st] = (in && mass && i % 3 == k % 3) ? C.mf_k[576 + (i / 3) * 8 + k / 3] : 0.f;
        }
    // Loads for the next layer and the row updates are issued before the
    // products of the current layer and consumed after them, so their
    // latency overlaps the matrix-core work: a thread holds up to 3 plane
    // nodes, 2 cells and 2 owned nodes (nnode <= 585, ncell <= 512,
    // nown <= 455).
    float pv[3][3], ev[2], fb[2][3], fd[2][3], fa[2][3], fx[2][3];
    // the thread's layer-invariant indices: plane nodes (block index in the
    // plane, -1 outside the box or past the plane), cells (index in the
    // layer, -1 past it), owned nodes (block index in the plane, -1 past)
    // and their cells in the layer (-1: no such cell); quotients by the
    // float reciprocal (exact here: t < 1024, divisors <= 65)
    auto divmod = [](int t, int d, float inv, int &qq, int &rr) {
        qq = (int)(((float)t + 0.5f) * inv);
        rr = t - qq * d;
    };
    const float inx = 1.f / (float)NNX, icx = 1.f / (float)NCX, iox = 1.f / (float)NOX;
    int pnode[3], pcell[2], obr[2], oq[2][4];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int t = tid + 256 * i;
        int ly, lx;
        divmod(t, NNX, inx, ly, lx);
        const int gx = cxl + lx, gy = cyl + ly;
        pnode[i] = (t < nnode && gx >= b0x && gx <= b1x && gy >= b0y && gy <= b1y)
                       ? (gx - b0x) + sy * (gy - b0y) : -1;
    }
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int t = tid + 256 * i;
        int ly, lx;
        divmod(t, NCX, icx, ly, lx);
        pcell[i] = t < ncell ? (cxl + lx) + (nx - 1) * (cyl + ly) : -1;
        divmod(t, NOX, iox, ly, lx);
        const int gx = b0x + lx, gy = y0 + ly;
        obr[i] = t < nown ? lx + sy * (y0 - b0y + ly) : -1;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int cx = gx - 1 + (c & 1), cy = gy - 1 + (c >> 1);
            oq[i][c] = (t < nown && cx >= cxl && cx <= cxh && cy >= cyl && cy <= cyh)
                           ? (cx - cxl) + NCX * (cy - cyl) : -1;
        }
    }
    auto fetch_plane = [&](int zz) {
        const bool zin = zz >= b0z && zz <= b1z;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            pv[i][0] = pv[i][1] = pv[i][2] = 0.f;
            if (zin && pnode[i] >= 0) {
                const vec3<float> w = ld3(src, (size_t)pnode[i] + (size_t)sz * (zz - b0z));
                pv[i][0] = w.x;
                pv[i][1] = w.y;
                pv[i][2] = w.z;
            }
        }
    };
    auto store_plane = [&](int slot) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int t = tid + 256 * i;
            if (t < nnode) {
                Ul[slot][3 * t] = pv[i][0];
                Ul[slot][3 * t + 1] = pv[i][1];
                Ul[slot][3 * t + 2] = pv[i][2];
            }
        }
    };
    const int lstride = (nx - 1) * (ny - 1);
    auto fetch_E = [&](int cz) {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            ev[i] = 0.f;
            if (pcell[i] >= 0) {
                const int cell = pcell[i] + lstride * cz;
                ev[i] = (float)Es[cellid ? cell : celltab[cell]] * cK;
            }
        }
    };
    auto store_E = [&]() {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int t = tid + 256 * i;
            if (t < ncell) El[t] = ev[i];
        }
    };
    auto fetch_rows = [&](int zz) {
        if (zz < b0z || zz > b1z) return;
#pragma unroll
        for (int i = 0; i < 2; i++) {
            if (obr[i] < 0) continue;
            const int br = obr[i] + sz * (zz - b0z);
            const vec3<float> vb3 = ld3(b, br);
            fb[i][0] = vb3.x;
            fb[i][1] = vb3.y;
            fb[i][2] = vb3.z;
            if constexpr (STEP != MG_RESID) {
                const vec3<float> w = ld3(dinv, br);
                fd[i][0] = w.x;
                fd[i][1] = w.y;
                fd[i][2] = w.z;
            }
            if constexpr (STEP != MG_RESID) {
                const vec3<float> w = ld3(xsa, br);
                fa[i][0] = w.x;
                fa[i][1] = w.y;
                fa[i][2] = w.z;
            }
            if constexpr (STEP == MG_POST2) {
                const vec3<float> w = ld3(B, br);
                fx[i][0] = w.x;
                fx[i][1] = w.y;
                fx[i][2] = w.z;
            }
        }
    };
    float rz = 0.f;
    // the owned nodes of plane zz: the sums of the cells below (Al) and, when
    // `above`, of the current layer's cells, then the row update (the
    // arithmetic of mg_finish1); `keep` stores the current layer's sums for
    // the plane above
    auto finalize = [&](int zz, bool above, bool keep) {
        const bool own = zz >= b0z && zz <= b1z;
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int t = tid + 256 * i;
            if (obr[i] < 0) continue;
            float yv[3] = {Al[3 * t], Al[3 * t + 1], Al[3 * t + 2]};
            float at[3] = {0.f, 0.f, 0.f};
            if (above) {
#pragma unroll
                for (int c = 0; c < 4; c++) {  // cells (x - 1 + dx, y - 1 + dy), c = dx + 2 dy
                    if (oq[i][c] < 0) continue;
                    const float *yq = Yl + oq[i][c] * kMfYS;
                    const int ob = (1 - (c & 1)) + 2 * (1 - (c >> 1));  // the node's corner, below
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        yv[d] += yq[3 * ob + d];
                        at[d] += yq[3 * (ob + 4) + d];  // as the cell's top corner
                    }
                }
            }
            if (own) {
                float out[3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const float bv = fb[i][d];
                    if constexpr (STEP == MG_PRE) {
                        const float x1 = fa[i][d] * ith;
                        out[d] = x1 + c1 * x1 + c2 * fd[i][d] * (bv - yv[d] * ith);
                    } else if constexpr (STEP == MG_RESID) {
                        out[d] = bv - yv[d];
                    } else if constexpr (STEP == MG_POST1) {
                        out[d] = fa[i][d] + ith * fd[i][d] * (bv - yv[d]);
                    } else {
                        out[d] = fx[i][d] + c1 * (fx[i][d] - fa[i][d]) + c2 * fd[i][d] * (bv - yv[d]);
                        rz += bv * out[d];
                    }
                }
                st3(dst, obr[i] + sz * (zz - b0z), out[0], out[1], out[2]);
            }
            if (keep) {
                Al[3 * t] = at[0];
                Al[3 * t + 1] = at[1];
                Al[3 * t + 2] = at[2];
            }
        }
    };
    // products of one layer's cells: groups of 16 cells, two groups per
    // wave in flight (four independent accumulators)
    auto products = [&](int cur, int nxt) {
        for (int gi = wave; gi * 16 < ncell; gi += 8) {
            const int q0 = gi * 16 + j16, q1 = q0 + 64;
            const bool v0 = q0 < ncell, v1 = q1 < ncell;
            float w0[6], w1[6];
            const float e0 = v0 ? El[q0] : 0.f, e1 = v1 ? El[q1] : 0.f;
            int y0c, x0, y1c, x1;
            divmod(v0 ? q0 : 0, NCX, icx, y0c, x0);
            divmod(v1 ? q1 : 0, NCX, icx, y1c, x1);
#pragma unroll
            for (int st = 0; st < 6; st++) {
                const int k = 4 * st + g, o = k / 3, p = k - 3 * (k / 3);
                const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
                const float *U = Ul[oz ? nxt : cur];
                w0[st] = v0 ? U[((y0c + oy) * NNX + x0 + ox) * 3 + p] : 0.f;
                w1[st] = v1 ? U[((y1c + oy) * NNX + x1 + ox) * 3 + p] : 0.f;
            }
            f32x4 d00 = {0.f, 0.f, 0.f, 0.f}, d01 = d00, d10 = d00, d11 = d00;
#pragma unroll
            for (int st = 0; st < 6; st++) {
                const float a = e0 * w0[st], c = e1 * w1[st];
                d00 = __builtin_amdgcn_mfma_f32_16x16x4f32(ak[0][st], a, d00, 0, 0, 0);
                d01 = __builtin_amdgcn_mfma_f32_16x16x4f32(ak[1][st], a, d01, 0, 0, 0);
                d10 = __builtin_amdgcn_mfma_f32_16x16x4f32(ak[0][st], c, d10, 0, 0, 0);
                d11 = __builtin_amdgcn_mfma_f32_16x16x4f32(ak[1][st], c, d11, 0, 0, 0);
            }
            if (mass) {
#pragma unroll
                for (int st = 0; st < 6; st++) {
                    const float a = cM * w0[st], c = cM * w1[st];
                    d00 = __builtin_amdgcn_mfma_f32_16x16x4f32(am[0][st], a, d00, 0, 0, 0);
                    d01 = __builtin_amdgcn_mfma_f32_16x16x4f32(am[1][st], a, d01, 0, 0, 0);
                    d10 = __builtin_amdgcn_mfma_f32_16x16x4f32(am[0][st], c, d10, 0, 0, 0);
                    d11 = __builtin_amdgcn_mfma_f32_16x16x4f32(am[1][st], c, d11, 0, 0, 0);
                }
            }
            if (v0) {
                float *yq = Yl + q0 * kMfYS;
#pragma unroll
                for (int r = 0; r < 4; r++) yq[4 * g + r] = d00[r];
                if (g < 2) {
#pragma unroll
                    for (int r = 0; r < 4; r++) yq[16 + 4 * g + r] = d01[r];
                }
            }
            if (v1) {
                float *yq = Yl + q1 * kMfYS;
#pragma unroll
                for (int r = 0; r < 4; r++) yq[4 * g + r] = d10[r];
                if (g < 2) {
#pragma unroll
                    for (int r = 0; r < 4; r++) yq[16 + 4 * g + r] = d11[r];
                }
            }
        }
    };
    fetch_plane(czl);
    store_plane(0);
    fetch_plane(czl + 1);
    store_plane(1);
    fetch_E(czl);
    store_E();
    for (int t = tid; t < nown * 3; t += 256) Al[t] = 0.f;
    __syncthreads();
    int cur = 0;
    for (int cz = czl; cz <= czh; cz++) {
        const int nxt = cur ^ 1;
        const bool more = cz + 1 <= czh;
        if (more) {
            fetch_plane(cz + 2);
            fetch_E(cz + 1);
        }
        fetch_rows(cz);
        products(cur, nxt);
        __syncthreads();
        finalize(cz, true, true);
        if (more) {
            store_plane(cur);  // plane cz + 2 takes plane cz's slot
            store_E();
        }
        __syncthreads();
        cur = nxt;
    }
    // the top plane of the grid has no cell layer above it
    if (czh + 1 <= b1z) {
        fetch_rows(czh + 1);
        finalize(czh + 1, false, false);
    }
    if (part) {
        float v = rz;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[tid >> 6] = v;
        __syncthreads();
        if (tid == 0)
            pout[(size_t)yb * 3] = (double)red[0] + (double)red[1] + (double)red[2] + (double)red[3];
    }
}

// Restriction to a coarse level: b_c = P^T (b - K A) of the finer level
// (its B buffer), free components only; z_c = D_c^-1 b_c starts the coarse
// smoother.  Items: the coarse level's.
template <typename real, int T>
__global__ __launch_bounds__(256) void k_mg_restrict(const ClassDev *__restrict__ cls,
                                                     const int4 *__restrict__ items, int nitems,
                                                     int pbuf, int mode)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const ClassDev &F = cls[C.mg_fine];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    if (sample >= C.nsamp || !mg_act(cls[C.mg_root].st[sample], mode)) return;
    const MSA_GLOBAL real *Bf = gbl((const real *)(pbuf ? F.p1 : F.p0) + (size_t)tile * F.tile_vec);
    const size_t vb = (size_t)tile * C.tile_vec;
    real *bc = (real *)C.r + vb, *zc = (real *)C.z + vb;
    const real *dinv = (const real *)C.dinv + vb;
    const MSA_GLOBAL int *rcol = gbl(C.mg_r_col);
    const MSA_GLOBAL float *rw = gbl(C.mg_r_w);
    const MSA_GLOBAL uint8_t *ffree = gbl(F.block_free);
    const int width = C.mg_r_width, nbc = C.nb;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    const bool st = __builtin_amdgcn_readfirstlane(C.mg_rstruct);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int cb = slice * R + rr;
        if (cb >= C.nb) continue;
        double acc[3] = {0, 0, 0};
        if (st) {
            // fine points 2I + d, d in {-1, 0, 1} per direction, weight 1/2
            // per odd offset; fine blocks = the fine box, lexicographic
            const int g = C.mf_bpos[cb];
            const int fi = 2 * (g & 1023), fj = 2 * ((g >> 10) & 1023), fk = 2 * (g >> 20);
            const int lx = F.mf_box_lo[0], ly = F.mf_box_lo[1], lz = F.mf_box_lo[2];
            const int hx = F.mf_box_hi[0], hy = F.mf_box_hi[1], hz = F.mf_box_hi[2];
            const int sy = hx - lx + 1, sz = sy * (hy - ly + 1);
            // branch-free: points outside the fine box are read past the
            // end of the vector (zeros)
            const int nbf = F.nb;
            const __amdgpu_buffer_rsrc_t fbuf = wave_buffer((const void *)Bf, (unsigned)nbf * 3u * sizeof(real));
            const unsigned fout = (unsigned)nbf * 3u * sizeof(real);
#pragma unroll
            for (int n = 0; n < 27; n++) {
                const int dx = n % 3 - 1, dy = (n / 3) % 3 - 1, dz = n / 9 - 1;
                const int x = fi + dx, y = fj + dy, z = fk + dz;
                const bool in = x >= lx && x <= hx && y >= ly && y <= hy && z >= lz && z <= hz;
                const unsigned off =
                    in ? (unsigned)((x - lx) + sy * (y - ly) + sz * (z - lz)) * 3u * (unsigned)sizeof(real)
                       : fout;
                const double w = (dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0);
                real v0, v1, v2;
                if constexpr (sizeof(real) == 4) {
                    const f32x3 v = buffer_load_f32x3(fbuf, (int)off, 0, 0);
                    v0 = v.x;
                    v1 = v.y;
                    v2 = v.z;
                } else {
                    v0 = buffer_load_f64(fbuf, (int)off, 0, 0);
                    v1 = buffer_load_f64(fbuf, (int)(off == fout ? fout : off + 8u), 0, 0);
                    v2 = buffer_load_f64(fbuf, (int)(off == fout ? fout : off + 16u), 0, 0);
                }
                acc[0] += w * (double)v0;
                acc[1] += w * (double)v1;
                acc[2] += w * (double)v2;
            }
        } else
#pragma unroll 3
        for (int k = 0; k < width; k++) {
            const int fb = rcol[(size_t)k * nbc + cb];
            const double w = rw[(size_t)k * nbc + cb];
            const uint8_t ff = ffree[fb];
            real v0, v1, v2;
            gather3<real, T>(Bf, fb, s, v0, v1, v2);
            acc[0] += (ff & 1) ? w * (double)v0 : 0.0;
            acc[1] += (ff & 2) ? w * (double)v1 : 0.0;
            acc[2] += (ff & 4) ? w * (double)v2 : 0.0;
        }
        const uint8_t cf = C.block_free[cb];
        if constexpr (T == 1) {
            const real v0 = (cf & 1) ? (real)acc[0] : (real)0;
            const real v1 = (cf & 2) ? (real)acc[1] : (real)0;
            const real v2 = (cf & 4) ? (real)acc[2] : (real)0;
            const vec3<real> dv = ld3(gbl(dinv), cb);
            st3(gbl(bc), cb, v0, v1, v2);
            st3(gbl(zc), cb, dv.x * v0, dv.y * v1, dv.z * v2);
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const size_t o = ((size_t)cb * 3 + d) * T + s;
                const real v = ((cf >> d) & 1) ? (real)acc[d] : (real)0;
                bc[o] = v;
                zc[o] = dinv[o] * v;
            }
        }
    }
}

// Coarse correction: A += P z_c on the free components.  Items: the finer
// level's.
template <typename real, int T>
__global__ __launch_bounds__(256) void k_mg_prolong(const ClassDev *__restrict__ cls,
                                                    const int4 *__restrict__ items, int nitems,
                                                    int mode)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const ClassDev &K = cls[C.mg_coarse];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    if (sample >= C.nsamp || !mg_act(cls[C.mg_root].st[sample], mode)) return;
    const MSA_GLOBAL real *zc = gbl((const real *)K.z + (size_t)tile * K.tile_vec);
    real *A = (real *)C.q + (size_t)tile * C.tile_vec;
    const MSA_GLOBAL int *pcol = gbl(C.mg_p_col);
    const MSA_GLOBAL float *pw = gbl(C.mg_p_w);
    const int width = C.mg_p_width, nbf = C.nb;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    const bool st = __builtin_amdgcn_readfirstlane(C.mg_pstruct);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int fb = slice * R + rr;
        if (fb >= C.nb) continue;
        double acc[3] = {0, 0, 0};
        if (st) {
            // coarse points: i/2 for even i, (i-1)/2 and (i+1)/2 with weight
            // 1/2 for odd i (per direction); coarse blocks = the coarse box.
            // Branch-free: the second point of an even direction has weight
            // 0, and points outside the coarse box are read past the end of
            // the vector (zeros).
            const int g = C.mf_bpos[fb];
            const int i = g & 1023, j = (g >> 10) & 1023, k2 = g >> 20;
            const int lx = K.mf_box_lo[0], ly = K.mf_box_lo[1], lz = K.mf_box_lo[2];
            const int hx = K.mf_box_hi[0], hy = K.mf_box_hi[1], hz = K.mf_box_hi[2];
            const int sy = hx - lx + 1, sz = sy * (hy - ly + 1);
            const int nbc = K.nb;
            const __amdgpu_buffer_rsrc_t cb = wave_buffer((const void *)zc, (unsigned)nbc * 3u * sizeof(real));
            const unsigned cout = (unsigned)nbc * 3u * sizeof(real);
            const double wx = (i & 1) ? 0.5 : 1.0, wy = (j & 1) ? 0.5 : 1.0, wz = (k2 & 1) ? 0.5 : 1.0;
#pragma unroll
            for (int n = 0; n < 8; n++) {
                const int ox = n & 1, oy = (n >> 1) & 1, oz = n >> 2;
                const int x = (i >> 1) + ox, y = (j >> 1) + oy, z = (k2 >> 1) + oz;
                // points of weight 0 (the second point of an even direction)
                // are read past the end too, so they fetch nothing
                const bool used = (!ox || (i & 1)) && (!oy || (j & 1)) && (!oz || (k2 & 1));
                const bool in = used && x >= lx && x <= hx && y >= ly && y <= hy && z >= lz && z <= hz;
                const unsigned off =
                    in ? (unsigned)((x - lx) + sy * (y - ly) + sz * (z - lz)) * 3u * (unsigned)sizeof(real)
                       : cout;
                const double w = wx * wy * wz;
                real v0, v1, v2;
                if constexpr (sizeof(real) == 4) {
                    const f32x3 v = buffer_load_f32x3(cb, (int)off, 0, 0);
                    v0 = v.x;
                    v1 = v.y;
                    v2 = v.z;
                } else {
                    v0 = buffer_load_f64(cb, (int)off, 0, 0);
                    v1 = buffer_load_f64(cb, (int)(off == cout ? cout : off + 8u), 0, 0);
                    v2 = buffer_load_f64(cb, (int)(off == cout ? cout : off + 16u), 0, 0);
                }
                acc[0] += w * (double)v0;
                acc[1] += w * (double)v1;
                acc[2] += w * (double)v2;
            }
        } else
#pragma unroll 4
        for (int k = 0; k < width; k++) {
            const int cb = pcol[(size_t)k * nbf + fb];
            const double w = pw[(size_t)k * nbf + fb];
            real v0, v1, v2;
            gather3<real, T>(zc, cb, s, v0, v1, v2);
            acc[0] += w * (double)v0;
            acc[1] += w * (double)v1;
            acc[2] += w * (double)v2;
        }
        const uint8_t ff = C.block_free[fb];
        if constexpr (T == 1) {
            // the node's three components move as one 12-byte access; fixed
            // components are written back unchanged
            MSA_GLOBAL real *Ag = gbl(A);
            const vec3<real> a = ld3(Ag, fb);
            st3(Ag, fb, (ff & 1) ? (real)((double)a.x + acc[0]) : a.x,
                (ff & 2) ? (real)((double)a.y + acc[1]) : a.y,
                (ff & 4) ? (real)((double)a.z + acc[2]) : a.z);
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++)
                if ((ff >> d) & 1) {
                    const size_t o = ((size_t)fb * 3 + d) * T + s;
                    A[o] = (real)((double)A[o] + acc[d]);
                }
        }
    }
}

// Modulus of a coarse element = mean of its eight children's (fixed order).
template <int T>
__global__ void k_mg_coarsen_E(const ClassDev *__restrict__ cp, const ClassDev *__restrict__ cls)
{
    const ClassDev &C = *cp;
    const ClassDev &F = cls[C.mg_fine];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)C.ntiles * C.n_elem * T;
    if (i >= n) return;
    const int s = (int)(i % T);
    const size_t te = i / T;
    const int e = (int)(te % C.n_elem), tile = (int)(te / C.n_elem);
    double acc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
        acc += F.E[((size_t)tile * F.n_elem + C.mg_child[8 * (size_t)e + k]) * T + s];
    C.E[i] = 0.125 * acc;
}

// lam = max over free rows of sum_j |K_ij| / K_ii (Gershgorin bound of the
// spectrum of D^-1 K); mg_lam must hold 0 before.  The maximum of
// non-negative doubles is taken on their bit patterns (order-independent).
template <typename real, int T>
__global__ __launch_bounds__(256) void k_mg_gersh(const ClassDev *__restrict__ cls,
                                                  const int4 *__restrict__ items, int nitems)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    if (C.mf_allfree) return;  // k_mg_gersh_mf
    const real *val = (const real *)C.val + (size_t)tile * C.tile_vals;
    const real *dinv = (const real *)C.dinv + (size_t)tile * C.tile_vec;
    double g = 0;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    if (sample < C.nsamp) {
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * R + rr;
            if (br >= C.nb) continue;
            const int base = C.slice_ptr[slice], W = C.slice_ptr[slice + 1] - base;
            double sum[3] = {0, 0, 0};
            for (int j = 0; j < W; j++) {
                real m[9];
                load_row_block<real, T, true>(C, gbl(val), base, j, C.cidx[(size_t)(base + j) * R + rr],
                                              rr, s, m);
#pragma unroll
                for (int p = 0; p < 3; p++)
                    sum[p] += fabs((double)m[3 * p]) + fabs((double)m[3 * p + 1]) +
                              fabs((double)m[3 * p + 2]);
            }
#pragma unroll
            for (int p = 0; p < 3; p++)
                g = fmax(g, sum[p] * fabs((double)dinv[((size_t)br * 3 + p) * T + s]));
        }
    }
#pragma unroll
    for (int off = T; off < 64; off <<= 1) g = fmax(g, __shfl_xor(g, off, 64));
    if (rr == 0 && sample < C.nsamp && g > 0)
        atomicMax((unsigned long long *)&C.mg_lam[sample], (unsigned long long)__double_as_longlong(g));
}

// The same bound matrix-free, on box grids of fully free blocks (one-sample
// tiles, mixed precision): per node, each of the 27 neighbour blocks is
// formed from the moduli of the cells the two nodes share and the FP32 unit
// element matrix (+ the consistent mass on the diagonal in the dynamics
// mode), its absolute row sums accumulated; lambda = max over rows of
// sum_j |K_ij| / K_ii.  Neighbours outside the box are constrained nodes.
__global__ __launch_bounds__(256) void k_mg_gersh_mf(const ClassDev *__restrict__ cls,
                                                     const int4 *__restrict__ items, int nitems)
{
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    if (!C.mf_allfree) return;
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sample = tile;
    float g = 0.f;
    if (sample < C.nsamp) {
        const int nx = __builtin_amdgcn_readfirstlane(C.mf_nx);
        const int ny = __builtin_amdgcn_readfirstlane(C.mf_ny);
        const int nz = __builtin_amdgcn_readfirstlane(C.mf_nz);
        const int b0x = C.mf_box_lo[0], b0y = C.mf_box_lo[1], b0z = C.mf_box_lo[2];
        const int b1x = C.mf_box_hi[0], b1y = C.mf_box_hi[1], b1z = C.mf_box_hi[2];
        const float cK = C.dyn ? (float)C.cK : 1.f, cM = C.dyn ? (float)C.cM : 0.f;
        const double *Es = C.E + (size_t)tile * C.n_elem;
        const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk * max(1, it.w), C.nslices);
        for (int slice = s0 + wave; slice < s1; slice += kWaves) {
            const int br = slice * 64 + lane;
            if (br >= C.nb) continue;
            const int gp = gbl(C.mf_bpos)[br];
            const int i = gp & 1023, j = (gp >> 10) & 1023, k = gp >> 20;
            const bool lo[3] = {i > 0, j > 0, k > 0}, hi[3] = {i < nx - 1, j < ny - 1, k < nz - 1};
            const bool blo[3] = {i > b0x, j > b0y, k > b0z}, bhi[3] = {i < b1x, j < b1y, k < b1z};
            float Ec[8], cw[8];
            mf_moduli(C, Es, i, j, k, nx, ny, lo, hi, cK, Ec, cw);
            float cm[8];  // mass weight per cell (zero without dynamics)
#pragma unroll
            for (int c = 0; c < 8; c++) cm[c] = cw[c] * cM;
            float rs[3] = {0.f, 0.f, 0.f}, dg[3] = {1.f, 1.f, 1.f};
#pragma unroll
            for (int n = 0; n < 27; n++) {
                // the element-matrix pointer redefined per neighbour: its
                // scalar loads stay with their neighbour
                cfloat *kh = mf_khat(C), *mh = kh + 576;
                const int dx = n % 3 - 1, dy = (n / 3) % 3 - 1, dz = n / 9 - 1;
                const bool in = (dx >= 0 || blo[0]) && (dx <= 0 || bhi[0]) && (dy >= 0 || blo[1]) &&
                                (dy <= 0 || bhi[1]) && (dz >= 0 || blo[2]) && (dz <= 0 || bhi[2]);
                float B[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
                    // cell c spans node offsets {c* - 1, c*} in each direction
                    if (dx < cx - 1 || dx > cx || dy < cy - 1 || dy > cy || dz < cz - 1 || dz > cz) continue;
                    const int a = (1 - cx) + 2 * (1 - cy) + 4 * (1 - cz);
                    const int bb = (dx - cx + 1) + 2 * (dy - cy + 1) + 4 * (dz - cz + 1);
                    cfloat *kab = kh + (a * 8 + bb) * 9;
#pragma unroll
                    for (int q = 0; q < 9; q++) B[q] = __builtin_fmaf(Ec[c], kab[q], B[q]);
                    const float mab = mh[a * 8 + bb];
                    B[0] = __builtin_fmaf(cm[c], mab, B[0]);
                    B[4] = __builtin_fmaf(cm[c], mab, B[4]);
                    B[8] = __builtin_fmaf(cm[c], mab, B[8]);
                }
                const float f = in ? 1.f : 0.f;
#pragma unroll
                for (int d = 0; d < 3; d++)
                    rs[d] = __builtin_fmaf(f, fabsf(B[3 * d]) + fabsf(B[3 * d + 1]) + fabsf(B[3 * d + 2]),
                                           rs[d]);
                if (n == 13) {
                    dg[0] = B[0];
                    dg[1] = B[4];
                    dg[2] = B[8];
                }
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int d = 0; d < 3; d++) g = fmaxf(g, rs[d] / fabsf(dg[d]));
        }
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) g = fmaxf(g, __shfl_xor(g, off, 64));
    if (lane == 0 && sample < C.nsamp && g > 0.f)
        atomicMax((unsigned long long *)&C.mg_lam[sample],
                  (unsigned long long)__double_as_longlong((double)g));
}

// One sample's level-0 right-hand side for a single cycle: r = src,
// z = D^-1 r (the state a PCG update leaves).
template <typename real, int T>
__global__ void k_mg_load(const ClassDev *__restrict__ cp, const double *__restrict__ src,
                          int tile, int lane)
{
    const ClassDev &C = *cp;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= 3 * C.nb) return;
    const size_t o = (size_t)tile * C.tile_vec + (size_t)row * T + lane;
    const real v = (real)src[row];
    ((real *)C.r)[o] = v;
    ((real *)C.z)[o] = ((const real *)C.dinv)[o] * v;
}

// Dense inverse X = (L L^T)^-1 of every sample's band Cholesky factor
// (coarsest multigrid levels, bn <= kDenseMax): one wavefront per column j
// of one sample; forward L y = e_j and backward L^T x = y with the column
// in LDS, each row's band sum split over the 64 lanes and reduced by
// shuffles.  X is symmetric: column j is written as row j (contiguous).
__global__ __launch_bounds__(64) void k_band_inverse(const ClassDev *__restrict__ cp)
{
    __shared__ double y[kDenseMax];
    const ClassDev &C = *cp;
    const int smp = blockIdx.y, j = blockIdx.x, lane = threadIdx.x;
    const int n = C.bn, W = C.bW;
    const MSA_GLOBAL double *L = gbl(C.band + (size_t)smp * C.band_ss);
    for (int i = lane; i < n; i += 64) y[i] = 0.0;
    __syncthreads();
    // forward: y_i = (delta_ij - sum_{k in [max(j, i-W+1), i)} L_ik y_k) / L_ii
    for (int i = j; i < n; i++) {
        double part = 0.0;
        for (int k2 = max(j, i - W + 1) + lane; k2 < i; k2 += 64)
            part += L[(size_t)k2 * W + (i - k2)] * y[k2];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
        const double yi = ((i == j ? 1.0 : 0.0) - part) / L[(size_t)i * W];
        __syncthreads();
        if (lane == 0) y[i] = yi;
        __syncthreads();
    }
    // backward: x_i = (y_i - sum_{k in (i, min(n, i+W))} L_ki x_k) / L_ii
    for (int i = n - 1; i >= 0; i--) {
        double part = 0.0;
        const int kend = min(n, i + W);
        for (int k2 = i + 1 + lane; k2 < kend; k2 += 64)
            part += L[(size_t)i * W + (k2 - i)] * y[k2];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
        const double xi = (y[i] - part) / L[(size_t)i * W];
        __syncthreads();
        if (lane == 0) y[i] = xi;
        __syncthreads();
    }
    MSA_GLOBAL double *X = gbl(C.binvd + (size_t)smp * n * n + (size_t)j * n);
    for (int i = lane; i < n; i += 64) X[i] = y[i];
}

// Coarsest solve by the dense inverse: z[bperm[i]] = sum_j X[j][i] b[bperm[j]]
// (X symmetric: lane i reads row j at column i, so loads are coalesced).
template <typename real, int T>
__global__ __launch_bounds__(64) void k_mg_dense_solve(const ClassDev *__restrict__ cp, int mode)
{
    const ClassDev &C = *cp;
    const int smp = blockIdx.y;
    if (!mg_act(C.mg_st[smp], mode)) return;
    const int n = C.bn;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int tile = smp / T, s = smp % T;
    const size_t vb = (size_t)tile * C.tile_vec;
    const real *b = (const real *)C.r + vb;
    const MSA_GLOBAL double *X = gbl(C.binvd + (size_t)smp * n * n);
    // the right-hand side in band order, staged once per workgroup, so the
    // loads of the inverse's columns do not wait on index loads
    __shared__ double bl[kDenseMax];
    for (int jj = threadIdx.x; jj < n; jj += 64) bl[jj] = (double)b[(size_t)C.bperm[jj] * T + s];
    __syncthreads();
    double acc = 0.0;
    if (i < n) {
#pragma unroll 8
        for (int jj = 0; jj < n; jj++) acc += X[(size_t)jj * n + i] * bl[jj];
        ((real *)C.z)[vb + (size_t)C.bperm[i] * T + s] = (real)acc;
    }
}

// Coarsest level: its right-hand side (r) into band order, and the band
// solution back into z.
template <typename real, int T>
__global__ void k_mg_band_in(const ClassDev *__restrict__ cp, int mode)
{
    const ClassDev &C = *cp;
    const int smp = blockIdx.y;
    if (!mg_act(C.mg_st[smp], mode)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.bn) return;
    const int tile = smp / T, s = smp % T;
    C.by[(size_t)smp * C.bn + i] =
        (double)((const real *)C.r)[(size_t)tile * C.tile_vec + (size_t)C.bperm[i] * T + s];
}

template <typename real, int T>
__global__ void k_mg_band_out(const ClassDev *__restrict__ cp, int mode)
{
    const ClassDev &C = *cp;
    const int smp = blockIdx.y;
    if (!mg_act(C.mg_st[smp], mode)) return;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= 3 * C.nb) return;
    const int tile = smp / T, s = smp % T;
    const int i = C.binv[row];
    ((real *)C.z)[(size_t)tile * C.tile_vec + (size_t)row * T + s] =
        i >= 0 ? (real)C.by[(size_t)smp * C.bn + i] : (real)0;
}

// First-sample direct preconditioner (the Monte Carlo calculator's
// LDL^T of the first sample's K, applied to every sample): z = the band
// solution of the solved class's rows (constrained components 0), and the
// r.z partials of the PCG (slot 0, as the multigrid cycle's last step
// writes them) for the samples the mode selects.
template <typename real, int T>
__global__ __launch_bounds__(256) void k_first_out(const ClassDev *__restrict__ cls,
                                                   const int4 *__restrict__ items, int nitems,
                                                   int mode)
{
    constexpr int R = 64 / T;
    const int item = xcd_item(nitems);
    if (item < 0) return;
    const int4 it = items[item];
    const ClassDev &C = cls[it.x];
    const int tile = it.y, chunk = it.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = lane / T, s = lane % T;
    const int sample = tile * T + s;
    const bool act = sample < C.nsamp && mg_act(C.st[sample], mode);
    const size_t vb = (size_t)tile * C.tile_vec;
    const real *r = (const real *)C.r + vb;
    real *z = (real *)C.z + vb;
    const double *y = C.by + (size_t)(act ? sample : 0) * C.bn;
    double rz = 0.0;
    const int s0 = chunk * C.chunk, s1 = min(s0 + C.chunk, C.nslices);
    for (int slice = s0 + wave; slice < s1; slice += kWaves) {
        const int br = slice * R + rr;
        if (!act || br >= C.nb) continue;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int row = 3 * br + d;
            const int i = C.binv[row];
            const size_t o = (size_t)row * T + s;
            const real zv = i >= 0 ? (real)y[i] : (real)0;
            z[o] = zv;
            rz += (double)(r[o] * zv);
        }
    }
    double acc[1] = {rz};
    block_partials<T, 1>(acc, C.partial + ((size_t)tile * C.nchunks + chunk) * 3 * T);
}

// ------------------------------------------------------------ utilities

template <typename V>
__global__ void k_fill(V *p, size_t n, V v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <typename real>
__global__ void k_extract(const real *__restrict__ src, size_t n, int T, int lane,
                          double *__restrict__ dst)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (double)src[i * T + lane];
}

__global__ void k_insert(const double *__restrict__ src, size_t n, int T, int lane,
                         double *__restrict__ dst)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i * T + lane] = src[i];
}

// ------------------------------------------------------------ launchers

static inline unsigned grid1d(size_t n, unsigned block)
{
    size_t g = (n + block - 1) / block;
    if (g > 1u << 20) g = 1u << 20;
    return (unsigned)(g ? g : 1);
}

hipError_t launch_fill_f64(hipStream_t s, double *p, size_t n, double v)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill<double>, dim3(grid1d(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_fill_f32(hipStream_t s, float *p, size_t n, float v)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill<float>, dim3(grid1d(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_fill_u32(hipStream_t s, unsigned *p, size_t n, unsigned v)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill<unsigned>, dim3(grid1d(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_ke_packed(hipStream_t s, int n_elem, const double *xyz8, const double *E,
                            double nu, double *ke300, int *status)
{
    if (n_elem <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ke_packed, dim3(n_elem), dim3(64), 0, s, n_elem, xyz8, E, nu, ke300,
                       status);
    return hipGetLastError();
}

hipError_t launch_gauss_matrices(hipStream_t s, const double *xyz, double *out)
{
    hipLaunchKernelGGL(k_gauss_matrices, dim3(1), dim3(64), 0, s, xyz, out);
    return hipGetLastError();
}

hipError_t launch_h8strains(hipStream_t s, const double *B, const double *u, double *strains)
{
    hipLaunchKernelGGL(k_h8strains, dim3(1), dim3(64), 0, s, B, u, strains);
    return hipGetLastError();
}

hipError_t launch_h8forces(hipStream_t s, const double *B, const double *w, const double *stresses,
                           double *forces)
{
    hipLaunchKernelGGL(k_h8forces, dim3(1), dim3(64), 0, s, B, w, stresses, forces);
    return hipGetLastError();
}

hipError_t launch_h8mlumped(hipStream_t s, double density, const double *w, double *M)
{
    hipLaunchKernelGGL(k_h8mlumped, dim3(1), dim3(64), 0, s, density, w, M);
    return hipGetLastError();
}

hipError_t launch_h8k(hipStream_t s, const double *E, const double *B, const double *w, double *K)
{
    hipLaunchKernelGGL(k_h8k, dim3(5), dim3(64), 0, s, E, B, w, K);
    return hipGetLastError();
}

#define MSA_DISPATCH_T(T, BODY)                                                   \
    switch (T) {                                                                  \
    case 1: { constexpr int TT = 1; BODY; } break;                                \
    case 2: { constexpr int TT = 2; BODY; } break;                                \
    case 4: { constexpr int TT = 4; BODY; } break;                                \
    case 8: { constexpr int TT = 8; BODY; } break;                                \
    case 16: { constexpr int TT = 16; BODY; } break;                              \
    case 32: { constexpr int TT = 32; BODY; } break;                              \
    case 64: { constexpr int TT = 64; BODY; } break;                              \
    default: return hipErrorInvalidValue;                                         \
    }

// Waves per element along blockIdx.y: enough to fill the chip even for the
// small colours of small meshes, at most one per tile.
static inline unsigned tile_waves(int nblocks, int ntiles)
{
    const int want = (32768 + nblocks - 1) / std::max(nblocks, 1);
    return (unsigned)std::max(1, std::min(std::min(ntiles, want), 1024));
}

hipError_t launch_assemble(hipStream_t s, int T, const ClassDev *d_cls, const ColourLaunch &cl,
                           int nblocks, double nu, int ntiles)
{
    if (nblocks <= 0) return hipSuccess;
    const dim3 grid(nblocks, tile_waves(nblocks, ntiles));
    MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_assemble<TT>), grid, dim3(64), 0, s, d_cls, cl, nu));
    return hipGetLastError();
}

hipError_t launch_element_results(hipStream_t s, int T, const ClassDev *d_cls_c, int n_elem,
                                  const double *u, int lane, int tile, double nu, double *strains,
                                  double *stresses, double *forces)
{
    if (n_elem <= 0) return hipSuccess;
    MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_element_results<TT>), dim3(n_elem), dim3(64), 0, s,
                                         d_cls_c, u, lane, tile, nu, strains, stresses, forces));
    return hipGetLastError();
}

hipError_t launch_khat_cache(hipStream_t s, const ClassDev *d_cls_c, int n_elem, double nu,
                             int with_mass)
{
    if (n_elem <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_khat_cache, dim3(n_elem), dim3(64), 0, s, d_cls_c, nu, with_mass);
    return hipGetLastError();
}


hipError_t launch_residual64(hipStream_t s, int T, const ClassDev *d_cls, const ColourLaunch &cl,
                             int nblocks, double nu, int need_due, int ntiles)
{
    if (nblocks <= 0) return hipSuccess;
    // replacements: up to 4 waves per element share the due list (each
    // stages the element's 4.6 KB Khat; waves without a due sample exit
    // before staging)
    const dim3 grid(nblocks, need_due ? (unsigned)std::max(1, std::min(4, ntiles))
                                      : tile_waves(nblocks, ntiles));
    MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_residual64<TT>), grid, dim3(64), 0, s, d_cls, cl, nu,
                                         need_due));
    return hipGetLastError();
}

#define MSA_PCG_LAUNCH(KERNEL, ...)                                                       \
    do {                                                                                  \
        if (nitems <= 0) return hipSuccess;                                               \
        const dim3 grid(xcd_grid(nitems)), block(256);                                    \
        if (mixed) {                                                                      \
            MSA_DISPATCH_T(T, hipLaunchKernelGGL((KERNEL<float, TT>), grid, block, 0, s,  \
                                                 __VA_ARGS__));                           \
        } else {                                                                          \
            MSA_DISPATCH_T(T, hipLaunchKernelGGL((KERNEL<double, TT>), grid, block, 0, s, \
                                                 __VA_ARGS__));                           \
        }                                                                                 \
        return hipGetLastError();                                                         \
    } while (0)

hipError_t launch_dinv(hipStream_t s, bool mixed, int T, const ClassDev *d_cls, const int4 *items,
                       int nitems)
{
    MSA_PCG_LAUNCH(k_dinv, d_cls, items, nitems);
}

hipError_t launch_pcg_init(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                           const int4 *items, int nitems, int mode, int *active_count)
{
    MSA_PCG_LAUNCH(k_pcg_init, d_cls, items, nitems, mode, active_count);
}

template <typename real>
static hipError_t pcg_a_unrolled(hipStream_t s, int U, dim3 grid, const ClassDev *d_cls,
                                 const int4 *items, int nitems, int pbuf, int *active_count)
{
    switch (U) {
    case 1: hipLaunchKernelGGL((k_pcg_a<real, 1, 1>), grid, dim3(256), 0, s, d_cls, items, nitems, pbuf, active_count); break;
    case 2: hipLaunchKernelGGL((k_pcg_a<real, 1, 2>), grid, dim3(256), 0, s, d_cls, items, nitems, pbuf, active_count); break;
    case 3: hipLaunchKernelGGL((k_pcg_a<real, 1, 3>), grid, dim3(256), 0, s, d_cls, items, nitems, pbuf, active_count); break;
    case 4: hipLaunchKernelGGL((k_pcg_a<real, 1, 4>), grid, dim3(256), 0, s, d_cls, items, nitems, pbuf, active_count); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_pcg_a(hipStream_t s, bool mixed, int T, int U, const ClassDev *d_cls,
                        const int4 *items, int nitems, int pbuf, int *active_count)
{
    if (T == 1 && U != 3) {
        if (nitems <= 0) return hipSuccess;
        const dim3 grid(xcd_grid(nitems));
        return mixed ? pcg_a_unrolled<float>(s, U, grid, d_cls, items, nitems, pbuf, active_count)
                     : pcg_a_unrolled<double>(s, U, grid, d_cls, items, nitems, pbuf, active_count);
    }
    if (U != 3) return hipErrorInvalidValue;  // lane tiles T > 1 use groups of 3 slots
    MSA_PCG_LAUNCH(k_pcg_a, d_cls, items, nitems, pbuf, active_count);
}

hipError_t launch_pcg_a_mf(hipStream_t s, const ClassDev *d_cls, const int4 *items, int nitems,
                           int pbuf)
{
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pcg_p, dim3(xcd_grid(nitems)), dim3(256), 0, s, d_cls, items, nitems,
                       pbuf);
    hipLaunchKernelGGL(k_pcg_a_mf, dim3(xcd_grid(nitems)), dim3(256), 0, s, d_cls, items, nitems,
                       pbuf);
    return hipGetLastError();
}

template <typename real, int U>
static hipError_t pcg_persist_u(hipStream_t s, const ClassDev *d_cls, const int4 *items, int nitems,
                                const int2 *tiles, int ntiles, const ReduceArgs &ra,
                                const ReduceArgs &rb, int iters, int *grid_out)
{
    const void *fn = (const void *)k_pcg_persist<real, 1, U>;
    int dev = 0, ncu = 0, coop = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0);
    if (e != hipSuccess) return e;
    if (!coop || per_cu <= 0) return hipErrorNotSupported;
    // co-resident workgroups, a multiple of the 8 XCDs, no more than the items need
    int G = (per_cu * ncu) & ~7;
    G = std::min(G, ((nitems + 7) / 8) * 8);
    if (G < 8) return hipErrorInvalidConfiguration;
    *grid_out = G;
    int pbuf0 = 0;
    ReduceArgs a = ra, b = rb;
    void *args[] = {(void *)&d_cls, (void *)&items, (void *)&nitems, (void *)&tiles, (void *)&ntiles,
                    (void *)&a,     (void *)&b,     (void *)&pbuf0,  (void *)&iters};
    return hipLaunchCooperativeKernel(fn, dim3(G), dim3(256), args, 0, s);
}

hipError_t launch_pcg_persist(hipStream_t s, bool mixed, int U, const ClassDev *d_cls,
                              const int4 *items, int nitems, const int2 *tiles, int ntiles,
                              const ReduceArgs &ra, const ReduceArgs &rb, int iters, int *grid_out)
{
    if (nitems <= 0) return hipSuccess;
#define MSA_PERSIST_U(real)                                                                   \
    switch (U) {                                                                             \
    case 1: return pcg_persist_u<real, 1>(s, d_cls, items, nitems, tiles, ntiles, ra, rb, iters, grid_out); \
    case 2: return pcg_persist_u<real, 2>(s, d_cls, items, nitems, tiles, ntiles, ra, rb, iters, grid_out); \
    case 3: return pcg_persist_u<real, 3>(s, d_cls, items, nitems, tiles, ntiles, ra, rb, iters, grid_out); \
    case 4: return pcg_persist_u<real, 4>(s, d_cls, items, nitems, tiles, ntiles, ra, rb, iters, grid_out); \
    default: return hipErrorInvalidValue;                                                    \
    }
    if (mixed) {
        MSA_PERSIST_U(float)
    } else {
        MSA_PERSIST_U(double)
    }
#undef MSA_PERSIST_U
}

hipError_t launch_pcg_b(hipStream_t s, bool mixed, int T, const ClassDev *d_cls, const int4 *items,
                        int nitems, int pbuf, double tol, int max_it, int *active_count)
{
    MSA_PCG_LAUNCH(k_pcg_b, d_cls, items, nitems, pbuf, tol, max_it, active_count);
}

hipError_t launch_spmv(hipStream_t s, bool mixed, int T, const ClassDev *d_cls, const int4 *items,
                       int nitems, const double *x, double *y)
{
    MSA_PCG_LAUNCH(k_spmv, d_cls, items, nitems, x, y);
}

hipError_t launch_assemble_gather(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                                  const int4 *items, int nitems)
{
    MSA_PCG_LAUNCH(k_assemble_gather, d_cls, items, nitems);
}

hipError_t launch_assemble_gather_g(hipStream_t s, bool mixed, const ClassDev *d_cls,
                                    const int4 *gitems, int nitems)
{
    if (nitems <= 0) return hipSuccess;
    const dim3 grid(xcd_grid(nitems)), block(256);
    if (mixed)
        hipLaunchKernelGGL((k_assemble_gather_g<float, kAsmGroup>), grid, block, 0, s, d_cls, gitems,
                           nitems);
    else
        hipLaunchKernelGGL((k_assemble_gather_g<double, kAsmGroup>), grid, block, 0, s, d_cls, gitems,
                           nitems);
    return hipGetLastError();
}

#define MSA_T_LAUNCH(KERNEL, ...)                                                          \
    do {                                                                                   \
        if (nitems <= 0) return hipSuccess;                                                \
        const dim3 grid(xcd_grid(nitems)), block(256);                                     \
        MSA_DISPATCH_T(T, hipLaunchKernelGGL((KERNEL<TT>), grid, block, 0, s, __VA_ARGS__)); \
        return hipGetLastError();                                                          \
    } while (0)

hipError_t launch_ir_begin(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                           int nitems, int *ir_count, int warm)
{
    MSA_T_LAUNCH(k_ir_begin, d_cls, items, nitems, ir_count, warm);
}

hipError_t launch_ir_update(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                            int nitems)
{
    MSA_T_LAUNCH(k_ir_update, d_cls, items, nitems);
}

hipError_t launch_ir_norm(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                          int nitems, double tol, int max_outer, int *ir_count)
{
    MSA_T_LAUNCH(k_ir_norm, d_cls, items, nitems, tol, max_outer, ir_count);
}

hipError_t launch_ru_fold(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                          int nitems)
{
    MSA_T_LAUNCH(k_ru_fold, d_cls, items, nitems);
}

hipError_t launch_ru_norm(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                          int nitems, double tol, int max_it, int *ir_count, int *active_count)
{
    MSA_T_LAUNCH(k_ru_norm, d_cls, items, nitems, tol, max_it, ir_count, active_count);
}

hipError_t launch_ru_restart(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                             int nitems)
{
    MSA_T_LAUNCH(k_ru_restart, d_cls, items, nitems);
}

hipError_t launch_reduce(hipStream_t s, int T, const ClassDev *d_cls, const int2 *tiles,
                         int ntiles, const ReduceArgs &a)
{
    if (ntiles <= 0) return hipSuccess;
    MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_reduce<TT>), dim3(ntiles), dim3(1024), 0, s, d_cls,
                                         tiles, a));
    return hipGetLastError();
}

hipError_t launch_band_scatter(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                               int nitems)
{
    const bool mixed = false;
    MSA_PCG_LAUNCH(k_band_scatter, d_cls, items, nitems);
}

hipError_t launch_band_scatter_any(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                                   const int4 *items, int nitems)
{
    MSA_PCG_LAUNCH(k_band_scatter, d_cls, items, nitems);
}

hipError_t launch_chol_diag(hipStream_t s, const ClassDev *cp, int nsamp, int k0, int nk)
{
    hipLaunchKernelGGL(k_chol_diag, dim3(1, nsamp), dim3(64), 0, s, cp, k0, nk);
    return hipGetLastError();
}

hipError_t launch_chol_panel(hipStream_t s, const ClassDev *cp, int nsamp, int k0, int nk, int r0,
                             int m)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_chol_panel, dim3((m + kBandNB - 1) / kBandNB, nsamp), dim3(256), 0, s, cp,
                       k0, nk, r0, m);
    return hipGetLastError();
}

hipError_t launch_chol_update(hipStream_t s, const ClassDev *cp, int nsamp, int k0, int nk, int r0,
                              int m)
{
    if (m <= 0) return hipSuccess;
    const int T = (m + kBandNB - 1) / kBandNB;
    hipLaunchKernelGGL(k_chol_update, dim3(T * (T + 1) / 2, nsamp), dim3(256), 0, s, cp, k0, nk,
                       r0, m);
    return hipGetLastError();
}

hipError_t launch_band_rhs(hipStream_t s, int T, const ClassDev *cp, int nsamp, int bn,
                           int from_fs)
{
    if (bn <= 0) return hipSuccess;
    MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_band_rhs<TT>), dim3((bn + 255) / 256, nsamp), dim3(256),
                                         0, s, cp, from_fs));
    return hipGetLastError();
}

hipError_t launch_band_fwd(hipStream_t s, const ClassDev *cp, int nsamp, int k0, int nk, int r0,
                           int m, int mode)
{
    hipLaunchKernelGGL(k_band_fwd, dim3(1, nsamp), dim3(64), 0, s, cp, k0, nk, r0, m, 0, mode);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || m <= 0) return e;
    hipLaunchKernelGGL(k_band_fwd, dim3((m + 63) / 64, nsamp), dim3(64), 0, s, cp, k0, nk, r0, m, 1,
                       mode);
    return hipGetLastError();
}

hipError_t launch_band_bwd(hipStream_t s, const ClassDev *cp, int nsamp, int k0, int nk, int r0,
                           int m, int mode)
{
    if (m > 0) {
        hipLaunchKernelGGL(k_band_bwd, dim3((m + kBandChunkRows - 1) / kBandChunkRows, nsamp),
                           dim3(64), 0, s, cp, k0, nk, r0, m, 0, mode);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_band_bwd, dim3(1, nsamp), dim3(64), 0, s, cp, k0, nk, r0, m, 1, mode);
    return hipGetLastError();
}

hipError_t launch_band_solve_one(hipStream_t s, const ClassDev *cp, int nsamp, int bn, int mode)
{
    if (bn <= 0) return hipSuccess;
    if (bn > kBandOneMax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_band_solve_one, dim3(nsamp), dim3(256), 0, s, cp, mode);
    return hipGetLastError();
}

hipError_t launch_band_out(hipStream_t s, int T, const ClassDev *cp, int nsamp, int nrows)
{
    if (nrows <= 0) return hipSuccess;
    MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_band_out<TT>), dim3((nrows + 255) / 256, nsamp),
                                         dim3(256), 0, s, cp));
    return hipGetLastError();
}

hipError_t launch_set_int(hipStream_t s, int *p, int n, int v)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_int, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_step_advance(hipStream_t s, int *step)
{
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, s, step);
    return hipGetLastError();
}

hipError_t launch_dyn_bacc(hipStream_t s, const ClassDev *d_cls, const ColourLaunch &cl,
                           int nblocks, double nu, int dir)
{
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dyn_bacc, dim3(nblocks), dim3(64), 0, s, d_cls, cl, nu, dir);
    return hipGetLastError();
}

hipError_t launch_dyn_reset(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                            const int4 *items, int nitems)
{
    MSA_T_LAUNCH(k_dyn_reset, d_cls, items, nitems, (int)mixed);
}

hipError_t launch_dyn_prep(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                           const int4 *items, int nitems, const DynCoef &dc)
{
    MSA_T_LAUNCH(k_dyn_prep, d_cls, items, nitems, (int)mixed, dc);
}

hipError_t launch_dyn_rhs(hipStream_t s, int T, const ClassDev *d_cls, const ColourLaunch &cl,
                          int nblocks, double nu, double betaR, int ntiles)
{
    if (nblocks <= 0) return hipSuccess;
    const int stiff = betaR != 0.0;
    const dim3 grid(nblocks, tile_waves(nblocks, T == 1 ? (ntiles + 1) / 2 : ntiles));
    MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_dyn_rhs<TT>), grid, dim3(64), 0, s, d_cls, cl, nu,
                                         stiff));
    return hipGetLastError();
}

hipError_t launch_fs_static(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                            int nitems)
{
    MSA_T_LAUNCH(k_fs_static, d_cls, items, nitems);
}

hipError_t launch_fs_norm(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                          int nitems)
{
    MSA_T_LAUNCH(k_fs_norm, d_cls, items, nitems);
}

hipError_t launch_pcg_init_warm(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                                int nitems)
{
    MSA_T_LAUNCH(k_pcg_init_warm, d_cls, items, nitems);
}

hipError_t launch_dyn_update(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                             const int4 *items, int nitems, const DynCoef &dc)
{
    MSA_T_LAUNCH(k_dyn_update, d_cls, items, nitems, (int)mixed, dc);
}

hipError_t launch_kl_basis(hipStream_t s, const ClassDev *d_cls, int n_elem, int M, int kmax,
                           const double *w1d, const double *a1d, const int *evpPos,
                           const double *eigs, const double *domain, double sigma)
{
    hipLaunchKernelGGL(k_kl_basis, dim3((n_elem + 127) / 128), dim3(128), 0, s, d_cls, n_elem, M,
                       kmax, w1d, a1d, evpPos, eigs, domain, sigma);
    return hipGetLastError();
}

hipError_t launch_kl_eval(hipStream_t s, int T, const ClassDev *d_cls, int ntiles, int n_elem,
                          int M)
{
    const size_t total = (size_t)((ntiles + kKlTiles - 1) / kKlTiles) * n_elem * T;
    const unsigned g = (unsigned)((total + 255) / 256);
    MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_kl_eval<TT>), dim3(g), dim3(256), 0, s, d_cls, ntiles,
                                         n_elem, M));
    return hipGetLastError();
}

hipError_t launch_extract_sample(hipStream_t s, bool mixed, int T, const void *src, size_t n_rows,
                                 int lane, double *dst)
{
    if (!n_rows) return hipSuccess;
    const unsigned g = (unsigned)((n_rows + 255) / 256);
    if (mixed)
        hipLaunchKernelGGL(k_extract<float>, dim3(g), dim3(256), 0, s, (const float *)src, n_rows,
                           T, lane, dst);
    else
        hipLaunchKernelGGL(k_extract<double>, dim3(g), dim3(256), 0, s, (const double *)src,
                           n_rows, T, lane, dst);
    return hipGetLastError();
}

hipError_t launch_insert_sample(hipStream_t s, int T, const double *src, size_t n_rows, int lane,
                                double *dst)
{
    if (!n_rows) return hipSuccess;
    const unsigned g = (unsigned)((n_rows + 255) / 256);
    hipLaunchKernelGGL(k_insert, dim3(g), dim3(256), 0, s, src, n_rows, T, lane, dst);
    return hipGetLastError();
}

hipError_t launch_extract_values(hipStream_t s, bool mixed, int T, const void *val, size_t n,
                                 int lane, double *dst)
{
    return launch_extract_sample(s, mixed, T, val, n, lane, dst);
}

template <int STEP>
static hipError_t mg_smooth_step(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                                 const int4 *items, int nitems, int pbuf, int mode, double ratio,
                                 int mf)
{
    if (nitems <= 0) return hipSuccess;
    const dim3 grid(xcd_grid(nitems)), block(256);
    if (mf != 0) {
        if (!mixed || T != 1) return hipErrorInvalidValue;
        if (mf == 2)
            hipLaunchKernelGGL((k_mg_smooth<float, 1, STEP, 2>), grid, block, 0, s, d_cls, items,
                               nitems, pbuf, mode, ratio);
        else
            hipLaunchKernelGGL((k_mg_smooth<float, 1, STEP, 1>), grid, block, 0, s, d_cls, items,
                               nitems, pbuf, mode, ratio);
    } else if (mixed) {
        MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_mg_smooth<float, TT, STEP>), grid, block, 0, s,
                                             d_cls, items, nitems, pbuf, mode, ratio));
    } else {
        MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_mg_smooth<double, TT, STEP>), grid, block, 0, s,
                                             d_cls, items, nitems, pbuf, mode, ratio));
    }
    return hipGetLastError();
}

hipError_t launch_mg_smooth(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                            const int4 *items, int nitems, int step, int pbuf, int mode,
                            double ratio, int mf)
{
    switch (step) {
    case MG_PRE: return mg_smooth_step<MG_PRE>(s, mixed, T, d_cls, items, nitems, pbuf, mode, ratio, mf);
    case MG_RESID: return mg_smooth_step<MG_RESID>(s, mixed, T, d_cls, items, nitems, pbuf, mode, ratio, mf);
    case MG_POST1: return mg_smooth_step<MG_POST1>(s, mixed, T, d_cls, items, nitems, pbuf, mode, ratio, mf);
    case MG_POST2: return mg_smooth_step<MG_POST2>(s, mixed, T, d_cls, items, nitems, pbuf, mode, ratio, mf);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_mg_smooth_mfma(hipStream_t s, const ClassDev *d_cls, const int4 *items,
                                 int nitems, int step, int pbuf, int mode, double ratio)
{
    if (nitems <= 0) return hipSuccess;
    const dim3 grid(xcd_grid(nitems)), block(256);
    switch (step) {
    case MG_PRE: hipLaunchKernelGGL(k_mg_smooth_mfma<MG_PRE>, grid, block, 0, s, d_cls, items, nitems, pbuf, mode, ratio); break;
    case MG_RESID: hipLaunchKernelGGL(k_mg_smooth_mfma<MG_RESID>, grid, block, 0, s, d_cls, items, nitems, pbuf, mode, ratio); break;
    case MG_POST1: hipLaunchKernelGGL(k_mg_smooth_mfma<MG_POST1>, grid, block, 0, s, d_cls, items, nitems, pbuf, mode, ratio); break;
    case MG_POST2: hipLaunchKernelGGL(k_mg_smooth_mfma<MG_POST2>, grid, block, 0, s, d_cls, items, nitems, pbuf, mode, ratio); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_mg_restrict(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                              const int4 *items, int nitems, int pbuf, int mode)
{
    MSA_PCG_LAUNCH(k_mg_restrict, d_cls, items, nitems, pbuf, mode);
}

hipError_t launch_mg_prolong(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                             const int4 *items, int nitems, int mode)
{
    MSA_PCG_LAUNCH(k_mg_prolong, d_cls, items, nitems, mode);
}

hipError_t launch_mg_coarsen_E(hipStream_t s, int T, const ClassDev *cp, const ClassDev *d_cls,
                               int ntiles, int n_elem)
{
    const size_t n = (size_t)ntiles * n_elem * T;
    if (!n) return hipSuccess;
    MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_mg_coarsen_E<TT>), dim3((unsigned)((n + 255) / 256)),
                                         dim3(256), 0, s, cp, d_cls));
    return hipGetLastError();
}

hipError_t launch_mg_gersh(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                           const int4 *items, int nitems)
{
    MSA_PCG_LAUNCH(k_mg_gersh, d_cls, items, nitems);
}

hipError_t launch_mg_gersh_mf(hipStream_t s, const ClassDev *d_cls, const int4 *items, int nitems)
{
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mg_gersh_mf, dim3(xcd_grid(nitems)), dim3(256), 0, s, d_cls, items, nitems);
    return hipGetLastError();
}

hipError_t launch_mg_band_in(hipStream_t s, bool mixed, int T, const ClassDev *cp, int nsamp,
                             int bn, int mode)
{
    if (bn <= 0) return hipSuccess;
    const dim3 grid((bn + 255) / 256, nsamp);
    if (mixed) {
        MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_mg_band_in<float, TT>), grid, dim3(256), 0, s, cp,
                                             mode));
    } else {
        MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_mg_band_in<double, TT>), grid, dim3(256), 0, s, cp,
                                             mode));
    }
    return hipGetLastError();
}

hipError_t launch_residual64_mf(hipStream_t s, const ClassDev *d_cls, const int4 *items, int nitems)
{
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_residual64_mf, dim3(xcd_grid(nitems)), dim3(256), 0, s, d_cls, items, nitems);
    return hipGetLastError();
}

hipError_t launch_residual64_mf_box(hipStream_t s, const ClassDev *d_cls, const int4 *items,
                                   int nitems)
{
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_residual64_mf_box, dim3(xcd_grid(nitems)), dim3(256), 0, s, d_cls, items,
                       nitems);
    return hipGetLastError();
}

hipError_t launch_mf_khat(hipStream_t s, const ClassDev *cp)
{
    hipLaunchKernelGGL(k_mf_khat, dim3(3), dim3(256), 0, s, cp);
    return hipGetLastError();
}

hipError_t launch_band_inverse(hipStream_t s, const ClassDev *cp, int nsamp, int bn)
{
    if (bn <= 0) return hipSuccess;
    if (bn > kDenseMax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_band_inverse, dim3(bn, nsamp), dim3(64), 0, s, cp);
    return hipGetLastError();
}

hipError_t launch_mg_dense_solve(hipStream_t s, bool mixed, int T, const ClassDev *cp, int nsamp,
                                 int bn, int mode)
{
    if (bn <= 0) return hipSuccess;
    const dim3 grid((bn + 63) / 64, nsamp);
    if (mixed) {
        MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_mg_dense_solve<float, TT>), grid, dim3(64), 0, s, cp,
                                             mode));
    } else {
        MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_mg_dense_solve<double, TT>), grid, dim3(64), 0, s, cp,
                                             mode));
    }
    return hipGetLastError();
}

hipError_t launch_mg_load(hipStream_t s, bool mixed, int T, const ClassDev *cp, const double *src,
                          size_t nrows, int tile, int lane)
{
    if (!nrows) return hipSuccess;
    const dim3 grid((unsigned)((nrows + 255) / 256));
    if (mixed) {
        MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_mg_load<float, TT>), grid, dim3(256), 0, s, cp, src,
                                             tile, lane));
    } else {
        MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_mg_load<double, TT>), grid, dim3(256), 0, s, cp, src,
                                             tile, lane));
    }
    return hipGetLastError();
}

hipError_t launch_first_out(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                           const int4 *items, int nitems, int mode)
{
    MSA_PCG_LAUNCH(k_first_out, d_cls, items, nitems, mode);
}

hipError_t launch_mg_band_out(hipStream_t s, bool mixed, int T, const ClassDev *cp, int nsamp,
                              int nrows, int mode)
{
    if (nrows <= 0) return hipSuccess;
    const dim3 grid((nrows + 255) / 256, nsamp);
    if (mixed) {
        MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_mg_band_out<float, TT>), grid, dim3(256), 0, s, cp,
                                             mode));
    } else {
        MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_mg_band_out<double, TT>), grid, dim3(256), 0, s, cp,
                                             mode));
    }
    return hipGetLastError();
}

hipError_t launch_lz_step(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                          const int4 *items, int nitems, int op, int k)
{
    MSA_PCG_LAUNCH(k_lz_step, d_cls, items, nitems, op, k);
}

hipError_t launch_lz_sum(hipStream_t s, int T, const ClassDev *d_cls, const int *list, int nlist,
                         int max_samp, int which, int k)
{
    if (nlist <= 0 || max_samp <= 0) return hipSuccess;
    MSA_DISPATCH_T(T, hipLaunchKernelGGL((k_lz_sum<TT>), dim3((max_samp + 3) / 4, nlist), dim3(256),
                                         0, s, d_cls, list, which, k));
    return hipGetLastError();
}

hipError_t launch_lz_eig(hipStream_t s, const ClassDev *d_cls, const int *list, int nlist,
                         int max_samp, int steps, double margin)
{
    if (nlist <= 0 || max_samp <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lz_eig, dim3((max_samp + 255) / 256, nlist), dim3(256), 0, s, d_cls, list,
                       steps, margin);
    return hipGetLastError();
}

}  // namespace msa
