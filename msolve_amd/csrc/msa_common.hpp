// Internal declarations shared by the msolve_amd sources (not part of the
// C ABI).  Device data layout of a batch class:
//
//   block rows      node blocks with at least one free DOF, in node order;
//                   internal vector index = 3*block + component
//   sliced ELL      slices of R = 64/T block rows; slice s owns slots
//                   [slice_ptr[s], slice_ptr[s+1]) per row (width = longest
//                   row of the slice); cidx[slot*R + rr] = block column or -1
//   values          val[tile][((slot*9 + e)*R + rr)*T + s], e = 3p+q of the
//                   3x3 block, s = sample lane within the tile
//   vectors         v[tile][(block*3 + d)*T + s]
//
// One wavefront covers R block rows x T samples, so a value load is always
// 64 consecutive words and a gathered vector load is T consecutive words.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/msolve_amd.h"

namespace msa {

constexpr int kMaxColours = 64;
constexpr int kWaves = 4;  // waves per workgroup of the PCG kernels

// Position of entry e (row-major in the 3x3 block) of stored slot `slot`,
// block row lane rr and sample lane s within one tile's values.
// One-sample tiles (T = 1): per slot, entries 0-7 in 16-byte groups per
// row lane (4 floats or 2 doubles), then entry 8 per row lane, so that a
// lane reads its block with 16-byte loads; T > 1: entry-major, then row
// lane, then sample lane.  Either way a slot holds 9 x 64 values.
template <typename real, int T>
__host__ __device__ constexpr size_t val_at(size_t slot, int e, int rr, int s)
{
    if constexpr (T == 1) {
        constexpr int W = 16 / (int)sizeof(real);
        return slot * 576 + (e < 8 ? (size_t)(e / W) * 64 * W + (size_t)rr * W + e % W
                                   : (size_t)512 + rr);
    } else {
        constexpr int R = 64 / T;
        return ((slot * 9 + e) * R + rr) * T + s;
    }
}

void set_error(const char *fmt, ...);

#define MSA_HIP(call)                                                         \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            ::msa::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,       \
                             hipGetErrorString(e_));                          \
            return MSA_ERR_HIP;                                               \
        }                                                                     \
    } while (0)

#define MSA_CHECK(cond, code, ...)                                            \
    do {                                                                      \
        if (!(cond)) {                                                        \
            ::msa::set_error(__VA_ARGS__);                                    \
            return (code);                                                    \
        }                                                                     \
    } while (0)

// Per-sample solver state (device).  One per sample lane of every tile.
struct SampleState {
    double rz, pq, alpha, beta;  // PCG scalars
    double rr, bnorm, relres;    // |r|^2, |b|, |r|/|b|
    double r64norm, scale;       // mixed: |f - K u|, inner right-hand-side scale
    double fnorm;                // |f| of a per-sample right-hand side (dynamics)
    int iters, status, active;   // inner PCG
    int outer, total_iters, ir_active, ru_due;  // ru_due: FP64 replacement pending
    int cum_iters;               // dynamics: iterations summed over the time steps
};

// Host-side mesh analysis (mesh.cpp).
struct MeshHost {
    int n_nodes = 0, n_elem = 0, n_free = 0;
    int nb = 0;                          // free node blocks
    std::vector<double> xyz;             // [n_nodes*3]
    std::vector<int> conn;               // [n_elem*8]
    std::vector<int> dof;                // [n_nodes*3] reference free-DOF ids
    std::vector<int> node_block;         // [n_nodes] block or -1
    std::vector<uint8_t> node_free;      // [n_nodes] bit d = DOF d free
    std::vector<int> block_node;         // [nb]
    std::vector<int> ref2int;            // [n_free] -> 3*block + d
    std::vector<int> brow_ptr, bcol;     // block adjacency, sorted
    std::vector<int> diag_pos;           // [nb] position of the diagonal
    std::vector<uint8_t> emap;           // [n_elem*64] slot of (a,b) or 255
    int max_row = 0;
    int ncolours = 0;
    std::vector<int> colour_ptr;         // [ncolours+1]
    std::vector<int> colour_elems;       // elements sorted by colour
    long long nnz_scalar = 0;
};

int build_mesh_host(MeshHost &m, int n_nodes, const double *xyz, int n_elem,
                    const int *conn, const int *dof, int n_free);
// Free DOFs in band order for the direct solver: perm[band] = internal row,
// inv[internal row] = band index or -1; bw = half bandwidth.
int band_order(const MeshHost &m, std::vector<int> &perm, std::vector<int> &inv, long long &bw);

// Grid transfer between a structured mesh and its coarsening by two in
// every direction (trilinear interpolation of node values).
struct MgTransfer {
    std::vector<int> child;                // [coarse n_elem*8] fine elements of each coarse element
    std::vector<int> p_ptr, p_col;         // per fine block: coarse blocks it interpolates from
    std::vector<double> p_w;               //   and their weights
    std::vector<int> r_ptr, r_col;         // per coarse block: fine blocks (transpose of p)
    std::vector<double> r_w;
};
// Coarse mesh c (every other grid line) of a tensor-product grid mesh f
// whose element counts are even; MSA_ERR_ARG (with the reason) otherwise.
// Coarse nodes inherit the constraints of the fine nodes they sit on; each
// coarse element keeps the local node order of its low-corner child.
int mg_coarsen(const MeshHost &f, MeshHost &c, MgTransfer &t);

// A tensor-product grid whose elements are equal cells with one local node
// order (all elements translates of element 0, so they share one element
// matrix): grid point -> block (-1 constrained node), cell -> element,
// block -> packed grid position i | j << 10 | k << 20, corner o = ox + 2 oy
// + 4 oz -> local node.  MSA_ERR_ARG when the mesh is not such a grid.
struct GridInfo {
    int nx = 0, ny = 0, nz = 0;
    std::vector<int> gblock, cell_elem, bpos;
    int corner_node[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int cell_identity = 0;  // element number == cell number
    int box = 0;            // the blocks are exactly the grid points of [box_lo, box_hi],
    int box_lo[3] = {0, 0, 0}, box_hi[3] = {0, 0, 0};  // numbered lexicographically
};
int grid_info(const MeshHost &m, GridInfo &g);

// Device view of one class of a batch (copied to device memory).
// FP64 unit element matrix of matrix-free classes: Khat (576) + Mhat (64),
// then Khat transposed by input component (576)
constexpr int kMfK64T = 640;
constexpr int kMfK64 = 640 + 576;

struct ClassDev {
    int nb, nslices, ntiles, nsamp;
    int n_elem, nchunks, chunk, M;
    long long tile_vals, tile_vec;
    double fnorm, c0;
    const int *slice_ptr, *cidx, *diag_pos;
    // box grids (blocks = the points of a box, numbered lexicographically
    // with strides 1, sym_sy, sym_sz): stencil slot order, 27 slots per
    // row, slot s = (dz+1)*9 + (dy+1)*3 + (dx+1) holding the neighbour at
    // offset (dx, dy, dz) or a zero block, so the diagonal is slot 13 and
    // slots 13..26 hold the upper triangle.  sym: one-sample tiles whose PCG
    // product reads only those 14 slots (the lower blocks are the
    // transposes of upper blocks stored in the rows above).
    int stencil, sym, sym_sy, sym_sz;
    const uint8_t *block_free;
    const double *xyz;
    const int *conn, *node_block;
    const uint8_t *node_free, *emap;
    const int *colour_elems;
    void *val;
    void *x, *r, *z, *p0, *p1, *q, *dinv;
    double *u64, *r64;
    const double *f;      // [nb*3] internal numbering, shared by samples
    double *E;            // [ntiles][n_elem][T]
    double *phi;          // [n_elem][M]
    double *xi;           // [ntiles*T][M]
    SampleState *st;      // [ntiles*T]
    double *partial;      // [ntiles][nchunks][3][T]
    int *negdet;          // [1]
    int *due;             // [1 + nsamp]: count, then samples with a pending replacement
    // gather assembly: for every stored block (slot, rr) the contributing
    // elements as e*64 + a*8 + b, at most 8, ascending e, -1 terminated:
    // gmap[(slot*8 + g)*R + rr]
    const int *gmap;
    double *kc;           // [n_elem][576] unit-modulus Khat, block layout (a*8+b)*9 + 3p+q
    double *mc;           // [n_elem][64] unit-density consistent mass (dynamics)
    int colour_ptr[kMaxColours + 1];
    // linear Newmark dynamics (dyn = 1): the stored matrix is the effective
    // stiffness cK K + cM Mhat (cM includes the density); per-sample state
    int dyn, probe;       // probe: internal row 3*block + d recorded each step, or -1
    int hist_len, pad_dyn;
    double cK, cM, rho;
    double *fs;           // [ntiles][tile_vec] right-hand side of the current step
    double *uold, *vel, *acc, *wm, *wk;  // [ntiles][tile_vec]
    double *bacc;         // [nb*3] mass-acceleration load of a unit base acceleration
    double *hist;         // [nsamp][hist_len] probe displacement per step
    // direct solver: lower band of the Cholesky factor per sample,
    // band[s][j*bW + (i - j)] for 0 <= i - j < bW, in band order
    int bn, bW;           // band dimension (free DOFs), storage width
    const int *bperm;     // [bn] band index -> internal row
    const int *binv;      // [nb*3] internal row -> band index or -1
    double *band;         // [nsamp][bn*bW]
    long long band_ss;    // doubles between two samples' factors: bn*bW, or 0 when every
                          // sample uses the one factor at band (first-sample preconditioner)
    double *by;           // [nsamp][bn] right-hand side / solution in band order
    double *bpart;        // [nsamp][kBandChunks][64] partial sums of the backward solve
    int *bfail;           // [nsamp] non-positive pivot met
    // zero-energy modes of the direct solver's factorization (zv null: off):
    // a pivot of magnitude under zem_tol is set to 1 and its row and column
    // of L cleared; zv[s][k] holds the k-th mode in band order (first the
    // cleared row, then L^-T of it), zcol[s][k] its band column
    double zem_tol;
    int zem_max;          // modes kept per sample (more: bfail)
    double *zv;           // [nsamp][zem_max][bn]
    int *zcol;            // [nsamp][zem_max]
    int *zcount;          // [nsamp] modes found
    unsigned long long *zmask;  // [nsamp][ceil(bn/64)] mode columns of each 64-column block
    double *binvd;        // coarsest multigrid level: [nsamp][bn][bn] dense inverse, band order
    // geometric multigrid preconditioner (msa_batch_set_preconditioner).
    // Every class of a hierarchy is an entry of the class table; mg_* are
    // table indices (-1: none).  root: the solved class whose sample state
    // drives the cycle; coarse: the next coarser level; fine: the finer one.
    int mg_root, mg_coarse, mg_fine, mg_depth;
    const SampleState *mg_st;  // sample states of the root class
    double *mg_lam;       // [ntiles*T] upper end of the smoothing interval (spectrum of D^-1 K)
    double *mg_lz;        // [ntiles*T][2*kLzMax+1] Lanczos alpha, beta, start scale
    // transfer rows in ELL layout (entry k of row i at k*nb + i, padding:
    // column 0 with weight 0): prolongation rows of this class's blocks
    // (columns: blocks of mg_coarse), restriction rows of this coarse
    // class's blocks (columns: blocks of mg_fine)
    int mg_p_width, mg_r_width;
    const int *mg_p_col, *mg_r_col;
    const float *mg_p_w, *mg_r_w;
    const int *mg_child;  // [n_elem*8] elements of the finer level in each element
    // transfers from the grid structure instead of the ELL lists (uniform
    // box grids on both sides, every block fully free): prolongation of
    // this class (mg_pstruct) / restriction into it (mg_rstruct)
    int mg_pstruct, mg_rstruct;
    // matrix-free smoothing products (mixed precision, one-sample tiles) on
    // a uniform structured grid: K v from the cell moduli and the shared
    // unit element matrix mf_k[576] (+ unit mass mf_k[576 + 64])
    // (mf_k64: the same in FP64, then Khat again at kMfK64T with the 3x3
    // blocks transposed and grouped by (a, input component q, b):
    // [((a * 3 + q) * 8 + b) * 3 + p] = Khat[a][b][p][q])
    int mf, mf_nx, mf_ny, mf_nz;
    int mf_corner[8];
    // index shortcuts: element = cell (mf_cell_id); blocks = a lexicographic
    // sub-box of the grid (mf_box; neighbour block = own block + offset)
    int mf_cell_id, mf_box;
    int mf_box_lo[3], mf_box_hi[3];
    // the PCG product q = K p of this box-grid class matrix-free as well
    // (set per solve: MIXED, one-sample tiles, multigrid set up)
    int mf_pcg;
    // FP64 copy of the stored values (same layout as a one-sample-tile FP64
    // batch) for the coarsest multigrid level of a MIXED batch: the band
    // factor is computed from these, not from the FP32 values (null: none)
    double *val64;
    // box grid whose blocks are all fully free (every in-box neighbour
    // couples all three components): the Gershgorin bound is formed
    // matrix-free (k_mg_gersh_mf)
    int mf_allfree;
    const int *mf_gblock, *mf_cell, *mf_bpos;
    float *mf_k;
    double *mf_k64;       // [640] the same in FP64 (replacement residuals)
};

constexpr int kLzMax = 16;         // Lanczos steps of the smoothing-interval estimate, at most
constexpr int kBandNB = 64;        // block size of the band Cholesky
constexpr int kBandChunkRows = 512;  // rows per partial of the backward solve
constexpr int kDenseMax = 256;       // coarsest multigrid levels up to this size use a dense inverse
constexpr int kBandOneMax = 2048;    // bands up to this many rows: both sweeps in one launch
constexpr int kKlTiles = 8;          // tiles per thread of the KL field evaluation
constexpr int kNoallocMaxElem = 4096; // elements per call of msa_ke_packed_noalloc

// The no-allocation entry point's device array (noalloc.hip): looks up the
// current device's address of the array once; msa_ctx_create calls it.
int noalloc_setup(int device);
size_t noalloc_scratch_bytes();

// Newmark coefficients (NewmarkDynamicAnalyzer.cs:53-77) and the load history.
struct DynCoef {
    double a0, a1, a2, a3, a4, a5, a6, a7;
    double alphaR, betaR;  // Rayleigh damping C = alphaR M + betaR K
    double dt;
    const double *acc;     // [nsteps] base acceleration per step (device)
    int *step;             // current step (device counter)
    int start;             // PCG start of a step: 0 zero, 1 previous displacement,
                           // 2 predictor u + dt v + dt^2/2 a
};

// Elementwise launch over one colour of every class: blocks [prefix[c],
// prefix[c+1]) belong to class c.
struct ColourLaunch {
    int ncls;
    int colour;
    int prefix[33];
};

// State updates applied by the reduction kernel after each partial-sum kernel.
enum ReduceOp {
    OP_INIT0 = 0, OP_INIT1, OP_A, OP_B, OP_IRNORM, OP_RUNORM, OP_RURESTART,
    OP_FNORM,  // |f| of the per-sample right-hand side
    OP_INIT2,  // start of a warm-started FP64 solve (relres against |f|)
    OP_DIRECT  // after a direct solve: |f - K x|/|f| and the factorization status
};

struct ReduceArgs {
    int op;
    int max_it;
    int max_outer;
    int pad;
    double tol;
    double ru_drop;  // > 0: reliable-update mode of OP_B (trigger factor)
    double ru_tol;   // FP64 stopping tolerance seen by OP_B in that mode
    int *active_count;
    int *ir_count;
};

enum KernelKind { KK_SPMV = 0, KK_UPDATE = 1, KK_ASSEMBLE = 2, KK_RES64 = 3, KK_OTHER = 4, KK_MG = 5 };
constexpr int kKernelKinds = 6;

// Launch wrappers (kernels.hip).  T is the lane tile of the batch.
hipError_t launch_fill_f64(hipStream_t s, double *p, size_t n, double v);
hipError_t launch_fill_f32(hipStream_t s, float *p, size_t n, float v);
hipError_t launch_fill_u32(hipStream_t s, unsigned *p, size_t n, unsigned v);
hipError_t launch_ke_packed(hipStream_t s, int n_elem, const double *xyz8, const double *E,
                            double nu, double *ke300, int *status);
hipError_t launch_gauss_matrices(hipStream_t s, const double *xyz, double *out);
hipError_t launch_h8k(hipStream_t s, const double *E, const double *B, const double *w, double *K);
hipError_t launch_h8strains(hipStream_t s, const double *B, const double *u, double *strains);
hipError_t launch_h8forces(hipStream_t s, const double *B, const double *w, const double *stresses,
                           double *forces);
hipError_t launch_h8mlumped(hipStream_t s, double density, const double *w, double *M);
hipError_t launch_element_results(hipStream_t s, int T, const ClassDev *d_cls_c, int n_elem,
                                  const double *u, int lane, int tile, double nu, double *strains,
                                  double *stresses, double *forces);
// FP32-integration scatter assembly; ntiles: the most tiles of any class
// (sizes blockIdx.y)
hipError_t launch_assemble(hipStream_t s, int T, const ClassDev *d_cls, const ColourLaunch &cl,
                           int nblocks, double nu, int ntiles);
hipError_t launch_khat_cache(hipStream_t s, const ClassDev *d_cls_c, int n_elem, double nu,
                             int with_mass);
hipError_t launch_assemble_gather(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                                  const int4 *items, int nitems);
// one-sample tiles: gather assembly of kAsmGroup samples per work item
// (class, tile group, chunk)
#ifndef MSA_ASM_GROUP
#define MSA_ASM_GROUP 8
#endif
constexpr int kAsmGroup = MSA_ASM_GROUP;
hipError_t launch_assemble_gather_g(hipStream_t s, bool mixed, const ClassDev *d_cls,
                                    const int4 *gitems, int nitems);
hipError_t launch_residual64(hipStream_t s, int T, const ClassDev *d_cls, const ColourLaunch &cl,
                             int nblocks, double nu, int need_due, int ntiles);
hipError_t launch_dinv(hipStream_t s, bool mixed, int T, const ClassDev *d_cls, const int4 *items,
                       int nitems);
hipError_t launch_pcg_init(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                           const int4 *items, int nitems, int mode, int *active_count);
// U: slots per group of a row (1..4 at T = 1, 3 for T > 1)
hipError_t launch_pcg_a(hipStream_t s, bool mixed, int T, int U, const ClassDev *d_cls,
                        const int4 *items, int nitems, int pbuf, int *active_count);
hipError_t launch_residual64_mf_box(hipStream_t s, const ClassDev *d_cls, const int4 *items,
                                   int nitems);
hipError_t launch_pcg_a_mf(hipStream_t s, const ClassDev *d_cls, const int4 *items, int nitems,
                           int pbuf);
hipError_t launch_pcg_b(hipStream_t s, bool mixed, int T, const ClassDev *d_cls, const int4 *items,
                        int nitems, int pbuf, double tol, int max_it, int *active_count);
hipError_t launch_ir_begin(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                           int nitems, int *ir_count, int warm);
hipError_t launch_ir_update(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                            int nitems);
hipError_t launch_ir_norm(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                          int nitems, double tol, int max_outer, int *ir_count);
hipError_t launch_ru_fold(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                          int nitems);
hipError_t launch_ru_norm(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                          int nitems, double tol, int max_it, int *ir_count, int *active_count);
hipError_t launch_ru_restart(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                             int nitems);
hipError_t launch_spmv(hipStream_t s, bool mixed, int T, const ClassDev *d_cls, const int4 *items,
                       int nitems, const double *x, double *y);
hipError_t launch_reduce(hipStream_t s, int T, const ClassDev *d_cls, const int2 *tiles,
                         int ntiles, const ReduceArgs &a);
// direct solver (band Cholesky); cp = the class's device ClassDev
hipError_t launch_band_scatter(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                               int nitems);
hipError_t launch_chol_diag(hipStream_t s, const ClassDev *cp, int nsamp, int k0, int nk);
hipError_t launch_chol_panel(hipStream_t s, const ClassDev *cp, int nsamp, int k0, int nk, int r0,
                             int m);
hipError_t launch_chol_update(hipStream_t s, const ClassDev *cp, int nsamp, int k0, int nk, int r0,
                              int m);
hipError_t launch_band_rhs(hipStream_t s, int T, const ClassDev *cp, int nsamp, int bn,
                           int from_fs);
// mode >= 0: coarsest multigrid level, only the samples mg_act serves
hipError_t launch_band_fwd(hipStream_t s, const ClassDev *cp, int nsamp, int k0, int nk, int r0,
                           int m, int mode = -1);
hipError_t launch_band_solve_one(hipStream_t s, const ClassDev *cp, int nsamp, int bn, int mode);
hipError_t launch_band_bwd(hipStream_t s, const ClassDev *cp, int nsamp, int k0, int nk, int r0,
                           int m, int mode = -1);
hipError_t launch_band_out(hipStream_t s, int T, const ClassDev *cp, int nsamp, int nrows);
// multigrid preconditioner: items of one level of every hierarchy
enum MgStep { MG_PRE = 0, MG_RESID = 1, MG_POST1 = 2, MG_POST2 = 3 };
// which samples a cycle serves: 0 every sample, 1 active, 2 due for a
// residual replacement (ir_active && ru_due), 3 refining (ir_active)
// mf: 0 stored values; the items' classes use the matrix-free product
// (mixed, T = 1) on general uniform grids (1) or on box grids (2)
hipError_t launch_mg_smooth(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                            const int4 *items, int nitems, int step, int pbuf, int mode,
                            double ratio, int mf);
// matrix-free smoothing with the element products on the matrix cores (box
// grids of fully free blocks, MIXED, one-sample tiles); items: (class,
// sample, y block of kMfmaRows node rows, 0)
hipError_t launch_mg_smooth_mfma(hipStream_t s, const ClassDev *d_cls, const int4 *items,
                                 int nitems, int step, int pbuf, int mode, double ratio);
#ifndef MSA_MFMA_ROWS
#define MSA_MFMA_ROWS 7
#endif
constexpr int kMfmaRows = MSA_MFMA_ROWS;  // node rows (y) per workgroup of the MFMA smoother
constexpr int kMfmaMaxCells = 64;  // cells in x it handles
hipError_t launch_mg_restrict(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                              const int4 *items, int nitems, int pbuf, int mode);
hipError_t launch_mg_prolong(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                             const int4 *items, int nitems, int mode);
hipError_t launch_mg_coarsen_E(hipStream_t s, int T, const ClassDev *cp, const ClassDev *d_cls,
                               int ntiles, int n_elem);
hipError_t launch_mg_gersh_mf(hipStream_t s, const ClassDev *d_cls, const int4 *items, int nitems);
hipError_t launch_mg_gersh(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                           const int4 *items, int nitems);
hipError_t launch_mg_band_in(hipStream_t s, bool mixed, int T, const ClassDev *cp, int nsamp,
                             int bn, int mode);
// first-sample direct preconditioner: z from the band solution + r.z partials
hipError_t launch_first_out(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                           const int4 *items, int nitems, int mode);
hipError_t launch_mg_band_out(hipStream_t s, bool mixed, int T, const ClassDev *cp, int nsamp,
                              int nrows, int mode);
hipError_t launch_mf_khat(hipStream_t s, const ClassDev *cp);
// replacement residual of the due samples of matrix-free classes (T = 1)
hipError_t launch_residual64_mf(hipStream_t s, const ClassDev *d_cls, const int4 *items, int nitems);
// coarsest multigrid level: dense inverse from the band factor, and its use
hipError_t launch_band_inverse(hipStream_t s, const ClassDev *cp, int nsamp, int bn);
hipError_t launch_mg_dense_solve(hipStream_t s, bool mixed, int T, const ClassDev *cp, int nsamp,
                                 int bn, int mode);
hipError_t launch_mg_load(hipStream_t s, bool mixed, int T, const ClassDev *cp, const double *src,
                          size_t nrows, int tile, int lane);
hipError_t launch_band_scatter_any(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                                   const int4 *items, int nitems);
// dynamics
hipError_t launch_set_int(hipStream_t s, int *p, int n, int v);
hipError_t launch_dyn_bacc(hipStream_t s, const ClassDev *d_cls, const ColourLaunch &cl,
                           int nblocks, double nu, int dir);
hipError_t launch_dyn_reset(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                            const int4 *items, int nitems);
hipError_t launch_dyn_prep(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                           const int4 *items, int nitems, const DynCoef &dc);
hipError_t launch_dyn_rhs(hipStream_t s, int T, const ClassDev *d_cls, const ColourLaunch &cl,
                          int nblocks, double nu, double betaR, int ntiles);
hipError_t launch_fs_static(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                            int nitems);
hipError_t launch_fs_norm(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                          int nitems);
hipError_t launch_pcg_init_warm(hipStream_t s, int T, const ClassDev *d_cls, const int4 *items,
                                int nitems);
hipError_t launch_dyn_update(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                             const int4 *items, int nitems, const DynCoef &dc);
hipError_t launch_step_advance(hipStream_t s, int *step);
hipError_t launch_kl_basis(hipStream_t s, const ClassDev *d_cls, int n_elem, int M, int kmax,
                           const double *w1d, const double *a1d, const int *evpPos,
                           const double *eigs, const double *domain, double sigma);
hipError_t launch_kl_eval(hipStream_t s, int T, const ClassDev *d_cls, int ntiles, int n_elem,
                          int M);
hipError_t launch_extract_sample(hipStream_t s, bool mixed, int T, const void *src, size_t n_rows,
                                 int lane, double *dst);
// The whole Jacobi-PCG (one-sample tiles) in one cooperative launch of at
// most `iters` iterations; *grid_out = workgroups launched.
hipError_t launch_pcg_persist(hipStream_t s, bool mixed, int U, const ClassDev *d_cls,
                              const int4 *items, int nitems, const int2 *tiles, int ntiles,
                              const ReduceArgs &ra, const ReduceArgs &rb, int iters, int *grid_out);
// Lanczos estimate of the top of the spectrum of D^-1 K per sample and
// level (the upper end of the Chebyshev smoothing interval).  op: 0 start
// vector, 2 alpha partials, 3 next vector and its D-norm partials; sum
// (over the classes in list): which 0 start norm, 1 alpha_k, 2 n_k+1.
hipError_t launch_lz_step(hipStream_t s, bool mixed, int T, const ClassDev *d_cls,
                          const int4 *items, int nitems, int op, int k);
hipError_t launch_lz_sum(hipStream_t s, int T, const ClassDev *d_cls, const int *list, int nlist,
                         int max_samp, int which, int k);
hipError_t launch_lz_eig(hipStream_t s, const ClassDev *d_cls, const int *list, int nlist,
                         int max_samp, int steps, double margin);
hipError_t launch_insert_sample(hipStream_t s, int T, const double *src, size_t n_rows, int lane,
                                double *dst);
hipError_t launch_extract_values(hipStream_t s, bool mixed, int T, const void *val, size_t nslots9R,
                                 int lane, double *dst);

}  // namespace msa

// Opaque handle types of the C ABI.
struct msa_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int *h_pinned = nullptr;  // small pinned host scratch for polling
};

struct msa_mesh {
    msa_ctx *ctx = nullptr;
    msa::MeshHost h;
    double *d_xyz = nullptr;
    int *d_conn = nullptr, *d_node_block = nullptr, *d_colour_elems = nullptr;
    uint8_t *d_node_free = nullptr, *d_emap = nullptr;
};
