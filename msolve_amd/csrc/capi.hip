// C ABI of msolve_amd (include/msolve_amd.h): contexts, meshes, Monte Carlo
// batches and their solve drivers.  All device work runs on the context's
// stream; host syncs happen only where a result must reach the host (the
// convergence poll between captured iteration blocks, downloads).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <random>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "msa_common.hpp"

namespace msa {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

}  // namespace msa

using namespace msa;

// ------------------------------------------------------------------ batch

namespace {

struct ClassHost {
    msa_mesh *mesh = nullptr;
    long long hist_cap = 0;  // probe history capacity (doubles)
    double kc_nu = -2.0;     // Poisson ratio of the element cache (-2: not built)
    bool kc_mass = false;    // cache holds the consistent mass too
    // direct solver: band order (host) and storage sizes
    std::vector<int> bperm, binv;
    long long bw = -1;       // half bandwidth, -1: order not built
    size_t band_cap = 0;     // doubles allocated for the factors
    // zero-energy modes of the direct solver (msa_batch_set_zem): buffers
    // sized for zem_alloc modes per sample
    double *zv = nullptr;
    int *zcol = nullptr, *zcount = nullptr;
    unsigned long long *zmask = nullptr;
    int zem_alloc = 0;
    int nsamp = 0, ntiles = 0, nslices = 0, nchunks = 0, chunk = 0;
    std::vector<int> slice_ptr, cidx;
    // slot within its row's slice of every adjacency entry (brow_ptr/bcol
    // order): its rank in the row, or its stencil slot on box grids
    std::vector<uint8_t> srank;
    int R = 64;  // block rows per slice
    int slot_of(const MeshHost &h, int br, int j) const
    {
        return slice_ptr[br / R] + srank[h.brow_ptr[br] + j];
    }
    long long tile_vals = 0, tile_vec = 0;
    ClassDev dev{};
    bool has_rhs = false, has_E = false, has_kl = false, has_xi = false;
    int M = 0;
    double fnorm = 0;
};

struct Timing {
    bool on = false;
    double ms[kKernelKinds] = {0, 0, 0, 0, 0, 0};
    long long n[kKernelKinds] = {0, 0, 0, 0, 0, 0};
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

}  // namespace

struct msa_batch {
    msa_ctx *ctx = nullptr;
    int precision = MSA_FP64, T = 1, R = 64;
    int unroll = 3;  // slots per group of k_pcg_a's rows (msa_batch_set_schedule)
    std::vector<ClassHost> cls;
    ClassDev *d_cls = nullptr;
    std::vector<int4> items;
    int4 *d_items = nullptr;
    int4 *d_gitems = nullptr;  // one-sample tiles: (class, group of kAsmGroup tiles, chunk)
    int n_gitems = 0;
    std::vector<int2> tiles;  // (class, tile) pairs, one reduction workgroup each
    int2 *d_tiles = nullptr;
    int *d_counts = nullptr;  // [0] PCG active samples, [1] refinement active samples
    std::vector<void *> allocs;
    long long bytes = 0;
    bool assembled = false;
    bool loaded = false;      // values set by msa_batch_set_matrix_csr, not assembled
    bool factored = false;    // band Cholesky factors match the assembled values
    double zem_tol = 0.0;     // zero-energy-mode pivot tolerance (0: off)
    bool persistent = false;  // Jacobi PCG as one cooperative launch (msa_batch_set_persistent)
    long long persist_launches = 0;
    int persist_grid = 0;
    int zem_max = 0;          // modes kept per sample
    int solver = MSA_SOLVER_PCG;
    double nu = 0.3;
    // mixed-precision refinement: restarted inner solves or periodic
    // residual replacement
    int refine = MSA_REFINE_RELIABLE;
    int period = 8;       // iterations between replacement checks
    double drop = 0.1;    // FP32 residual reduction that triggers a replacement
    bool refine_set = false;  // set by msa_batch_set_refinement (else: preconditioner defaults)
    // element integration precision of MIXED batches (values stored in FP32)
    int integration = MSA_FP64;
    // linear Newmark dynamics (msa_batch_set_dynamics)
    bool dyn = false;
    DynCoef dc{};
    double rho = 0;
    double *d_acc = nullptr;  // base acceleration history of the current run
    int acc_cap = 0;
    int *d_step = nullptr;    // device step counter
    // captured blocks of PCG iterations (kind 0 plain, 1 with replacement):
    // the caller's block and the short tail block replayed once a sample
    // has converged
    hipGraphExec_t graph = nullptr, graph_tail = nullptr;
    int graph_iters = 0, graph_tail_iters = 0, graph_kind = -1;
    double graph_tol = -1;
    int graph_maxit = -1;
    long long graph_kernels = 0;   // kernel nodes of the current graph
    long long graph_captures = 0;  // graphs captured (and validated) so far
    long long graph_replays = 0;   // launches of captured graphs so far
    Timing timing;
    // geometric multigrid preconditioner (msa_batch_set_preconditioner)
    int precond = MSA_PRECOND_JACOBI;
    double mg_ratio = 12.0;             // smoothing interval [lam/ratio, lam]
    int mg_lz_steps = 4;                // Lanczos steps of lam (0: the Gershgorin bound)
    double mg_lz_margin = 1.3;          // lam = min(Gershgorin, margin x top Ritz value)
    int *d_lz_list = nullptr;           // the smoothed classes (device), for the Lanczos sums
    std::vector<int> lz_list;
    std::vector<ClassHost> lv;          // coarse levels: class table entry cls.size() + i
    std::vector<msa_mesh *> mg_meshes;  // their meshes (owned by the batch)
    std::vector<int> mg_nlev;           // per solved class: coarse levels (0: Jacobi)
    int mg_depth = 0;                   // deepest level of any hierarchy
    std::vector<int4 *> mg_smooth;      // per depth: items of the classes that smooth there
    std::vector<int> mg_nsmooth;
    std::vector<int4 *> mg_sm_a, mg_sm_f, mg_sm_b;  // the same split: stored-value / matrix-free
                                                   // products on general / box grids
    std::vector<int> mg_nsm_a, mg_nsm_f, mg_nsm_b;
    // box grids whose products can run on the matrix cores: their items for
    // the MFMA kernel (class, sample, y block) and for the VALU box kernel
    std::vector<int4 *> mg_sm_m, mg_sm_e;
    std::vector<int> mg_nsm_m, mg_nsm_e;
    bool matrix_cores = false;          // msa_batch_set_matrix_cores
    bool full_storage = false;          // msa_batch_set_symmetric_product(b, 0)
    bool mf_product = true;             // msa_batch_set_matrix_free_product
    // PCG product work items while some class takes the matrix-free product:
    // the stored-value classes, and the matrix-free ones (runs of chunks)
    int4 *d_items_st = nullptr, *d_items_mf = nullptr;
    int n_items_st = 0, n_items_mf = 0;
    std::vector<int4 *> mg_restr;       // per depth >= 1: items of every level there
    std::vector<int> mg_nrestr;
    int4 *mg_lv_items = nullptr;        // every coarse level
    int mg_nlv_items = 0;
    int4 *mg_lv_gitems = nullptr;       // the same in sample groups (one-sample tiles)
    int mg_nlv_gitems = 0;
    int4 *mg_co_items = nullptr;        // the coarsest levels
    int mg_nco_items = 0;
    int4 *mf_items = nullptr;           // solved classes with matrix-free products
    int n_mf_items = 0;                 //   (general grids; box grids: runs of chunks below)
    int4 *mf_box_items = nullptr;
    int n_mf_box_items = 0;
    std::vector<int> mg_coarsest;       // class table indices of the coarsest levels
    int mg_max_levels = 0;              // max_levels the hierarchy was built with
    size_t d_cls_cap = 0;               // ClassDev entries allocated at d_cls
    bool mg_ready = false;              // level matrices match the assembled classes
};

namespace {

bool is_mixed(const msa_batch *b) { return b->precision == MSA_MIXED; }

size_t real_size(const msa_batch *b) { return is_mixed(b) ? sizeof(float) : sizeof(double); }

int dalloc(msa_batch *b, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc(%zu bytes) failed: %s (batch holds %lld bytes)", bytes,
                  hipGetErrorString(e), b->bytes);
        *p = nullptr;
        return MSA_ERR_OOM;
    }
    b->allocs.push_back(*p);
    b->bytes += (long long)bytes;
    return MSA_OK;
}

#define MSA_TRY(x)                      \
    do {                                \
        int rc_ = (x);                  \
        if (rc_ != MSA_OK) return rc_;  \
    } while (0)

// Slot of every adjacency entry within its row (ClassHost::srank): the
// stencil slot on a box grid (ClassDev::stencil), with sy, sz the block
// strides in y and z; otherwise the entry's rank in the row, and false.
// A box thinner than 3 points in x or y keeps the ranks: there the stencil
// order is not the column order.
bool stencil_ranks(const MeshHost &h, std::vector<uint8_t> &srank, int &sy, int &sz)
{
    srank.resize(h.bcol.size());
    for (int br = 0; br < h.nb; br++)
        for (int j = h.brow_ptr[br]; j < h.brow_ptr[br + 1]; j++)
            srank[j] = (uint8_t)(j - h.brow_ptr[br]);
    sy = sz = 0;
    GridInfo g;
    if (grid_info(h, g) != MSA_OK || !g.box) {
        set_error("");
        return false;
    }
    const int nx = g.box_hi[0] - g.box_lo[0] + 1, ny = g.box_hi[1] - g.box_lo[1] + 1;
    if (nx < 3 || ny < 3) return false;
    std::vector<uint8_t> st(h.bcol.size());
    for (int br = 0; br < h.nb; br++) {
        const int pr = g.bpos[br];
        for (int j = h.brow_ptr[br]; j < h.brow_ptr[br + 1]; j++) {
            const int bc = h.bcol[j], pc = g.bpos[bc];
            const int dx = (pc & 1023) - (pr & 1023);
            const int dy = ((pc >> 10) & 1023) - ((pr >> 10) & 1023);
            const int dz = (pc >> 20) - (pr >> 20);
            if (std::abs(dx) > 1 || std::abs(dy) > 1 || std::abs(dz) > 1) return false;
            if (bc != br + dx + nx * dy + nx * ny * dz) return false;
            const int s = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
            if ((s > 13) != (bc > br) || (s == 13) != (bc == br)) return false;
            st[j] = (uint8_t)s;
        }
    }
    srank.swap(st);
    sy = nx;
    sz = nx * ny;
    return true;
}

// Device arrays and block-ELL layout of one class (a solved class or a
// multigrid level) for nsamp samples.
int setup_class(msa_batch *b, ClassHost &C, msa_mesh *mesh, int nsamp)
{
    const int T = b->T, R = b->R;
    const size_t rs = real_size(b);
    int rc = MSA_OK;
    {
        const MeshHost &h = mesh->h;
        C.mesh = mesh;
        C.nsamp = nsamp;
        C.ntiles = (C.nsamp + T - 1) / T;
        C.nslices = (h.nb + R - 1) / R;
        C.chunk = kWaves * std::max(1, 16 / R);
        C.nchunks = (C.nslices + C.chunk - 1) / C.chunk;
        C.R = R;
        int sy = 0, sz = 0;
        const bool stencil = stencil_ranks(h, C.srank, sy, sz);
        C.slice_ptr.assign(C.nslices + 1, 0);
        for (int sl = 0; sl < C.nslices; sl++) {
            int wmax = 0;
            for (int rr = 0; rr < R; rr++) {
                int br = sl * R + rr;
                if (br < h.nb) wmax = std::max(wmax, h.brow_ptr[br + 1] - h.brow_ptr[br]);
            }
            C.slice_ptr[sl + 1] = C.slice_ptr[sl] + (stencil ? 27 : wmax);
        }
        const long long nslots = C.slice_ptr[C.nslices];
        C.cidx.assign((size_t)nslots * R, -1);
        for (int br = 0; br < h.nb; br++) {
            const int rr = br % R;
            for (int j = 0; j < h.brow_ptr[br + 1] - h.brow_ptr[br]; j++)
                C.cidx[(size_t)C.slot_of(h, br, j) * R + rr] = h.bcol[h.brow_ptr[br] + j];
        }
        // the class's diagonal slots and element-pair slots (the mesh's
        // ranks unless the stencil order replaces them)
        std::vector<int> dpos(h.nb);
        for (int br = 0; br < h.nb; br++) dpos[br] = C.srank[h.brow_ptr[br] + h.diag_pos[br]];
        std::vector<uint8_t> emap_c;
        if (stencil) {
            emap_c.assign(h.emap.size(), 255);
            for (int e = 0; e < h.n_elem; e++)
                for (int a = 0; a < 8; a++) {
                    const int br = h.node_block[h.conn[8 * e + a]];
                    for (int bb = 0; bb < 8; bb++) {
                        const uint8_t j = h.emap[(size_t)e * 64 + a * 8 + bb];
                        if (j != 255 && br >= 0)
                            emap_c[(size_t)e * 64 + a * 8 + bb] = C.srank[h.brow_ptr[br] + j];
                    }
                }
        }
        std::vector<uint8_t> bfree(h.nb);
        for (int br = 0; br < h.nb; br++) bfree[br] = h.node_free[h.block_node[br]];
        C.tile_vals = nslots * 9 * R * T;
        C.tile_vec = (long long)h.nb * 3 * T;

        ClassDev &D = C.dev;
        memset(&D, 0, sizeof(D));
        D.nb = h.nb;
        D.nslices = C.nslices;
        D.ntiles = C.ntiles;
        D.nsamp = C.nsamp;
        D.n_elem = h.n_elem;
        D.nchunks = C.nchunks;
        D.chunk = C.chunk;
        D.tile_vals = C.tile_vals;
        D.tile_vec = C.tile_vec;
        D.xyz = mesh->d_xyz;
        D.conn = mesh->d_conn;
        D.node_block = mesh->d_node_block;
        D.node_free = mesh->d_node_free;
        D.emap = mesh->d_emap;
        D.colour_elems = mesh->d_colour_elems;
        D.stencil = stencil ? 1 : 0;
        D.sym = stencil && T == 1 && !b->full_storage ? 1 : 0;
        D.sym_sy = sy;
        D.sym_sz = sz;
        MSA_CHECK(h.ncolours <= kMaxColours, MSA_ERR_ARG, "too many colours");
        for (int k = 0; k <= h.ncolours; k++) D.colour_ptr[k] = h.colour_ptr[k];
        auto A = [&](void **p, size_t bytes) {
            if (rc == MSA_OK) rc = dalloc(b, p, bytes);
        };
        void *p = nullptr;
        A(&p, C.slice_ptr.size() * sizeof(int));
        D.slice_ptr = (const int *)p;
        if (rc == MSA_OK) MSA_HIP(hipMemcpy(p, C.slice_ptr.data(), C.slice_ptr.size() * sizeof(int), hipMemcpyHostToDevice));
        A(&p, C.cidx.size() * sizeof(int));
        D.cidx = (const int *)p;
        if (rc == MSA_OK) MSA_HIP(hipMemcpy(p, C.cidx.data(), C.cidx.size() * sizeof(int), hipMemcpyHostToDevice));
        A(&p, dpos.size() * sizeof(int));
        D.diag_pos = (const int *)p;
        if (rc == MSA_OK) MSA_HIP(hipMemcpy(p, dpos.data(), dpos.size() * sizeof(int), hipMemcpyHostToDevice));
        if (stencil) {
            A(&p, emap_c.size());
            D.emap = (const uint8_t *)p;
            if (rc == MSA_OK) MSA_HIP(hipMemcpy(p, emap_c.data(), emap_c.size(), hipMemcpyHostToDevice));
        }
        A(&p, bfree.size());
        D.block_free = (const uint8_t *)p;
        if (rc == MSA_OK) MSA_HIP(hipMemcpy(p, bfree.data(), bfree.size(), hipMemcpyHostToDevice));
        {
            // gather lists of the assembly: element contributions per stored
            // block, in ascending element order
            std::vector<int> gmap((size_t)nslots * 8 * R, -1);
            std::vector<uint8_t> cnt((size_t)nslots * R, 0);
            for (int e = 0; e < h.n_elem && rc == MSA_OK; e++)
                for (int a = 0; a < 8; a++) {
                    const int br = h.node_block[h.conn[8 * e + a]];
                    if (br < 0) continue;
                    for (int bb = 0; bb < 8; bb++) {
                        const int j = h.emap[(size_t)e * 64 + a * 8 + bb];
                        if (j == 255) continue;
                        const size_t slot = (size_t)C.slot_of(h, br, j);
                        const size_t sr = slot * R + br % R;
                        const int k = cnt[sr]++;
                        if (k >= 8) {
                            set_error("mesh: more than 8 elements share a node pair");
                            rc = MSA_ERR_ARG;
                            break;
                        }
                        gmap[(slot * 8 + k) * R + br % R] = e * 64 + a * 8 + bb;
                    }
                }
            A(&p, gmap.size() * sizeof(int));
            D.gmap = (const int *)p;
            if (rc == MSA_OK)
                MSA_HIP(hipMemcpy(p, gmap.data(), gmap.size() * sizeof(int), hipMemcpyHostToDevice));
            A((void **)&D.kc, (size_t)h.n_elem * 576 * sizeof(double));
        }
        const size_t nvec = (size_t)C.ntiles * C.tile_vec;
        A(&D.val, (size_t)C.ntiles * C.tile_vals * rs);
        // upper-triangle classes never write their lower slots: zeros, so
        // that a consumer reading a padding slot sees a zero block (every
        // stencil class, as msa_batch_set_symmetric_product may switch it)
        if (rc == MSA_OK && D.stencil) MSA_HIP(hipMemset(D.val, 0, (size_t)C.ntiles * C.tile_vals * rs));
        A(&D.x, nvec * rs);
        A(&D.r, nvec * rs);
        A(&D.z, nvec * rs);
        A(&D.p0, nvec * rs);
        A(&D.p1, nvec * rs);
        A(&D.q, nvec * rs);
        A(&D.dinv, nvec * rs);
        if (is_mixed(b)) {
            A((void **)&D.u64, nvec * sizeof(double));
            A((void **)&D.r64, nvec * sizeof(double));
        }
        A(&p, (size_t)h.nb * 3 * sizeof(double));
        D.f = (const double *)p;
        A((void **)&D.E, (size_t)C.ntiles * h.n_elem * T * sizeof(double));
        A((void **)&D.st, (size_t)C.ntiles * T * sizeof(SampleState));
        A((void **)&D.partial, (size_t)C.ntiles * C.nchunks * 3 * T * sizeof(double));
        A((void **)&D.negdet, sizeof(int));
        A((void **)&D.due, (1 + (size_t)C.ntiles * T) * sizeof(int));
        if (rc != MSA_OK) return rc;
        hipStream_t s = b->ctx->stream;
        MSA_HIP(hipMemsetAsync(D.st, 0, (size_t)C.ntiles * T * sizeof(SampleState), s));
        MSA_HIP(hipMemsetAsync(D.negdet, 0, sizeof(int), s));
        MSA_HIP(hipMemsetAsync(D.due, 0, sizeof(int), s));
        MSA_HIP(hipMemsetAsync((void *)D.f, 0, (size_t)h.nb * 3 * sizeof(double), s));
        MSA_HIP(hipMemsetAsync(D.E, 0, (size_t)C.ntiles * h.n_elem * T * sizeof(double), s));
        D.mg_root = D.mg_coarse = D.mg_fine = -1;
    }
    return rc;
}


// Timed launch helper: with timing on, bracket the launch with events on the
// context stream and wait for it.
template <typename F>
int timed(msa_batch *b, int kind, F &&launch)
{
    hipStream_t s = b->ctx->stream;
    if (b->timing.on) MSA_HIP(hipEventRecord(b->timing.e0, s));
    hipError_t e = launch();
    if (e != hipSuccess) {
        set_error("kernel launch failed: %s", hipGetErrorString(e));
        return MSA_ERR_HIP;
    }
    if (b->timing.on) {
        MSA_HIP(hipEventRecord(b->timing.e1, s));
        MSA_HIP(hipEventSynchronize(b->timing.e1));
        float ms = 0;
        MSA_HIP(hipEventElapsedTime(&ms, b->timing.e0, b->timing.e1));
        b->timing.ms[kind] += ms;
        b->timing.n[kind] += 1;
    }
    return MSA_OK;
}

int reduce(msa_batch *b, int op, double tol = 0.0, int max_it = 0, int max_outer = 0,
           double ru_drop = 0.0, double ru_tol = 0.0)
{
    ReduceArgs a{};
    a.op = op;
    a.tol = tol;
    a.max_it = max_it;
    a.max_outer = max_outer;
    a.ru_drop = ru_drop;
    a.ru_tol = ru_tol;
    a.active_count = b->d_counts;
    a.ir_count = b->d_counts + 1;
    return timed(b, KK_OTHER, [&] {
        return launch_reduce(b->ctx->stream, b->T, b->d_cls, b->d_tiles, (int)b->tiles.size(), a);
    });
}

void drop_graph(msa_batch *b)
{
    for (hipGraphExec_t *g : {&b->graph, &b->graph_tail})
        if (*g) {
            (void)hipGraphExecDestroy(*g);
            *g = nullptr;
        }
    b->graph_tail_iters = 0;
}

// Class table: the solved classes, then the multigrid levels.  Growing it
// moves d_cls, which invalidates a captured graph.
int upload_class_table(msa_batch *b)
{
    const size_t n = b->cls.size() + b->lv.size();
    std::vector<ClassDev> h(n);
    for (size_t c = 0; c < b->cls.size(); c++) h[c] = b->cls[c].dev;
    for (size_t l = 0; l < b->lv.size(); l++) h[b->cls.size() + l] = b->lv[l].dev;
    if (n > b->d_cls_cap) {
        void *p = nullptr;
        MSA_TRY(dalloc(b, &p, sizeof(ClassDev) * n));
        b->d_cls = (ClassDev *)p;
        b->d_cls_cap = n;
        drop_graph(b);
    }
    MSA_HIP(hipMemcpyAsync(b->d_cls, h.data(), sizeof(ClassDev) * h.size(),
                           hipMemcpyHostToDevice, b->ctx->stream));
    MSA_HIP(hipStreamSynchronize(b->ctx->stream));
    return MSA_OK;
}

int read_count(msa_batch *b, int which, int *out)
{
    MSA_HIP(hipMemcpyAsync(b->ctx->h_pinned, b->d_counts + which, sizeof(int),
                           hipMemcpyDeviceToHost, b->ctx->stream));
    MSA_HIP(hipStreamSynchronize(b->ctx->stream));
    *out = *b->ctx->h_pinned;
    return MSA_OK;
}

int mg_vcycle(msa_batch *b, int pbuf, int mode);
int mg_values(msa_batch *b);
int mg_lanczos(msa_batch *b);
int precond_values(msa_batch *b);

// q = K p with p = z + beta p_old for every class: the stored-value product
// and, for the classes whose product is matrix-free, k_pcg_a_mf.
hipError_t launch_product(msa_batch *b, int pbuf)
{
    hipStream_t s = b->ctx->stream;
    if (b->n_items_mf == 0)
        return launch_pcg_a(s, is_mixed(b), b->T, b->unroll, b->d_cls, b->d_items,
                            (int)b->items.size(), pbuf, b->d_counts);
    const hipError_t e = launch_pcg_a(s, is_mixed(b), b->T, b->unroll, b->d_cls, b->d_items_st,
                                      b->n_items_st, pbuf, b->d_counts);
    return e != hipSuccess ? e : launch_pcg_a_mf(s, b->d_cls, b->d_items_mf, b->n_items_mf, pbuf);
}

// ru_tol > 0: reliable-update mode (inner test off, replacement triggers on).
int launch_iteration_pair(msa_batch *b, int pbuf, double tol, int max_it, double ru_tol = 0.0)
{
    const bool mx = is_mixed(b);
    hipStream_t s = b->ctx->stream;
    MSA_TRY(timed(b, KK_SPMV, [&] { return launch_product(b, pbuf); }));
    MSA_TRY(reduce(b, OP_A));
    MSA_TRY(timed(b, KK_UPDATE, [&] {
        return launch_pcg_b(s, mx, b->T, b->d_cls, b->d_items, (int)b->items.size(), pbuf, tol,
                            max_it, b->d_counts);
    }));
    MSA_TRY(mg_vcycle(b, pbuf, 1));
    if (ru_tol > 0)
        MSA_TRY(reduce(b, OP_B, 0.0, max_it, 0, b->drop, ru_tol));
    else
        MSA_TRY(reduce(b, OP_B, tol, max_it));
    return MSA_OK;
}

ColourLaunch colour_launch(const msa_batch *b, int colour, int *nblocks, bool skip_mf = false);
int max_colours(const msa_batch *b);

int max_tiles(const msa_batch *b)
{
    int m = 1;
    for (auto &c : b->cls) m = std::max(m, c.ntiles);
    return m;
}

// Fold the FP32 correction into u64, form the FP64 residual, test it and
// restart the FP32 recurrence from it (reliable-update refinement).
int launch_replacement(msa_batch *b, double tol, int max_it)
{
    hipStream_t s = b->ctx->stream;
    const int T = b->T, ni = (int)b->items.size();
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_ru_fold(s, T, b->d_cls, b->d_items, ni); }));
    // matrix-free residual of the uniform-grid classes (multigrid set up),
    // the coloured element scatter for the others
    const bool mfres = b->precond == MSA_PRECOND_MULTIGRID && b->mg_ready &&
                       b->n_mf_items + b->n_mf_box_items > 0;
    if (mfres)
        MSA_TRY(timed(b, KK_RES64, [&] {
            const hipError_t e = launch_residual64_mf(s, b->d_cls, b->mf_items, b->n_mf_items);
            return e != hipSuccess
                       ? e
                       : launch_residual64_mf_box(s, b->d_cls, b->mf_box_items, b->n_mf_box_items);
        }));
    const int ncol = max_colours(b);
    for (int k = 0; k < ncol; k++) {
        int nblk = 0;
        ColourLaunch cl = colour_launch(b, k, &nblk, mfres);
        if (nblk == 0) continue;
        MSA_TRY(timed(b, KK_RES64, [&] { return launch_residual64(s, T, b->d_cls, cl, nblk, b->nu, 1, max_tiles(b)); }));
    }
    MSA_TRY(timed(b, KK_OTHER, [&] {
        return launch_ru_norm(s, T, b->d_cls, b->d_items, ni, tol, max_it, b->d_counts + 1,
                              b->d_counts);
    }));
    MSA_TRY(reduce(b, OP_RUNORM, tol, max_it));
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_ru_restart(s, T, b->d_cls, b->d_items, ni); }));
    MSA_TRY(mg_vcycle(b, 1, 2));  // the restart keeps the search direction in p0
    MSA_TRY(reduce(b, OP_RURESTART));
    return MSA_OK;
}

// One block of `iters` PCG iterations; kind 1 adds a residual replacement
// after every b->period iterations (inner stopping test disabled).
int launch_block(msa_batch *b, int kind, int iters, double tol, int max_it)
{
    for (int k = 0; k < iters; k++) {
        MSA_TRY(launch_iteration_pair(b, k & 1, kind == 1 ? 0.0 : tol, max_it, kind == 1 ? tol : 0.0));
        if (kind == 1 && (k + 1) % b->period == 0) MSA_TRY(launch_replacement(b, tol, max_it));
    }
    return MSA_OK;
}

// Host-side check of a captured graph before it is instantiated: only kernel
// nodes (no memset or copy node: those have given wrong results on replays
// here), each with a non-empty grid and workgroup, at most 2^32 - 1
// work-items per dimension (the dispatch packet's 32-bit grid size), a
// workgroup within the kernel's own limit (its __launch_bounds__), and LDS
// within the 160 KB of a CU.
int validate_graph(hipGraph_t g, long long *n_kernels)
{
    size_t nn = 0;
    MSA_HIP(hipGraphGetNodes(g, nullptr, &nn));
    std::vector<hipGraphNode_t> nodes(nn);
    if (nn) MSA_HIP(hipGraphGetNodes(g, nodes.data(), &nn));
    long long k = 0;
    for (size_t i = 0; i < nn; i++) {
        hipGraphNodeType t;
        MSA_HIP(hipGraphNodeGetType(nodes[i], &t));
        if (t == hipGraphNodeTypeEmpty) continue;
        MSA_CHECK(t == hipGraphNodeTypeKernel, MSA_ERR_STATE,
                  "captured graph: node %zu has type %d; only kernel nodes may be recorded", i,
                  (int)t);
        hipKernelNodeParams p{};
        MSA_HIP(hipGraphKernelNodeGetParams(nodes[i], &p));
        const dim3 G = p.gridDim, B = p.blockDim;
        MSA_CHECK(G.x >= 1 && G.y >= 1 && G.z >= 1 && B.x >= 1 && B.y >= 1 && B.z >= 1,
                  MSA_ERR_STATE, "captured graph: kernel node %zu has grid (%u,%u,%u) block (%u,%u,%u)",
                  i, G.x, G.y, G.z, B.x, B.y, B.z);
        const unsigned long long lim = 0xffffffffull;
        MSA_CHECK((unsigned long long)G.x * B.x <= lim && (unsigned long long)G.y * B.y <= lim &&
                      (unsigned long long)G.z * B.z <= lim,
                  MSA_ERR_STATE, "captured graph: kernel node %zu exceeds 2^32-1 work-items in a "
                  "dimension", i);
        hipFuncAttributes fa{};
        MSA_HIP(hipFuncGetAttributes(&fa, p.func));
        const long long threads = (long long)B.x * B.y * B.z;
        MSA_CHECK(threads <= fa.maxThreadsPerBlock, MSA_ERR_STATE,
                  "captured graph: kernel node %zu has %lld threads per workgroup, its kernel "
                  "allows %d", i, threads, fa.maxThreadsPerBlock);
        MSA_CHECK(fa.sharedSizeBytes + p.sharedMemBytes <= 160u * 1024u, MSA_ERR_STATE,
                  "captured graph: kernel node %zu needs %zu bytes of LDS", i,
                  (size_t)(fa.sharedSizeBytes + p.sharedMemBytes));
        k++;
    }
    *n_kernels = k;
    return MSA_OK;
}

// Records the launches of fn() on the context stream into an instantiated,
// host-validated graph (n_kernels: its kernel nodes).
template <typename F>
int capture(msa_batch *b, hipGraphExec_t *out, F &&fn, long long *n_kernels = nullptr)
{
    hipStream_t s = b->ctx->stream;
    MSA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = fn();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    long long nk = 0;
    if (rc == MSA_OK && e == hipSuccess) rc = validate_graph(g, &nk);
    if (rc == MSA_OK && e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (rc != MSA_OK) return rc;
    MSA_HIP(e);
    if (n_kernels) *n_kernels = nk;
    return MSA_OK;
}

// Smallest block the iteration loop replays: two iterations (the product
// buffers alternate), or one replacement period.
int tail_iters(const msa_batch *b, int kind) { return kind == 1 ? b->period : 2; }

// The caller's block of `iters` iterations and, when shorter, the tail
// block, recorded together (one capture in the batch's count).
int build_graph(msa_batch *b, int kind, int iters, double tol, int max_it)
{
    drop_graph(b);
    MSA_TRY(capture(b, &b->graph, [&] { return launch_block(b, kind, iters, tol, max_it); },
                    &b->graph_kernels));
    const int tail = tail_iters(b, kind);
    if (tail < iters) {
        MSA_TRY(capture(b, &b->graph_tail, [&] { return launch_block(b, kind, tail, tol, max_it); }));
        b->graph_tail_iters = tail;
    }
    b->graph_iters = iters;
    b->graph_kind = kind;
    b->graph_tol = tol;
    b->graph_maxit = max_it;
    b->graph_captures++;
    return MSA_OK;
}

// Iterate until the polled counter (0: PCG active samples, 1: samples still
// refining) reaches zero.  kind 1 interleaves residual replacements.
int pcg_loop(msa_batch *b, int kind, double tol, int max_it, int graph_iters)
{
    if (b->persistent && kind == 0 && b->precond == MSA_PRECOND_JACOBI && !b->timing.on) {
        MSA_CHECK(b->T == 1, MSA_ERR_STATE,
                  "the persistent launch takes one-sample tiles (the batch has lane tile %d)", b->T);
        ReduceArgs ra{}, rb{};
        ra.op = OP_A;
        rb.op = OP_B;
        rb.tol = tol;
        rb.max_it = max_it;
        ra.active_count = rb.active_count = b->d_counts;
        ra.ir_count = rb.ir_count = b->d_counts + 1;
        int G = 0;
        MSA_HIP(launch_pcg_persist(b->ctx->stream, is_mixed(b), b->unroll, b->d_cls, b->d_items,
                                   (int)b->items.size(), b->d_tiles, (int)b->tiles.size(), ra, rb,
                                   max_it + 1, &G));
        b->persist_launches++;
        b->persist_grid = G;
        int active = 0;
        MSA_TRY(read_count(b, 0, &active));
        MSA_CHECK(active <= 0, MSA_ERR_STATE, "persistent PCG: %d samples still active", active);
        return MSA_OK;
    }
    const bool use_graph = graph_iters > 0 && !b->timing.on;
    int block = graph_iters > 0 ? graph_iters : 8;
    if (kind == 1) block = ((block + b->period - 1) / b->period) * b->period;
    if (use_graph && (b->graph == nullptr || b->graph_iters != block || b->graph_kind != kind ||
                      b->graph_tol != tol || b->graph_maxit != max_it))
        MSA_TRY(build_graph(b, kind, block, tol, max_it));
    const int which = kind == 1 ? 1 : 0;
    // Once a sample has converged, the rest follow within a few iterations
    // in a Monte Carlo batch, and every launch of a block still dispatches
    // the tiles of the converged samples: the loop then replays the tail
    // block, which stops at most tail - 1 iterations after the last sample.
    const int tail = tail_iters(b, kind);
    int total = 0;
    for (auto &C : b->cls) total += C.nsamp;
    long long done = 0;
    for (;;) {
        int active = 0;
        MSA_TRY(read_count(b, which, &active));
        if (active <= 0) break;
        if (done > (long long)max_it + block) {
            set_error("PCG driver: samples still active after %lld iterations", done);
            return MSA_ERR_STATE;
        }
        const int it = active < total && tail < block ? tail : block;
        if (use_graph) {
            MSA_HIP(hipGraphLaunch(it == block ? b->graph : b->graph_tail, b->ctx->stream));
            b->graph_replays++;
        } else {
            MSA_TRY(launch_block(b, kind, it, tol, max_it));
        }
        done += it;
    }
    return MSA_OK;
}

// skip_mf: leave out the classes whose replacement residual is matrix-free
ColourLaunch colour_launch(const msa_batch *b, int colour, int *nblocks, bool skip_mf)
{
    ColourLaunch cl{};
    cl.ncls = (int)b->cls.size();
    cl.colour = colour;
    cl.prefix[0] = 0;
    for (size_t c = 0; c < b->cls.size(); c++) {
        const MeshHost &m = b->cls[c].mesh->h;
        int cnt = colour < m.ncolours ? m.colour_ptr[colour + 1] - m.colour_ptr[colour] : 0;
        if (skip_mf && b->cls[c].dev.mf) cnt = 0;
        cl.prefix[c + 1] = cl.prefix[c] + cnt;
    }
    *nblocks = cl.prefix[cl.ncls];
    return cl;
}

int max_colours(const msa_batch *b)
{
    int m = 0;
    for (auto &c : b->cls) m = std::max(m, c.mesh->h.ncolours);
    return m;
}

msa_ctx *g_default_ctx = nullptr;
std::once_flag g_default_once;
std::mutex g_default_mutex;
double *g_drop_buf = nullptr;  // device scratch of the Fortran drop-ins

}  // namespace

// ================================================================ C ABI

namespace {

// Device copies of a mesh's host analysis.
int upload_mesh(msa_mesh *m)
{
    msa_ctx *ctx = m->ctx;
    const MeshHost &h = m->h;
    hipError_t e = hipSetDevice(ctx->device);
    auto up = [&](void **dst, const void *src, size_t bytes) {
        if (e != hipSuccess) return;
        e = hipMalloc(dst, bytes ? bytes : 16);
        if (e == hipSuccess && bytes)
            e = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    };
    up((void **)&m->d_xyz, h.xyz.data(), h.xyz.size() * sizeof(double));
    up((void **)&m->d_conn, h.conn.data(), h.conn.size() * sizeof(int));
    up((void **)&m->d_node_block, h.node_block.data(), h.node_block.size() * sizeof(int));
    up((void **)&m->d_node_free, h.node_free.data(), h.node_free.size());
    up((void **)&m->d_emap, h.emap.data(), h.emap.size());
    up((void **)&m->d_colour_elems, h.colour_elems.data(), h.colour_elems.size() * sizeof(int));
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        set_error("mesh upload failed: %s", hipGetErrorString(e));
        return MSA_ERR_HIP;
    }
    return MSA_OK;
}

}  // namespace

// Direct solver helpers (band Cholesky).
namespace {

// Band order, factor storage and solve buffers of one class; *changed is
// set when its ClassDev changed.  shared: one factor for all samples (the
// first-sample preconditioner), else one per sample (the direct solver).
int band_setup(msa_batch *b, ClassHost &C, bool *changed, bool shared = false)
{
    hipStream_t s = b->ctx->stream;
    bool upload = false;
    if (!C.dev.bperm) {
        if (C.bw < 0) MSA_TRY(band_order(C.mesh->h, C.bperm, C.binv, C.bw));
        void *p = nullptr;
        MSA_TRY(dalloc(b, &p, C.bperm.size() * sizeof(int)));
        MSA_HIP(hipMemcpyAsync(p, C.bperm.data(), C.bperm.size() * sizeof(int), hipMemcpyHostToDevice, s));
        C.dev.bperm = (const int *)p;
        MSA_TRY(dalloc(b, &p, C.binv.size() * sizeof(int)));
        MSA_HIP(hipMemcpyAsync(p, C.binv.data(), C.binv.size() * sizeof(int), hipMemcpyHostToDevice, s));
        C.dev.binv = (const int *)p;
        C.dev.bn = (int)C.bperm.size();
        // storage width: bandwidth + one block, rounded to 64
        C.dev.bW = (int)(((C.bw + 1 + kBandNB + 63) / 64) * 64);
        upload = true;
    }
    const size_t per = (size_t)C.dev.bn * C.dev.bW;
    const size_t need = shared ? per : (size_t)C.nsamp * per;
    if (C.band_cap < need) {
        void *p = nullptr;
        MSA_TRY(dalloc(b, &p, need * sizeof(double)));
        C.dev.band = (double *)p;
        C.band_cap = need;
        upload = true;
    }
    if (!C.dev.by) {
        void *p = nullptr;
        MSA_TRY(dalloc(b, &p, (size_t)C.nsamp * C.dev.bn * sizeof(double)));
        C.dev.by = (double *)p;
        const int nchunk = (C.dev.bW + kBandChunkRows - 1) / kBandChunkRows;
        MSA_TRY(dalloc(b, &p, (size_t)C.nsamp * nchunk * kBandNB * sizeof(double)));
        C.dev.bpart = (double *)p;
        MSA_TRY(dalloc(b, &p, (size_t)C.nsamp * sizeof(int)));
        C.dev.bfail = (int *)p;
        upload = true;
    }
    const long long ss = shared ? 0 : (long long)per;
    if (C.dev.band_ss != ss) {
        C.dev.band_ss = ss;
        upload = true;
    }
    if (upload) *changed = true;
    return MSA_OK;
}

int prepare_band(msa_batch *b)
{
    bool upload = false;
    for (auto &C : b->cls) MSA_TRY(band_setup(b, C, &upload));
    if (upload) MSA_TRY(upload_class_table(b));
    return MSA_OK;
}

int band_backward(msa_batch *b, const ClassDev *cp, int n, int W, int ns, int mode);

// Forward and backward sweeps with the band factor of every sample (in
// place on by): one launch for bands of at most kBandOneMax rows, else two
// launch pairs per 64-column block.  mode >= 0: coarsest multigrid level,
// samples selected by the cycle's mode.
int band_sweeps(msa_batch *b, const ClassDev *cp, int n, int W, int ns, int mode)
{
    hipStream_t s = b->ctx->stream;
    if (n <= kBandOneMax)
        return timed(b, KK_OTHER, [&] { return launch_band_solve_one(s, cp, ns, n, mode); });
    for (int k0 = 0; k0 < n; k0 += kBandNB) {
        const int nk = std::min(kBandNB, n - k0), r0 = k0 + nk, m = std::min(n - r0, W - kBandNB);
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_band_fwd(s, cp, ns, k0, nk, r0, m, mode); }));
    }
    return band_backward(b, cp, n, W, ns, mode);
}

// The backward sweep L^T x = y alone (in place on by), block by block.
int band_backward(msa_batch *b, const ClassDev *cp, int n, int W, int ns, int mode)
{
    hipStream_t s = b->ctx->stream;
    const int last = ((n - 1) / kBandNB) * kBandNB;
    for (int k0 = last; k0 >= 0; k0 -= kBandNB) {
        const int nk = std::min(kBandNB, n - k0), r0 = k0 + nk, m = std::min(n - r0, W - kBandNB);
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_band_bwd(s, cp, ns, k0, nk, r0, m, mode); }));
    }
    return MSA_OK;
}

// Zero-energy-mode buffers and switches of one class for the next
// factorization (on when b->zem_tol > 0); *changed when its ClassDev changed.
int zem_setup(msa_batch *b, ClassHost &C, bool *changed)
{
    hipStream_t s = b->ctx->stream;
    ClassDev &D = C.dev;
    if (b->zem_tol <= 0.0) {
        if (D.zv) {
            D.zv = nullptr;
            D.zcol = D.zcount = nullptr;
            D.zmask = nullptr;
            *changed = true;
        }
        return MSA_OK;
    }
    const int nblk = (D.bn + kBandNB - 1) / kBandNB;
    if (C.zem_alloc < b->zem_max) {
        void *p = nullptr;
        MSA_TRY(dalloc(b, &p, (size_t)C.nsamp * b->zem_max * D.bn * sizeof(double)));
        C.zv = (double *)p;
        MSA_TRY(dalloc(b, &p, (size_t)C.nsamp * b->zem_max * sizeof(int)));
        C.zcol = (int *)p;
        MSA_TRY(dalloc(b, &p, (size_t)C.nsamp * sizeof(int)));
        C.zcount = (int *)p;
        MSA_TRY(dalloc(b, &p, (size_t)C.nsamp * nblk * sizeof(unsigned long long)));
        C.zmask = (unsigned long long *)p;
        C.zem_alloc = b->zem_max;
    }
    if (D.zv != C.zv || D.zem_tol != b->zem_tol || D.zem_max != b->zem_max) {
        D.zv = C.zv;
        D.zcol = C.zcol;
        D.zcount = C.zcount;
        D.zmask = C.zmask;
        D.zem_tol = b->zem_tol;
        D.zem_max = b->zem_max;
        *changed = true;
    }
    MSA_HIP(hipMemsetAsync(C.zv, 0, (size_t)C.nsamp * b->zem_max * D.bn * sizeof(double), s));
    MSA_HIP(hipMemsetAsync(C.zcol, 0xff, (size_t)C.nsamp * b->zem_max * sizeof(int), s));
    MSA_HIP(hipMemsetAsync(C.zcount, 0, (size_t)C.nsamp * sizeof(int), s));
    MSA_HIP(hipMemsetAsync(C.zmask, 0, (size_t)C.nsamp * nblk * sizeof(unsigned long long), s));
    return MSA_OK;
}

// The modes of every sample of class c after its factorization: mode k of a
// sample is L^-T of its saved row (SkylineMatrix2D.cs:138-157), by the
// backward sweep on by (free after a factorization), mode by mode.
int zem_modes(msa_batch *b, size_t c)
{
    hipStream_t s = b->ctx->stream;
    ClassHost &C = b->cls[c];
    const int n = C.dev.bn, ns = C.nsamp, zmax = b->zem_max;
    std::vector<int> cnt(ns);
    MSA_HIP(hipMemcpyAsync(cnt.data(), C.zcount, ns * sizeof(int), hipMemcpyDeviceToHost, s));
    MSA_HIP(hipStreamSynchronize(s));
    int most = 0;
    for (int v : cnt) most = std::max(most, std::min(v, zmax));
    const size_t row = (size_t)n * sizeof(double);
    for (int k = 0; k < most; k++) {
        MSA_HIP(hipMemcpy2DAsync(C.dev.by, row, C.zv + (size_t)k * n, zmax * row, row, ns,
                                 hipMemcpyDeviceToDevice, s));
        MSA_TRY(band_backward(b, b->d_cls + c, n, C.dev.bW, ns, -1));
        MSA_HIP(hipMemcpy2DAsync(C.zv + (size_t)k * n, zmax * row, C.dev.by, row, row, ns,
                                 hipMemcpyDeviceToDevice, s));
    }
    return MSA_OK;
}

// Forward and backward band sweeps of class c for every sample, then the
// solution into the PCG solution vector x.
int band_solve_class(msa_batch *b, size_t c, bool from_fs)
{
    hipStream_t s = b->ctx->stream;
    ClassHost &C = b->cls[c];
    const ClassDev *cp = b->d_cls + c;
    const int n = C.dev.bn, W = C.dev.bW, ns = C.nsamp;
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_band_rhs(s, b->T, cp, ns, n, from_fs ? 1 : 0); }));
    MSA_TRY(band_sweeps(b, cp, n, W, ns, -1));
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_band_out(s, b->T, cp, ns, 3 * C.dev.nb); }));
    return MSA_OK;
}

// Residual |f - K x|/|f| and status of a direct solve (OP_DIRECT).
int direct_stats(msa_batch *b)
{
    hipStream_t s = b->ctx->stream;
    // a fill kernel, not a memset: the Newmark step graph records this
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_set_int(s, b->d_counts, 1, 0); }));
    MSA_TRY(timed(b, KK_OTHER, [&] {
        return launch_pcg_init_warm(s, b->T, b->d_cls, b->d_items, (int)b->items.size());
    }));
    return reduce(b, OP_DIRECT);
}

// ---------------------------------------------------------------- multigrid

constexpr int kMgCoarsestMax = 12000;  // free DOFs of a coarsest level (band Cholesky per sample)

ClassHost &tab(msa_batch *b, int i)
{
    return i < (int)b->cls.size() ? b->cls[i] : b->lv[i - b->cls.size()];
}

template <typename V>
int upload_vec(msa_batch *b, const std::vector<V> &v, const V **out)
{
    void *p = nullptr;
    MSA_TRY(dalloc(b, &p, std::max<size_t>(v.size(), 1) * sizeof(V)));
    if (!v.empty()) MSA_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
    *out = (const V *)p;
    return MSA_OK;
}

int items_to_device(msa_batch *b, const std::vector<int4> &v, int4 **out, int *n)
{
    const int4 *p = nullptr;
    MSA_TRY(upload_vec(b, v, &p));
    *out = (int4 *)p;
    *n = (int)v.size();
    return MSA_OK;
}

// CSR rows -> ELL (entry k of row i at k*nrows + i; padding column 0,
// weight 0); returns the width.
int to_ell(const std::vector<int> &ptr, const std::vector<int> &col, const std::vector<double> &w,
           std::vector<int> &ecol, std::vector<float> &ew)
{
    const int n = (int)ptr.size() - 1;
    int width = 0;
    for (int i = 0; i < n; i++) width = std::max(width, ptr[i + 1] - ptr[i]);
    ecol.assign((size_t)width * n, 0);
    ew.assign((size_t)width * n, 0.f);
    for (int i = 0; i < n; i++)
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            ecol[(size_t)(k - ptr[i]) * n + i] = col[k];
            ew[(size_t)(k - ptr[i]) * n + i] = (float)w[k];
        }
    return width;
}

void class_items(const ClassHost &C, int ci, std::vector<int4> &out)
{
    for (int t = 0; t < C.ntiles; t++)
        for (int k = 0; k < C.nchunks; k++) out.push_back(make_int4(ci, t, k, 0));
}

// Work items of the box-grid smoothing lists: runs of kMfRun chunks of one
// tile, the run length in the item's w.
constexpr int kMfRun = 4;
void class_run_items(const ClassHost &C, int ci, std::vector<int4> &out)
{
    for (int t = 0; t < C.ntiles; t++)
        for (int k = 0; k < C.nchunks; k += kMfRun)
            out.push_back(make_int4(ci, t, k, std::min(kMfRun, C.nchunks - k)));
}

// Which box-grid classes take the matrix-free PCG product (the setting,
// MIXED, one-sample tiles, the multigrid preconditioner set up: its unit
// element matrix and grid tables serve the product), and their work items.
// Called after the preconditioner values of a solve; a change re-uploads
// the class table and drops the captured graphs (they hold the items).
int sync_mf_pcg(msa_batch *b)
{
    bool changed = false, any = false;
    for (auto &C : b->cls) {
        const int want = b->mf_product && is_mixed(b) && b->T == 1 &&
                         b->precond == MSA_PRECOND_MULTIGRID && b->mg_ready && C.dev.mf &&
                         C.dev.mf_box;
        changed |= C.dev.mf_pcg != want;
        C.dev.mf_pcg = want;
        any |= want != 0;
    }
    if (!changed) return MSA_OK;
    if (any) {
        std::vector<int4> st, mf;
        for (int c = 0; c < (int)b->cls.size(); c++) {
            if (b->cls[c].dev.mf_pcg)
                class_run_items(b->cls[c], c, mf);
            else
                class_items(b->cls[c], c, st);
        }
        MSA_TRY(items_to_device(b, st, &b->d_items_st, &b->n_items_st));
        MSA_TRY(items_to_device(b, mf, &b->d_items_mf, &b->n_items_mf));
    } else {
        b->n_items_st = b->n_items_mf = 0;
    }
    drop_graph(b);
    return upload_class_table(b);
}

// Work items of the grouped gather assembly: tiles in groups of kAsmGroup.
void class_group_items(const ClassHost &C, int ci, std::vector<int4> &out)
{
    for (int g = 0; g < (C.ntiles + kAsmGroup - 1) / kAsmGroup; g++)
        for (int k = 0; k < C.nchunks; k++) out.push_back(make_int4(ci, g, k, 0));
}

// Gather assembly of the classes of `items` (one-sample tiles: grouped).
int assemble_items(msa_batch *b, const int4 *items, int nitems, const int4 *gitems, int ngitems)
{
    hipStream_t s = b->ctx->stream;
    const bool mx = is_mixed(b);
    return timed(b, KK_ASSEMBLE, [&] {
        return b->T == 1 ? launch_assemble_gather_g(s, mx, b->d_cls, gitems, ngitems)
                         : launch_assemble_gather(s, mx, b->T, b->d_cls, items, nitems);
    });
}

// Matrix-free smoothing products for a class on a uniform structured grid
// (mixed precision, one-sample tiles); other classes keep the stored values.
int mf_setup(msa_batch *b, ClassHost &C)
{
    if (!is_mixed(b) || b->T != 1 || C.dev.mf) return MSA_OK;
    GridInfo g;
    if (grid_info(C.mesh->h, g) != MSA_OK) return MSA_OK;
    ClassDev &D = C.dev;
    MSA_TRY(upload_vec(b, g.gblock, &D.mf_gblock));
    MSA_TRY(upload_vec(b, g.cell_elem, &D.mf_cell));
    MSA_TRY(upload_vec(b, g.bpos, &D.mf_bpos));
    void *p = nullptr;
    MSA_TRY(dalloc(b, &p, 640 * sizeof(float)));
    D.mf_k = (float *)p;
    MSA_TRY(dalloc(b, &p, kMfK64 * sizeof(double)));
    D.mf_k64 = (double *)p;
    D.mf_nx = g.nx;
    D.mf_ny = g.ny;
    D.mf_nz = g.nz;
    for (int o = 0; o < 8; o++) D.mf_corner[o] = g.corner_node[o];
    D.mf_cell_id = g.cell_identity;
    D.mf_box = g.box;
    for (int d = 0; d < 3; d++) {
        D.mf_box_lo[d] = g.box_lo[d];
        D.mf_box_hi[d] = g.box_hi[d];
    }
    D.mf = 1;
    return MSA_OK;
}

// A box grid of fully free blocks whose smoothing products fit the MFMA
// kernel: one-sample MIXED tiles, at most kMfmaMaxCells cells in x, and at
// most as many y blocks as the class has PCG partial slots per sample.
bool mfma_ok(const msa_batch *b, const ClassHost &C)
{
    const ClassDev &D = C.dev;
    if (!is_mixed(b) || b->T != 1 || !D.mf || !D.mf_box || !D.mf_allfree) return false;
    const int cxl = std::max(D.mf_box_lo[0] - 1, 0), cxh = std::min(D.mf_box_hi[0], D.mf_nx - 2);
    const int nyb = (D.mf_box_hi[1] - D.mf_box_lo[1] + kMfmaRows) / kMfmaRows;
    return cxh - cxl + 1 <= kMfmaMaxCells && nyb <= C.nchunks;
}

// Items of the MFMA smoother: (class, sample, y block of kMfmaRows rows).
void mfma_items(const ClassHost &C, int ci, std::vector<int4> &out)
{
    const int nyb = (C.dev.mf_box_hi[1] - C.dev.mf_box_lo[1] + kMfmaRows) / kMfmaRows;
    for (int smp = 0; smp < C.nsamp; smp++)
        for (int y = 0; y < nyb; y++) out.push_back(make_int4(ci, smp, y, 0));
}

// Coarse hierarchies of every class whose mesh is a structured grid that
// halves at least once and whose coarsest level stays small; the other
// classes keep the Jacobi preconditioner.
int mg_build(msa_batch *b, int max_levels)
{
    const int ncls = (int)b->cls.size();
    b->mg_nlev.assign(ncls, 0);
    std::vector<std::vector<int>> depth_classes(1);  // per depth: table indices
    for (int c = 0; c < ncls; c++) {
        // host coarsening first: commit only a usable hierarchy
        std::vector<MeshHost> meshes;
        std::vector<MgTransfer> tr;
        const MeshHost *f = &b->cls[c].mesh->h;
        while (max_levels <= 0 || (int)meshes.size() < max_levels) {
            MeshHost ch;
            MgTransfer t;
            if (mg_coarsen(*f, ch, t) != MSA_OK) break;
            meshes.push_back(std::move(ch));
            tr.push_back(std::move(t));
            f = &meshes.back();
        }
        set_error("");
        if (meshes.empty() || meshes.back().n_free > kMgCoarsestMax) continue;
        b->mg_nlev[c] = (int)meshes.size();
        depth_classes[0].push_back(c);
        MSA_TRY(mf_setup(b, b->cls[c]));
        int parent = c;
        for (size_t l = 0; l < meshes.size(); l++) {
            msa_mesh *m = new msa_mesh();
            m->ctx = b->ctx;
            m->h = std::move(meshes[l]);
            b->mg_meshes.push_back(m);
            MSA_TRY(upload_mesh(m));
            const int ci = (int)(b->cls.size() + b->lv.size());
            b->lv.emplace_back();
            ClassHost &L = b->lv.back();
            MSA_TRY(setup_class(b, L, m, b->cls[c].nsamp));
            L.has_E = true;
            MSA_TRY(mf_setup(b, L));
            ClassDev &D = L.dev;
            D.mg_root = c;
            D.mg_st = b->cls[c].dev.st;
            D.mg_fine = parent;
            D.mg_depth = (int)l + 1;
            const MgTransfer &t = tr[l];
            MSA_TRY(upload_vec(b, t.child, &D.mg_child));
            std::vector<int> col;
            std::vector<float> w;
            D.mg_r_width = to_ell(t.r_ptr, t.r_col, t.r_w, col, w);
            MSA_TRY(upload_vec(b, col, &D.mg_r_col));
            MSA_TRY(upload_vec(b, w, &D.mg_r_w));
            ClassDev &P = tab(b, parent).dev;
            P.mg_coarse = ci;
            P.mg_p_width = to_ell(t.p_ptr, t.p_col, t.p_w, col, w);
            MSA_TRY(upload_vec(b, col, &P.mg_p_col));
            MSA_TRY(upload_vec(b, w, &P.mg_p_w));
            void *p = nullptr;
            MSA_TRY(dalloc(b, &p, (size_t)tab(b, parent).ntiles * b->T * sizeof(double)));
            P.mg_lam = (double *)p;
            MSA_TRY(dalloc(b, &p, (size_t)tab(b, parent).ntiles * b->T * (2 * kLzMax + 1) *
                                      sizeof(double)));
            P.mg_lz = (double *)p;
            if ((int)depth_classes.size() <= D.mg_depth) depth_classes.resize(D.mg_depth + 1);
            depth_classes[D.mg_depth].push_back(ci);
            if (l + 1 == meshes.size()) {
                bool changed = false;
                ClassHost &Cc = b->lv.back();
                MSA_TRY(band_setup(b, Cc, &changed));
                // MIXED batches factor the coarsest level from an FP64 copy
                // of its values (the FP32 rounding, amplified by the
                // condition number of a thin structure's coarse operator,
                // would weaken the cycle)
                if (is_mixed(b) && b->T == 1) {
                    void *p = nullptr;
                    MSA_TRY(dalloc(b, &p, (size_t)Cc.ntiles * Cc.tile_vals * sizeof(double)));
                    Cc.dev.val64 = (double *)p;
                    if (Cc.dev.stencil)  // lower slots may not be written (as the FP32 values)
                        MSA_HIP(hipMemset(p, 0, (size_t)Cc.ntiles * Cc.tile_vals * sizeof(double)));
                }
                // small coarsest operators: dense inverse, one launch per cycle
                if (Cc.dev.bn <= kDenseMax) {
                    void *p = nullptr;
                    MSA_TRY(dalloc(b, &p, (size_t)Cc.nsamp * Cc.dev.bn * Cc.dev.bn * sizeof(double)));
                    Cc.dev.binvd = (double *)p;
                }
                b->mg_coarsest.push_back(ci);
            }
            parent = ci;
        }
    }
    // structured transfers where both sides are uniform box grids of fully
    // free blocks
    auto all_free = [](const MeshHost &m) {
        for (int bb = 0; bb < m.nb; bb++)
            if (m.node_free[m.block_node[bb]] != 7) return false;
        return true;
    };
    for (int ci = 0; ci < ncls + (int)b->lv.size(); ci++) {
        ClassDev &D = tab(b, ci).dev;
        D.mf_allfree = D.mf && D.mf_box && all_free(tab(b, ci).mesh->h) ? 1 : 0;
    }
    for (auto &L : b->lv) {
        ClassHost &P = tab(b, L.dev.mg_fine);
        if (L.dev.mf && P.dev.mf && L.dev.mf_box && P.dev.mf_box && all_free(L.mesh->h) &&
            all_free(P.mesh->h)) {
            L.dev.mg_rstruct = 1;
            P.dev.mg_pstruct = 1;
        }
    }
    b->mg_depth = (int)depth_classes.size() - 1;
    b->mg_smooth.assign(b->mg_depth, nullptr);
    b->mg_nsmooth.assign(b->mg_depth, 0);
    b->mg_sm_a.assign(b->mg_depth, nullptr);
    b->mg_nsm_a.assign(b->mg_depth, 0);
    b->mg_sm_f.assign(b->mg_depth, nullptr);
    b->mg_sm_b.assign(b->mg_depth, nullptr);
    b->mg_nsm_b.assign(b->mg_depth, 0);
    b->mg_sm_m.assign(b->mg_depth, nullptr);
    b->mg_nsm_m.assign(b->mg_depth, 0);
    b->mg_sm_e.assign(b->mg_depth, nullptr);
    b->mg_nsm_e.assign(b->mg_depth, 0);
    b->mg_nsm_f.assign(b->mg_depth, 0);
    b->mg_restr.assign(b->mg_depth + 1, nullptr);
    b->mg_nrestr.assign(b->mg_depth + 1, 0);
    std::vector<int4> all, co;
    for (int d = 0; d <= b->mg_depth; d++) {
        std::vector<int4> sm, sma, smf, smb, smm, sme, re;
        for (int ci : depth_classes[d]) {
            ClassHost &C = tab(b, ci);
            if (d > 0) {
                class_items(C, ci, re);
                class_items(C, ci, all);
            }
            if (C.dev.mg_coarse >= 0) {
                class_items(C, ci, sm);
                if (mfma_ok(b, C)) {
                    mfma_items(C, ci, smm);
                    class_run_items(C, ci, sme);
                } else if (C.dev.mf && C.dev.mf_box && b->T == 1) {
                    class_run_items(C, ci, smb);
                } else {
                    class_items(C, ci, !C.dev.mf ? sma : (C.dev.mf_box ? smb : smf));
                }
            } else if (d > 0) {
                class_items(C, ci, co);
            }
        }
        if (d < b->mg_depth) {
            MSA_TRY(items_to_device(b, sm, &b->mg_smooth[d], &b->mg_nsmooth[d]));
            MSA_TRY(items_to_device(b, sma, &b->mg_sm_a[d], &b->mg_nsm_a[d]));
            MSA_TRY(items_to_device(b, smf, &b->mg_sm_f[d], &b->mg_nsm_f[d]));
            MSA_TRY(items_to_device(b, smb, &b->mg_sm_b[d], &b->mg_nsm_b[d]));
            MSA_TRY(items_to_device(b, smm, &b->mg_sm_m[d], &b->mg_nsm_m[d]));
            MSA_TRY(items_to_device(b, sme, &b->mg_sm_e[d], &b->mg_nsm_e[d]));
        }
        if (d > 0) MSA_TRY(items_to_device(b, re, &b->mg_restr[d], &b->mg_nrestr[d]));
    }
    MSA_TRY(items_to_device(b, all, &b->mg_lv_items, &b->mg_nlv_items));
    {
        std::vector<int4> mfi, mfb;
        for (int c = 0; c < ncls; c++)
            if (b->cls[c].dev.mf) {
                if (b->cls[c].dev.mf_box)
                    class_run_items(b->cls[c], c, mfb);
                else
                    class_items(b->cls[c], c, mfi);
            }
        MSA_TRY(items_to_device(b, mfi, &b->mf_items, &b->n_mf_items));
        MSA_TRY(items_to_device(b, mfb, &b->mf_box_items, &b->n_mf_box_items));
    }
    std::vector<int4> gall;
    for (size_t l = 0; l < b->lv.size(); l++)
        class_group_items(b->lv[l], (int)(b->cls.size() + l), gall);
    MSA_TRY(items_to_device(b, gall, &b->mg_lv_gitems, &b->mg_nlv_gitems));
    MSA_TRY(items_to_device(b, co, &b->mg_co_items, &b->mg_nco_items));
    MSA_TRY(upload_class_table(b));
    b->mg_ready = false;
    return MSA_OK;
}

// Level matrices from the assembled classes: moduli averaged level by level,
// the unit-modulus element cache, gather assembly, D^-1, Gershgorin bounds
// and the band Cholesky factors of the coarsest levels.
int mg_values(msa_batch *b)
{
    b->mg_ready = false;
    if (b->precond != MSA_PRECOND_MULTIGRID || b->mg_depth == 0) {
        b->mg_ready = true;
        return MSA_OK;
    }
    hipStream_t s = b->ctx->stream;
    const bool mx = is_mixed(b);
    const int T = b->T, nc = (int)b->cls.size();
    bool upload = false;
    for (auto &L : b->lv) {
        const ClassDev &R = b->cls[L.dev.mg_root].dev;
        if (L.dev.dyn != R.dyn || L.dev.cK != R.cK || L.dev.cM != R.cM) {
            L.dev.dyn = R.dyn;
            L.dev.cK = R.cK;
            L.dev.cM = R.cM;
            L.dev.rho = R.rho;
            upload = true;
        }
        if (L.dev.dyn && !L.dev.mc) {
            void *p = nullptr;
            MSA_TRY(dalloc(b, &p, (size_t)L.mesh->h.n_elem * 64 * sizeof(double)));
            L.dev.mc = (double *)p;
            upload = true;
        }
    }
    if (upload) MSA_TRY(upload_class_table(b));
    for (int d = 1; d <= b->mg_depth; d++)
        for (size_t l = 0; l < b->lv.size(); l++) {
            ClassHost &L = b->lv[l];
            if (L.dev.mg_depth != d) continue;
            MSA_TRY(timed(b, KK_OTHER, [&] {
                return launch_mg_coarsen_E(s, T, b->d_cls + nc + l, b->d_cls, L.ntiles, L.dev.n_elem);
            }));
        }
    for (size_t l = 0; l < b->lv.size(); l++) {
        ClassHost &L = b->lv[l];
        if (L.kc_nu == b->nu && L.kc_mass == (bool)L.dev.dyn) continue;
        MSA_HIP(hipMemsetAsync(L.dev.negdet, 0, sizeof(int), s));
        MSA_TRY(timed(b, KK_ASSEMBLE, [&] {
            return launch_khat_cache(s, b->d_cls + nc + l, L.dev.n_elem, b->nu, L.dev.dyn);
        }));
        L.kc_nu = b->nu;
        L.kc_mass = L.dev.dyn;
    }
    MSA_TRY(assemble_items(b, b->mg_lv_items, b->mg_nlv_items, b->mg_lv_gitems, b->mg_nlv_gitems));
    for (int ci = 0; ci < nc + (int)b->lv.size(); ci++)
        if (tab(b, ci).dev.mf)
            MSA_TRY(timed(b, KK_OTHER, [&] { return launch_mf_khat(s, b->d_cls + ci); }));
    MSA_TRY(timed(b, KK_OTHER, [&] {
        return launch_dinv(s, mx, T, b->d_cls, b->mg_lv_items, b->mg_nlv_items);
    }));
    for (int ci = 0; ci < nc + (int)b->lv.size(); ci++) {
        ClassHost &C = tab(b, ci);
        if (C.dev.mg_coarse < 0) continue;
        MSA_TRY(timed(b, KK_OTHER, [&] {
            return launch_fill_f64(s, C.dev.mg_lam, (size_t)C.ntiles * T, 0.0);
        }));
    }
    // Gershgorin bounds: matrix-free on box grids of fully free blocks (the
    // stored-value kernel skips those classes)
    for (int d = 0; d < b->mg_depth; d++) {
        MSA_TRY(timed(b, KK_OTHER, [&] {
            return launch_mg_gersh(s, mx, T, b->d_cls, b->mg_smooth[d], b->mg_nsmooth[d]);
        }));
        MSA_TRY(timed(b, KK_OTHER, [&] {
            return launch_mg_gersh_mf(s, b->d_cls, b->mg_sm_b[d], b->mg_nsm_b[d]);
        }));
        MSA_TRY(timed(b, KK_OTHER, [&] {
            return launch_mg_gersh_mf(s, b->d_cls, b->mg_sm_e[d], b->mg_nsm_e[d]);
        }));
    }
    MSA_TRY(mg_lanczos(b));
    for (int ci : b->mg_coarsest) {
        ClassHost &C = tab(b, ci);
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_fill_f64(s, C.dev.band, C.band_cap, 0.0); }));
        MSA_HIP(hipMemsetAsync(C.dev.bfail, 0, (size_t)C.nsamp * sizeof(int), s));
    }
    MSA_TRY(timed(b, KK_OTHER, [&] {
        return launch_band_scatter_any(s, mx, T, b->d_cls, b->mg_co_items, b->mg_nco_items);
    }));
    for (int ci : b->mg_coarsest) {
        ClassHost &C = tab(b, ci);
        const ClassDev *cp = b->d_cls + ci;
        const int n = C.dev.bn, W = C.dev.bW, ns = C.nsamp;
        for (int k0 = 0; k0 < n; k0 += kBandNB) {
            const int nk = std::min(kBandNB, n - k0), r0 = k0 + nk, m = std::min(n - r0, W - kBandNB);
            MSA_TRY(timed(b, KK_ASSEMBLE, [&] { return launch_chol_diag(s, cp, ns, k0, nk); }));
            MSA_TRY(timed(b, KK_ASSEMBLE, [&] { return launch_chol_panel(s, cp, ns, k0, nk, r0, m); }));
            MSA_TRY(timed(b, KK_ASSEMBLE, [&] { return launch_chol_update(s, cp, ns, k0, nk, r0, m); }));
        }
    }
    // a coarse operator that is not positive definite leaves the cycle undefined
    for (int ci : b->mg_coarsest) {
        ClassHost &C = tab(b, ci);
        std::vector<int> fail(C.nsamp);
        MSA_HIP(hipMemcpyAsync(fail.data(), C.dev.bfail, fail.size() * sizeof(int),
                               hipMemcpyDeviceToHost, s));
        MSA_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < C.nsamp; k++)
            MSA_CHECK(!fail[k], MSA_ERR_STATE,
                      "multigrid: coarsest operator of class %d, sample %d is not positive definite",
                      C.dev.mg_root, k);
    }
    // dense inverses of the coarsest operators; rows outside the band order
    // (constrained components) of their solution vectors stay zero
    for (int ci : b->mg_coarsest) {
        ClassHost &C = tab(b, ci);
        if (C.dev.binvd)
            MSA_TRY(timed(b, KK_ASSEMBLE, [&] { return launch_band_inverse(s, b->d_cls + ci, C.nsamp, C.dev.bn); }));
        const size_t nz = (size_t)C.ntiles * C.tile_vec;
        MSA_TRY(timed(b, KK_OTHER, [&] {
            return mx ? launch_fill_f32(s, (float *)C.dev.z, nz, 0.f)
                      : launch_fill_f64(s, (double *)C.dev.z, nz, 0.0);
        }));
    }
    b->mg_ready = true;
    return MSA_OK;
}

// First-sample direct preconditioner (the Monte Carlo calculator of
// MonteCarloSolverPCGDirectMatrixCalculator.cs:38-80: the LDL^T of the first
// sample's K, reused for every sample): the band Cholesky of sample 0's
// assembled values of every class, one factor shared by the class's samples.
int first_values(msa_batch *b)
{
    hipStream_t s = b->ctx->stream;
    const bool mx = is_mixed(b);
    bool upload = false;
    for (auto &C : b->cls) {
        MSA_TRY(band_setup(b, C, &upload, true));
        if (C.dev.zv) {  // no zero-energy modes in the shared preconditioner factor
            C.dev.zv = nullptr;
            C.dev.zcol = C.dev.zcount = nullptr;
            C.dev.zmask = nullptr;
            upload = true;
        }
        if (C.dev.mg_st != C.dev.st) {
            C.dev.mg_st = C.dev.st;  // the sweeps select samples by the cycle's mode
            upload = true;
        }
    }
    if (upload) MSA_TRY(upload_class_table(b));
    for (auto &C : b->cls) {
        MSA_TRY(timed(b, KK_OTHER, [&] {
            return launch_fill_f64(s, C.dev.band, (size_t)C.dev.bn * C.dev.bW, 0.0);
        }));
        MSA_HIP(hipMemsetAsync(C.dev.bfail, 0, sizeof(int), s));
    }
    MSA_TRY(timed(b, KK_OTHER, [&] {
        return launch_band_scatter_any(s, mx, b->T, b->d_cls, b->d_items, (int)b->items.size());
    }));
    for (size_t c = 0; c < b->cls.size(); c++) {
        ClassHost &C = b->cls[c];
        const ClassDev *cp = b->d_cls + c;
        const int n = C.dev.bn, W = C.dev.bW;
        for (int k0 = 0; k0 < n; k0 += kBandNB) {
            const int nk = std::min(kBandNB, n - k0), r0 = k0 + nk, m = std::min(n - r0, W - kBandNB);
            MSA_TRY(timed(b, KK_ASSEMBLE, [&] { return launch_chol_diag(s, cp, 1, k0, nk); }));
            MSA_TRY(timed(b, KK_ASSEMBLE, [&] { return launch_chol_panel(s, cp, 1, k0, nk, r0, m); }));
            MSA_TRY(timed(b, KK_ASSEMBLE, [&] { return launch_chol_update(s, cp, 1, k0, nk, r0, m); }));
        }
    }
    for (size_t c = 0; c < b->cls.size(); c++) {
        MSA_HIP(hipMemcpyAsync(b->ctx->h_pinned, b->cls[c].dev.bfail, sizeof(int),
                               hipMemcpyDeviceToHost, s));
        MSA_HIP(hipStreamSynchronize(s));
        MSA_CHECK(*b->ctx->h_pinned == 0, MSA_ERR_STATE,
                  "first-sample preconditioner: the first sample's K of class %zu is not "
                  "positive definite", c);
    }
    b->factored = false;  // a direct solver's factor of sample 0 was overwritten
    return MSA_OK;
}

// Values of the chosen preconditioner from the assembled classes.
int precond_values(msa_batch *b)
{
    MSA_CHECK(!(b->loaded && b->precond == MSA_PRECOND_MULTIGRID && b->mg_depth > 0),
              MSA_ERR_STATE,
              "the multigrid levels are built from the moduli: a batch of loaded matrices takes "
              "the Jacobi or first-sample preconditioner");
    if (b->precond != MSA_PRECOND_FIRST_SAMPLE) return mg_values(b);
    b->mg_ready = false;
    MSA_TRY(first_values(b));
    b->mg_ready = true;
    return MSA_OK;
}

// z = L^-T L^-1 r with the shared first-sample factor, for the samples the
// mode selects (see k_mg_smooth), and the r.z partials of the PCG.
int first_apply(msa_batch *b, int mode)
{
    hipStream_t s = b->ctx->stream;
    const bool mx = is_mixed(b);
    const int T = b->T;
    for (size_t c = 0; c < b->cls.size(); c++) {
        ClassHost &C = b->cls[c];
        const ClassDev *cp = b->d_cls + c;
        const int n = C.dev.bn, W = C.dev.bW, ns = C.nsamp;
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_mg_band_in(s, mx, T, cp, ns, n, mode); }));
        MSA_TRY(band_sweeps(b, cp, n, W, ns, mode));
    }
    return timed(b, KK_OTHER, [&] {
        return launch_first_out(s, mx, T, b->d_cls, b->d_items, (int)b->items.size(), mode);
    });
}

// z = B r for the classes with a hierarchy (B: one V-cycle); their z held
// D^-1 r on entry.  pbuf selects the free search-direction buffer, mode the
// samples served (see k_mg_smooth).
// One smoothing step (MG_PRE ... MG_POST2) of every class at depth d.
int mg_smooth_level(msa_batch *b, int d, int step, int pbuf, int mode)
{
    hipStream_t s = b->ctx->stream;
    const bool mx = is_mixed(b);
    const int T = b->T;
    const double ratio = b->mg_ratio;
    // timing kind KK_MG: the finest level's smoothing products only
    if (b->mg_nsm_a[d] > 0)
        MSA_TRY(timed(b, d == 0 ? KK_MG : KK_OTHER, [&] {
            return launch_mg_smooth(s, mx, T, b->d_cls, b->mg_sm_a[d], b->mg_nsm_a[d], step, pbuf,
                                    mode, ratio, 0);
        }));
    if (b->mg_nsm_f[d] > 0)
        MSA_TRY(timed(b, d == 0 ? KK_MG : KK_OTHER, [&] {
            return launch_mg_smooth(s, mx, T, b->d_cls, b->mg_sm_f[d], b->mg_nsm_f[d], step, pbuf,
                                    mode, ratio, 1);
        }));
    if (b->mg_nsm_b[d] > 0)
        MSA_TRY(timed(b, d == 0 ? KK_MG : KK_OTHER, [&] {
            return launch_mg_smooth(s, mx, T, b->d_cls, b->mg_sm_b[d], b->mg_nsm_b[d], step, pbuf,
                                    mode, ratio, 2);
        }));
    if (b->matrix_cores && b->mg_nsm_m[d] > 0)
        MSA_TRY(timed(b, d == 0 ? KK_MG : KK_OTHER, [&] {
            return launch_mg_smooth_mfma(s, b->d_cls, b->mg_sm_m[d], b->mg_nsm_m[d], step, pbuf,
                                         mode, ratio);
        }));
    if (!b->matrix_cores && b->mg_nsm_e[d] > 0)
        MSA_TRY(timed(b, d == 0 ? KK_MG : KK_OTHER, [&] {
            return launch_mg_smooth(s, mx, T, b->d_cls, b->mg_sm_e[d], b->mg_nsm_e[d], step, pbuf,
                                    mode, ratio, 2);
        }));
    return MSA_OK;
}

// The top of each sample's spectrum of D^-1 K on every smoothed level by
// mg_lz_steps Lanczos steps (k_lz_step), all levels in lock-step: each
// step's product is the RESID smoothing step with b = 0 (so the product is
// the level's own, matrix-free on box grids); mg_lam then holds
// min(Gershgorin, margin x top Ritz value).  Uses the levels' q, z, r and
// p0, which are free before a solve.
int mg_lanczos(msa_batch *b)
{
    const int K = b->mg_lz_steps;
    if (K <= 0) return MSA_OK;
    hipStream_t s = b->ctx->stream;
    const bool mx = is_mixed(b);
    const int T = b->T;
    std::vector<int> smoothed;
    int max_samp = 0;
    for (int ci = 0; ci < (int)(b->cls.size() + b->lv.size()); ci++)
        if (tab(b, ci).dev.mg_coarse >= 0) {
            smoothed.push_back(ci);
            max_samp = std::max(max_samp, tab(b, ci).nsamp);
        }
    if (smoothed.empty()) return MSA_OK;
    if (smoothed != b->lz_list) {
        void *p = nullptr;
        MSA_TRY(dalloc(b, &p, smoothed.size() * sizeof(int)));
        MSA_HIP(hipMemcpyAsync(p, smoothed.data(), smoothed.size() * sizeof(int),
                               hipMemcpyHostToDevice, s));
        b->d_lz_list = (int *)p;
        b->lz_list = smoothed;
    }
    const int nl = (int)smoothed.size();
    auto step = [&](int op, int k) {
        for (int d = 0; d < b->mg_depth; d++)
            MSA_TRY(timed(b, KK_OTHER, [&] {
                return launch_lz_step(s, mx, T, b->d_cls, b->mg_smooth[d], b->mg_nsmooth[d], op, k);
            }));
        return (int)MSA_OK;
    };
    auto sums = [&](int which, int k) {
        return timed(b, KK_OTHER, [&] {
            return launch_lz_sum(s, T, b->d_cls, b->d_lz_list, nl, max_samp, which, k);
        });
    };
    MSA_TRY(step(0, 0));
    MSA_TRY(sums(0, 0));
    for (int k = 0; k < K; k++) {
        for (int d = 0; d < b->mg_depth; d++) MSA_TRY(mg_smooth_level(b, d, MG_RESID, 0, 0));
        MSA_TRY(step(2, k));
        MSA_TRY(sums(1, k));
        if (k + 1 == K) break;
        MSA_TRY(step(3, k));
        MSA_TRY(sums(2, k));
    }
    return timed(b, KK_OTHER, [&] {
        return launch_lz_eig(s, b->d_cls, b->d_lz_list, nl, max_samp, K, b->mg_lz_margin);
    });
}

int mg_vcycle(msa_batch *b, int pbuf, int mode)
{
    if (b->precond == MSA_PRECOND_FIRST_SAMPLE) return first_apply(b, mode);
    if (b->precond != MSA_PRECOND_MULTIGRID || b->mg_depth == 0) return MSA_OK;
    hipStream_t s = b->ctx->stream;
    const bool mx = is_mixed(b);
    const int T = b->T;
    auto smooth = [&](int d, int step) { return mg_smooth_level(b, d, step, pbuf, mode); };
    for (int d = 0; d < b->mg_depth; d++) {
        MSA_TRY(smooth(d, MG_PRE));
        MSA_TRY(smooth(d, MG_RESID));
        MSA_TRY(timed(b, KK_OTHER, [&] {
            return launch_mg_restrict(s, mx, T, b->d_cls, b->mg_restr[d + 1], b->mg_nrestr[d + 1],
                                      pbuf, mode);
        }));
    }
    for (int ci : b->mg_coarsest) {
        ClassHost &C = tab(b, ci);
        const ClassDev *cp = b->d_cls + ci;
        const int n = C.dev.bn, W = C.dev.bW, ns = C.nsamp;
        if (C.dev.binvd) {
            MSA_TRY(timed(b, KK_OTHER, [&] { return launch_mg_dense_solve(s, mx, T, cp, ns, n, mode); }));
            continue;
        }
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_mg_band_in(s, mx, T, cp, ns, n, mode); }));
        MSA_TRY(band_sweeps(b, cp, n, W, ns, mode));
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_mg_band_out(s, mx, T, cp, ns, 3 * C.dev.nb, mode); }));
    }
    for (int d = b->mg_depth - 1; d >= 0; d--) {
        MSA_TRY(timed(b, KK_OTHER, [&] {
            return launch_mg_prolong(s, mx, T, b->d_cls, b->mg_smooth[d], b->mg_nsmooth[d], mode);
        }));
        MSA_TRY(smooth(d, MG_POST1));
        MSA_TRY(smooth(d, MG_POST2));
    }
    return MSA_OK;
}

}  // namespace

extern "C" {

const char *msa_last_error(void) { return g_last_error.c_str(); }
int msa_version(void) { return 100; }

int msa_ctx_create(int device, msa_ctx **out)
{
    MSA_CHECK(out, MSA_ERR_ARG, "msa_ctx_create: out is null");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        set_error("no HIP device available (%s)", hipGetErrorString(e));
        return MSA_ERR_NO_DEVICE;
    }
    MSA_CHECK(device >= 0 && device < n, MSA_ERR_ARG, "device %d out of range (%d devices)",
              device, n);
    MSA_HIP(hipSetDevice(device));
    // this device's address of the no-allocation entry point's array
    MSA_TRY(noalloc_setup(device));
    msa_ctx *c = new msa_ctx();
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_pinned, 64, hipHostMallocDefault);
    if (e != hipSuccess) {
        set_error("context setup failed: %s", hipGetErrorString(e));
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return MSA_ERR_HIP;
    }
    *out = c;
    return MSA_OK;
}

int msa_ctx_destroy(msa_ctx *ctx)
{
    if (!ctx) return MSA_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MSA_OK;
}

void *msa_ctx_stream(msa_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int msa_ctx_synchronize(msa_ctx *ctx)
{
    MSA_CHECK(ctx, MSA_ERR_ARG, "null context");
    MSA_HIP(hipStreamSynchronize(ctx->stream));
    return MSA_OK;
}

int msa_enumerate_dofs(int n_nodes, const unsigned char *constrained, int *dof, int *n_free)
{
    MSA_CHECK(n_nodes >= 0 && constrained && dof && n_free, MSA_ERR_ARG,
              "msa_enumerate_dofs: bad arguments");
    int total = 0;
    for (int n = 0; n < n_nodes; n++)
        for (int d = 0; d < 3; d++) dof[3 * n + d] = constrained[3 * n + d] ? -1 : total++;
    *n_free = total;
    return MSA_OK;
}

// ---------------------------------------------------------------- mesh

int msa_mesh_create(msa_ctx *ctx, int n_nodes, const double *xyz, int n_elem, const int *conn,
                    const int *dof, int n_free, msa_mesh **out)
{
    MSA_CHECK(ctx && out, MSA_ERR_ARG, "msa_mesh_create: null argument");
    *out = nullptr;
    msa_mesh *m = new msa_mesh();
    m->ctx = ctx;
    int rc = build_mesh_host(m->h, n_nodes, xyz, n_elem, conn, dof, n_free);
    if (rc != MSA_OK) {
        delete m;
        return rc;
    }
    rc = upload_mesh(m);
    if (rc != MSA_OK) {
        msa_mesh_destroy(m);
        return rc;
    }
    *out = m;
    return MSA_OK;
}


int msa_mesh_info(const msa_mesh *mesh, long long *info)
{
    MSA_CHECK(mesh && info, MSA_ERR_ARG, "msa_mesh_info: null argument");
    const MeshHost &h = mesh->h;
    info[0] = h.nb;
    info[1] = (long long)h.bcol.size();
    info[2] = h.nnz_scalar;
    info[3] = h.ncolours;
    info[4] = h.max_row;
    info[5] = h.n_free;
    info[6] = h.n_elem;
    info[7] = h.n_nodes;
    return MSA_OK;
}

int msa_mesh_destroy(msa_mesh *m)
{
    if (!m) return MSA_OK;
    (void)hipSetDevice(m->ctx->device);
    for (void *p : {(void *)m->d_xyz, (void *)m->d_conn, (void *)m->d_node_block,
                    (void *)m->d_node_free, (void *)m->d_emap, (void *)m->d_colour_elems})
        if (p) (void)hipFree(p);
    delete m;
    return MSA_OK;
}

// ---------------------------------------------------------------- elements

int msa_ke_packed(msa_ctx *ctx, int n_elem, const double *xyz8, const double *E, double nu,
                  double *ke300, int *elem_status)
{
    MSA_CHECK(ctx && xyz8 && E && ke300 && n_elem >= 0, MSA_ERR_ARG, "msa_ke_packed: bad args");
    if (n_elem == 0) return MSA_OK;
    MSA_HIP(hipSetDevice(ctx->device));
    double *dx = nullptr, *dE = nullptr, *dk = nullptr;
    int *ds = nullptr;
    hipStream_t s = ctx->stream;
    hipError_t e = hipMalloc(&dx, sizeof(double) * 24 * (size_t)n_elem);
    if (e == hipSuccess) e = hipMalloc(&dE, sizeof(double) * (size_t)n_elem);
    if (e == hipSuccess) e = hipMalloc(&dk, sizeof(double) * 300 * (size_t)n_elem);
    if (e == hipSuccess) e = hipMalloc(&ds, sizeof(int) * (size_t)n_elem);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dx, xyz8, sizeof(double) * 24 * (size_t)n_elem, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dE, E, sizeof(double) * (size_t)n_elem, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_ke_packed(s, n_elem, dx, dE, nu, dk, ds);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ke300, dk, sizeof(double) * 300 * (size_t)n_elem, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && elem_status)
        e = hipMemcpyAsync(elem_status, ds, sizeof(int) * (size_t)n_elem, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (void *p : {(void *)dx, (void *)dE, (void *)dk, (void *)ds})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) {
        set_error("msa_ke_packed: %s", hipGetErrorString(e));
        return MSA_ERR_HIP;
    }
    return MSA_OK;
}

static int default_ctx(msa_ctx **out)
{
    std::call_once(g_default_once, [] {
        msa_ctx *c = nullptr;
        if (msa_ctx_create(0, &c) == MSA_OK) {
            if (hipMalloc(&g_drop_buf, sizeof(double) * 4096) == hipSuccess)
                g_default_ctx = c;
            else
                msa_ctx_destroy(c);
        }
    });
    *out = g_default_ctx;
    if (!g_default_ctx) {
        set_error("no default HIP context for the Fortran drop-ins");
        return MSA_ERR_NO_DEVICE;
    }
    return MSA_OK;
}

static void fill_nan(double *p, size_t n)
{
    if (!p) return;
    for (size_t i = 0; i < n; i++) p[i] = NAN;
}

void CALCH8GAUSSMATRICES(const int *iInt, const double *xyz, double *w, double *S, double *DS,
                         double *J, double *detJ, double *B)
{
    msa_ctx *ctx = nullptr;
    bool ok = iInt && *iInt == 2 && xyz && default_ctx(&ctx) == MSA_OK;
    if (ok) {
        std::lock_guard<std::mutex> lock(g_default_mutex);
        double out[1496];
        hipError_t e = hipSetDevice(ctx->device);
        if (e == hipSuccess)
            e = hipMemcpyAsync(g_drop_buf + 2048, xyz, 24 * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream);
        if (e == hipSuccess) e = launch_gauss_matrices(ctx->stream, g_drop_buf + 2048, g_drop_buf);
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, g_drop_buf, sizeof(out), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) {
            memcpy(w, out, 8 * sizeof(double));
            memcpy(S, out + 8, 64 * sizeof(double));
            memcpy(DS, out + 72, 192 * sizeof(double));
            memcpy(J, out + 264, 72 * sizeof(double));
            memcpy(detJ, out + 336, 8 * sizeof(double));
            memcpy(B, out + 344, 1152 * sizeof(double));
            return;
        }
        set_error("CALCH8GAUSSMATRICES: %s", hipGetErrorString(e));
    } else if (iInt && *iInt != 2) {
        set_error("CALCH8GAUSSMATRICES: iInt=%d unsupported (only 2)", *iInt);
    }
    fill_nan(w, 8);
    fill_nan(S, 64);
    fill_nan(DS, 192);
    fill_nan(J, 72);
    fill_nan(detJ, 8);
    fill_nan(B, 1152);
}

void CALCH8K(const int *iInt, const double *E, const double *B, const double *w, double *K)
{
    msa_ctx *ctx = nullptr;
    bool ok = iInt && *iInt == 2 && E && B && w && default_ctx(&ctx) == MSA_OK;
    if (ok) {
        std::lock_guard<std::mutex> lock(g_default_mutex);
        double *dE = g_drop_buf, *dB = g_drop_buf + 288, *dw = g_drop_buf + 1440,
               *dK = g_drop_buf + 1448;
        hipError_t e = hipSetDevice(ctx->device);
        if (e == hipSuccess) e = hipMemcpyAsync(dE, E, 288 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dB, B, 1152 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dw, w, 8 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = launch_h8k(ctx->stream, dE, dB, dw, dK);
        if (e == hipSuccess) e = hipMemcpyAsync(K, dK, 300 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) return;
        set_error("CALCH8K: %s", hipGetErrorString(e));
    } else if (iInt && *iInt != 2) {
        set_error("CALCH8K: iInt=%d unsupported (only 2)", *iInt);
    }
    fill_nan(K, 300);
}

void CALCH8STRAINS(const int *iInt, const double *B, const double *u, double *strains)
{
    msa_ctx *ctx = nullptr;
    bool ok = iInt && *iInt == 2 && B && u && default_ctx(&ctx) == MSA_OK;
    if (ok) {
        std::lock_guard<std::mutex> lock(g_default_mutex);
        double *dB = g_drop_buf, *du = g_drop_buf + 1152, *dS = g_drop_buf + 1176;
        hipError_t e = hipSetDevice(ctx->device);
        if (e == hipSuccess) e = hipMemcpyAsync(dB, B, 1152 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(du, u, 24 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = launch_h8strains(ctx->stream, dB, du, dS);
        if (e == hipSuccess) e = hipMemcpyAsync(strains, dS, 48 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) return;
        set_error("CALCH8STRAINS: %s", hipGetErrorString(e));
    } else if (iInt && *iInt != 2) {
        set_error("CALCH8STRAINS: iInt=%d unsupported (only 2)", *iInt);
    }
    fill_nan(strains, 48);
}

void CALCH8FORCES(const int *iInt, const double *B, const double *w, const double *stresses,
                  double *forces)
{
    msa_ctx *ctx = nullptr;
    bool ok = iInt && *iInt == 2 && B && w && stresses && default_ctx(&ctx) == MSA_OK;
    if (ok) {
        std::lock_guard<std::mutex> lock(g_default_mutex);
        double *dB = g_drop_buf, *dw = g_drop_buf + 1152, *dS = g_drop_buf + 1160,
               *dF = g_drop_buf + 1208;
        hipError_t e = hipSetDevice(ctx->device);
        if (e == hipSuccess) e = hipMemcpyAsync(dB, B, 1152 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dw, w, 8 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dS, stresses, 48 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = launch_h8forces(ctx->stream, dB, dw, dS, dF);
        if (e == hipSuccess) e = hipMemcpyAsync(forces, dF, 24 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) return;
        set_error("CALCH8FORCES: %s", hipGetErrorString(e));
    } else if (iInt && *iInt != 2) {
        set_error("CALCH8FORCES: iInt=%d unsupported (only 2)", *iInt);
    }
    fill_nan(forces, 24);
}

void CALCH8MLUMPED(const int *iInt, const double *density, const double *w, double *M)
{
    msa_ctx *ctx = nullptr;
    bool ok = iInt && *iInt == 2 && density && w && default_ctx(&ctx) == MSA_OK;
    if (ok) {
        std::lock_guard<std::mutex> lock(g_default_mutex);
        double *dw = g_drop_buf, *dM = g_drop_buf + 8;
        hipError_t e = hipSetDevice(ctx->device);
        if (e == hipSuccess) e = hipMemcpyAsync(dw, w, 8 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = launch_h8mlumped(ctx->stream, *density, dw, dM);
        if (e == hipSuccess) e = hipMemcpyAsync(M, dM, 300 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) return;
        set_error("CALCH8MLUMPED: %s", hipGetErrorString(e));
    } else if (iInt && *iInt != 2) {
        set_error("CALCH8MLUMPED: iInt=%d unsupported (only 2)", *iInt);
    }
    fill_nan(M, 300);
}

// ---------------------------------------------------------------- batch

int msa_batch_destroy(msa_batch *b)
{
    if (!b) return MSA_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    if (b->timing.e0) (void)hipEventDestroy(b->timing.e0);
    if (b->timing.e1) (void)hipEventDestroy(b->timing.e1);
    for (void *p : b->allocs) (void)hipFree(p);
    for (msa_mesh *m : b->mg_meshes) msa_mesh_destroy(m);
    delete b;
    return MSA_OK;
}

int msa_batch_create(msa_ctx *ctx, int n_classes, msa_mesh *const *meshes, const int *n_samples,
                     int precision, int tile, msa_batch **out)
{
    MSA_CHECK(ctx && meshes && n_samples && out, MSA_ERR_ARG, "msa_batch_create: null argument");
    *out = nullptr;
    MSA_CHECK(n_classes >= 1 && n_classes <= 32, MSA_ERR_ARG, "n_classes=%d (1..32)", n_classes);
    MSA_CHECK(precision == MSA_FP64 || precision == MSA_MIXED, MSA_ERR_ARG, "precision=%d",
              precision);
    for (int c = 0; c < n_classes; c++) {
        MSA_CHECK(meshes[c] && meshes[c]->ctx == ctx, MSA_ERR_ARG, "mesh %d invalid", c);
        MSA_CHECK(n_samples[c] >= 1, MSA_ERR_ARG, "n_samples[%d]=%d", c, n_samples[c]);
        MSA_CHECK(meshes[c]->h.n_free >= 1, MSA_ERR_ARG, "mesh %d has no free DOF to solve for", c);
    }
    if (tile == 0) tile = 1;  // one sample per lane tile: no lane waits for a slower sample
    MSA_CHECK(tile == 1 || tile == 2 || tile == 4 || tile == 8 || tile == 16 || tile == 32 ||
                  tile == 64,
              MSA_ERR_ARG, "tile=%d (1,2,4,...,64)", tile);
    MSA_HIP(hipSetDevice(ctx->device));
    msa_batch *b = new msa_batch();
    b->ctx = ctx;
    b->precision = precision;
    b->T = tile;
    b->R = 64 / tile;
    b->cls.resize(n_classes);
    int rc = MSA_OK;
    auto fail = [&](int code) {
        msa_batch_destroy(b);
        return code;
    };
    for (int c = 0; c < n_classes && rc == MSA_OK; c++) {
        rc = setup_class(b, b->cls[c], meshes[c], n_samples[c]);
        if (rc != MSA_OK) break;
        ClassHost &C = b->cls[c];
        C.dev.mg_root = c;
        for (int t = 0; t < C.ntiles; t++) {
            b->tiles.push_back(make_int2(c, t));
            for (int k = 0; k < C.nchunks; k++) b->items.push_back(make_int4(c, t, k, 0));
        }
    }
    if (rc != MSA_OK) return fail(rc);
    void *p = nullptr;
    rc = dalloc(b, &p, sizeof(ClassDev) * n_classes);
    b->d_cls = (ClassDev *)p;
    b->d_cls_cap = n_classes;
    if (rc == MSA_OK) rc = dalloc(b, &p, sizeof(int4) * b->items.size());
    b->d_items = (int4 *)p;
    if (rc == MSA_OK) rc = dalloc(b, &p, 4 * sizeof(int));
    b->d_counts = (int *)p;
    if (rc == MSA_OK) rc = dalloc(b, &p, sizeof(int2) * b->tiles.size());
    b->d_tiles = (int2 *)p;
    if (rc != MSA_OK) return fail(rc);
    {
        std::vector<int4> g;
        for (int c = 0; c < n_classes; c++) class_group_items(b->cls[c], c, g);
        const int4 *gp = nullptr;
        rc = upload_vec(b, g, &gp);
        if (rc != MSA_OK) return fail(rc);
        b->d_gitems = (int4 *)gp;
        b->n_gitems = (int)g.size();
    }
    if (hipMemcpy(b->d_items, b->items.data(), sizeof(int4) * b->items.size(),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_tiles, b->tiles.data(), sizeof(int2) * b->tiles.size(),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b->d_counts, 0, 4 * sizeof(int)) != hipSuccess) {
        set_error("batch upload failed");
        return fail(MSA_ERR_HIP);
    }
    if (hipEventCreate(&b->timing.e0) != hipSuccess || hipEventCreate(&b->timing.e1) != hipSuccess) {
        set_error("event creation failed");
        return fail(MSA_ERR_HIP);
    }
    rc = upload_class_table(b);
    if (rc != MSA_OK) return fail(rc);
    *out = b;
    return MSA_OK;
}

long long msa_batch_bytes(const msa_batch *b) { return b ? b->bytes : 0; }

int msa_batch_set_modulus(msa_batch *b, int cls, const double *E)
{
    MSA_CHECK(b && E && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG,
              "msa_batch_set_modulus: bad arguments");
    ClassHost &C = b->cls[cls];
    const int T = b->T, ne = C.mesh->h.n_elem;
    std::vector<double> h((size_t)C.ntiles * ne * T, 0.0);
    for (int s = 0; s < C.nsamp; s++) {
        const int tile = s / T, lane = s % T;
        for (int e = 0; e < ne; e++) h[((size_t)tile * ne + e) * T + lane] = E[(size_t)s * ne + e];
    }
    MSA_HIP(hipSetDevice(b->ctx->device));
    MSA_HIP(hipMemcpy(C.dev.E, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    C.has_E = true;
    b->assembled = false;
    return MSA_OK;
}

int msa_batch_set_kl(msa_batch *b, int cls, int M, int kmax, const double *w1d, const double *a1d,
                     const int *evpPos, const double *eigs, const double *domain, double meanE,
                     double cov)
{
    MSA_CHECK(b && cls >= 0 && cls < (int)b->cls.size() && M > 0 && kmax > 0 && w1d && a1d &&
                  evpPos && eigs && domain && meanE > 0 && cov >= 0,
              MSA_ERR_ARG, "msa_batch_set_kl: bad arguments");
    for (int i = 0; i < M; i++)
        for (int d = 0; d < 3; d++)
            MSA_CHECK(evpPos[3 * i + d] >= 1 && evpPos[3 * i + d] <= kmax, MSA_ERR_ARG,
                      "evpPos[%d][%d]=%d outside 1..%d", i, d, evpPos[3 * i + d], kmax);
    ClassHost &C = b->cls[cls];
    MSA_HIP(hipSetDevice(b->ctx->device));
    const int ne = C.mesh->h.n_elem;
    if (C.M != M) {
        MSA_CHECK(C.M == 0, MSA_ERR_STATE, "KL order of class %d already set to %d", cls, C.M);
        MSA_TRY(dalloc(b, (void **)&C.dev.phi, (size_t)ne * M * sizeof(double)));
        MSA_TRY(dalloc(b, (void **)&C.dev.xi, (size_t)C.ntiles * b->T * M * sizeof(double)));
        MSA_HIP(hipMemset(C.dev.xi, 0, (size_t)C.ntiles * b->T * M * sizeof(double)));
        C.M = M;
        C.dev.M = M;
    }
    const double sigma = sqrt(log(pow(cov / 1.0, 2) + 1.0));
    C.dev.c0 = log(meanE) - 0.5 * pow(sigma, 2);
    MSA_TRY(upload_class_table(b));
    std::vector<double> tab;
    tab.insert(tab.end(), w1d, w1d + 3 * kmax);
    tab.insert(tab.end(), a1d, a1d + 3 * kmax);
    tab.insert(tab.end(), eigs, eigs + M);
    tab.insert(tab.end(), domain, domain + 3);
    double *dt = nullptr;
    int *dp = nullptr;
    MSA_HIP(hipMalloc(&dt, tab.size() * sizeof(double)));
    hipError_t e = hipMalloc(&dp, 3 * (size_t)M * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(dt, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp, evpPos, 3 * (size_t)M * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_kl_basis(b->ctx->stream, b->d_cls + cls, ne, M, kmax, dt, dt + 3 * kmax, dp,
                            dt + 6 * kmax, dt + 6 * kmax + M, sigma);
    if (e == hipSuccess) e = hipStreamSynchronize(b->ctx->stream);
    (void)hipFree(dt);
    if (dp) (void)hipFree(dp);
    MSA_HIP(e);
    C.has_kl = true;
    return MSA_OK;
}

int msa_batch_set_xi(msa_batch *b, int cls, const double *xi)
{
    MSA_CHECK(b && xi && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG,
              "msa_batch_set_xi: bad arguments");
    ClassHost &C = b->cls[cls];
    MSA_CHECK(C.M > 0, MSA_ERR_STATE, "class %d has no KL table (msa_batch_set_kl)", cls);
    MSA_HIP(hipSetDevice(b->ctx->device));
    MSA_HIP(hipMemcpy(C.dev.xi, xi, (size_t)C.nsamp * C.M * sizeof(double), hipMemcpyHostToDevice));
    C.has_xi = true;
    return MSA_OK;
}

int msa_batch_eval_kl(msa_batch *b)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_HIP(hipSetDevice(b->ctx->device));
    for (size_t c = 0; c < b->cls.size(); c++) {
        ClassHost &C = b->cls[c];
        if (!C.has_kl) continue;
        MSA_CHECK(C.has_xi, MSA_ERR_STATE, "class %zu: KL coefficients not set", c);
        MSA_TRY(timed(b, KK_OTHER, [&] {
            return launch_kl_eval(b->ctx->stream, b->T, b->d_cls + c, C.ntiles, C.mesh->h.n_elem,
                                  C.M);
        }));
        C.has_E = true;
    }
    b->assembled = false;
    return MSA_OK;
}

int msa_batch_set_rhs(msa_batch *b, int cls, const double *f)
{
    MSA_CHECK(b && f && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG,
              "msa_batch_set_rhs: bad arguments");
    ClassHost &C = b->cls[cls];
    const MeshHost &h = C.mesh->h;
    std::vector<double> fi((size_t)h.nb * 3, 0.0);
    double ss = 0;
    for (int g = 0; g < h.n_free; g++) {
        fi[h.ref2int[g]] = f[g];
        ss += f[g] * f[g];
    }
    C.fnorm = sqrt(ss);
    C.dev.fnorm = C.fnorm;
    MSA_HIP(hipSetDevice(b->ctx->device));
    MSA_HIP(hipMemcpy((void *)C.dev.f, fi.data(), fi.size() * sizeof(double), hipMemcpyHostToDevice));
    C.has_rhs = true;
    return upload_class_table(b);
}

int msa_batch_assemble(msa_batch *b, double nu)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_CHECK(nu > -1.0 && nu < 0.5, MSA_ERR_ARG, "Poisson ratio %g outside (-1, 0.5)", nu);
    for (size_t c = 0; c < b->cls.size(); c++)
        MSA_CHECK(b->cls[c].has_E, MSA_ERR_STATE, "class %zu: no modulus field", c);
    MSA_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    const bool mx = is_mixed(b);
    // FP64 element cache (rebuilt when nu or the dynamics mode changes): the
    // gather assembly and the FP64 residual/right-hand-side kernels read it
    {
        bool upload = false;
        for (size_t c = 0; c < b->cls.size(); c++) {
            ClassHost &C = b->cls[c];
            if (C.kc_nu == nu && C.kc_mass == (bool)C.dev.dyn) continue;
            if (C.dev.dyn && !C.dev.mc) {
                void *p = nullptr;
                MSA_TRY(dalloc(b, &p, (size_t)C.mesh->h.n_elem * 64 * sizeof(double)));
                C.dev.mc = (double *)p;
                upload = true;
            }
            if (upload) MSA_TRY(upload_class_table(b));
            MSA_HIP(hipMemsetAsync(C.dev.negdet, 0, sizeof(int), s));
            MSA_TRY(timed(b, KK_ASSEMBLE, [&] {
                return launch_khat_cache(s, b->d_cls + c, C.mesh->h.n_elem, nu, C.dev.dyn);
            }));
            C.kc_nu = nu;
            C.kc_mass = C.dev.dyn;
        }
    }
    if (!mx || b->integration == MSA_FP64) {
        // every stored value summed in FP64 and written once
        MSA_TRY(assemble_items(b, b->d_items, (int)b->items.size(), b->d_gitems, b->n_gitems));
    } else {
        // FP32 element integration: coloured FP32 scatter-add
        for (auto &C : b->cls) {
            const size_t n = (size_t)C.ntiles * C.tile_vals;
            MSA_TRY(timed(b, KK_OTHER, [&] { return launch_fill_f32(s, (float *)C.dev.val, n, 0.f); }));
        }
        const int ncol = max_colours(b);
        for (int k = 0; k < ncol; k++) {
            int nblk = 0;
            ColourLaunch cl = colour_launch(b, k, &nblk);
            MSA_TRY(timed(b, KK_ASSEMBLE, [&] {
                return launch_assemble(s, b->T, b->d_cls, cl, nblk, nu, max_tiles(b));
            }));
        }
    }
    MSA_TRY(timed(b, KK_OTHER, [&] {
        return launch_dinv(s, mx, b->T, b->d_cls, b->d_items, (int)b->items.size());
    }));
    b->nu = nu;
    b->assembled = true;
    b->loaded = false;
    b->factored = false;
    b->mg_ready = false;  // preconditioner values follow at the solve or msa_batch_prepare_preconditioner
    return MSA_OK;
}

int msa_batch_prepare_preconditioner(msa_batch *b)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_CHECK(b->assembled, MSA_ERR_STATE, "batch not assembled");
    MSA_HIP(hipSetDevice(b->ctx->device));
    if (!b->mg_ready) MSA_TRY(precond_values(b));
    MSA_HIP(hipStreamSynchronize(b->ctx->stream));
    return MSA_OK;
}

int msa_batch_solve(msa_batch *b, double tol, int max_it, double inner_tol, int max_outer,
                    int graph_iters)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    if (b->solver == MSA_SOLVER_DIRECT) return msa_batch_direct_solve(b);
    MSA_CHECK(b->assembled, MSA_ERR_STATE, "batch not assembled");
    for (size_t c = 0; c < b->cls.size(); c++)
        MSA_CHECK(b->cls[c].has_rhs, MSA_ERR_STATE, "class %zu: no load vector", c);
    MSA_CHECK(tol > 0 && max_it > 0, MSA_ERR_ARG, "tol=%g max_it=%d", tol, max_it);
    MSA_CHECK(graph_iters >= 0 && graph_iters % 2 == 0, MSA_ERR_ARG,
              "graph_iters=%d must be even", graph_iters);
    MSA_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    const bool mx = is_mixed(b);
    const int T = b->T, ni = (int)b->items.size();
    if (!b->mg_ready) MSA_TRY(precond_values(b));
    MSA_TRY(sync_mf_pcg(b));
    if (b->dyn) {
        // static solve with the effective matrix: per-sample right-hand side = f
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_fs_static(s, T, b->d_cls, b->d_items, ni); }));
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_fs_norm(s, T, b->d_cls, b->d_items, ni); }));
        MSA_TRY(reduce(b, OP_FNORM));
    }
    if (!mx) {
        MSA_HIP(hipMemsetAsync(b->d_counts, 0, sizeof(int), s));
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_pcg_init(s, false, T, b->d_cls, b->d_items, ni, 0, b->d_counts); }));
        MSA_TRY(mg_vcycle(b, 1, 0));
        MSA_TRY(reduce(b, OP_INIT0));
        return pcg_loop(b, 0, tol, max_it, graph_iters);
    }
    if (b->refine == MSA_REFINE_RELIABLE) {
        MSA_HIP(hipMemsetAsync(b->d_counts, 0, 2 * sizeof(int), s));
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_ir_begin(s, T, b->d_cls, b->d_items, ni, b->d_counts + 1, 0); }));
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_pcg_init(s, true, T, b->d_cls, b->d_items, ni, 1, b->d_counts); }));
        MSA_TRY(mg_vcycle(b, 1, 3));
        MSA_TRY(reduce(b, OP_INIT1));
        MSA_TRY(pcg_loop(b, 1, tol, max_it, graph_iters));
        MSA_HIP(hipStreamSynchronize(s));
        return MSA_OK;
    }
    MSA_CHECK(inner_tol > 0 && inner_tol < 1 && max_outer >= 1, MSA_ERR_ARG,
              "inner_tol=%g max_outer=%d", inner_tol, max_outer);
    MSA_HIP(hipMemsetAsync(b->d_counts + 1, 0, sizeof(int), s));
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_ir_begin(s, T, b->d_cls, b->d_items, ni, b->d_counts + 1, 0); }));
    const int ncol = max_colours(b);
    for (int outer = 0; outer < max_outer; outer++) {
        int left = 0;
        MSA_TRY(read_count(b, 1, &left));
        if (left <= 0) break;
        MSA_HIP(hipMemsetAsync(b->d_counts, 0, sizeof(int), s));
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_pcg_init(s, true, T, b->d_cls, b->d_items, ni, 1, b->d_counts); }));
        MSA_TRY(mg_vcycle(b, 1, 3));
        MSA_TRY(reduce(b, OP_INIT1));
        MSA_TRY(pcg_loop(b, 0, inner_tol, max_it, graph_iters));
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_ir_update(s, T, b->d_cls, b->d_items, ni); }));
        for (int k = 0; k < ncol; k++) {
            int nblk = 0;
            ColourLaunch cl = colour_launch(b, k, &nblk);
            MSA_TRY(timed(b, KK_RES64, [&] { return launch_residual64(s, T, b->d_cls, cl, nblk, b->nu, 0, max_tiles(b)); }));
        }
        MSA_TRY(timed(b, KK_OTHER, [&] {
            return launch_ir_norm(s, T, b->d_cls, b->d_items, ni, tol, max_outer, b->d_counts + 1);
        }));
        MSA_TRY(reduce(b, OP_IRNORM, tol, 0, max_outer));
    }
    MSA_HIP(hipStreamSynchronize(s));
    return MSA_OK;
}

int msa_batch_run(msa_batch *b, double nu, double tol, int max_it, double inner_tol, int max_outer,
                  int graph_iters)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    bool any_kl = false;
    for (auto &C : b->cls) any_kl |= C.has_kl;
    if (any_kl) MSA_TRY(msa_batch_eval_kl(b));
    MSA_TRY(msa_batch_assemble(b, nu));
    return msa_batch_solve(b, tol, max_it, inner_tol, max_outer, graph_iters);
}

static int extract_lane(msa_batch *b, bool real_is_float, const void *tile_base, size_t n, int lane,
                        std::vector<double> &host)
{
    double *d = nullptr;
    MSA_HIP(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(double)));
    host.resize(n);
    hipError_t e = launch_extract_sample(b->ctx->stream, real_is_float, b->T, tile_base, n, lane, d);
    if (e == hipSuccess)
        e = hipMemcpyAsync(host.data(), d, n * sizeof(double), hipMemcpyDeviceToHost, b->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->ctx->stream);
    (void)hipFree(d);
    MSA_HIP(e);
    return MSA_OK;
}

int msa_batch_get_solution(msa_batch *b, int cls, int sample, double *u)
{
    MSA_CHECK(b && u && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    ClassHost &C = b->cls[cls];
    MSA_CHECK(sample >= 0 && sample < C.nsamp, MSA_ERR_ARG, "sample %d out of range", sample);
    MSA_HIP(hipSetDevice(b->ctx->device));
    const int tile = sample / b->T, lane = sample % b->T;
    std::vector<double> h;
    const size_t n = (size_t)C.mesh->h.nb * 3;
    if (is_mixed(b))
        MSA_TRY(extract_lane(b, false, C.dev.u64 + (size_t)tile * C.tile_vec, n, lane, h));
    else
        MSA_TRY(extract_lane(b, false, (double *)C.dev.x + (size_t)tile * C.tile_vec, n, lane, h));
    const MeshHost &m = C.mesh->h;
    for (int g = 0; g < m.n_free; g++) u[g] = h[m.ref2int[g]];
    return MSA_OK;
}

// ---------------------------------------------------------------- dynamics

int msa_batch_set_dynamics(msa_batch *b, double rho, double rayleigh_alpha, double rayleigh_beta,
                           double newmark_alpha, double newmark_delta, double dt)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    // NewmarkDynamicAnalyzer.InitializeCoefficients (NewmarkDynamicAnalyzer.cs:55-57)
    MSA_CHECK(newmark_delta >= 0.5, MSA_ERR_ARG,
              "Newmark delta has to be bigger than 0.5 (delta = %g)", newmark_delta);
    const double alim = 0.25 * (0.5 + newmark_delta) * (0.5 + newmark_delta);
    MSA_CHECK(newmark_alpha >= alim, MSA_ERR_ARG, "Newmark alpha has to be bigger than %g (alpha = %g)",
              alim, newmark_alpha);
    MSA_CHECK(dt > 0 && std::isfinite(dt) && rho >= 0 && std::isfinite(rho) &&
                  std::isfinite(rayleigh_alpha) && std::isfinite(rayleigh_beta),
              MSA_ERR_ARG, "dt=%g rho=%g rayleigh=(%g, %g)", dt, rho, rayleigh_alpha, rayleigh_beta);
    MSA_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    const double al = newmark_alpha, de = newmark_delta;
    DynCoef &c = b->dc;
    c.a0 = 1 / (al * dt * dt);
    c.a1 = de / (al * dt);
    c.a2 = 1 / (al * dt);
    c.a3 = 1 / (2 * al) - 1;
    c.a4 = de / al - 1;
    c.a5 = dt * 0.5 * (de / al - 2);
    c.a6 = dt * (1 - de);
    c.a7 = de * dt;
    c.alphaR = rayleigh_alpha;
    c.betaR = rayleigh_beta;
    c.dt = dt;
    b->rho = rho;
    if (!b->d_step) {
        void *p = nullptr;
        MSA_TRY(dalloc(b, &p, sizeof(int)));
        b->d_step = (int *)p;
    }
    for (auto &C : b->cls) {
        ClassDev &D = C.dev;
        const size_t nvec = (size_t)C.ntiles * C.tile_vec;
        if (!D.dyn) {
            double **vecs[6] = {&D.fs, &D.uold, &D.vel, &D.acc, &D.wm, &D.wk};
            for (double **v : vecs) {
                void *p = nullptr;
                MSA_TRY(dalloc(b, &p, nvec * sizeof(double)));
                *v = (double *)p;
                MSA_HIP(hipMemsetAsync(p, 0, nvec * sizeof(double), s));
            }
            void *p = nullptr;
            MSA_TRY(dalloc(b, &p, (size_t)D.nb * 3 * sizeof(double)));
            D.bacc = (double *)p;
            MSA_HIP(hipMemsetAsync(p, 0, (size_t)D.nb * 3 * sizeof(double), s));
            D.probe = -1;
            D.hist = nullptr;
            D.hist_len = 0;
        }
        D.dyn = 1;
        D.cK = 1.0 + c.a1 * rayleigh_beta;   // K + a1 betaR K
        D.cM = (c.a0 + c.a1 * rayleigh_alpha) * rho;  // a0 M + a1 alphaR M
        D.rho = rho;
    }
    b->dyn = true;
    b->assembled = false;
    return upload_class_table(b);
}

int msa_batch_set_probe(msa_batch *b, int cls, int dof)
{
    MSA_CHECK(b && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    MSA_CHECK(b->dyn, MSA_ERR_STATE, "msa_batch_set_dynamics first");
    ClassHost &C = b->cls[cls];
    const MeshHost &m = C.mesh->h;
    MSA_CHECK(dof >= -1 && dof < m.n_free, MSA_ERR_ARG, "dof %d out of range", dof);
    C.dev.probe = dof < 0 ? -1 : m.ref2int[dof];
    return upload_class_table(b);
}

}  // extern "C"

namespace {

// One Newmark step up to its PCG iterations (NewmarkDynamicAnalyzer.Solve,
// NewmarkDynamicAnalyzer.cs:280-299): the effective right-hand side
// (:189-206, the mass-acceleration load of the step), its norm, and the
// start of the step's solve from the previous displacement (or zero), or
// the direct solve.  Kernels only (no memset), so it can be captured.
int newmark_head(msa_batch *b, const DynCoef &dc, double tol, int max_it)
{
    hipStream_t s = b->ctx->stream;
    const bool mx = is_mixed(b);
    const int T = b->T, ni = (int)b->items.size(), ncol = max_colours(b);
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_dyn_prep(s, mx, T, b->d_cls, b->d_items, ni, dc); }));
    for (int k = 0; k < ncol; k++) {
        int nblk = 0;
        ColourLaunch cl = colour_launch(b, k, &nblk);
        MSA_TRY(timed(b, KK_RES64, [&] { return launch_dyn_rhs(s, T, b->d_cls, cl, nblk, b->nu, dc.betaR, max_tiles(b)); }));
    }
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_fs_norm(s, T, b->d_cls, b->d_items, ni); }));
    MSA_TRY(reduce(b, OP_FNORM));
    if (b->solver == MSA_SOLVER_DIRECT) {
        for (size_t c = 0; c < b->cls.size(); c++) MSA_TRY(band_solve_class(b, c, true));
        return direct_stats(b);
    }
    if (!mx) {
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_set_int(s, b->d_counts, 1, 0); }));
        if (dc.start) {
            MSA_TRY(timed(b, KK_OTHER, [&] { return launch_pcg_init_warm(s, T, b->d_cls, b->d_items, ni); }));
            MSA_TRY(mg_vcycle(b, 1, 0));
            return reduce(b, OP_INIT2, tol);
        }
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_pcg_init(s, false, T, b->d_cls, b->d_items, ni, 0, b->d_counts); }));
        MSA_TRY(mg_vcycle(b, 1, 0));
        return reduce(b, OP_INIT0);
    }
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_set_int(s, b->d_counts, 2, 0); }));
    const int warm = dc.start ? 1 : 0;
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_ir_begin(s, T, b->d_cls, b->d_items, ni, b->d_counts + 1, warm); }));
    for (int k = 0; k < ncol; k++) {
        int nblk = 0;
        ColourLaunch cl = colour_launch(b, k, &nblk);
        MSA_TRY(timed(b, KK_RES64, [&] { return launch_residual64(s, T, b->d_cls, cl, nblk, b->nu, 0, max_tiles(b)); }));
    }
    MSA_TRY(timed(b, KK_OTHER, [&] {
        return launch_ir_norm(s, T, b->d_cls, b->d_items, ni, tol, 1 << 30, b->d_counts + 1);
    }));
    MSA_TRY(reduce(b, OP_IRNORM, tol, 0, 1 << 30));
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_pcg_init(s, true, T, b->d_cls, b->d_items, ni, 1, b->d_counts); }));
    MSA_TRY(mg_vcycle(b, 1, 3));
    return reduce(b, OP_INIT1);
}

// The end of a Newmark step: velocity and acceleration update
// (NewmarkDynamicAnalyzer.cs:301-315), probe record, step counter.
int newmark_tail(msa_batch *b, const DynCoef &dc)
{
    hipStream_t s = b->ctx->stream;
    MSA_TRY(timed(b, KK_OTHER, [&] {
        return launch_dyn_update(s, is_mixed(b), b->T, b->d_cls, b->d_items, (int)b->items.size(), dc);
    }));
    return timed(b, KK_OTHER, [&] { return launch_step_advance(s, b->d_step); });
}

}  // namespace

extern "C" {

int msa_batch_newmark(msa_batch *b, int nsteps, const double *acc, int acc_dir, double tol,
                      int max_it, int graph_iters, int warm_start)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_CHECK(b->dyn, MSA_ERR_STATE, "msa_batch_set_dynamics first");
    MSA_CHECK(b->assembled, MSA_ERR_STATE, "batch not assembled");
    MSA_CHECK(nsteps >= 1, MSA_ERR_ARG, "nsteps=%d", nsteps);
    MSA_CHECK(acc_dir >= -1 && acc_dir <= 2, MSA_ERR_ARG, "acc_dir=%d", acc_dir);
    MSA_CHECK(acc_dir < 0 || acc, MSA_ERR_ARG, "acceleration history is null");
    MSA_CHECK(tol > 0 && max_it > 0, MSA_ERR_ARG, "tol=%g max_it=%d", tol, max_it);
    MSA_CHECK(graph_iters >= 0 && graph_iters % 2 == 0, MSA_ERR_ARG,
              "graph_iters=%d must be even", graph_iters);
    MSA_CHECK(warm_start >= 0 && warm_start <= 2, MSA_ERR_ARG, "warm_start=%d", warm_start);
    const bool mx = is_mixed(b);
    MSA_CHECK(!mx || b->refine == MSA_REFINE_RELIABLE, MSA_ERR_STATE,
              "dynamics of a MIXED batch uses MSA_REFINE_RELIABLE");
    if (b->solver == MSA_SOLVER_DIRECT && !b->factored) MSA_TRY(msa_batch_factorize(b));
    MSA_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    const int T = b->T, ni = (int)b->items.size();
    if (b->solver != MSA_SOLVER_DIRECT && !b->mg_ready) MSA_TRY(precond_values(b));
    MSA_TRY(sync_mf_pcg(b));
    if (acc_dir >= 0) {
        if (b->acc_cap < nsteps) {
            void *p = nullptr;
            MSA_TRY(dalloc(b, &p, (size_t)nsteps * sizeof(double)));
            b->d_acc = (double *)p;
            b->acc_cap = nsteps;
        }
        MSA_HIP(hipMemcpyAsync(b->d_acc, acc, (size_t)nsteps * sizeof(double),
                               hipMemcpyHostToDevice, s));
    }
    for (auto &C : b->cls) {
        if (C.dev.probe < 0) {
            C.dev.hist_len = 0;
            continue;
        }
        const long long need = (long long)C.nsamp * nsteps;
        if (C.hist_cap < need) {
            void *p = nullptr;
            MSA_TRY(dalloc(b, &p, (size_t)need * sizeof(double)));
            C.dev.hist = (double *)p;
            C.hist_cap = need;
        }
        C.dev.hist_len = nsteps;
        MSA_HIP(hipMemsetAsync(C.dev.hist, 0, (size_t)need * sizeof(double), s));
    }
    MSA_TRY(upload_class_table(b));
    const int ncol = max_colours(b);
    // mass-acceleration load of a unit base acceleration
    for (auto &C : b->cls)
        MSA_HIP(hipMemsetAsync(C.dev.bacc, 0, (size_t)C.dev.nb * 3 * sizeof(double), s));
    if (acc_dir >= 0)
        for (int k = 0; k < ncol; k++) {
            int nblk = 0;
            ColourLaunch cl = colour_launch(b, k, &nblk);
            MSA_TRY(timed(b, KK_OTHER, [&] { return launch_dyn_bacc(s, b->d_cls, cl, nblk, b->nu, acc_dir); }));
        }
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_dyn_reset(s, mx, T, b->d_cls, b->d_items, ni); }));
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_set_int(s, b->d_step, 1, 0); }));
    DynCoef dc = b->dc;
    dc.acc = acc_dir >= 0 ? b->d_acc : nullptr;
    dc.step = b->d_step;
    dc.start = warm_start;
    const bool direct = b->solver == MSA_SOLVER_DIRECT;
    const int kind = mx ? 1 : 0;
    if (graph_iters <= 0 || b->timing.on) {
        for (int step = 0; step < nsteps; step++) {
            MSA_TRY(newmark_head(b, dc, tol, max_it));
            if (!direct) MSA_TRY(pcg_loop(b, kind, tol, max_it, graph_iters));
            MSA_TRY(newmark_tail(b, dc));
        }
        MSA_HIP(hipStreamSynchronize(s));
        return MSA_OK;
    }
    // Captured steps: G0 = head + first iteration block (step 0), G1 = the
    // previous step's update + head + first block (steps 1..n-1), G2 = the
    // last update.  One graph launch and one poll per step; further blocks
    // (the PCG block graph) only while samples are still iterating.
    int block = graph_iters;
    if (kind == 1) block = ((block + b->period - 1) / b->period) * b->period;
    if (!direct && (b->graph == nullptr || b->graph_iters != block || b->graph_kind != kind ||
                    b->graph_tol != tol || b->graph_maxit != max_it))
        MSA_TRY(build_graph(b, kind, block, tol, max_it));
    hipGraphExec_t G[3] = {nullptr, nullptr, nullptr};
    auto destroy = [&] {
        for (auto &g : G)
            if (g) (void)hipGraphExecDestroy(g);
    };
    int rc = MSA_OK;
    for (int k = 0; k < 3 && rc == MSA_OK; k++) {
        rc = capture(b, &G[k], [&]() -> int {
            if (k >= 1) MSA_TRY(newmark_tail(b, dc));
            if (k == 2) return MSA_OK;
            MSA_TRY(newmark_head(b, dc, tol, max_it));
            return direct ? MSA_OK : launch_block(b, kind, block, tol, max_it);
        });
        if (rc == MSA_OK) b->graph_captures++;
    }
    for (int step = 0; step < nsteps && rc == MSA_OK; step++) {
        hipError_t e = hipGraphLaunch(G[step == 0 ? 0 : 1], s);
        if (e != hipSuccess) {
            set_error("Newmark step graph launch: %s", hipGetErrorString(e));
            rc = MSA_ERR_HIP;
            break;
        }
        b->graph_replays++;
        if (!direct) rc = pcg_loop(b, kind, tol, max_it, graph_iters);
    }
    if (rc == MSA_OK && hipGraphLaunch(G[2], s) != hipSuccess) {
        set_error("Newmark update graph launch failed");
        rc = MSA_ERR_HIP;
    }
    if (rc == MSA_OK && hipStreamSynchronize(s) != hipSuccess) rc = MSA_ERR_HIP;
    destroy();
    return rc;
}

// ---------------------------------------------------------------- direct solver

int msa_batch_factorize(msa_batch *b)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_CHECK(!is_mixed(b), MSA_ERR_STATE, "the direct solver needs an FP64 batch");
    MSA_CHECK(b->assembled, MSA_ERR_STATE, "batch not assembled");
    MSA_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    MSA_TRY(prepare_band(b));
    {
        bool upload = false;
        for (auto &C : b->cls) MSA_TRY(zem_setup(b, C, &upload));
        if (upload) MSA_TRY(upload_class_table(b));
    }
    if (b->precond == MSA_PRECOND_FIRST_SAMPLE) b->mg_ready = false;  // its shared factor is overwritten
    for (auto &C : b->cls) {
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_fill_f64(s, C.dev.band, C.band_cap, 0.0); }));
        MSA_HIP(hipMemsetAsync(C.dev.bfail, 0, (size_t)C.nsamp * sizeof(int), s));
    }
    MSA_TRY(timed(b, KK_OTHER, [&] {
        return launch_band_scatter(s, b->T, b->d_cls, b->d_items, (int)b->items.size());
    }));
    for (size_t c = 0; c < b->cls.size(); c++) {
        ClassHost &C = b->cls[c];
        const ClassDev *cp = b->d_cls + c;
        const int n = C.dev.bn, W = C.dev.bW, ns = C.nsamp;
        for (int k0 = 0; k0 < n; k0 += kBandNB) {
            const int nk = std::min(kBandNB, n - k0), r0 = k0 + nk,
                      m = std::min(n - r0, W - kBandNB);
            MSA_TRY(timed(b, KK_ASSEMBLE, [&] { return launch_chol_diag(s, cp, ns, k0, nk); }));
            MSA_TRY(timed(b, KK_ASSEMBLE, [&] { return launch_chol_panel(s, cp, ns, k0, nk, r0, m); }));
            MSA_TRY(timed(b, KK_ASSEMBLE, [&] { return launch_chol_update(s, cp, ns, k0, nk, r0, m); }));
        }
        if (b->zem_tol > 0.0) MSA_TRY(zem_modes(b, c));
    }
    MSA_HIP(hipStreamSynchronize(s));
    b->factored = true;
    return MSA_OK;
}

int msa_batch_set_persistent(msa_batch *b, int enable)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_CHECK(!enable || b->T == 1, MSA_ERR_STATE,
              "the persistent launch takes one-sample tiles (the batch has lane tile %d)", b->T);
    b->persistent = enable != 0;
    return MSA_OK;
}

int msa_batch_persistent_info(msa_batch *b, long long *info)
{
    MSA_CHECK(b && info, MSA_ERR_ARG, "bad arguments");
    info[0] = b->persist_launches;
    info[1] = b->persist_grid;
    return MSA_OK;
}

int msa_batch_set_zem(msa_batch *b, double tol, int max_modes)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_CHECK(tol >= 0.0 && std::isfinite(tol), MSA_ERR_ARG, "zero-energy tolerance %g", tol);
    MSA_CHECK(tol == 0.0 || (max_modes >= 1 && max_modes <= 64), MSA_ERR_ARG,
              "max_modes=%d outside [1, 64]", max_modes);
    b->zem_tol = tol;
    b->zem_max = tol > 0.0 ? max_modes : 0;
    b->factored = false;
    return MSA_OK;
}

int msa_batch_get_zem(msa_batch *b, int cls, int sample, int max_modes, int *n_modes, int *cols,
                      double *modes)
{
    MSA_CHECK(b && n_modes && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    ClassHost &C = b->cls[cls];
    MSA_CHECK(sample >= 0 && sample < C.nsamp, MSA_ERR_ARG, "sample %d out of range", sample);
    MSA_CHECK(b->factored && b->zem_tol > 0.0 && C.zcount && C.dev.zv == C.zv, MSA_ERR_STATE,
              "no factorization with zero-energy modes (msa_batch_set_zem, msa_batch_factorize)");
    MSA_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    int cnt = 0;
    MSA_HIP(hipMemcpyAsync(&cnt, C.zcount + sample, sizeof(int), hipMemcpyDeviceToHost, s));
    MSA_HIP(hipStreamSynchronize(s));
    *n_modes = cnt;
    MSA_CHECK(cnt <= b->zem_max, MSA_ERR_STATE,
              "sample %d: %d zero-energy modes, more than the %d kept", sample, cnt, b->zem_max);
    const int n = C.dev.bn, k = std::min(cnt, max_modes);
    if (k <= 0) return MSA_OK;
    std::vector<double> h((size_t)k * n);
    std::vector<int> hc(k);
    MSA_HIP(hipMemcpyAsync(h.data(), C.zv + (size_t)sample * b->zem_max * n, h.size() * sizeof(double),
                           hipMemcpyDeviceToHost, s));
    MSA_HIP(hipMemcpyAsync(hc.data(), C.zcol + (size_t)sample * b->zem_max, k * sizeof(int),
                           hipMemcpyDeviceToHost, s));
    MSA_HIP(hipStreamSynchronize(s));
    // band index -> reference free DOF: internal row bperm[i], ref2int^-1
    const MeshHost &m = C.mesh->h;
    std::vector<int> int2ref(3 * (size_t)m.nb, -1);
    for (int g = 0; g < m.n_free; g++) int2ref[m.ref2int[g]] = g;
    for (int j = 0; j < k; j++) {
        if (cols) cols[j] = int2ref[C.bperm[hc[j]]];
        if (modes)
            for (int i = 0; i < n; i++) modes[(size_t)j * m.n_free + int2ref[C.bperm[i]]] = h[(size_t)j * n + i];
    }
    return MSA_OK;
}

int msa_batch_direct_solve(msa_batch *b)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    for (size_t c = 0; c < b->cls.size(); c++)
        MSA_CHECK(b->cls[c].has_rhs || b->dyn, MSA_ERR_STATE, "class %zu: no load vector", c);
    if (!b->factored) MSA_TRY(msa_batch_factorize(b));
    MSA_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    const int T = b->T, ni = (int)b->items.size();
    if (b->dyn) {
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_fs_static(s, T, b->d_cls, b->d_items, ni); }));
        MSA_TRY(timed(b, KK_OTHER, [&] { return launch_fs_norm(s, T, b->d_cls, b->d_items, ni); }));
        MSA_TRY(reduce(b, OP_FNORM));
    }
    for (size_t c = 0; c < b->cls.size(); c++) MSA_TRY(band_solve_class(b, c, b->dyn));
    MSA_TRY(direct_stats(b));
    MSA_HIP(hipStreamSynchronize(s));
    return MSA_OK;
}

int msa_batch_set_preconditioner(msa_batch *b, int kind, int max_levels, double ratio)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_CHECK(kind == MSA_PRECOND_JACOBI || kind == MSA_PRECOND_MULTIGRID ||
                  kind == MSA_PRECOND_FIRST_SAMPLE,
              MSA_ERR_ARG, "preconditioner kind=%d", kind);
    MSA_CHECK(ratio == 0.0 || ratio > 1.0, MSA_ERR_ARG, "smoothing ratio %g (0 or > 1)", ratio);
    MSA_HIP(hipSetDevice(b->ctx->device));
    if (kind == MSA_PRECOND_MULTIGRID) {
        if (b->mg_nlev.empty()) {
            MSA_TRY(mg_build(b, max_levels));
            b->mg_max_levels = max_levels;
        } else {
            MSA_CHECK(max_levels == b->mg_max_levels, MSA_ERR_STATE,
                      "the batch's hierarchy was built with max_levels=%d", b->mg_max_levels);
        }
    }
    b->precond = kind;
    b->mg_ratio = ratio > 0 ? ratio : 12.0;
    if (!b->refine_set) {
        // a V-cycle iteration cuts the residual about 5x: check for a
        // replacement every 2 iterations, replace after a 1e-4 drop
        // (the same for the first-sample factor: a strong preconditioner)
        b->period = kind != MSA_PRECOND_JACOBI ? 2 : 8;
        b->drop = kind != MSA_PRECOND_JACOBI ? 1e-4 : 0.1;
    }
    drop_graph(b);  // the captured iteration blocks hold the preconditioner
    b->mg_ready = false;
    if (b->assembled) MSA_TRY(precond_values(b));
    return MSA_OK;
}

int msa_batch_mg_info(msa_batch *b, int cls, long long *info)
{
    MSA_CHECK(b && info && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    info[0] = info[1] = info[2] = info[3] = info[4] = info[5] = 0;
    if (b->mg_nlev.empty()) return MSA_OK;
    info[0] = b->mg_nlev[cls];
    info[4] = b->cls[cls].dev.mf;
    for (int ci = b->cls[cls].dev.mg_coarse; ci >= 0; ci = tab(b, ci).dev.mg_coarse)
        info[5] += tab(b, ci).dev.mf;
    for (int ci = b->cls[cls].dev.mg_coarse; ci >= 0; ci = tab(b, ci).dev.mg_coarse) {
        const ClassHost &L = tab(b, ci);
        info[2] += L.mesh->h.n_free;
        if (L.dev.mg_coarse < 0) {
            info[1] = L.mesh->h.n_free;
            info[3] = L.bw;
        }
    }
    return MSA_OK;
}

int msa_batch_set_smoothing_interval(msa_batch *b, int lanczos_steps, double margin)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_CHECK(lanczos_steps == 0 || (lanczos_steps >= 2 && lanczos_steps <= kLzMax), MSA_ERR_ARG,
              "lanczos_steps=%d (0, or 2..%d)", lanczos_steps, kLzMax);
    MSA_CHECK(margin >= 1.0 && margin <= 2.0, MSA_ERR_ARG, "margin %g outside [1, 2]", margin);
    b->mg_lz_steps = lanczos_steps;
    b->mg_lz_margin = margin;
    b->mg_ready = false;
    return MSA_OK;
}

int msa_batch_get_smoothing_interval(msa_batch *b, int *lanczos_steps, double *margin)
{
    MSA_CHECK(b && lanczos_steps && margin, MSA_ERR_ARG, "bad arguments");
    *lanczos_steps = b->mg_lz_steps;
    *margin = b->mg_lz_margin;
    return MSA_OK;
}

int msa_batch_mg_lambda(msa_batch *b, int cls, int sample, int max_levels, double *lam)
{
    MSA_CHECK(b && lam && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    MSA_CHECK(sample >= 0 && sample < b->cls[cls].nsamp, MSA_ERR_ARG, "sample %d out of range", sample);
    MSA_CHECK(b->precond == MSA_PRECOND_MULTIGRID && b->cls[cls].dev.mg_coarse >= 0, MSA_ERR_STATE,
              "class %d has no multigrid hierarchy", cls);
    MSA_CHECK(b->assembled, MSA_ERR_STATE, "batch not assembled");
    MSA_HIP(hipSetDevice(b->ctx->device));
    if (!b->mg_ready) MSA_TRY(mg_values(b));
    int l = 0;
    for (int ci = cls; ci >= 0 && tab(b, ci).dev.mg_coarse >= 0 && l < max_levels; ci = tab(b, ci).dev.mg_coarse, l++)
        MSA_HIP(hipMemcpy(lam + l, tab(b, ci).dev.mg_lam + sample, sizeof(double), hipMemcpyDeviceToHost));
    for (int k = l; k < max_levels; k++) lam[k] = 0.0;
    return MSA_OK;
}

int msa_batch_mg_apply(msa_batch *b, int cls, int sample, const double *r, double *z)
{
    MSA_CHECK(b && r && z && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    ClassHost &C = b->cls[cls];
    MSA_CHECK(sample >= 0 && sample < C.nsamp, MSA_ERR_ARG, "sample %d out of range", sample);
    MSA_CHECK(b->assembled, MSA_ERR_STATE, "batch not assembled");
    MSA_CHECK(b->precond == MSA_PRECOND_MULTIGRID && C.dev.mg_coarse >= 0, MSA_ERR_STATE,
              "class %d has no multigrid hierarchy", cls);
    MSA_HIP(hipSetDevice(b->ctx->device));
    if (!b->mg_ready) MSA_TRY(mg_values(b));
    const MeshHost &m = C.mesh->h;
    const int tile = sample / b->T, lane = sample % b->T;
    const size_t n = (size_t)m.nb * 3;
    const bool mx = is_mixed(b);
    std::vector<double> xi(n, 0.0);
    for (int g = 0; g < m.n_free; g++) xi[m.ref2int[g]] = r[g];
    hipStream_t s = b->ctx->stream;
    double *din = nullptr;
    MSA_HIP(hipMalloc(&din, n * sizeof(double)));
    hipError_t e = hipMemcpyAsync(din, xi.data(), n * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_mg_load(s, mx, b->T, b->d_cls + cls, din, n, tile, lane);
    int rc = e == hipSuccess ? mg_vcycle(b, 1, 0) : MSA_OK;
    const size_t rsz = real_size(b);
    const char *zt = (const char *)C.dev.z + (size_t)tile * C.tile_vec * rsz;
    if (e == hipSuccess && rc == MSA_OK) e = launch_extract_sample(s, mx, b->T, zt, n, lane, din);
    if (e == hipSuccess && rc == MSA_OK)
        e = hipMemcpyAsync(xi.data(), din, n * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(din);
    MSA_HIP(e);
    MSA_TRY(rc);
    for (int g = 0; g < m.n_free; g++) z[g] = xi[m.ref2int[g]];
    return MSA_OK;
}

int msa_batch_set_solver(msa_batch *b, int solver)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_CHECK(solver == MSA_SOLVER_PCG || solver == MSA_SOLVER_DIRECT, MSA_ERR_ARG,
              "solver=%d", solver);
    MSA_CHECK(solver == MSA_SOLVER_PCG || !is_mixed(b), MSA_ERR_STATE,
              "the direct solver needs an FP64 batch");
    b->solver = solver;
    return MSA_OK;
}

int msa_batch_direct_info(msa_batch *b, int cls, long long *info)
{
    MSA_CHECK(b && info && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    ClassHost &C = b->cls[cls];
    if (C.bw < 0) MSA_TRY(band_order(C.mesh->h, C.bperm, C.binv, C.bw));
    const long long n = (long long)C.bperm.size();
    const long long W = ((C.bw + 1 + kBandNB + 63) / 64) * 64;
    info[0] = n;
    info[1] = C.bw;
    info[2] = W;
    info[3] = n * W * (long long)sizeof(double);
    return MSA_OK;
}

int msa_batch_get_history(msa_batch *b, int cls, int sample, double *hist)
{
    MSA_CHECK(b && hist && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    ClassHost &C = b->cls[cls];
    MSA_CHECK(sample >= 0 && sample < C.nsamp, MSA_ERR_ARG, "sample %d out of range", sample);
    MSA_CHECK(C.dev.hist && C.dev.hist_len > 0, MSA_ERR_STATE, "no probe history recorded");
    MSA_HIP(hipSetDevice(b->ctx->device));
    MSA_HIP(hipMemcpyAsync(hist, C.dev.hist + (size_t)sample * C.dev.hist_len,
                           (size_t)C.dev.hist_len * sizeof(double), hipMemcpyDeviceToHost,
                           b->ctx->stream));
    MSA_HIP(hipStreamSynchronize(b->ctx->stream));
    return MSA_OK;
}

int msa_batch_get_dynamic_state(msa_batch *b, int cls, int sample, double *u, double *v,
                                double *a, long long *iters)
{
    MSA_CHECK(b && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    MSA_CHECK(b->dyn, MSA_ERR_STATE, "not a dynamics batch");
    ClassHost &C = b->cls[cls];
    MSA_CHECK(sample >= 0 && sample < C.nsamp, MSA_ERR_ARG, "sample %d out of range", sample);
    MSA_HIP(hipSetDevice(b->ctx->device));
    const int tile = sample / b->T, lane = sample % b->T;
    const size_t n = (size_t)C.mesh->h.nb * 3, off = (size_t)tile * C.tile_vec;
    const MeshHost &m = C.mesh->h;
    std::vector<double> h;
    const double *src[3] = {is_mixed(b) ? C.dev.u64 + off : (const double *)C.dev.x + off,
                            C.dev.vel + off, C.dev.acc + off};
    double *dst[3] = {u, v, a};
    for (int k = 0; k < 3; k++) {
        if (!dst[k]) continue;
        MSA_TRY(extract_lane(b, false, src[k], n, lane, h));
        for (int g = 0; g < m.n_free; g++) dst[k][g] = h[m.ref2int[g]];
    }
    if (iters) {
        SampleState st;
        MSA_HIP(hipMemcpyAsync(&st, C.dev.st + sample, sizeof(SampleState), hipMemcpyDeviceToHost,
                               b->ctx->stream));
        MSA_HIP(hipStreamSynchronize(b->ctx->stream));
        *iters = st.cum_iters;
    }
    return MSA_OK;
}

int msa_batch_element_results(msa_batch *b, int cls, int sample, double *strains,
                              double *stresses, double *forces)
{
    MSA_CHECK(b && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    MSA_CHECK(b->assembled, MSA_ERR_STATE, "batch not assembled");
    ClassHost &C = b->cls[cls];
    MSA_CHECK(sample >= 0 && sample < C.nsamp, MSA_ERR_ARG, "sample %d out of range", sample);
    MSA_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    const int ne = C.mesh->h.n_elem, tile = sample / b->T, lane = sample % b->T;
    const double *u = (is_mixed(b) ? C.dev.u64 : (const double *)C.dev.x) + (size_t)tile * C.tile_vec;
    double *d = nullptr;
    const size_t n = (size_t)ne * 120;
    MSA_HIP(hipMalloc(&d, n * sizeof(double)));
    hipError_t e = launch_element_results(s, b->T, b->d_cls + cls, ne, u, lane, tile, b->nu,
                                          strains ? d : nullptr, stresses ? d + (size_t)ne * 48 : nullptr,
                                          forces ? d + (size_t)ne * 96 : nullptr);
    if (e == hipSuccess && strains)
        e = hipMemcpyAsync(strains, d, (size_t)ne * 48 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && stresses)
        e = hipMemcpyAsync(stresses, d + (size_t)ne * 48, (size_t)ne * 48 * sizeof(double),
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && forces)
        e = hipMemcpyAsync(forces, d + (size_t)ne * 96, (size_t)ne * 24 * sizeof(double),
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    MSA_HIP(e);
    return MSA_OK;
}

int msa_batch_get_stats(msa_batch *b, int cls, int *iters, int *outer, double *relres, int *status)
{
    MSA_CHECK(b && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    ClassHost &C = b->cls[cls];
    MSA_HIP(hipSetDevice(b->ctx->device));
    std::vector<SampleState> st((size_t)C.ntiles * b->T);
    int negdet = 0;
    MSA_HIP(hipMemcpyAsync(st.data(), C.dev.st, st.size() * sizeof(SampleState), hipMemcpyDeviceToHost, b->ctx->stream));
    MSA_HIP(hipMemcpyAsync(b->ctx->h_pinned, C.dev.negdet, sizeof(int), hipMemcpyDeviceToHost, b->ctx->stream));
    MSA_HIP(hipStreamSynchronize(b->ctx->stream));
    negdet = *b->ctx->h_pinned;
    const bool mx = is_mixed(b);
    for (int s = 0; s < C.nsamp; s++) {
        const SampleState &S = st[s];
        if (iters) iters[s] = mx ? S.total_iters : S.iters;
        if (outer) outer[s] = mx ? S.outer : 0;
        if (relres) relres[s] = S.relres;
        if (status) status[s] = (negdet && S.status == MSA_OK) ? MSA_NEG_DETJ : S.status;
    }
    return MSA_OK;
}

int msa_batch_get_modulus(msa_batch *b, int cls, int sample, double *E)
{
    MSA_CHECK(b && E && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    ClassHost &C = b->cls[cls];
    MSA_CHECK(sample >= 0 && sample < C.nsamp, MSA_ERR_ARG, "sample %d out of range", sample);
    MSA_HIP(hipSetDevice(b->ctx->device));
    const int tile = sample / b->T, lane = sample % b->T, ne = C.mesh->h.n_elem;
    std::vector<double> h;
    MSA_TRY(extract_lane(b, false, C.dev.E + (size_t)tile * ne * b->T, ne, lane, h));
    memcpy(E, h.data(), ne * sizeof(double));
    return MSA_OK;
}

int msa_batch_get_matrix_csr(msa_batch *b, int cls, int sample, long long *rowptr, int *col,
                             double *val)
{
    MSA_CHECK(b && rowptr && col && val && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG,
              "bad arguments");
    ClassHost &C = b->cls[cls];
    MSA_CHECK(sample >= 0 && sample < C.nsamp, MSA_ERR_ARG, "sample %d out of range", sample);
    MSA_CHECK(b->assembled, MSA_ERR_STATE, "batch not assembled");
    MSA_HIP(hipSetDevice(b->ctx->device));
    const int tile = sample / b->T, lane = sample % b->T, R = b->R;
    const size_t rs = real_size(b);
    std::vector<double> hv;
    const size_t n = (size_t)C.slice_ptr[C.nslices] * 9 * R;
    MSA_TRY(extract_lane(b, is_mixed(b), (const char *)C.dev.val + (size_t)tile * C.tile_vals * rs, n,
                         lane, hv));
    const MeshHost &m = C.mesh->h;
    long long k = 0;
    rowptr[0] = 0;
    // rows in reference order: free DOF g = (node, p)
    std::vector<std::pair<int, int>> rows(m.n_free);
    for (int nd = 0; nd < m.n_nodes; nd++)
        for (int p = 0; p < 3; p++)
            if (m.dof[3 * nd + p] >= 0) rows[m.dof[3 * nd + p]] = {nd, p};
    for (int g = 0; g < m.n_free; g++) {
        const int nd = rows[g].first, p = rows[g].second;
        const int br = m.node_block[nd], rr = br % R;
        for (int j = 0; j < m.brow_ptr[br + 1] - m.brow_ptr[br]; j++) {
            const int bc = m.bcol[m.brow_ptr[br] + j];
            const int mc = m.block_node[bc];
            for (int q = 0; q < 3; q++) {
                const int gc = m.dof[3 * mc + q];
                if (gc < 0) continue;
                col[k] = gc;
                size_t slot = (size_t)C.slot_of(m, br, j);
                int rl = rr, e = 3 * p + q;
                // upper-triangle classes: a lower block is the transpose of
                // the upper block of its column's row
                const int sj = C.srank[m.brow_ptr[br] + j];
                if (C.dev.sym && sj < 13) {
                    slot = (size_t)C.slice_ptr[bc / R] + (26 - sj);
                    rl = bc % R;
                    e = 3 * q + p;
                }
                // the extracted lane keeps the one-sample-tile layout at
                // T = 1, and is entry-major per row lane otherwise
                val[k] = hv[b->T > 1 ? (slot * 9 + e) * R + rl
                            : (rs == 4 ? val_at<float, 1>(slot, e, rl, 0)
                                       : val_at<double, 1>(slot, e, rl, 0))];
                k++;
            }
        }
        rowptr[g + 1] = k;
    }
    return MSA_OK;
}

int msa_batch_set_matrix_csr(msa_batch *b, int cls, int sample, const long long *rowptr,
                             const int *col, const double *val)
{
    MSA_CHECK(b && rowptr && col && val && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG,
              "bad arguments");
    ClassHost &C = b->cls[cls];
    MSA_CHECK(sample >= 0 && sample < C.nsamp, MSA_ERR_ARG, "sample %d out of range", sample);
    MSA_CHECK(!is_mixed(b), MSA_ERR_STATE,
              "a loaded matrix needs an MSA_FP64 batch (the mixed path's FP64 residual is "
              "formed from the moduli)");
    MSA_CHECK(!b->dyn, MSA_ERR_STATE, "a loaded matrix is a static stiffness: dynamics is set");
    const MeshHost &m = C.mesh->h;
    const int nf = m.n_free;
    MSA_CHECK(rowptr[0] == 0, MSA_ERR_ARG, "rowptr[0] = %lld", rowptr[0]);
    for (int g = 0; g < nf; g++)
        MSA_CHECK(rowptr[g + 1] >= rowptr[g], MSA_ERR_ARG, "rowptr decreases at row %d", g);
    MSA_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    const int tile = sample / b->T, lane = sample % b->T, R = b->R;
    const size_t n = (size_t)C.slice_ptr[C.nslices] * 9 * R;
    std::vector<double> hv(n, 0.0);
    std::vector<std::pair<int, int>> rows(nf);
    for (int nd = 0; nd < m.n_nodes; nd++)
        for (int p = 0; p < 3; p++)
            if (m.dof[3 * nd + p] >= 0) rows[m.dof[3 * nd + p]] = {nd, p};
    // position of every column of a row in the stored image (the same
    // addressing as msa_batch_get_matrix_csr); -1 outside the row's blocks
    std::vector<long long> where(nf, -1);
    for (int g = 0; g < nf; g++) {
        const int nd = rows[g].first, p = rows[g].second;
        const int br = m.node_block[nd], rr = br % R;
        const int nbj = m.brow_ptr[br + 1] - m.brow_ptr[br];
        for (int j = 0; j < nbj; j++) {
            const int mc = m.block_node[m.bcol[m.brow_ptr[br] + j]];
            const size_t slot = (size_t)C.slot_of(m, br, j);
            for (int q = 0; q < 3; q++) {
                const int gc = m.dof[3 * mc + q];
                if (gc >= 0)
                    where[gc] = b->T > 1 ? (long long)((slot * 9 + 3 * p + q) * R + rr)
                                         : (long long)val_at<double, 1>(slot, 3 * p + q, rr, 0);
            }
        }
        for (long long k = rowptr[g]; k < rowptr[g + 1]; k++) {
            const int c = col[k];
            MSA_CHECK(c >= 0 && c < nf, MSA_ERR_ARG, "row %d: column %d outside [0, %d)", g, c, nf);
            MSA_CHECK(std::isfinite(val[k]), MSA_ERR_ARG, "entry (%d, %d) is not finite", g, c);
            if (where[c] < 0) {  // zeros of a profile may lie outside the pattern
                MSA_CHECK(val[k] == 0.0, MSA_ERR_ARG,
                          "entry (%d, %d) = %g lies outside the mesh's sparsity pattern", g, c,
                          val[k]);
                continue;
            }
            hv[(size_t)where[c]] += val[k];
        }
        for (int j = 0; j < nbj; j++) {
            const int mc = m.block_node[m.bcol[m.brow_ptr[br] + j]];
            for (int q = 0; q < 3; q++)
                if (m.dof[3 * mc + q] >= 0) where[m.dof[3 * mc + q]] = -1;
        }
    }
    if (!b->assembled && !b->loaded)  // first values of a new batch: unset samples hold zeros
        for (auto &Ck : b->cls)
            MSA_HIP(hipMemsetAsync(Ck.dev.val, 0, (size_t)Ck.ntiles * Ck.tile_vals * sizeof(double), s));
    double *d = nullptr;
    MSA_HIP(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d, hv.data(), n * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = launch_insert_sample(s, b->T, d, n, lane, (double *)C.dev.val + (size_t)tile * C.tile_vals);
    if (e == hipSuccess)
        e = launch_dinv(s, false, b->T, b->d_cls, b->d_items, (int)b->items.size());
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    MSA_HIP(e);
    b->assembled = true;
    b->loaded = true;
    b->factored = false;
    b->mg_ready = false;
    return MSA_OK;
}

int msa_batch_spmv(msa_batch *b, int cls, int sample, const double *x, double *y)
{
    MSA_CHECK(b && x && y && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    ClassHost &C = b->cls[cls];
    MSA_CHECK(sample >= 0 && sample < C.nsamp, MSA_ERR_ARG, "sample %d out of range", sample);
    MSA_CHECK(b->assembled, MSA_ERR_STATE, "batch not assembled");
    MSA_HIP(hipSetDevice(b->ctx->device));
    const MeshHost &m = C.mesh->h;
    const int tile = sample / b->T, lane = sample % b->T;
    const size_t n = (size_t)m.nb * 3;
    std::vector<double> xi(n, 0.0);
    for (int g = 0; g < m.n_free; g++) xi[m.ref2int[g]] = x[g];
    std::vector<int4> its;
    for (auto &it : b->items)
        if (it.x == cls && it.y == tile) its.push_back(it);
    double *dx = nullptr, *dy = nullptr, *din = nullptr;
    int4 *di = nullptr;
    hipStream_t s = b->ctx->stream;
    hipError_t e = hipMalloc(&dx, C.tile_vec * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dy, C.tile_vec * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&din, n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&di, its.size() * sizeof(int4));
    if (e == hipSuccess) e = hipMemsetAsync(dx, 0, C.tile_vec * sizeof(double), s);
    if (e == hipSuccess) e = hipMemsetAsync(dy, 0, C.tile_vec * sizeof(double), s);
    if (e == hipSuccess) e = hipMemcpyAsync(din, xi.data(), n * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(di, its.data(), its.size() * sizeof(int4), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_insert_sample(s, b->T, din, n, lane, dx);
    if (e == hipSuccess) e = launch_spmv(s, is_mixed(b), b->T, b->d_cls, di, (int)its.size(), dx, dy);
    if (e == hipSuccess) e = launch_extract_sample(s, false, b->T, dy, n, lane, din);
    if (e == hipSuccess) e = hipMemcpyAsync(xi.data(), din, n * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (void *p : {(void *)dx, (void *)dy, (void *)din, (void *)di})
        if (p) (void)hipFree(p);
    MSA_HIP(e);
    for (int g = 0; g < m.n_free; g++) y[g] = xi[m.ref2int[g]];
    return MSA_OK;
}

int msa_batch_set_refinement(msa_batch *b, int method, int period, double drop)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_CHECK(method == MSA_REFINE_RESTART || method == MSA_REFINE_RELIABLE, MSA_ERR_ARG,
              "refinement method %d", method);
    MSA_CHECK(period >= 2 && period % 2 == 0 && period <= 4096, MSA_ERR_ARG,
              "replacement check period %d must be even, 2..4096", period);
    MSA_CHECK(drop > 0 && drop <= 1, MSA_ERR_ARG, "replacement trigger %g outside (0, 1]", drop);
    b->refine = method;
    b->period = period;
    b->drop = drop;
    b->refine_set = true;
    drop_graph(b);  // the captured iteration blocks hold the period and the trigger
    return MSA_OK;
}

int msa_batch_time_spmv(msa_batch *b, int reps, double *ms_per_launch)
{
    MSA_CHECK(b && ms_per_launch && reps >= 1, MSA_ERR_ARG, "msa_batch_time_spmv: bad arguments");
    MSA_CHECK(b->assembled, MSA_ERR_STATE, "batch not assembled");
    MSA_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    const bool mx = is_mixed(b);
    const int ni = (int)b->items.size();
    // the start of a PCG solve: every sample active, p = 0, z = D^-1 f
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_set_int(s, b->d_counts, 1, 0); }));
    MSA_TRY(timed(b, KK_OTHER, [&] { return launch_pcg_init(s, mx, b->T, b->d_cls, b->d_items, ni, 0, b->d_counts); }));
    MSA_TRY(reduce(b, OP_INIT0));
    // pbuf 0 reads z and p1, writes p0 and q: every launch does the same work
    MSA_HIP(launch_product(b, 0));
    MSA_HIP(hipEventRecord(b->timing.e0, s));
    for (int r = 0; r < reps; r++) MSA_HIP(launch_product(b, 0));
    MSA_HIP(hipEventRecord(b->timing.e1, s));
    MSA_HIP(hipEventSynchronize(b->timing.e1));
    float ms = 0;
    MSA_HIP(hipEventElapsedTime(&ms, b->timing.e0, b->timing.e1));
    *ms_per_launch = ms / reps;
    return MSA_OK;
}

int msa_batch_set_matrix_cores(msa_batch *b, int enable)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    const bool on = enable != 0;
    if (b->matrix_cores != on) drop_graph(b);  // the captured iteration blocks hold the kernels
    b->matrix_cores = on;
    return MSA_OK;
}

int msa_batch_set_symmetric_product(msa_batch *b, int enable)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    b->full_storage = enable == 0;
    bool changed = false;
    auto apply = [&](ClassHost &C) {
        const int want = C.dev.stencil && b->T == 1 && !b->full_storage ? 1 : 0;
        if (C.dev.sym != want) {
            C.dev.sym = want;
            changed = true;
        }
    };
    for (auto &C : b->cls) apply(C);
    for (auto &L : b->lv) apply(L);
    if (changed) {
        // the stored image changes form: the values are assembled again
        // (padding slots of either form hold zeros, so the other form's
        // leftovers are never read)
        b->assembled = false;
        b->mg_ready = false;
        MSA_TRY(upload_class_table(b));
    }
    return MSA_OK;
}

int msa_batch_set_matrix_free_product(msa_batch *b, int enable)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    b->mf_product = enable != 0;  // takes effect at the next solve
    return MSA_OK;
}

int msa_batch_product_form(msa_batch *b, int cls, int *form)
{
    MSA_CHECK(b && form, MSA_ERR_ARG, "bad arguments");
    MSA_CHECK(cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "class %d", cls);
    const ClassDev &D = b->cls[cls].dev;
    *form = D.mf_pcg ? 2 : (D.sym ? 1 : 0);
    return MSA_OK;
}

int msa_batch_graph_info(msa_batch *b, long long *info)
{
    MSA_CHECK(b && info, MSA_ERR_ARG, "bad arguments");
    info[0] = b->graph_captures;
    info[1] = b->graph ? b->graph_kernels : 0;
    info[2] = b->graph ? b->graph_iters : 0;
    info[3] = b->graph_replays;
    return MSA_OK;
}

int msa_batch_layout_info(msa_batch *b, int cls, long long *info)
{
    MSA_CHECK(b && info && cls >= 0 && cls < (int)b->cls.size(), MSA_ERR_ARG, "bad arguments");
    const ClassHost &C = b->cls[cls];
    info[0] = C.dev.stencil;
    info[1] = C.dev.sym;
    info[2] = C.slice_ptr[C.nslices];
    info[3] = C.dev.nb;
    return MSA_OK;
}

int msa_batch_set_integration(msa_batch *b, int precision)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_CHECK(precision == MSA_FP64 || precision == MSA_MIXED, MSA_ERR_ARG,
              "integration precision %d", precision);
    b->integration = precision;
    b->assembled = false;
    return MSA_OK;
}

int msa_batch_set_kernel_timing(msa_batch *b, int enable)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    b->timing.on = enable != 0;
    for (int k = 0; k < kKernelKinds; k++) {
        b->timing.ms[k] = 0;
        b->timing.n[k] = 0;
    }
    return MSA_OK;
}

int msa_batch_get_kernel_timing(msa_batch *b, double *ms_total, long long *launches)
{
    MSA_CHECK(b && ms_total && launches, MSA_ERR_ARG, "bad arguments");
    for (int k = 0; k < kKernelKinds; k++) {
        ms_total[k] = b->timing.ms[k];
        launches[k] = b->timing.n[k];
    }
    return MSA_OK;
}

int msa_batch_spmv_bytes(msa_batch *b, double *bytes)
{
    MSA_CHECK(b && bytes, MSA_ERR_ARG, "bad arguments");
    const double rs = (double)real_size(b);
    double kernel = 0, survey = 0;
    for (auto &C : b->cls) {
        const MeshHost &m = C.mesh->h;
        const double S = C.nsamp, nnz = (double)m.nnz_scalar, n = 3.0 * m.nb;
        const double idx = 4.0 * (double)m.bcol.size() + 4.0 * (C.nslices + 1);
        if (C.dev.mf_pcg)  // matrix-free: the p pass and the product's vectors, the moduli
            kernel += S * 5.0 * rs * n + S * 8.0 * m.n_elem;
        else if (C.dev.sym)  // the upper triangle with the diagonal; offsets, not indices
            kernel += S * rs * 0.5 * (nnz + n) + S * 4.0 * rs * n;
        else
            kernel += S * rs * nnz + idx + S * 4.0 * rs * n;
        survey += S * rs * nnz + idx + S * 2.0 * rs * n;
    }
    bytes[0] = kernel;
    bytes[1] = survey;
    return MSA_OK;
}

int msa_autotune_tile(msa_ctx *ctx, msa_mesh *mesh, int n_samples, int precision, int *best_tile,
                      double *ms_per_tile)
{
    MSA_CHECK(ctx && mesh && best_tile && n_samples >= 1, MSA_ERR_ARG, "bad arguments");
    const int tiles[7] = {1, 2, 4, 8, 16, 32, 64};
    double best = 1e300;
    *best_tile = 0;
    const MeshHost &h = mesh->h;
    std::vector<double> E((size_t)n_samples * h.n_elem, 1.0);
    std::vector<double> f(h.n_free, 1.0);
    for (int k = 0; k < 7; k++) {
        if (ms_per_tile) ms_per_tile[k] = -1;
        msa_batch *b = nullptr;
        int rc = msa_batch_create(ctx, 1, &mesh, &n_samples, precision, tiles[k], &b);
        if (rc != MSA_OK) continue;
        rc = msa_batch_set_modulus(b, 0, E.data());
        if (rc == MSA_OK) rc = msa_batch_set_rhs(b, 0, f.data());
        if (rc == MSA_OK) rc = msa_batch_assemble(b, 0.3);
        hipStream_t s = ctx->stream;
        const int ni = (int)b->items.size();
        if (rc == MSA_OK) {
            (void)hipMemsetAsync(b->d_counts, 0, sizeof(int), s);
            if (launch_pcg_init(s, precision == MSA_MIXED, b->T, b->d_cls, b->d_items, ni, 0,
                                b->d_counts) != hipSuccess)
                rc = MSA_ERR_HIP;
            if (rc == MSA_OK) rc = reduce(b, OP_INIT0);
        }
        if (rc == MSA_OK) {
            // warm-up, then 10 timed products (pbuf 0 keeps every sample active)
            const int reps = 10;
            hipError_t e = launch_pcg_a(s, precision == MSA_MIXED, b->T, 3, b->d_cls, b->d_items, ni, 0,
                                        b->d_counts);
            if (e == hipSuccess) e = hipEventRecord(b->timing.e0, s);
            for (int r = 0; r < reps && e == hipSuccess; r++)
                e = launch_pcg_a(s, precision == MSA_MIXED, b->T, 3, b->d_cls, b->d_items, ni, 0,
                                 b->d_counts);
            if (e == hipSuccess) e = hipEventRecord(b->timing.e1, s);
            if (e == hipSuccess) e = hipEventSynchronize(b->timing.e1);
            float ms = 0;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, b->timing.e0, b->timing.e1);
            if (e == hipSuccess) {
                const double per = ms / reps;
                if (ms_per_tile) ms_per_tile[k] = per;
                if (per < best) {
                    best = per;
                    *best_tile = tiles[k];
                }
            }
        }
        msa_batch_destroy(b);
    }
    MSA_CHECK(*best_tile > 0, MSA_ERR_STATE, "autotune: no tile could be timed (%s)",
              g_last_error.c_str());
    return MSA_OK;
}

int msa_batch_set_schedule(msa_batch *b, int unroll)
{
    MSA_CHECK(b, MSA_ERR_ARG, "null batch");
    MSA_CHECK(unroll >= 1 && unroll <= 4, MSA_ERR_ARG, "unroll=%d (1..4)", unroll);
    MSA_CHECK(b->T == 1 || unroll == 3, MSA_ERR_ARG,
              "lane tiles T > 1 use groups of 3 slots (unroll=%d, T=%d)", unroll, b->T);
    if (b->unroll != unroll) drop_graph(b);  // the captured iteration blocks hold the kernel
    b->unroll = unroll;
    return MSA_OK;
}

// Candidate schedules of msa_autotune_schedule: (lane tile T, slot unroll U).
static const int kTuneCand[MSA_TUNE_CANDIDATES][2] = {{1, 1}, {1, 2}, {1, 3}, {1, 4},
                                                      {2, 3}, {4, 3}, {16, 3}, {64, 3}};

int msa_autotune_schedule(msa_ctx *ctx, int n_classes, msa_mesh *const *meshes,
                          int samples_per_class, int precision, int precond, int *best_tile,
                          int *best_unroll, double *ms_per_sample)
{
    MSA_CHECK(ctx && meshes && best_tile && best_unroll && n_classes >= 1 && n_classes <= 32 &&
                  samples_per_class >= 1,
              MSA_ERR_ARG, "msa_autotune_schedule: bad arguments");
    MSA_CHECK(precond == MSA_PRECOND_JACOBI || precond == MSA_PRECOND_MULTIGRID ||
                  precond == MSA_PRECOND_FIRST_SAMPLE,
              MSA_ERR_ARG, "preconditioner kind=%d", precond);
    *best_tile = *best_unroll = 0;
    std::vector<int> ns(n_classes, samples_per_class);
    // a fixed lognormal modulus field (COV 0.3) per sample and a unit load on
    // every free DOF: samples converge in different iteration counts, as in
    // a Monte Carlo batch
    std::vector<std::vector<double>> E(n_classes), f(n_classes);
    {
        std::mt19937_64 rng(20261202);
        std::normal_distribution<double> nd(0.0, 1.0);
        const double sg = sqrt(log(1.0 + 0.09));
        for (int c = 0; c < n_classes; c++) {
            const MeshHost &h = meshes[c]->h;
            E[c].resize((size_t)samples_per_class * h.n_elem);
            for (double &e : E[c]) e = 2.1e5 * exp(-0.5 * sg * sg + sg * nd(rng));
            f[c].assign(h.n_free, 1.0);
        }
    }
    const double tol = precision == MSA_MIXED ? 1e-9 : 1e-10;
    double best = 1e300;
    for (int k = 0; k < MSA_TUNE_CANDIDATES; k++) {
        if (ms_per_sample) ms_per_sample[k] = -1;
        const int T = kTuneCand[k][0], U = kTuneCand[k][1];
        msa_batch *b = nullptr;
        int rc = msa_batch_create(ctx, n_classes, meshes, ns.data(), precision, T, &b);
        if (rc != MSA_OK) continue;
        if (precond != MSA_PRECOND_JACOBI) rc = msa_batch_set_preconditioner(b, precond, 0, 0.0);
        if (rc == MSA_OK) rc = msa_batch_set_schedule(b, U);
        for (int c = 0; c < n_classes && rc == MSA_OK; c++) {
            rc = msa_batch_set_modulus(b, c, E[c].data());
            if (rc == MSA_OK) rc = msa_batch_set_rhs(b, c, f[c].data());
        }
        // warm-up (graph capture), then one timed assembly + solve
        double t = 0;
        for (int rep = 0; rep < 2 && rc == MSA_OK; rep++) {
            const auto t0 = std::chrono::steady_clock::now();
            rc = msa_batch_assemble(b, 0.3);
            if (rc == MSA_OK) rc = msa_batch_solve(b, tol, 20000, 1e-4, 30, 16);
            if (rc == MSA_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = MSA_ERR_HIP;
            t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        bool ok = rc == MSA_OK;
        for (int c = 0; c < n_classes && ok; c++) {
            std::vector<int> st(samples_per_class);
            ok = msa_batch_get_stats(b, c, nullptr, nullptr, nullptr, st.data()) == MSA_OK;
            for (int v : st) ok = ok && v == MSA_OK;
        }
        msa_batch_destroy(b);
        if (!ok) continue;
        const double per = 1e3 * t / ((double)n_classes * samples_per_class);
        if (ms_per_sample) ms_per_sample[k] = per;
        if (per < best) {
            best = per;
            *best_tile = T;
            *best_unroll = U;
        }
    }
    MSA_CHECK(*best_tile > 0, MSA_ERR_STATE, "autotune: no schedule could be timed (%s)",
              g_last_error.c_str());
    set_error("");
    return MSA_OK;
}

int msa_autotune_candidates(int *tiles, int *unrolls)
{
    MSA_CHECK(tiles && unrolls, MSA_ERR_ARG, "bad arguments");
    for (int k = 0; k < MSA_TUNE_CANDIDATES; k++) {
        tiles[k] = kTuneCand[k][0];
        unrolls[k] = kTuneCand[k][1];
    }
    return MSA_OK;
}

}  // extern "C"
