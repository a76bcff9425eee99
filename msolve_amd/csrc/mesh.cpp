// Host-side analysis of a Hexa8 mesh, run once per mesh (not per sample).
//
// The reference rebuilds its skyline profile for every sample
// (GlobalMatrixAssemblerSkyline.CalculateRowIndex, GMAS.cs:14-44) and finds
// each scatter target through two dictionary lookups and
// SymmetricMatrix2D.IndexOf (GMAS.cs:68-98).  Here the sparsity is fixed per
// mesh: node blocks (3x3) of the free-DOF K, sorted block columns per block
// row, the slot of every element node pair (a, b) inside block row
// block(a), and an element colouring in which no two elements of a colour
// share a node, so the device scatter is race-free and deterministic.
#include <algorithm>
#include <cstring>

#include "msa_common.hpp"

namespace msa {

int build_mesh_host(MeshHost &m, int n_nodes, const double *xyz, int n_elem,
                    const int *conn, const int *dof, int n_free)
{
    MSA_CHECK(n_nodes > 0 && n_elem > 0 && n_free >= 0, MSA_ERR_ARG,
              "mesh: n_nodes=%d n_elem=%d n_free=%d", n_nodes, n_elem, n_free);
    MSA_CHECK(xyz && conn && dof, MSA_ERR_ARG, "mesh: null array");
    m.n_nodes = n_nodes;
    m.n_elem = n_elem;
    m.n_free = n_free;
    m.xyz.assign(xyz, xyz + 3 * (size_t)n_nodes);
    m.conn.assign(conn, conn + 8 * (size_t)n_elem);
    m.dof.assign(dof, dof + 3 * (size_t)n_nodes);
    for (size_t i = 0; i < m.conn.size(); i++)
        MSA_CHECK(m.conn[i] >= 0 && m.conn[i] < n_nodes, MSA_ERR_ARG,
                  "mesh: conn[%zu]=%d out of range", i, m.conn[i]);
    for (int e = 0; e < n_elem; e++)
        for (int a = 0; a < 8; a++)
            for (int b = a + 1; b < 8; b++)
                MSA_CHECK(m.conn[8 * e + a] != m.conn[8 * e + b], MSA_ERR_ARG,
                          "mesh: element %d repeats node %d", e, m.conn[8 * e + a]);

    // blocks and the reference -> internal permutation
    m.node_free.assign(n_nodes, 0);
    m.node_block.assign(n_nodes, -1);
    m.block_node.clear();
    m.ref2int.assign(n_free, -1);
    for (int n = 0; n < n_nodes; n++) {
        uint8_t mask = 0;
        for (int d = 0; d < 3; d++) {
            int g = m.dof[3 * n + d];
            MSA_CHECK(g >= -1 && g < n_free, MSA_ERR_ARG, "mesh: dof[%d]=%d out of range",
                      3 * n + d, g);
            if (g >= 0) mask |= (uint8_t)(1u << d);
        }
        m.node_free[n] = mask;
        if (mask) {
            m.node_block[n] = (int)m.block_node.size();
            m.block_node.push_back(n);
        }
    }
    m.nb = (int)m.block_node.size();
    for (int n = 0; n < n_nodes; n++)
        for (int d = 0; d < 3; d++) {
            int g = m.dof[3 * n + d];
            if (g < 0) continue;
            MSA_CHECK(m.ref2int[g] < 0, MSA_ERR_ARG, "mesh: free DOF %d assigned twice", g);
            m.ref2int[g] = 3 * m.node_block[n] + d;
        }
    for (int g = 0; g < n_free; g++)
        MSA_CHECK(m.ref2int[g] >= 0, MSA_ERR_ARG, "mesh: free DOF %d not assigned", g);

    // node -> incident elements
    std::vector<int> inc_ptr(n_nodes + 1, 0), inc;
    for (int e = 0; e < n_elem; e++)
        for (int a = 0; a < 8; a++) inc_ptr[m.conn[8 * e + a] + 1]++;
    for (int n = 0; n < n_nodes; n++) inc_ptr[n + 1] += inc_ptr[n];
    inc.resize(inc_ptr[n_nodes]);
    {
        std::vector<int> fill(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int e = 0; e < n_elem; e++)
            for (int a = 0; a < 8; a++) inc[fill[m.conn[8 * e + a]]++] = e;
    }

    // block adjacency
    m.brow_ptr.assign(m.nb + 1, 0);
    m.bcol.clear();
    m.diag_pos.assign(m.nb, -1);
    m.max_row = 0;
    m.nnz_scalar = 0;
    std::vector<int> tmp;
    for (int br = 0; br < m.nb; br++) {
        int n = m.block_node[br];
        tmp.clear();
        for (int p = inc_ptr[n]; p < inc_ptr[n + 1]; p++)
            for (int a = 0; a < 8; a++) {
                int bc = m.node_block[m.conn[8 * inc[p] + a]];
                if (bc >= 0) tmp.push_back(bc);
            }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        for (size_t j = 0; j < tmp.size(); j++) {
            if (tmp[j] == br) m.diag_pos[br] = (int)j;
            m.nnz_scalar += (long long)__builtin_popcount(m.node_free[n]) *
                            __builtin_popcount(m.node_free[m.block_node[tmp[j]]]);
        }
        m.bcol.insert(m.bcol.end(), tmp.begin(), tmp.end());
        m.brow_ptr[br + 1] = (int)m.bcol.size();
        m.max_row = std::max<int>(m.max_row, (int)tmp.size());
        MSA_CHECK(m.diag_pos[br] >= 0, MSA_ERR_STATE, "mesh: block %d lacks a diagonal", br);
    }
    MSA_CHECK(m.max_row < 255, MSA_ERR_ARG, "mesh: a node couples to %d nodes (max 254)",
              m.max_row);

    // element pair -> slot
    m.emap.assign((size_t)n_elem * 64, 255);
    for (int e = 0; e < n_elem; e++)
        for (int a = 0; a < 8; a++) {
            int br = m.node_block[m.conn[8 * e + a]];
            if (br < 0) continue;
            const int *row = m.bcol.data() + m.brow_ptr[br];
            int len = m.brow_ptr[br + 1] - m.brow_ptr[br];
            for (int b = 0; b < 8; b++) {
                int bc = m.node_block[m.conn[8 * e + b]];
                if (bc < 0) continue;
                int j = (int)(std::lower_bound(row, row + len, bc) - row);
                m.emap[(size_t)e * 64 + a * 8 + b] = (uint8_t)j;
            }
        }

    // greedy colouring in element order: smallest colour unused at all 8 nodes
    std::vector<uint64_t> used(n_nodes, 0);
    std::vector<int> colour(n_elem);
    m.ncolours = 0;
    for (int e = 0; e < n_elem; e++) {
        uint64_t u = 0;
        for (int a = 0; a < 8; a++) u |= used[m.conn[8 * e + a]];
        MSA_CHECK(~u != 0, MSA_ERR_ARG, "mesh: more than 64 element colours needed");
        int c = __builtin_ctzll(~u);
        MSA_CHECK(c < kMaxColours, MSA_ERR_ARG, "mesh: more than %d colours", kMaxColours);
        colour[e] = c;
        for (int a = 0; a < 8; a++) used[m.conn[8 * e + a]] |= (1ull << c);
        m.ncolours = std::max(m.ncolours, c + 1);
    }
    m.colour_ptr.assign(m.ncolours + 1, 0);
    for (int e = 0; e < n_elem; e++) m.colour_ptr[colour[e] + 1]++;
    for (int c = 0; c < m.ncolours; c++) m.colour_ptr[c + 1] += m.colour_ptr[c];
    m.colour_elems.resize(n_elem);
    {
        std::vector<int> fill(m.colour_ptr.begin(), m.colour_ptr.end() - 1);
        for (int e = 0; e < n_elem; e++) m.colour_elems[fill[colour[e]]++] = e;
    }
    return MSA_OK;
}

// Band ordering of the free DOFs for the direct solver.  Candidates: the
// block order itself, a reverse Cuthill-McKee order of the block graph
// (BFS from a pseudo-peripheral block, neighbours by ascending degree,
// reversed), and the six lexicographic coordinate orders (exact planes of
// structured meshes); the one with the smallest half bandwidth wins.  Within
// a block its free components keep the order X, Y, Z.
static long long dof_bandwidth(const MeshHost &m, const std::vector<int> &pos_block)
{
    // first DOF index of every block in the candidate order
    std::vector<int> order(m.nb);
    for (int b = 0; b < m.nb; b++) order[pos_block[b]] = b;
    std::vector<long long> first(m.nb);
    long long k = 0;
    for (int q = 0; q < m.nb; q++) {
        first[order[q]] = k;
        k += __builtin_popcount(m.node_free[m.block_node[order[q]]]);
    }
    long long bw = 0;
    for (int br = 0; br < m.nb; br++)
        for (int j = m.brow_ptr[br]; j < m.brow_ptr[br + 1]; j++) {
            const int bc = m.bcol[j];
            if (pos_block[bc] > pos_block[br]) continue;
            // rows of br minus columns of bc: last row of br - first column of bc
            const long long d = first[br] + __builtin_popcount(m.node_free[m.block_node[br]]) - 1 -
                                first[bc];
            bw = std::max(bw, d);
        }
    return bw;
}

int band_order(const MeshHost &m, std::vector<int> &perm, std::vector<int> &inv, long long &bw)
{
    std::vector<std::vector<int>> cand;
    {   // natural
        std::vector<int> p(m.nb);
        for (int b = 0; b < m.nb; b++) p[b] = b;
        cand.push_back(p);
    }
    {   // reverse Cuthill-McKee
        std::vector<int> deg(m.nb);
        for (int b = 0; b < m.nb; b++) deg[b] = m.brow_ptr[b + 1] - m.brow_ptr[b] - 1;
        std::vector<int> level(m.nb, -1), order;
        order.reserve(m.nb);
        auto bfs = [&](int start, std::vector<int> &out) {
            std::vector<int> q{start};
            level[start] = 0;
            for (size_t h = 0; h < q.size(); h++) {
                const int b = q[h];
                std::vector<int> nb;
                for (int j = m.brow_ptr[b]; j < m.brow_ptr[b + 1]; j++)
                    if (level[m.bcol[j]] < 0) nb.push_back(m.bcol[j]);
                std::sort(nb.begin(), nb.end(), [&](int x, int y) {
                    return deg[x] != deg[y] ? deg[x] < deg[y] : x < y;
                });
                for (int c : nb) {
                    level[c] = level[b] + 1;
                    q.push_back(c);
                }
            }
            out = q;
        };
        for (int s0 = 0; s0 < m.nb; s0++) {
            if (level[s0] >= 0) continue;
            // pseudo-peripheral start: minimum degree block of the component,
            // then the last block of a BFS from it
            std::vector<int> comp;
            bfs(s0, comp);
            int best = comp[0];
            for (int b : comp)
                if (deg[b] < deg[best]) best = b;
            for (int b : comp) level[b] = -1;
            std::vector<int> tmp;
            bfs(best, tmp);
            const int far = tmp.back();
            for (int b : tmp) level[b] = -1;
            bfs(far, tmp);
            order.insert(order.end(), tmp.begin(), tmp.end());
        }
        std::reverse(order.begin(), order.end());
        std::vector<int> p(m.nb);
        for (int q = 0; q < m.nb; q++) p[order[q]] = q;
        cand.push_back(p);
    }
    {   // lexicographic coordinate orders (slowest axis first)
        const int axes[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        for (auto &ax : axes) {
            std::vector<int> order(m.nb);
            for (int b = 0; b < m.nb; b++) order[b] = b;
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
                const double *px = &m.xyz[3 * (size_t)m.block_node[x]];
                const double *py = &m.xyz[3 * (size_t)m.block_node[y]];
                for (int k = 0; k < 3; k++)
                    if (px[ax[k]] != py[ax[k]]) return px[ax[k]] < py[ax[k]];
                return x < y;
            });
            std::vector<int> p(m.nb);
            for (int q = 0; q < m.nb; q++) p[order[q]] = q;
            cand.push_back(p);
        }
    }
    long long best_bw = -1;
    size_t best = 0;
    for (size_t c = 0; c < cand.size(); c++) {
        const long long b = dof_bandwidth(m, cand[c]);
        if (best_bw < 0 || b < best_bw) {
            best_bw = b;
            best = c;
        }
    }
    const std::vector<int> &pos = cand[best];
    std::vector<int> order(m.nb);
    for (int b = 0; b < m.nb; b++) order[pos[b]] = b;
    perm.clear();
    inv.assign((size_t)m.nb * 3, -1);
    for (int q = 0; q < m.nb; q++) {
        const int b = order[q];
        const uint8_t fr = m.node_free[m.block_node[b]];
        for (int d = 0; d < 3; d++)
            if ((fr >> d) & 1) {
                inv[3 * (size_t)b + d] = (int)perm.size();
                perm.push_back(3 * b + d);
            }
    }
    bw = best_bw;
    return MSA_OK;
}

}  // namespace msa

// ------------------------------------------------------------ multigrid levels

namespace msa {

// Sorted distinct coordinates of one axis (merged within tol) and the index
// of every node in them.
static void axis_grid(const MeshHost &m, int d, std::vector<double> &vals, std::vector<int> &idx)
{
    std::vector<double> c(m.n_nodes);
    for (int n = 0; n < m.n_nodes; n++) c[n] = m.xyz[3 * n + d];
    std::vector<double> s(c);
    std::sort(s.begin(), s.end());
    const double tol = 1e-9 * std::max(1e-300, s.back() - s.front());
    vals.clear();
    for (double v : s)
        if (vals.empty() || v - vals.back() > tol) vals.push_back(v);
    idx.resize(m.n_nodes);
    for (int n = 0; n < m.n_nodes; n++) {
        auto it = std::lower_bound(vals.begin(), vals.end(), c[n] - tol);
        idx[n] = (int)(it - vals.begin());
    }
}

int mg_coarsen(const MeshHost &f, MeshHost &c, MgTransfer &t)
{
    std::vector<double> gx[3];
    std::vector<int> gi[3];
    for (int d = 0; d < 3; d++) axis_grid(f, d, gx[d], gi[d]);
    const int nx = (int)gx[0].size(), ny = (int)gx[1].size(), nz = (int)gx[2].size();
    MSA_CHECK((long long)nx * ny * nz == f.n_nodes, MSA_ERR_ARG,
              "multigrid: the nodes do not form a tensor-product grid");
    const int ex = nx - 1, ey = ny - 1, ez = nz - 1;
    MSA_CHECK(ex >= 4 && ey >= 4 && ez >= 4 && ex % 2 == 0 && ey % 2 == 0 && ez % 2 == 0,
              MSA_ERR_ARG, "multigrid: %d x %d x %d elements cannot be halved", ex, ey, ez);
    MSA_CHECK((long long)ex * ey * ez == f.n_elem, MSA_ERR_ARG,
              "multigrid: element count does not match the grid");
    std::vector<int> gnode((size_t)nx * ny * nz, -1);
    for (int n = 0; n < f.n_nodes; n++) {
        const size_t g = gi[0][n] + (size_t)nx * (gi[1][n] + (size_t)ny * gi[2][n]);
        MSA_CHECK(gnode[g] < 0, MSA_ERR_ARG, "multigrid: two nodes share grid point");
        gnode[g] = n;
    }
    // every element must be one grid cell
    std::vector<int> cell_elem((size_t)ex * ey * ez, -1);
    for (int e = 0; e < f.n_elem; e++) {
        int lo[3] = {1 << 30, 1 << 30, 1 << 30};
        for (int a = 0; a < 8; a++)
            for (int d = 0; d < 3; d++) lo[d] = std::min(lo[d], gi[d][f.conn[8 * e + a]]);
        int seen = 0;
        for (int a = 0; a < 8; a++) {
            const int n = f.conn[8 * e + a];
            int off = 0;
            for (int d = 0; d < 3; d++) {
                const int o = gi[d][n] - lo[d];
                MSA_CHECK(o == 0 || o == 1, MSA_ERR_ARG, "multigrid: element %d is not a grid cell", e);
                off |= o << d;
            }
            seen |= 1 << off;
        }
        MSA_CHECK(seen == 255, MSA_ERR_ARG, "multigrid: element %d is not a grid cell", e);
        const size_t cell = lo[0] + (size_t)ex * (lo[1] + (size_t)ey * lo[2]);
        MSA_CHECK(cell_elem[cell] < 0, MSA_ERR_ARG, "multigrid: two elements share a cell");
        cell_elem[cell] = e;
    }
    // coarse grid: every other grid line
    const int cx = ex / 2 + 1, cy = ey / 2 + 1, cz = ez / 2 + 1;
    const int cn_nodes = cx * cy * cz, cn_elem = (ex / 2) * (ey / 2) * (ez / 2);
    std::vector<double> cxyz(3 * (size_t)cn_nodes);
    std::vector<uint8_t> cfix(3 * (size_t)cn_nodes);
    for (int K = 0; K < cz; K++)
        for (int J = 0; J < cy; J++)
            for (int I = 0; I < cx; I++) {
                const int cn = I + cx * (J + cy * K);
                const int fn = gnode[2 * I + (size_t)nx * (2 * J + (size_t)ny * 2 * K)];
                for (int d = 0; d < 3; d++) {
                    cxyz[3 * cn + d] = f.xyz[3 * fn + d];
                    cfix[3 * cn + d] = f.dof[3 * fn + d] < 0;
                }
            }
    std::vector<int> cdof(3 * (size_t)cn_nodes);
    int cn_free = 0;
    for (size_t i = 0; i < cdof.size(); i++) cdof[i] = cfix[i] ? -1 : cn_free++;
    std::vector<int> cconn(8 * (size_t)cn_elem);
    t.child.assign(8 * (size_t)cn_elem, -1);
    for (int K = 0; K < ez / 2; K++)
        for (int J = 0; J < ey / 2; J++)
            for (int I = 0; I < ex / 2; I++) {
                const int ce = I + (ex / 2) * (J + (ey / 2) * K);
                for (int o = 0; o < 8; o++) {
                    const size_t cell = (2 * I + (o & 1)) +
                                        (size_t)ex * ((2 * J + ((o >> 1) & 1)) +
                                                      (size_t)ey * (2 * K + (o >> 2)));
                    t.child[8 * (size_t)ce + o] = cell_elem[cell];
                }
                // local node order of the child at the cell's low corner
                const int e0 = t.child[8 * (size_t)ce];
                for (int a = 0; a < 8; a++) {
                    const int n = f.conn[8 * e0 + a];
                    const int oi = gi[0][n] - 2 * I, oj = gi[1][n] - 2 * J, ok = gi[2][n] - 2 * K;
                    cconn[8 * (size_t)ce + a] = (I + oi) + cx * ((J + oj) + cy * (K + ok));
                }
            }
    {
        const int rc = build_mesh_host(c, cn_nodes, cxyz.data(), cn_elem, cconn.data(), cdof.data(), cn_free);
        if (rc != MSA_OK) return rc;
    }
    // trilinear interpolation: a fine node on an odd grid line takes the two
    // neighbouring coarse lines, weighted by distance
    auto axis_w = [&](int d, int i, int *ci, double *w) {
        if (i % 2 == 0) {
            ci[0] = i / 2;
            w[0] = 1.0;
            return 1;
        }
        const double x0 = gx[d][i - 1], x1 = gx[d][i], x2 = gx[d][i + 1];
        ci[0] = (i - 1) / 2;
        w[0] = (x2 - x1) / (x2 - x0);
        ci[1] = (i + 1) / 2;
        w[1] = (x1 - x0) / (x2 - x0);
        return 2;
    };
    t.p_ptr.assign(f.nb + 1, 0);
    t.p_col.clear();
    t.p_w.clear();
    for (int fb = 0; fb < f.nb; fb++) {
        const int n = f.block_node[fb];
        int ci[3][2];
        double w[3][2];
        const int k0 = axis_w(0, gi[0][n], ci[0], w[0]);
        const int k1 = axis_w(1, gi[1][n], ci[1], w[1]);
        const int k2 = axis_w(2, gi[2][n], ci[2], w[2]);
        for (int c2 = 0; c2 < k2; c2++)
            for (int c1 = 0; c1 < k1; c1++)
                for (int c0 = 0; c0 < k0; c0++) {
                    const int cn = ci[0][c0] + cx * (ci[1][c1] + cy * ci[2][c2]);
                    const int cb = c.node_block[cn];
                    if (cb < 0) continue;
                    t.p_col.push_back(cb);
                    t.p_w.push_back(w[0][c0] * w[1][c1] * w[2][c2]);
                }
        t.p_ptr[fb + 1] = (int)t.p_col.size();
    }
    // restriction = transpose, fine blocks ascending per coarse block
    t.r_ptr.assign(c.nb + 1, 0);
    for (int cb : t.p_col) t.r_ptr[cb + 1]++;
    for (int i = 0; i < c.nb; i++) t.r_ptr[i + 1] += t.r_ptr[i];
    t.r_col.assign(t.p_col.size(), 0);
    t.r_w.assign(t.p_col.size(), 0.0);
    std::vector<int> fill(t.r_ptr.begin(), t.r_ptr.end() - 1);
    for (int fb = 0; fb < f.nb; fb++)
        for (int k = t.p_ptr[fb]; k < t.p_ptr[fb + 1]; k++) {
            const int pos = fill[t.p_col[k]]++;
            t.r_col[pos] = fb;
            t.r_w[pos] = t.p_w[k];
        }
    return MSA_OK;
}

}  // namespace msa

namespace msa {

int grid_info(const MeshHost &m, GridInfo &gi_out)
{
    std::vector<double> gx[3];
    std::vector<int> gi[3];
    for (int d = 0; d < 3; d++) axis_grid(m, d, gx[d], gi[d]);
    const int nx = (int)gx[0].size(), ny = (int)gx[1].size(), nz = (int)gx[2].size();
    if ((long long)nx * ny * nz != m.n_nodes || nx < 2 || ny < 2 || nz < 2) return MSA_ERR_ARG;
    if (nx > 1024 || ny > 1024 || nz > 1024) return MSA_ERR_ARG;
    if ((long long)(nx - 1) * (ny - 1) * (nz - 1) != m.n_elem) return MSA_ERR_ARG;
    // equal spacing per direction: every element has the same shape
    for (int d = 0; d < 3; d++) {
        const double h = gx[d][1] - gx[d][0];
        for (size_t i = 1; i + 1 < gx[d].size(); i++)
            if (std::abs((gx[d][i + 1] - gx[d][i]) - h) > 1e-9 * h) return MSA_ERR_ARG;
    }
    GridInfo g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.gblock.assign((size_t)nx * ny * nz, -2);
    for (int n = 0; n < m.n_nodes; n++) {
        const size_t p = gi[0][n] + (size_t)nx * (gi[1][n] + (size_t)ny * gi[2][n]);
        if (g.gblock[p] != -2) return MSA_ERR_ARG;
        g.gblock[p] = m.node_block[n];
    }
    g.bpos.assign(m.nb, 0);
    for (int b = 0; b < m.nb; b++) {
        const int n = m.block_node[b];
        g.bpos[b] = gi[0][n] | (gi[1][n] << 10) | (gi[2][n] << 20);
    }
    // every element one cell, all with element 0's local node order
    g.cell_elem.assign((size_t)(nx - 1) * (ny - 1) * (nz - 1), -1);
    int pattern[8] = {0};
    for (int e = 0; e < m.n_elem; e++) {
        int lo[3] = {1 << 30, 1 << 30, 1 << 30};
        for (int a = 0; a < 8; a++)
            for (int d = 0; d < 3; d++) lo[d] = std::min(lo[d], gi[d][m.conn[8 * e + a]]);
        for (int a = 0; a < 8; a++) {
            const int n = m.conn[8 * e + a];
            int off = 0;
            for (int d = 0; d < 3; d++) {
                const int o = gi[d][n] - lo[d];
                if (o != 0 && o != 1) return MSA_ERR_ARG;
                off |= o << d;
            }
            if (e == 0)
                pattern[a] = off;
            else if (pattern[a] != off)
                return MSA_ERR_ARG;
        }
        const size_t cell = lo[0] + (size_t)(nx - 1) * (lo[1] + (size_t)(ny - 1) * lo[2]);
        if (cell >= g.cell_elem.size() || g.cell_elem[cell] >= 0) return MSA_ERR_ARG;
        g.cell_elem[cell] = e;
    }
    int seen = 0;
    for (int a = 0; a < 8; a++) {
        g.corner_node[pattern[a]] = a;
        seen |= 1 << pattern[a];
    }
    if (seen != 255) return MSA_ERR_ARG;
    // element number = cell number?
    g.cell_identity = 1;
    for (size_t c = 0; c < g.cell_elem.size(); c++)
        if (g.cell_elem[c] != (int)c) {
            g.cell_identity = 0;
            break;
        }
    // free nodes = a sub-box of the grid, blocks numbered lexicographically in it?
    int blo[3] = {nx, ny, nz}, bhi[3] = {-1, -1, -1};
    for (int b = 0; b < m.nb; b++) {
        const int p[3] = {g.bpos[b] & 1023, (g.bpos[b] >> 10) & 1023, g.bpos[b] >> 20};
        for (int d = 0; d < 3; d++) {
            blo[d] = std::min(blo[d], p[d]);
            bhi[d] = std::max(bhi[d], p[d]);
        }
    }
    g.box = m.nb > 0;
    const int bx = bhi[0] - blo[0] + 1, by = bhi[1] - blo[1] + 1;
    for (int k = 0; k < nz && g.box; k++)
        for (int j = 0; j < ny && g.box; j++)
            for (int i = 0; i < nx && g.box; i++) {
                const bool inside = i >= blo[0] && i <= bhi[0] && j >= blo[1] && j <= bhi[1] &&
                                    k >= blo[2] && k <= bhi[2];
                const int want =
                    inside ? (i - blo[0]) + bx * ((j - blo[1]) + by * (k - blo[2])) : -1;
                if (g.gblock[i + (size_t)nx * (j + (size_t)ny * k)] != want) g.box = 0;
            }
    for (int d = 0; d < 3; d++) {
        g.box_lo[d] = blo[d];
        g.box_hi[d] = bhi[d];
    }
    gi_out = std::move(g);
    return MSA_OK;
}

}  // namespace msa
