// Host-side topology of the loop solve: the fundamental loop basis, the sheet resistance of the
// RWG edges and the loop resistance Rl = C^T R C. No device code: these run on the host cores
// beside the device pipeline (the Python caller releases the interpreter during each call, so a
// batch of boards prepares its topologies on several host threads at once).
//
// Semantics follow the reference's MQS model (src/pcbqs/mqs.py):
//   spanning tree   a breadth-first forest over the copper branches from a central root per island
//                   (double BFS sweep), ports never in it, other branches joining islands
//                   (spanning_tree, mqs.py:48-129)
//   loops           one per non-tree branch: the branch plus the tree path back between its ends
//                   (mqs.py:592-637); an edge inside one graph node carries no loop (ideal terminals,
//                   mqs.py:603-609)
//   sheet R         the RWG Gram matrix at 1 / (sigma t), 3-point rule, in total edge currents
//                   (mqs.py:478-515)
// The loop basis is one basis of the cycle space (port impedances and branch currents do not depend on
// the choice). This one is fixed completely: BFS visits neighbours in ascending node order, the
// tree edge of a node pair is its lowest-numbered copper branch, loops are numbered in branch order
// and then reordered with the free loops along a Morton curve of their centroids (so consecutive
// 256-loop blocks are spatially compact) and the port loops last.
#include "pq_common.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <unordered_map>
#include <vector>

#pragma clang fp contract(off)

using pq::guard;

namespace {

using i64 = int64_t;

// An undirected simple graph in CSR form with ascending neighbours.
struct Graph {
    std::vector<i64> ptr, adj;
};

// BFS from `root` over g: visit order, and parent of each visited node (-1 for root / unvisited
// when `parent` is given). dist (when given) is the BFS depth, -1 where unreached.
void bfs(const Graph &g, i64 root, std::vector<i64> &order, std::vector<i64> *parent, std::vector<i64> *dist,
         std::vector<char> &seen) {
    order.clear();
    order.push_back(root);
    seen[root] = 1;
    if (dist) (*dist)[root] = 0;
    for (size_t h = 0; h < order.size(); ++h) {
        const i64 u = order[h];
        for (i64 k = g.ptr[u]; k < g.ptr[u + 1]; ++k) {
            const i64 v = g.adj[k];
            if (seen[v]) continue;
            seen[v] = 1;
            order.push_back(v);
            if (parent) (*parent)[v] = u;
            if (dist) (*dist)[v] = (*dist)[u] + 1;
        }
    }
    for (i64 v : order) seen[v] = 0;
}

// The member of one island minimising the larger of its distances to the two ends of a double BFS
// sweep (the first such member in ascending order); members ascending.
i64 central_root(const Graph &g, const std::vector<i64> &members, std::vector<char> &seen, std::vector<i64> &d1,
                 std::vector<i64> &d2, std::vector<i64> &order) {
    if (members.size() <= 2) return members[0];
    bfs(g, members[0], order, nullptr, nullptr, seen);
    const i64 end = order.back();
    bfs(g, end, order, nullptr, &d1, seen);
    i64 far = -1, best = -1;
    for (i64 v : order) best = std::max(best, d1[v]);
    for (i64 v : members)                       // the smallest node at the largest distance
        if (d1[v] == best) { far = v; break; }
    bfs(g, far, order, nullptr, &d2, seen);
    i64 root = members[0], ecc = -1;
    for (i64 v : members) {
        const i64 e = std::max(d1[v], d2[v]);
        if (ecc < 0 || e < ecc) { ecc = e; root = v; }
    }
    for (i64 v : order) { d1[v] = -1; d2[v] = -1; }
    return root;
}

// Morton (Z-order) code of 2-D points, 16 bits per axis, over the points' bounding square.
std::vector<i64> morton(const std::vector<double> &px, const std::vector<double> &py) {
    const size_t n = px.size();
    std::vector<i64> code(n, 0);
    if (!n) return code;
    double xmin = px[0], xmax = px[0], ymin = py[0], ymax = py[0];
    for (size_t i = 1; i < n; ++i) {
        xmin = std::min(xmin, px[i]); xmax = std::max(xmax, px[i]);
        ymin = std::min(ymin, py[i]); ymax = std::max(ymax, py[i]);
    }
    double span = std::max(xmax - xmin, ymax - ymin);
    if (!(span > 0)) span = 1.0;
    for (size_t i = 0; i < n; ++i) {
        const i64 qx = (i64)((px[i] - xmin) / span * 65535), qy = (i64)((py[i] - ymin) / span * 65535);
        i64 c = 0;
        for (int b = 0; b < 16; ++b) c |= ((qx >> b) & 1) << (2 * b) | ((qy >> b) & 1) << (2 * b + 1);
        code[i] = c;
    }
    return code;
}

}  // namespace

struct pq_loop_basis {
    i64 n_nodes = 0, nb = 0, nl = 0, n_free = 0;
    std::vector<uint8_t> tree;
    std::vector<i64> parent, upb, ups, depth;
    std::vector<i64> cot;                 // closing branch of each loop, final order
    std::vector<i64> colptr, rows;        // C in CSC, rows ascending within a column
    std::vector<double> vals, pos;        // C's +-1 entries; loop centroids (nl x 2)
    std::vector<i64> port_branch;         // closing branch of each port loop, in loop order
};

struct pq_csr {
    i64 nrows = 0, ncols = 0;
    std::vector<i64> ptr, idx;
    std::vector<double> val;
};

namespace {

void build_basis(pq_loop_basis &B, i64 n_nodes, i64 nb, const i64 *ends, const uint8_t *is_src, i64 n_copper,
                 i64 n_rwg, const double *bpos, bool ideal) {
    B.n_nodes = n_nodes;
    B.nb = nb;
    auto A = [&](i64 k) { return ends[2 * k]; };
    auto Bn = [&](i64 k) { return ends[2 * k + 1]; };
    for (i64 k = 0; k < nb; ++k)
        PQ_REQUIRE(A(k) >= 0 && A(k) < n_nodes && Bn(k) >= 0 && Bn(k) < n_nodes,
                   "pq_loop_basis_create: branch end outside the node range");
    // the tree edge of each node pair: its lowest-numbered copper branch
    std::unordered_map<i64, i64> pair_branch;
    pair_branch.reserve((size_t)n_copper * 2 + 16);
    std::vector<i64> deg(n_nodes + 1, 0);
    std::vector<std::pair<i64, i64>> edges;
    for (i64 k = 0; k < std::min(n_copper, nb); ++k) {
        if (is_src[k] || A(k) == Bn(k)) continue;
        const i64 lo = std::min(A(k), Bn(k)), hi = std::max(A(k), Bn(k));
        if (pair_branch.emplace(lo * n_nodes + hi, k).second) {
            edges.emplace_back(lo, hi);
            ++deg[lo]; ++deg[hi];
        }
    }
    Graph g;
    g.ptr.assign(n_nodes + 1, 0);
    for (i64 v = 0; v < n_nodes; ++v) g.ptr[v + 1] = g.ptr[v] + deg[v];
    g.adj.resize(g.ptr[n_nodes]);
    {
        std::vector<i64> fill(g.ptr.begin(), g.ptr.end() - 1);
        for (auto &e : edges) { g.adj[fill[e.first]++] = e.second; g.adj[fill[e.second]++] = e.first; }
        for (i64 v = 0; v < n_nodes; ++v) std::sort(g.adj.begin() + g.ptr[v], g.adj.begin() + g.ptr[v + 1]);
    }
    // islands, labelled in order of their smallest node
    std::vector<i64> label(n_nodes, -1), order;
    std::vector<char> seen(n_nodes, 0);
    i64 ncomp = 0;
    std::vector<std::vector<i64>> members;
    for (i64 v = 0; v < n_nodes; ++v) {
        if (label[v] >= 0) continue;
        bfs(g, v, order, nullptr, nullptr, seen);
        std::vector<i64> m(order);
        std::sort(m.begin(), m.end());
        for (i64 u : m) label[u] = ncomp;
        members.push_back(std::move(m));
        ++ncomp;
    }
    B.parent.assign(n_nodes, -1);
    B.upb.assign(n_nodes, -1);
    B.tree.assign(nb, 0);
    std::vector<i64> d1(n_nodes, -1), d2(n_nodes, -1);
    for (i64 c = 0; c < ncomp; ++c) {
        const i64 root = central_root(g, members[c], seen, d1, d2, order);
        bfs(g, root, order, &B.parent, nullptr, seen);
    }
    for (i64 v = 0; v < n_nodes; ++v) {
        const i64 p = B.parent[v];
        if (p < 0) continue;
        const i64 k = pair_branch.at(std::min(v, p) * n_nodes + std::max(v, p));
        B.upb[v] = k;
        B.tree[k] = 1;
    }
    // the remaining non-source branches join islands the copper leaves apart
    std::vector<i64> uf(ncomp);
    std::iota(uf.begin(), uf.end(), 0);
    auto find = [&](i64 x) {
        while (uf[x] != x) { uf[x] = uf[uf[x]]; x = uf[x]; }
        return x;
    };
    for (i64 k = n_copper; k < nb; ++k) {
        if (is_src[k] || A(k) == Bn(k)) continue;
        const i64 ra = find(label[A(k)]), rb = find(label[Bn(k)]);
        if (ra == rb) continue;
        uf[rb] = ra;
        B.tree[k] = 1;
        // hang b's island under a through branch k: reverse the parent chain from b to its root
        i64 prev = A(k), prev_branch = k, node = Bn(k);
        while (node >= 0) {
            const i64 nxt = B.parent[node], nxt_branch = B.upb[node];
            B.parent[node] = prev; B.upb[node] = prev_branch;
            prev = node; prev_branch = nxt_branch; node = nxt;
        }
    }
    std::vector<i64> comp(n_nodes);
    for (i64 v = 0; v < n_nodes; ++v) comp[v] = find(label[v]);
    // depth below the root and orientation of the parent branch
    B.depth.assign(n_nodes, -1);
    std::vector<i64> path;
    for (i64 v = 0; v < n_nodes; ++v) {
        i64 u = v;
        path.clear();
        while (B.depth[u] < 0 && B.parent[u] >= 0) {
            path.push_back(u);
            u = B.parent[u];
            PQ_REQUIRE((i64)path.size() <= n_nodes, "pq_loop_basis_create: the tree has a cycle");
        }
        if (B.depth[u] < 0) B.depth[u] = 0;                  // a root
        for (auto it = path.rbegin(); it != path.rend(); ++it) B.depth[*it] = B.depth[B.parent[*it]] + 1;
    }
    B.ups.assign(n_nodes, 0);
    for (i64 v = 0; v < n_nodes; ++v)
        if (B.parent[v] >= 0) B.ups[v] = A(B.upb[v]) == B.parent[v] ? 1 : -1;
    // loops: every non-tree branch whose ends are joined through the tree
    std::vector<i64> cot;
    for (i64 k = 0; k < nb; ++k) {
        if (B.tree[k]) continue;
        if (!(A(k) != Bn(k) || k >= n_rwg || !ideal)) continue;
        if (comp[A(k)] != comp[Bn(k)]) continue;
        cot.push_back(k);
    }
    const i64 nl = (i64)cot.size();
    // each loop's branches: the closing branch (+1), the tree path from b up to the common
    // ancestor (-ups) and from a up to it (+ups)
    std::vector<i64> cptr(nl + 1, 0), crow;
    std::vector<double> cval;
    std::vector<std::pair<i64, double>> col;
    for (i64 l = 0; l < nl; ++l) {
        const i64 k = cot[l];
        col.clear();
        col.emplace_back(k, 1.0);
        i64 u = Bn(k), v = A(k);
        while (u != v) {
            const i64 du = B.depth[u], dv = B.depth[v];
            if (du >= dv) { col.emplace_back(B.upb[u], (double)-B.ups[u]); u = B.parent[u]; }
            if (dv >= du) { col.emplace_back(B.upb[v], (double)B.ups[v]); v = B.parent[v]; }
        }
        std::sort(col.begin(), col.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
        for (size_t i = 0; i + 1 < col.size(); ++i)
            PQ_REQUIRE(col[i].first != col[i + 1].first, "pq_loop_basis_create: a branch repeats in a loop");
        for (auto &e : col) { crow.push_back(e.first); cval.push_back(e.second); }
        cptr[l + 1] = (i64)crow.size();
    }
    // centroids: the mean position of each loop's branches (ascending branch order)
    std::vector<double> px(nl), py(nl);
    for (i64 l = 0; l < nl; ++l) {
        double sx = 0, sy = 0;
        for (i64 j = cptr[l]; j < cptr[l + 1]; ++j) { sx += bpos[2 * crow[j]]; sy += bpos[2 * crow[j] + 1]; }
        const double cnt = (double)std::max<i64>(cptr[l + 1] - cptr[l], 1);
        px[l] = sx / cnt; py[l] = sy / cnt;
    }
    // final order: free loops along the Morton curve (stable), then the port loops
    std::vector<i64> free_cols, src_cols;
    for (i64 l = 0; l < nl; ++l) (is_src[cot[l]] ? src_cols : free_cols).push_back(l);
    std::vector<double> fx(free_cols.size()), fy(free_cols.size());
    for (size_t i = 0; i < free_cols.size(); ++i) { fx[i] = px[free_cols[i]]; fy[i] = py[free_cols[i]]; }
    const std::vector<i64> code = morton(fx, fy);
    std::vector<i64> rank(free_cols.size());
    std::iota(rank.begin(), rank.end(), 0);
    std::stable_sort(rank.begin(), rank.end(), [&](i64 x, i64 y) { return code[x] < code[y]; });
    std::vector<i64> perm;
    perm.reserve(nl);
    for (i64 r : rank) perm.push_back(free_cols[r]);
    for (i64 l : src_cols) perm.push_back(l);
    B.nl = nl;
    B.n_free = (i64)free_cols.size();
    B.colptr.assign(nl + 1, 0);
    B.rows.clear(); B.vals.clear();
    B.rows.reserve(crow.size()); B.vals.reserve(cval.size());
    B.cot.resize(nl);
    B.pos.resize(2 * nl);
    for (i64 j = 0; j < nl; ++j) {
        const i64 l = perm[j];
        for (i64 t = cptr[l]; t < cptr[l + 1]; ++t) { B.rows.push_back(crow[t]); B.vals.push_back(cval[t]); }
        B.colptr[j + 1] = (i64)B.rows.size();
        B.cot[j] = cot[l];
        B.pos[2 * j] = px[l]; B.pos[2 * j + 1] = py[l];
    }
    B.port_branch.clear();
    for (i64 l : src_cols) B.port_branch.push_back(cot[l]);
}

}  // namespace

extern "C" {

int pq_loop_basis_create(int64_t n_nodes, int64_t n_branches, const int64_t *ends, const uint8_t *is_source,
                         int64_t n_copper, int64_t n_rwg, const double *branch_pos, int ideal_terminals,
                         pq_loop_basis **out) {
    return guard([&] {
        PQ_REQUIRE(out && (n_branches == 0 || (ends && is_source && branch_pos)), "pq_loop_basis_create: null argument");
        PQ_REQUIRE(n_nodes > 0 && n_branches >= 0 && n_copper >= 0 && n_rwg >= 0,
                   "pq_loop_basis_create: negative size");
        auto *b = new pq_loop_basis();
        try {
            build_basis(*b, n_nodes, n_branches, ends, is_source, n_copper, n_rwg, branch_pos, ideal_terminals != 0);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int pq_loop_basis_sizes(const pq_loop_basis *b, int64_t info[4]) {
    return guard([&] {
        PQ_REQUIRE(b && info, "pq_loop_basis_sizes: null argument");
        info[0] = b->nl;
        info[1] = (int64_t)b->rows.size();
        info[2] = b->n_free;
        info[3] = (int64_t)b->port_branch.size();
    });
}

int pq_loop_basis_copy(const pq_loop_basis *b, uint8_t *tree, int64_t *parent, int64_t *upb, int64_t *ups,
                       int64_t *depth, int64_t *cotree, int64_t *colptr, int64_t *rows, double *vals,
                       double *positions, int64_t *port_branch) {
    return guard([&] {
        PQ_REQUIRE(b, "pq_loop_basis_copy: null handle");
        auto put = [](auto *dst, const auto &src) {
            if (dst && !src.empty()) std::copy(src.begin(), src.end(), dst);
        };
        put(tree, b->tree); put(parent, b->parent); put(upb, b->upb); put(ups, b->ups); put(depth, b->depth);
        put(cotree, b->cot); put(colptr, b->colptr); put(rows, b->rows); put(vals, b->vals);
        put(positions, b->pos); put(port_branch, b->port_branch);
    });
}

int pq_loop_basis_destroy(pq_loop_basis *b) {
    delete b;
    return PQ_OK;
}

// Sheet resistance of the RWG edges in total edge currents: per triangle t and each pair (m, n) of
// its edges, alpha_m alpha_n (A_t / 3) sum_q (P_q - v_m).(P_q - v_n) / (sigma thickness_t) / (l_m l_n),
// with alpha = +-l / (2 A) (T+ / T-), P_q the 3-point rule (mqs.py:478-515). Summed over the
// triangles an edge pair shares.
int pq_sheet_resistance(int64_t ne, const int64_t *tp, const int64_t *tm, const double *length, const double *vp,
                        const double *vm, int64_t nt, const double *area, const double *points,
                        const double *thickness, double sigma, pq_csr **out) {
    return guard([&] {
        PQ_REQUIRE(out && (ne == 0 || (tp && tm && length && vp && vm && area && points && thickness)),
                   "pq_sheet_resistance: null argument");
        PQ_REQUIRE(ne >= 0 && nt >= 0 && sigma > 0, "pq_sheet_resistance: invalid size or conductivity");
        // each triangle's edge terms: T+ terms by ascending edge, then T- terms
        std::vector<i64> cnt(nt + 1, 0);
        for (i64 e = 0; e < ne; ++e) {
            PQ_REQUIRE(tp[e] >= 0 && tp[e] < nt && tm[e] >= 0 && tm[e] < nt, "pq_sheet_resistance: triangle index");
            ++cnt[tp[e] + 1]; ++cnt[tm[e] + 1];
        }
        for (i64 t = 0; t < nt; ++t) cnt[t + 1] += cnt[t];
        std::vector<i64> term_edge(2 * ne), fill(cnt.begin(), cnt.end() - 1);
        std::vector<double> term_alpha(2 * ne);
        std::vector<const double *> term_vert(2 * ne);
        for (i64 e = 0; e < ne; ++e) {
            const i64 s = fill[tp[e]]++;
            term_edge[s] = e; term_alpha[s] = length[e] / (2 * area[tp[e]]); term_vert[s] = vp + 2 * e;
        }
        for (i64 e = 0; e < ne; ++e) {
            const i64 s = fill[tm[e]]++;
            term_edge[s] = e; term_alpha[s] = -length[e] / (2 * area[tm[e]]); term_vert[s] = vm + 2 * e;
        }
        std::vector<std::vector<std::pair<i64, double>>> rowv(ne);
        for (i64 t = 0; t < nt; ++t) {
            const i64 s0 = cnt[t], n = cnt[t + 1] - s0;
            PQ_REQUIRE(n <= 3, "pq_sheet_resistance: a triangle carries more than three edges");
            const double w = area[t] / 3.0 / (sigma * thickness[t]);
            const double *P = points + 6 * t;
            for (i64 i = 0; i < n; ++i)
                for (i64 j = 0; j < n; ++j) {
                    const i64 a = s0 + i, b = s0 + j;
                    const double *va = term_vert[a], *vb = term_vert[b];
                    double d = 0;
                    for (int q = 0; q < 3; ++q)
                        for (int c = 0; c < 2; ++c) d += (P[2 * q + c] - va[c]) * (P[2 * q + c] - vb[c]);
                    const double v = term_alpha[a] * term_alpha[b] * w * d /
                                     (length[term_edge[a]] * length[term_edge[b]]);
                    rowv[term_edge[a]].emplace_back(term_edge[b], v);
                }
        }
        auto *m = new pq_csr();
        m->nrows = m->ncols = ne;
        m->ptr.assign(ne + 1, 0);
        for (i64 e = 0; e < ne; ++e) {
            auto &r = rowv[e];
            std::stable_sort(r.begin(), r.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
            for (size_t i = 0; i < r.size(); ++i) {
                if (i > 0 && r[i].first == r[i - 1].first) { m->val.back() += r[i].second; continue; }
                m->idx.push_back(r[i].first);
                m->val.push_back(r[i].second);
            }
            m->ptr[e + 1] = (i64)m->idx.size();
        }
        *out = m;
    });
}

// Rl = C^T R C of a loop basis with R = blockdiag(Re, diag(rv), 0): Re the (ne x ne) sheet
// resistance (CSR), rv the via barrels' resistances (branches ne .. ne + nv), the remaining
// branches (ports) without resistance. Column j's entries on and above the diagonal are summed
// over the branches of loop j (ascending), their R rows and the loops through each branch, and
// mirrored below it, so Rl is exactly symmetric. Rows and columns ascending; entries that cancel
// exactly are not stored.
int pq_loop_resistance(const pq_loop_basis *b, const pq_csr *re, int64_t nv, const double *rv, pq_csr **out) {
    return guard([&] {
        PQ_REQUIRE(b && re && out && (nv == 0 || rv), "pq_loop_resistance: null argument");
        const i64 ne = re->nrows, nb = b->nb, nl = b->nl;
        PQ_REQUIRE(ne + nv <= nb, "pq_loop_resistance: more resistive branches than branches");
        // C by rows (branch -> the loops through it, ascending, with signs)
        std::vector<i64> rptr(nb + 1, 0);
        for (i64 r : b->rows) ++rptr[r + 1];
        for (i64 k = 0; k < nb; ++k) rptr[k + 1] += rptr[k];
        std::vector<i64> rloop(b->rows.size()), fill(rptr.begin(), rptr.end() - 1);
        std::vector<double> rsign(b->rows.size());
        for (i64 l = 0; l < nl; ++l)
            for (i64 t = b->colptr[l]; t < b->colptr[l + 1]; ++t) {
                const i64 s = fill[b->rows[t]]++;
                rloop[s] = l; rsign[s] = b->vals[t];
            }
        // the upper triangle by columns: (i, j, value) with i <= j
        std::vector<i64> uptr(nl + 1, 0), urow;
        std::vector<double> uval;
        std::vector<double> acc(nl, 0.0);
        std::vector<i64> stamp(nl, -1), touched;
        for (i64 j = 0; j < nl; ++j) {
            touched.clear();
            auto add = [&](i64 k, double y) {
                for (i64 s = rptr[k]; s < rptr[k + 1]; ++s) {
                    const i64 i = rloop[s];
                    if (i > j) break;
                    if (stamp[i] != j) { stamp[i] = j; acc[i] = 0.0; touched.push_back(i); }
                    acc[i] += rsign[s] * y;
                }
            };
            for (i64 t = b->colptr[j]; t < b->colptr[j + 1]; ++t) {
                const i64 k2 = b->rows[t];
                const double c2 = b->vals[t];
                if (k2 < ne) {
                    for (i64 s = re->ptr[k2]; s < re->ptr[k2 + 1]; ++s) add(re->idx[s], re->val[s] * c2);
                } else if (k2 < ne + nv) {
                    add(k2, rv[k2 - ne] * c2);
                }
            }
            std::sort(touched.begin(), touched.end());
            for (i64 i : touched)
                if (acc[i] != 0.0) { urow.push_back(i); uval.push_back(acc[i]); }
            uptr[j + 1] = (i64)urow.size();
        }
        // full CSR: row i holds (i, j) for j >= i from column j's upper part and (i, j) for j < i
        // from column i's upper part; both come out ascending in j
        auto *m = new pq_csr();
        try {
            m->nrows = m->ncols = nl;
            m->ptr.assign(nl + 1, 0);
            for (i64 j = 0; j < nl; ++j)
                for (i64 t = uptr[j]; t < uptr[j + 1]; ++t) {
                    ++m->ptr[urow[t] + 1];                       // (i, j) in row i
                    if (urow[t] != j) ++m->ptr[j + 1];           // (j, i) in row j
                }
            for (i64 i = 0; i < nl; ++i) m->ptr[i + 1] += m->ptr[i];
            m->idx.resize(m->ptr[nl]);
            m->val.resize(m->ptr[nl]);
            std::vector<i64> pos(m->ptr.begin(), m->ptr.end() - 1);
            // row j's entries below the diagonal (columns i < j) first, in ascending i ...
            for (i64 j = 0; j < nl; ++j)
                for (i64 t = uptr[j]; t < uptr[j + 1]; ++t)
                    if (urow[t] != j) { m->idx[pos[j]] = urow[t]; m->val[pos[j]] = uval[t]; ++pos[j]; }
            // ... then the diagonal and above (columns j >= i), visited in ascending j
            for (i64 j = 0; j < nl; ++j)
                for (i64 t = uptr[j]; t < uptr[j + 1]; ++t) {
                    const i64 i = urow[t];
                    m->idx[pos[i]] = j; m->val[pos[i]] = uval[t]; ++pos[i];
                }
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int pq_csr_sizes(const pq_csr *m, int64_t info[3]) {
    return guard([&] {
        PQ_REQUIRE(m && info, "pq_csr_sizes: null argument");
        info[0] = m->nrows; info[1] = m->ncols; info[2] = (int64_t)m->idx.size();
    });
}

int pq_csr_copy(const pq_csr *m, int64_t *ptr, int64_t *idx, double *val) {
    return guard([&] {
        PQ_REQUIRE(m && ptr && (m->idx.empty() || (idx && val)), "pq_csr_copy: null argument");
        std::copy(m->ptr.begin(), m->ptr.end(), ptr);
        std::copy(m->idx.begin(), m->idx.end(), idx);
        std::copy(m->val.begin(), m->val.end(), val);
    });
}

int pq_csr_destroy(pq_csr *m) {
    delete m;
    return PQ_OK;
}

}  // extern "C"
