"""Host-side topology of the loop solve: branches, spanning tree, loop basis, sparse R, vias.

What is built follows the reference's MQS model (src/pcbqs/mqs.py):
  branches     RWG edges (node of T+ -> node of T-), via segments, ports (mqs.py:404-416)
  tree         a breadth-first forest over the copper from a central root per island (double
               sweep), ports never in it, parts joining islands (spanning_tree, mqs.py:48-129)
  loops        one per non-tree branch: the branch plus the tree path back (mqs.py:592-637); an
               edge inside one graph node carries no loop (ideal_terminals, mqs.py:603-609)
  R            the RWG Gram matrix at 1/(sigma t), 3-point rule, in total edge currents
               (mqs.py:478-515); via barrels' DC resistance (mqs.py:132-137)
  L_vias       parallel-filament partial inductances (Grover; mqs.py:40-45,140-153)
The loop basis is any basis of the cycle space: port impedances and branch currents do not
depend on the choice. LoopModel builds it with numpy and scipy (loop_basis, sheet_resistance and
scipy's C^T R C). native=True builds the same tree and loops (bitwise) and the resistances within
rounding in the library's host code instead (pq_loop_basis_create, pq_sheet_resistance,
pq_loop_resistance), 1.5-2x faster per board and without holding the interpreter; it is opt-in: a
sweep whose worker threads ran it while other threads made their first device calls ended in a
host segmentation fault inside it whose cause is not known.

The loops are then ordered for the device: free loops first, sorted along a Morton curve of
their centroids (so that consecutive blocks of loops are spatially compact: block-Jacobi
blocks), the port loops last.
"""
import ctypes
from dataclasses import dataclass

import numpy as np
from scipy import sparse
from scipy.sparse import csgraph

from .geometry import BAR3, MU0, RWG


def filament_mutual(z0a, z1a, z0b, z1b, s):
    """Partial mutual inductance of parallel z filaments s apart (Grover)."""
    g = lambda u: u * np.arcsinh(u / s) - np.sqrt(u * u + s * s)
    return MU0 / (4 * np.pi) * (g(z1a - z0b) + g(z0a - z1b) - g(z1a - z1b) - g(z0a - z0b))


def via_matrices(vseg, sigma, plating):
    """(R diag (nv,), L (nv, nv)) of via segments (x, y, r, z0, z1)."""
    v = np.asarray(vseg, float).reshape(-1, 5)
    if not len(v):
        return np.zeros(0), np.zeros((0, 0))
    x, y, rr, z0, z1 = v.T
    ri = np.maximum(rr - plating, 0.0)
    R = (z1 - z0) / (sigma * np.pi * (rr ** 2 - ri ** 2))
    s = np.hypot(x[:, None] - x[None], y[:, None] - y[None])
    s = np.where(s < 1e-9, np.maximum(rr[:, None], rr[None]), s)
    return R, filament_mutual(z0[:, None], z1[:, None], z0[None], z1[None], s)


def sheet_resistance(rwg, t_of, sigma):
    """Sparse (Ne, Ne) R in total edge currents: per triangle t the Gram entries
    alpha_m alpha_n (A/3) sum_q (P_q - v_m).(P_q - v_n) / (sigma t), scaled by 1/(l_m l_n)."""
    nodes, tris = rwg.nodes, rwg.tris
    Pq = np.einsum('qk,tkc->tqc', BAR3, nodes[tris])                       # (nt, 3, 2)
    tri = np.r_[rwg.tp, rwg.tm]
    edge = np.r_[np.arange(rwg.Ne), np.arange(rwg.Ne)]
    alpha = np.r_[rwg.l / (2 * rwg.area[rwg.tp]), -rwg.l / (2 * rwg.area[rwg.tm])]
    vert = np.vstack([rwg.vp, rwg.vm])
    order = np.argsort(tri, kind='stable')
    tri, edge, alpha, vert = tri[order], edge[order], alpha[order], vert[order]
    starts = np.r_[0, np.flatnonzero(np.diff(tri)) + 1]
    counts = np.diff(np.r_[starts, len(tri)])
    Dq = Pq[tri] - vert[:, None]                                            # (terms, 3, 2): P_q - v
    wt = rwg.area[tri] / 3.0 / (sigma * t_of[tri])
    lt = rwg.l[edge]
    rows, cols, vals = [], [], []
    for i in range(3):
        for j in range(3):
            sel = counts > max(i, j)
            a, b = starts[sel] + i, starts[sel] + j
            d = (Dq[a] * Dq[b]).sum((1, 2))
            rows.append(edge[a]); cols.append(edge[b])
            vals.append(alpha[a] * alpha[b] * wt[a] * d / (lt[a] * lt[b]))
    return sparse.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(rwg.Ne, rwg.Ne)).tocsr()


def _central_root(adj, members):
    """The node minimising the larger of its distances to the two ends of a double BFS sweep."""
    if len(members) <= 2:
        return members[0]
    order = csgraph.breadth_first_order(adj, members[0], directed=False, return_predecessors=False)
    d1 = csgraph.shortest_path(adj, unweighted=True, directed=False, indices=order[-1])
    far = np.argmax(np.where(np.isfinite(d1), d1, -1))
    d2 = csgraph.shortest_path(adj, unweighted=True, directed=False, indices=far)
    ecc = np.maximum(d1[members], d2[members])
    return members[int(np.argmin(ecc))]


@dataclass
class LoopBasis:
    C: sparse.csc_matrix          # (nb, nloops) +-1, columns in device order
    perm_branches: np.ndarray     # unused placeholder for symmetry with C's rows (identity)
    n_free: int                   # free loops come first; the ports' loops follow
    port_cols: np.ndarray         # column of each port's loop
    tree: np.ndarray              # (nb,) bool
    positions: np.ndarray         # (nloops, 2) loop centroids
    cotree: np.ndarray            # (nloops,) the non-tree branch that closes each loop (coefficient +1)
    parent: np.ndarray = None     # (n_nodes,) tree parent of each node, -1 at a root
    upb: np.ndarray = None        # (n_nodes,) branch to the parent, -1 at a root
    ups: np.ndarray = None        # (n_nodes,) +1 where that branch runs from the parent to the node
    depth: np.ndarray = None      # (n_nodes,) depth below the root
    cot_ends: np.ndarray = None   # (nloops, 2) the closing branch's ends (a, b): loop = a -> b, tree b -> a


def loop_basis(n_nodes, ends, is_src, n_copper, n_rwg, branch_pos, ideal_terminals=True):
    """The fundamental loop basis of the branch graph (see the module docstring)."""
    ends = np.asarray(ends, np.int64).reshape(-1, 2)
    nb = len(ends)
    a, b = ends[:, 0], ends[:, 1]
    copper = (np.arange(nb) < n_copper) & ~is_src & (a != b)
    ks = np.flatnonzero(copper)
    # one branch per node pair carries the tree edge (the lowest-numbered)
    lo, hi = np.minimum(a[ks], b[ks]), np.maximum(a[ks], b[ks])
    key = lo * n_nodes + hi
    ukey, first = np.unique(key, return_index=True)          # node pairs ascending, each one's first branch
    pair_branch = ks[first]
    adj = sparse.coo_matrix((np.ones(len(first)), (lo[first], hi[first])), shape=(n_nodes, n_nodes)).tocsr()
    adj = (adj + adj.T).tocsr()
    ncomp, label = csgraph.connected_components(adj, directed=False)
    tree = np.zeros(nb, bool)
    parent = np.full(n_nodes, -1)
    by_comp = np.argsort(label, kind='stable')
    bounds = np.r_[0, np.flatnonzero(np.diff(label[by_comp])) + 1, n_nodes]
    for c in range(ncomp):
        members = by_comp[bounds[c]:bounds[c + 1]]
        root = _central_root(adj, members)
        order, pred = csgraph.breadth_first_order(adj, root, directed=False, return_predecessors=True)
        child = order[1:]
        parent[child] = pred[child]
    has = parent >= 0
    child = np.flatnonzero(has)
    pk = np.minimum(child, parent[child]) * n_nodes + np.maximum(child, parent[child])
    upb = np.full(n_nodes, -1)
    upb[child] = pair_branch[np.searchsorted(ukey, pk)]        # every tree edge is a copper node pair
    tree[upb[child]] = True
    # parts (non-source branches beyond the copper) join islands the copper leaves apart
    comp = label.copy()
    uf = np.arange(ncomp)

    def find(x):
        while uf[x] != x:
            uf[x] = uf[uf[x]]
            x = uf[x]
        return x
    for k in range(n_copper, nb):
        if not is_src[k] and a[k] != b[k]:
            ra, rb = find(comp[a[k]]), find(comp[b[k]])
            if ra != rb:
                uf[rb] = ra
                tree[k] = True
                # re-root b's island under a through branch k
                _attach(parent, upb, a[k], b[k], k, adj)
    comp = np.array([find(x) for x in range(ncomp)], dtype=np.int64)[label] if n_nodes else comp
    # depth and orientation of the parent branch
    depth = np.full(n_nodes, -1)
    roots = np.flatnonzero(parent < 0)
    depth[roots] = 0
    pending = np.flatnonzero(parent >= 0)
    while len(pending):
        ready = pending[depth[parent[pending]] >= 0]
        if not len(ready):
            raise RuntimeError('loop basis: the tree has a cycle')
        depth[ready] = depth[parent[ready]] + 1
        pending = pending[depth[pending] < 0]
    ups = np.zeros(n_nodes, np.int64)
    ch = np.flatnonzero(parent >= 0)
    ups[ch] = np.where(a[upb[ch]] == parent[ch], 1, -1)
    # loops: every non-tree branch whose ends are joined through the tree
    cot = np.flatnonzero(~tree & ((a != b) | (np.arange(nb) >= n_rwg) | (not ideal_terminals)))
    cot = cot[comp[a[cot]] == comp[b[cot]]]
    if not ideal_terminals:
        pass
    nl = len(cot)
    col = np.arange(nl)
    rows, cols, vals = [cot], [col], [np.ones(nl)]
    # climb both ends to the common ancestor, the deeper end (both at equal depth) one step at a
    # time; loops leave the working set once their ends meet
    u, v = b[cot], a[cot]
    open_ = u != v
    idx, u, v = col[open_], u[open_], v[open_]
    while len(idx):
        du, dv = depth[u], depth[v]
        mu, mv = du >= dv, dv >= du
        uu, vv = u[mu], v[mv]
        rows += [upb[uu], upb[vv]]
        cols += [idx[mu], idx[mv]]
        vals += [-ups[uu].astype(float), ups[vv].astype(float)]
        u[mu] = parent[uu]
        v[mv] = parent[vv]
        open_ = u != v
        idx, u, v = idx[open_], u[open_], v[open_]
    rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    C = sparse.coo_matrix((vals, (rows, cols)), shape=(nb, nl)).tocsc()
    C.sum_duplicates()
    # loop centroids and order: free loops along a Morton curve, then the ports
    absC = abs(C)
    pos = (absC.T @ branch_pos) / np.maximum(np.asarray(absC.sum(0)).ravel(), 1)[:, None]
    src_cols = np.flatnonzero(is_src[cot])
    free_cols = np.flatnonzero(~is_src[cot])
    order_free = free_cols[np.argsort(morton(pos[free_cols]), kind='stable')]
    order = np.r_[order_free, src_cols]
    C = C[:, order].tocsc()
    port_branch = cot[src_cols]
    return LoopBasis(C, None, len(free_cols), np.arange(len(free_cols), nl), tree, pos[order], cot[order],
                     parent, upb, ups, depth, ends[cot[order]]), port_branch


def _attach(parent, upb, a, b, k, adj):
    """Hang the island of b under a through branch k: reverse the parent chain from b to its root."""
    prev, prev_branch, node = a, k, b
    while node >= 0:
        nxt, nxt_branch = parent[node], upb[node]
        parent[node], upb[node] = prev, prev_branch
        prev, prev_branch, node = node, nxt_branch, nxt


def morton(p):
    """Morton (Z-order) code of 2-D points, 16 bits per axis."""
    p = np.asarray(p, float)
    if not len(p):
        return np.zeros(0, np.int64)
    span = np.ptp(p, 0).max()
    q = ((p - p.min(0)) / (span if span > 0 else 1.0) * 65535).astype(np.int64)
    code = np.zeros(len(p), np.int64)
    for i in range(16):
        code |= ((q[:, 0] >> i) & 1) << (2 * i) | ((q[:, 1] >> i) & 1) << (2 * i + 1)
    return code


def _csr_from(handle):
    """A scipy CSR copy of a pq_csr handle."""
    from ._lib import call, ptr
    info = np.zeros(3, np.int64)
    call('pq_csr_sizes', handle, ptr(info))
    rows, cols, nnz = (int(v) for v in info)
    indptr, indices, data = np.empty(rows + 1, np.int64), np.empty(nnz, np.int64), np.empty(nnz)
    call('pq_csr_copy', handle, ptr(indptr), ptr(indices), ptr(data))
    return sparse.csr_matrix((data, indices, indptr), shape=(rows, cols))


def native_topology(n_nodes, ends, is_src, n_copper, n_rwg, branch_pos, ideal_terminals, rwg, t_of, sigma, Rv):
    """(LoopBasis, port branches, Re, Rl) from the library's host code: the loop basis as loop_basis
    builds it, the sheet resistance as sheet_resistance, Rl = C^T blockdiag(Re, diag(Rv), 0) C."""
    from ._lib import call, ptr
    ends = np.ascontiguousarray(np.asarray(ends, np.int64).reshape(-1, 2))
    src = np.ascontiguousarray(is_src, np.uint8)
    bpos = np.ascontiguousarray(branch_pos, np.float64)
    nb = len(ends)
    hb, hre, hrl = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    try:
        call('pq_loop_basis_create', int(n_nodes), nb, ptr(ends), ptr(src), int(n_copper), int(n_rwg), ptr(bpos),
             int(bool(ideal_terminals)), ctypes.byref(hb))
        info = np.zeros(4, np.int64)
        call('pq_loop_basis_sizes', hb, ptr(info))
        nl, nnz, n_free, n_port = (int(v) for v in info)
        tree = np.empty(nb, np.uint8)
        parent, upb, ups, depth = (np.empty(int(n_nodes), np.int64) for _ in range(4))
        cot, colptr, rows = np.empty(nl, np.int64), np.empty(nl + 1, np.int64), np.empty(nnz, np.int64)
        vals, pos, port_branch = np.empty(nnz), np.empty((nl, 2)), np.empty(n_port, np.int64)
        call('pq_loop_basis_copy', hb, ptr(tree), ptr(parent), ptr(upb), ptr(ups), ptr(depth), ptr(cot), ptr(colptr),
             ptr(rows), ptr(vals), ptr(pos), ptr(port_branch))
        # the sheet resistance: the 3-point rule points in numpy (the same products as sheet_resistance)
        points = np.ascontiguousarray(np.einsum('qk,tkc->tqc', BAR3, rwg.nodes[rwg.tris]))
        tp, tm = np.ascontiguousarray(rwg.tp, np.int64), np.ascontiguousarray(rwg.tm, np.int64)
        vp, vm = np.ascontiguousarray(rwg.vp, np.float64), np.ascontiguousarray(rwg.vm, np.float64)
        length, area = np.ascontiguousarray(rwg.l, np.float64), np.ascontiguousarray(rwg.area, np.float64)
        thick = np.ascontiguousarray(t_of, np.float64)
        call('pq_sheet_resistance', int(rwg.Ne), ptr(tp), ptr(tm), ptr(length), ptr(vp), ptr(vm), len(area), ptr(area),
             ptr(points), ptr(thick), float(sigma), ctypes.byref(hre))
        Re = _csr_from(hre)
        rv = np.ascontiguousarray(Rv, np.float64)
        call('pq_loop_resistance', hb, hre, len(rv), ptr(rv), ctypes.byref(hrl))
        Rl = _csr_from(hrl)
    finally:
        for h, name in ((hrl, 'pq_csr_destroy'), (hre, 'pq_csr_destroy'), (hb, 'pq_loop_basis_destroy')):
            if h.value:
                call(name, h)
    C = sparse.csc_matrix((vals, rows, colptr), shape=(nb, nl))
    basis = LoopBasis(C, None, n_free, np.arange(n_free, nl), tree.astype(bool), pos, cot, parent, upb, ups, depth,
                      ends[cot])
    return basis, port_branch, Re, Rl


class LoopModel:
    """A meshed board (pcbqs_amd.board.Board) as the loop solve sees it. native=True builds the
    topology and resistances with the library's host code (opt-in, see the module docstring)."""

    def __init__(self, board, ideal_terminals=True, native=False):
        self.board = board
        self.rwg = RWG(board.nodes, board.tris)
        r = self.rwg
        node_of = board.node_of
        self.ne, self.nv, self.np = r.Ne, len(board.vseg), len(board.ports)
        self.br = np.r_[np.c_[node_of[r.tp], node_of[r.tm]].reshape(-1, 2),
                        np.asarray(board.via_nodes, np.int64).reshape(-1, 2)]
        ports = np.array([(board.terminals[a], board.terminals[b]) for _, a, b in board.ports], np.int64).reshape(-1, 2)
        self.ends = np.r_[self.br, ports]
        self.nb = len(self.ends)
        self.is_src = np.r_[np.zeros(len(self.br), bool), np.ones(self.np, bool)]
        cen = board.nodes[board.tris].mean(1)
        bpos = np.zeros((self.nb, 2))
        bpos[:self.ne] = (cen[r.tp] + cen[r.tm]) / 2
        for k, (x, y, _, _, _) in enumerate(board.vseg):
            bpos[self.ne + k] = (x, y)
        for k, (_, a, b) in enumerate(board.ports):
            bpos[self.ne + self.nv + k] = bpos[:self.ne].mean(0) if self.ne else 0
        self.Rv, self.Lv = via_matrices(board.vseg, board.sigma, board.plating)
        if native:
            self.basis, port_branch, self.Re, self.Rl = native_topology(
                board.n_nodes, self.ends, self.is_src, len(self.br), self.ne, bpos, ideal_terminals, r, board.t_of,
                board.sigma, self.Rv)
        else:
            self.basis, port_branch = loop_basis(board.n_nodes, self.ends, self.is_src, len(self.br), self.ne, bpos,
                                                 ideal_terminals)
            self.Re = sheet_resistance(r, board.t_of, board.sigma)
            C = self.basis.C
            self.Rl = (C.T @ self.R @ C).tocsr()
        # the ports' loops in port order
        where = {int(b): k for k, b in enumerate(port_branch)}
        self.port_cols = np.array([self.basis.port_cols[where[self.ne + self.nv + k]] for k in range(self.np)],
                                  dtype=np.int64)

    @property
    def R(self):
        """The branch resistance blockdiag(Re, diag(Rv), 0) (sparse CSR)."""
        return sparse.block_diag((self.Re, sparse.diags(self.Rv), sparse.csr_matrix((self.np, self.np))),
                                 format='csr')

    @property
    def n_loops(self):
        return self.basis.C.shape[1]

    @property
    def n_free(self):
        return self.basis.n_free
