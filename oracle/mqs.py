"""Oracle for the loop/port solve (test infrastructure only; see oracle/__init__.py).

Restates, in numpy/scipy:
  RWG                  src/pcbqs/_numerics/base.py:10-41 (edge order: insertion order of a dict over
                       the triangles' edges)
  combine, qform       src/pcbqs/_numerics/base.py:51-60
  sheet_resistance     src/pcbqs/mqs.py:478-515 (conductor_loss='dc' branch)
  via_resistance       src/pcbqs/mqs.py:132-137
  filament_mutual      src/pcbqs/mqs.py:40-45
  via_inductance       src/pcbqs/mqs.py:140-153
  spanning_tree        src/pcbqs/mqs.py:48-129
  loops                src/pcbqs/mqs.py:592-637
  direct_ports         src/pcbqs/mqs_ports.py:52-69
  rl_equivalent        src/pcbqs/mqs_ports.py:72-86
and builds the copper model of the reference's copper-reference.npz fixture (trace_board of
tests/test_mqs.py:14-24 with the options of tests/test_standalone.py:18-31) from its stored mesh.

Pinned by: tests/test_copper_reference.py reproduces the fixture's port impedance (rtol 1e-11),
branch currents (rtol 1e-10) and field (rtol 1e-10) from the stored nodes and triangles.
"""
import collections

import numpy as np
from scipy import sparse
from scipy.linalg import lu_factor, lu_solve

from . import fill as F

MU0 = 4e-7 * np.pi
SIGMA_CU = 5.8e7


class RWG:
    def __init__(self, nodes, tris):
        self.nodes, self.tris = nodes, tris
        V = nodes[tris]
        e1, e2 = V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]
        self.area = 0.5 * (e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
        assert (self.area > 0).all(), 'triangles must be counter-clockwise'
        self.cent = V.mean(1)
        seen = {}
        for t, (a, b, c) in enumerate(tris):
            for i, (u, v) in enumerate(((a, b), (b, c), (c, a))):
                seen.setdefault((min(u, v), max(u, v)), []).append((t, i))
        E = [(k, v) for k, v in seen.items() if len(v) == 2]
        self.Ne = Ne = len(E)
        u, v = np.array([k for k, _ in E], int).reshape(-1, 2).T
        (t1, i1), (t2, i2) = np.array([s for _, s in E], int).reshape(-1, 2, 2).transpose(1, 2, 0)
        self.tp, self.tm = t1, t2
        self.vp, self.vm = nodes[tris[t1, (i1 + 2) % 3]], nodes[tris[t2, (i2 + 2) % 3]]
        pu, pv = nodes[u], nodes[v]
        self.l = np.linalg.norm(pv - pu, axis=1)
        alpha = np.concatenate([self.l / (2 * self.area[t1]), -self.l / (2 * self.area[t2])])
        vv = np.vstack([self.vp, self.vm])
        rows, cols = np.r_[t1, t2], np.r_[np.arange(Ne), np.arange(Ne)]
        sp = lambda vals: sparse.csr_matrix((vals, (rows, cols)), shape=(len(tris), Ne))
        self.Aal, self.Bx, self.By = sp(alpha), sp(alpha * vv[:, 0]), sp(alpha * vv[:, 1])


def qform(S1, M, S2):
    return (S2.T @ (S1.T @ M).T).T


def combine(r, Mrr, Mrx, Mry, M0):
    A, Bx, By = r.Aal, r.Bx, r.By
    return (qform(A, Mrr, A) - qform(A, Mrx, Bx) - qform(A, Mry, By) - qform(Bx, Mrx.T, A) - qform(By, Mry.T, A)
            + qform(Bx, M0, Bx) + qform(By, M0, By))


def sheet_resistance(r, t_of, sigma=SIGMA_CU):
    """scale R scale, scale = diag(1/l): the 3-point Gram matrix of the RWG functions at 1/(sigma t)."""
    Ar, Pq = r.area, np.einsum('qk,tkc->tqc', F.BAR3, r.nodes[r.tris])
    rs = 1.0 / (sigma * t_of)
    m = lambda v: sparse.diags(Ar * rs * (F.W3[None, :] * v).sum(1))
    R = combine(r, m((Pq ** 2).sum(-1)), m(Pq[..., 0]), m(Pq[..., 1]), sparse.diags(Ar * rs)).tocsr()
    scale = sparse.diags(1.0 / r.l)
    return (scale @ R @ scale).tocsr()


def via_resistance(vseg, sigma=SIGMA_CU, plating=0.025e-3):
    v = np.asarray(vseg, float).reshape(-1, 5)
    rr = v[:, 2]
    ri = np.maximum(rr - plating, 0.0)
    return (v[:, 4] - v[:, 3]) / (sigma * np.pi * (rr ** 2 - ri ** 2))


def filament_mutual(z0a, z1a, z0b, z1b, s):
    """Grover: mu0/4pi [g(a1-b0) + g(a0-b1) - g(a1-b1) - g(a0-b0)], g(u) = u asinh(u/s) - sqrt(u^2+s^2)."""
    g = lambda u: u * np.arcsinh(u / s) - np.sqrt(u * u + s * s)
    return MU0 / (4 * np.pi) * (g(z1a - z0b) + g(z0a - z1b) - g(z1a - z1b) - g(z0a - z0b))


def via_inductance(vseg):
    v = np.asarray(vseg, float).reshape(-1, 5)
    x, y, rr, z0, z1 = v.T
    s = np.hypot(x[:, None] - x[None], y[:, None] - y[None])
    s = np.where(s < 1e-9, np.maximum(rr[:, None], rr[None]), s)
    return filament_mutual(z0[:, None], z1[:, None], z0[None], z1[None], s)


def spanning_tree(n_nodes, ends, is_src, first=None):
    nb = len(ends)
    first = nb if first is None else first
    adj = [[] for _ in range(n_nodes)]
    for k in range(first):
        a, b = ends[k]
        if not is_src[k] and a != b:
            adj[a].append((b, k))
            adj[b].append((a, k))
    tree = np.zeros(nb, bool)

    def sweep(root):
        dist = {root: 0}
        queue = collections.deque([root])
        order = [root]
        while queue:
            u = queue.popleft()
            for v, _ in adj[u]:
                if v not in dist:
                    dist[v] = dist[u] + 1
                    queue.append(v)
                    order.append(v)
        return order, dist
    seen = np.zeros(n_nodes, bool)
    for s in range(n_nodes):
        if seen[s]:
            continue
        order, _ = sweep(s)
        seen[order] = True
        root = s
        if len(order) > 2:
            _, d1 = sweep(order[-1])
            _, d2 = sweep(max(d1, key=d1.get))
            root = min(order, key=lambda v: max(d1[v], d2[v]))
        mark = {root}
        queue = collections.deque([root])
        while queue:
            u = queue.popleft()
            for v, k in adj[u]:
                if v not in mark:
                    mark.add(v)
                    tree[k] = True
                    queue.append(v)
    parent = list(range(n_nodes))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    for k in np.flatnonzero(tree):
        a, b = ends[k]
        parent[find(b)] = find(a)
    for k in range(first, nb):
        a, b = ends[k]
        if not is_src[k] and find(a) != find(b):
            parent[find(b)] = find(a)
            tree[k] = True
    tadj = [[] for _ in range(n_nodes)]
    for k in np.flatnonzero(tree):
        a, b = ends[k]
        tadj[a].append((b, k))
        tadj[b].append((a, k))
    up = np.full(n_nodes, -1)
    upb = np.full(n_nodes, -1)
    ups = np.zeros(n_nodes, int)
    depth = np.full(n_nodes, -1)
    comp = np.full(n_nodes, -1)
    for root in range(n_nodes):
        if depth[root] >= 0:
            continue
        depth[root] = 0
        comp[root] = root
        stack = [root]
        while stack:
            u = stack.pop()
            for v, k in tadj[u]:
                if depth[v] < 0:
                    depth[v], up[v], upb[v], comp[v] = depth[u] + 1, u, k, root
                    ups[v] = 1 if ends[k][0] == u else -1
                    stack.append(v)
    return tree, comp, up, upb, ups, depth


def loops(n_nodes, ends, is_src, n_copper, n_rwg, ideal_terminals=True):
    """(C (nb x nloops) csc, {source branch: column}, tree)."""
    nb = len(ends)
    tree, comp, up, upb, ups, depth = spanning_tree(n_nodes, ends, is_src, n_copper)
    rows, cols, vals, src_cols = [], [], [], {}
    col = 0
    for k in np.flatnonzero(~tree):
        a, b = ends[k]
        if a == b and k < n_rwg and ideal_terminals:
            continue
        if comp[a] != comp[b]:
            continue
        u, v = b, a
        pu, pv = [], []
        while depth[u] > depth[v]:
            pu.append((upb[u], -ups[u])); u = up[u]
        while depth[v] > depth[u]:
            pv.append((upb[v], ups[v])); v = up[v]
        while u != v:
            pu.append((upb[u], -ups[u])); u = up[u]
            pv.append((upb[v], ups[v])); v = up[v]
        rows.append(k); cols.append(col); vals.append(1.0)
        for kb, sg in pu + pv[::-1]:
            rows.append(kb); cols.append(col); vals.append(float(sg))
        if is_src[k]:
            src_cols[k] = col
        col += 1
    C = sparse.coo_matrix((vals, (rows, cols)), shape=(nb, col)).tocsc()
    return C, src_cols, tree


def direct_ports(C, Rl, Ll, sc, frequencies):
    """Port Z (F,P,P) and branch currents (F,nb,P): the free loops eliminated by LU per frequency."""
    free = np.setdiff1d(np.arange(C.shape[1]), sc)
    C = C.tocsc()
    Zs, responses = [], []
    for fi in frequencies:
        jw = 2j * np.pi * fi
        Z = lambda rows, cols: Rl[np.ix_(rows, cols)] + jw * Ll[np.ix_(rows, cols)]
        X = lu_solve(lu_factor(Z(free, free)), -Z(free, sc)) if len(free) else np.zeros((0, len(sc)), complex)
        Zs.append(Z(sc, sc) + Z(sc, free) @ X)
        responses.append(-(C[:, sc].toarray() + C[:, free] @ X))
    return np.asarray(Zs), np.asarray(responses)


def rl_equivalent(frequencies, impedance, index=-1, tolerance=1e-7):
    f = frequencies[index]
    if f <= 0:
        raise ValueError('RL extraction needs a positive reference frequency')
    Z = impedance[index]
    mats = []
    for A in (Z.real, Z.imag / (2 * np.pi * f)):
        A = (A + A.T) / 2
        e, V = np.linalg.eigh(A)
        scale = max(abs(e).max(), 1e-30)
        if e.min() < -tolerance * scale:
            raise ValueError('non-passive copper reduction; refine/check topology')
        mats.append((V * np.maximum(e, scale * 1e-12)) @ V.T)
    return mats[0], mats[1]


class CopperModel:
    """A meshed board as the loop solve sees it: triangles with heights and thicknesses, the
    graph nodes (terminals and via fans merged), branches (RWG edges, via segments, ports)."""

    def __init__(self, nodes, tris, z_of, t_of, node_of, terminals, vseg, via_nodes, ports,
                 near=3.0, outer_levels=1, ideal_terminals=True, sigma=SIGMA_CU, plating=0.025e-3):
        self.nodes, self.tris = np.asarray(nodes, float), np.asarray(tris)
        self.z_of, self.t_of = np.asarray(z_of, float), np.asarray(t_of, float)
        self.node_of = np.asarray(node_of)
        self.n_nodes = int(self.node_of.max()) + 1
        self.terminals = terminals                       # {name: node}
        self.vseg, self.via_nodes = list(vseg), list(via_nodes)
        self.ports = list(ports)                          # [(name, terminal a, terminal b)]
        self.near, self.outer_levels, self.ideal_terminals = near, outer_levels, ideal_terminals
        self.sigma, self.plating = sigma, plating
        self.rwg = RWG(self.nodes, self.tris)
        r = self.rwg
        self.br = [(int(self.node_of[a]), int(self.node_of[b])) for a, b in zip(r.tp, r.tm)] + \
                  [tuple(v) for v in self.via_nodes]
        self.ne, self.nv = r.Ne, len(self.vseg)

    def fill(self, workers=None):
        """(local nodes, M0, Mrx, Mry, Mrr) through the C restatement."""
        g = F.prepare(self.nodes, self.tris, self.z_of, self.outer_levels)
        (m0, mx, my, rr), pairs, analytic = F.cr_fill(g, self.near, self.outer_levels, False, workers)
        return g, m0, mx, my, rr

    def assemble(self, moments=None, projection='c'):
        """The loop matrices; projection 'qform' is the reference's sparse products (mqs.py:665-667),
        'c' the same C^T Le C by the OpenMP restatement cr_project (what large boards need)."""
        g, m0, mx, my, rr = moments if moments is not None else self.fill()
        r = RWG(g.nodes, self.tris)
        Le = F.cr_combine(r, rr, mx, my, m0, MU0 / (4 * np.pi), 1.0 / r.l)
        Re = sheet_resistance(r, self.t_of, self.sigma)
        Rv = sparse.diags(via_resistance(self.vseg, self.sigma, self.plating)) if self.nv else sparse.csr_matrix((0, 0))
        Lv = via_inductance(self.vseg) if self.nv else np.zeros((0, 0))
        npart = len(self.ports)
        nb = self.ne + self.nv + npart
        R = sparse.block_diag((Re, Rv, sparse.csr_matrix((npart, npart))), format='csr')
        ends = list(self.br) + [(self.terminals[a], self.terminals[b]) for _, a, b in self.ports]
        is_src = np.array([False] * len(self.br) + [True] * npart)
        C, src, tree = loops(self.n_nodes, ends, is_src, len(self.br), self.ne, self.ideal_terminals)
        Ce, Cv = C[:self.ne], C[self.ne:self.ne + self.nv]
        Rl = np.ascontiguousarray((C.T @ R @ C).toarray())
        Ll = (qform(Ce, Le, Ce) if projection == 'qform' else F.cr_project(Ce, Le)) + \
            (qform(Cv, Lv, Cv) if self.nv else 0)
        self.mats = dict(Le=Le, Lv=Lv, R=R, C=C, Rl=Rl, Ll=np.ascontiguousarray(Ll), src=src, nb=nb,
                         tree=tree, ends=ends, is_src=is_src)
        return self.mats

    def port_response(self, frequencies):
        m = self.mats
        offset = len(self.br)
        sc = np.array([m['src'][offset + i] for i in range(len(self.ports))])
        Z, branch = direct_ports(m['C'], m['Rl'], m['Ll'], sc, np.atleast_1d(frequencies))
        branch[:, offset:] = 0
        return Z, branch

    def sheets(self):
        """[(nodes, tris, layer triangle indices, z)] per distinct height."""
        out = []
        for zval in sorted(set(self.z_of.tolist())):
            idx = np.flatnonzero(self.z_of == zval)
            out.append((idx, zval))
        return out


def copper_reference_model(nodes, tris):
    """The fixture's model: trace_board (a 30x20 mm plane on B.Cu at z=0, a 0.5 mm trace and two
    pads on F.Cu at z=0.5 mm, both 35 um, two vias of r 0.15 mm), terminals s/g on F.Cu, one port
    'p' from s to g; ideal_terminals=False."""
    nodes, tris = np.asarray(nodes, float), np.asarray(tris)
    # the layers' triangles are stacked (B.Cu first: layers sorted by z), each on its own nodes
    split = next(k for k in range(1, len(tris)) if tris[k:].min() > tris[:k].max())
    layer = np.r_[np.zeros(split, int), np.ones(len(tris) - split, int)]
    z_of = np.where(layer == 0, 0.0, 0.5e-3)
    t_of = np.full(len(tris), 0.035e-3)
    cen = nodes[tris].mean(1) * 1e3
    boxes = {'s': (2, 9.6, 2.6, 10.4), 'g': (2, 7.7, 2.6, 8.5)}
    term_tris = {n: list(np.flatnonzero((layer == 1) & (cen[:, 0] >= b[0] - 1e-6) & (cen[:, 0] <= b[2] + 1e-6) &
                                         (cen[:, 1] >= b[1] - 1e-6) & (cen[:, 1] <= b[3] + 1e-6)))
                 for n, b in boxes.items()}
    vias = [(27.2, 10.0, 0.15, 0.0, 0.5), (2.5, 8.1, 0.15, 0.0, 0.5)]
    ring_tris = {}
    for lay in (0, 1):
        lt = np.flatnonzero(layer == lay)
        lnodes = np.unique(tris[lt])
        for k, (x, y, r, _, _) in enumerate(vias):
            d = np.hypot(nodes[lnodes, 0] * 1e3 - x, nodes[lnodes, 1] * 1e3 - y)
            c = lnodes[np.argmin(d)]
            ring_tris[(k, lay)] = list(lt[np.isin(tris[lt], [c]).any(1)])
    Nt = len(tris)
    parent = np.arange(Nt)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    def merge(ts):
        r0 = find(ts[0])
        for t in ts[1:]:
            parent[find(t)] = r0
    for ts in term_tris.values():
        merge(ts)
    for key in [(0, 0), (1, 0), (0, 1), (1, 1)]:
        if ring_tris[key]:
            merge(ring_tris[key])
    roots = np.array([find(t) for t in range(Nt)])
    _, node_of = np.unique(roots, return_inverse=True)
    terminals = {n: int(node_of[ts[0]]) for n, ts in term_tris.items()}
    vseg, via_nodes = [], []
    for k, (x, y, r, z0, z1) in enumerate(vias):
        na, nb = int(node_of[ring_tris[(k, 0)][0]]), int(node_of[ring_tris[(k, 1)][0]])
        via_nodes.append((na, nb))
        vseg.append((x * 1e-3, y * 1e-3, r * 1e-3, 0.0, 0.5e-3))
    return CopperModel(nodes, tris, z_of, t_of, node_of, terminals, vseg, via_nodes, [('p', 's', 'g')],
                       near=3.0, outer_levels=1, ideal_terminals=False)
