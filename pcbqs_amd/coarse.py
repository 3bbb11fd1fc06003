"""Coarse spaces for the two-level loop preconditioner.

A divergence-free sheet current is the curl of a stream function psi on the mesh vertices: the
total current across an interior edge (u, v) is +-(psi(v) - psi(u)), the sign that of the edge's
T+ -> T- normal against the rotated tangent (v - u) x z. Its loop coordinates are exactly its
values on the loops' cotree branches (every loop column of C has a single +1 on its cotree
branch and no other loop uses that branch). Coarse patterns:
  stream_patterns  piecewise-bilinear hat functions of a G x G grid over each layer (one hat per
                   layer dropped: all hats together sum to a constant, which carries no current)
  loop_restrict    branch patterns -> loop-space columns (the cotree rows of the free loops)
  unique_columns   repeated loop columns removed (tiny hats can restrict to the same vector)
The learned preconditioner (pcbqs_amd.learned) adds patterns predicted by a graph network.
"""
import numpy as np


def edge_signs(rwg):
    """+1 where the T+ -> T- normal of edge (u < v) equals the rotated tangent (t_y, -t_x)."""
    u, v = rwg.ends[:, 0], rwg.ends[:, 1]
    t = rwg.nodes[v] - rwg.nodes[u]
    nrm = np.c_[t[:, 1], -t[:, 0]]
    d = rwg.cent[rwg.tm] - rwg.cent[rwg.tp]
    return np.where((d * nrm).sum(1) >= 0, 1.0, -1.0)


def boundary_vertices(tris, n_nodes):
    """Vertices on an edge that belongs to one triangle only (outer boundaries and holes)."""
    e = np.sort(np.r_[tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]], axis=1)
    key = e[:, 0].astype(np.int64) * (n_nodes + 1) + e[:, 1]
    uniq, counts = np.unique(key, return_counts=True)
    lone = uniq[counts == 1]
    mask = np.zeros(n_nodes, bool)
    mask[lone // (n_nodes + 1)] = True
    mask[lone % (n_nodes + 1)] = True
    return mask


def stream_patterns(model, grid=8):
    """(nb, k) branch currents of coarse stream-function hats on every layer (vias carry none), a
    scipy CSC matrix. The stream functions vanish on boundary vertices: no current leaves the
    copper. Columns run over the layers, then i, then j of the hat at grid point (i, j), without the
    dropped hat (i = j = grid) and without hats that carry no current. A hat's support is the four
    grid cells around its point, so every node is evaluated for the 4 x 4 nearest hats only (each
    hat value from the same expression as a dense evaluation, zeros and rounding-level values at
    cell edges included)."""
    from scipy import sparse
    b, r = model.board, model.rwg
    s = edge_signs(r)
    u, v = r.ends[:, 0], r.ends[:, 1]
    interior = ~boundary_vertices(b.tris, len(b.nodes))
    rows, cols, vals = [], [], []
    base = 0
    for lay in np.unique(b.layer_of):
        lt = np.flatnonzero(b.layer_of == lay)
        ln = np.unique(b.tris[lt])
        x0, y0 = b.nodes[ln].min(0)
        x1, y1 = b.nodes[ln].max(0)
        gx, gy = np.linspace(x0, x1, grid + 1), np.linspace(y0, y1, grid + 1)
        hx, hy = max(gx[1] - gx[0], 1e-30), max(gy[1] - gy[0], 1e-30)
        on = np.zeros(len(b.nodes), bool)
        on[ln] = True
        edges = np.flatnonzero(on[u] & on[v])
        on &= interior
        # psi of the 4 x 4 hats nearest each node, for the nodes of this layer's edges
        nodes = np.unique(np.r_[u[edges], v[edges]])
        x, y = b.nodes[nodes, 0], b.nodes[nodes, 1]
        i0 = np.clip(np.floor((x - x0) / hx).astype(np.int64), 0, grid) - 1
        j0 = np.clip(np.floor((y - y0) / hy).astype(np.int64), 0, grid) - 1
        di, dj = np.meshgrid(np.arange(4), np.arange(4), indexing='ij')
        ii = i0[:, None] + di.ravel()[None]                       # (nodes, 16)
        jj = j0[:, None] + dj.ravel()[None]
        ok = (ii >= 0) & (ii <= grid) & (jj >= 0) & (jj <= grid) & ~((ii == grid) & (jj == grid))
        ic, jc = np.clip(ii, 0, grid), np.clip(jj, 0, grid)
        psi = (np.maximum(0, 1 - np.abs(x[:, None] - gx[ic]) / hx) *
               np.maximum(0, 1 - np.abs(y[:, None] - gy[jc]) / hy) * on[nodes][:, None])
        psi = np.where(ok, psi, 0.0)
        hat = ic * (grid + 1) + jc                                # hat index in this layer
        slot = np.full(len(b.nodes), -1)
        slot[nodes] = np.arange(len(nodes))
        # s (psi(v) - psi(u)) for every hat near either end: (edges, 32) entries, duplicates summed
        su, sv = slot[u[edges]], slot[v[edges]]
        e_rows = np.repeat(edges, 32)
        e_hats = np.c_[hat[sv], hat[su]].ravel()
        e_vals = (s[edges][:, None] * np.c_[psi[sv], -psi[su]]).ravel()
        keep = e_vals != 0
        rows.append(e_rows[keep])
        cols.append(base + e_hats[keep])
        vals.append(e_vals[keep])
        base += (grid + 1) ** 2
    if not rows:
        return sparse.csc_matrix((model.nb, 0))
    m = sparse.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(model.nb, base)).tocsc()
    m.sum_duplicates()
    m.eliminate_zeros()
    used = np.flatnonzero(np.diff(m.indptr) > 0)
    return m[:, used]


def loop_restrict(model, patterns, dense=True):
    """Branch patterns (nb, k), dense or sparse -> free-loop columns (n_free, k), each normalised to
    unit 2-norm (zero columns dropped); a scipy CSC matrix when dense=False and the patterns are
    sparse."""
    cot = model.basis.cotree[:model.n_free]
    if hasattr(patterns, 'tocsr'):
        sel = patterns.tocsr()[cot].tocsc()
        norm = np.sqrt(np.asarray(sel.multiply(sel).sum(0)).ravel())
        keep = np.flatnonzero(norm > 0)
        sel = sel[:, keep]
        # scale each stored entry by its column's norm, then densify once
        sel.data = sel.data / np.repeat(norm[keep], np.diff(sel.indptr))
        if not dense:
            sel.sort_indices()
            return sel
        return np.ascontiguousarray(sel.toarray(), np.float64)
    phi = np.ascontiguousarray(patterns[cot], np.float64)
    norm = np.linalg.norm(phi, axis=0)
    keep = norm > 0
    return np.ascontiguousarray(phi[:, keep] / norm[keep])


def unique_columns(m, tol=1e-12):
    """The columns of a sparse matrix m of unit columns without those that equal an earlier column up
    to sign (same nonzero rows, values within tol): such a column adds nothing to the span, and
    leaving it in would make the coarse space's Galerkin matrices singular. Tiny hats on small copper
    pieces can restrict to the same loop vector."""
    from scipy import sparse
    m = sparse.csc_matrix(m)
    m.sort_indices()
    seen, keep = {}, []
    for c in range(m.shape[1]):
        s, e = m.indptr[c], m.indptr[c + 1]
        v = m.data[s:e]
        if e > s and v[0] < 0:
            v = -v
        group = seen.setdefault(m.indices[s:e].tobytes(), [])
        if any(np.max(np.abs(w - v), initial=0.0) <= tol for w in group):
            continue
        group.append(v)
        keep.append(c)
    return m if len(keep) == m.shape[1] else m[:, keep]


def via_loop_columns(model, dense=True):
    """Unit loop-space columns of the free loops closed by a via segment: the inter-layer
    circulations (down one via, back up another) that no per-layer stream function carries. A scipy
    CSC matrix when dense=False."""
    cot = model.basis.cotree[:model.n_free]
    sel = np.flatnonzero((cot >= model.ne) & (cot < model.ne + model.nv))
    if not dense:
        from scipy import sparse
        return sparse.csc_matrix((np.ones(len(sel)), (sel, np.arange(len(sel)))), shape=(model.n_free, len(sel)))
    cols = np.zeros((model.n_free, len(sel)))
    cols[sel, np.arange(len(sel))] = 1.0
    return cols
