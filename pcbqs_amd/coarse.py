"""Coarse spaces for the two-level loop preconditioner.

A divergence-free sheet current is the curl of a stream function psi on the mesh vertices: the
total current across an interior edge (u, v) is +-(psi(v) - psi(u)), the sign that of the edge's
T+ -> T- normal against the rotated tangent (v - u) x z. Its loop coordinates are exactly its
values on the loops' cotree branches (every loop column of C has a single +1 on its cotree
branch and no other loop uses that branch). Coarse patterns:
  stream_patterns  piecewise-bilinear hat functions of a G x G grid over each layer (one hat per
                   layer dropped: all hats together sum to a constant, which carries no current)
  loop_restrict    branch patterns -> loop-space columns (the cotree rows of the free loops)
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
    """(nb, k) branch currents of coarse stream-function hats on every layer (vias carry none).
    The stream functions vanish on boundary vertices: no current leaves the copper."""
    b, r = model.board, model.rwg
    s = edge_signs(r)
    u, v = r.ends[:, 0], r.ends[:, 1]
    interior = ~boundary_vertices(b.tris, len(b.nodes))
    cols = []
    for lay in np.unique(b.layer_of):
        lt = np.flatnonzero(b.layer_of == lay)
        ln = np.unique(b.tris[lt])
        x0, y0 = b.nodes[ln].min(0)
        x1, y1 = b.nodes[ln].max(0)
        gx, gy = np.linspace(x0, x1, grid + 1), np.linspace(y0, y1, grid + 1)
        hx, hy = max(gx[1] - gx[0], 1e-30), max(gy[1] - gy[0], 1e-30)
        on = np.zeros(len(b.nodes), bool)
        on[ln] = True
        edges_on = on[u] & on[v]
        on &= interior
        count = 0
        for i in range(grid + 1):
            for j in range(grid + 1):
                if i == grid and j == grid:
                    continue                                    # the dropped hat
                psi = (np.maximum(0, 1 - np.abs(b.nodes[:, 0] - gx[i]) / hx) *
                       np.maximum(0, 1 - np.abs(b.nodes[:, 1] - gy[j]) / hy) * on)
                ib = np.zeros(model.nb)
                ib[:model.ne] = np.where(edges_on, s * (psi[v] - psi[u]), 0.0)
                if np.abs(ib).max() > 0:
                    cols.append(ib)
                    count += 1
    return np.array(cols).T if cols else np.zeros((model.nb, 0))


def loop_restrict(model, patterns):
    """Branch patterns (nb, k) -> free-loop columns (n_free, k), each normalised to unit 2-norm."""
    cot = model.basis.cotree[:model.n_free]
    phi = np.ascontiguousarray(patterns[cot], np.float64)
    norm = np.linalg.norm(phi, axis=0)
    keep = norm > 0
    return np.ascontiguousarray(phi[:, keep] / norm[keep])


def via_loop_columns(model):
    """Unit loop-space columns of the free loops closed by a via segment: the inter-layer
    circulations (down one via, back up another) that no per-layer stream function carries."""
    cot = model.basis.cotree[:model.n_free]
    sel = np.flatnonzero((cot >= model.ne) & (cot < model.ne + model.nv))
    cols = np.zeros((model.n_free, len(sel)))
    cols[sel, np.arange(len(sel))] = 1.0
    return cols
